// Host build of the gfx950 collision and its hand-written adjoint (vf_k17.h,
// vf_k17_adjoint.h): the same source the kernels inline, compiled for the CPU so the tests can
// compare it with the oracle without a GPU. Test infrastructure; not part of the library.
#include <cstring>

#include "vf_bc_adjoint.h"
#include "vf_k17_adjoint.h"
#include "vf_scaling.h"

using namespace vfa;

static K17Rates rates(float omega, const float* lim3, const float* force3)
{
    vf_lbm_parameter p{};
    p.omega = omega;
    for (int i = 0; i < 3; ++i) {
        p.quadricLimiters[i] = lim3[i];
        p.forces[i] = force3 ? force3[i] : 0.0f;
    }
    p.forceFactor = 1.0f;
    return make_k17_rates(p);
}

// force3 is the body force; the collision shifts the velocity by half of it
extern "C" void host_k17_collide(const float* fin, float* fout, long long n, float omega, const float* lim3,
                                 const float* force3)
{
    const K17Rates r = rates(omega, lim3, force3);
    for (long long i = 0; i < n; ++i) {
        float f[27];
        std::memcpy(f, fin + 27 * i, sizeof(f));
        MacroOut mo;
        k17_collide(f, r, mo);
        std::memcpy(fout + 27 * i, f, sizeof(f));
    }
}

extern "C" void host_k17_vjp(const float* fin, const float* g, long long n, float omega, const float* lim3,
                             const float* force3, float* fb, float* ub3n)
{
    const K17Rates r = rates(omega, lim3, force3);
    for (long long i = 0; i < n; ++i) {
        float f[27], gg[27], out[27];
        std::memcpy(f, fin + 27 * i, sizeof(f));
        std::memcpy(gg, g + 27 * i, sizeof(gg));
        k17_vjp(f, gg, r, out, ub3n ? ub3n + 3 * i : nullptr);
        std::memcpy(fb + 27 * i, out, sizeof(out));
    }
}

// one BC node: post (27), q (27, d-indexed, inactive < 0), G (27, adjoint of the written value of
// each incoming direction); fb (27) and Ub (3) out
extern "C" void host_bc_post_vjp(int kind, const float* post, const float* q, float omega, const float* G, float* fb,
                                 float* Ub)
{
    float p[27], qq[27], g[27], out[27], ub[3];
    std::memcpy(p, post, sizeof(p));
    std::memcpy(qq, q, sizeof(qq));
    std::memcpy(g, G, sizeof(g));
    switch (kind) {
    case kNoSlipBB: bc_post_vjp<kNoSlipBB>(p, qq, omega, g, out, ub); break;
    case kVelocityBB: bc_post_vjp<kVelocityBB>(p, qq, omega, g, out, ub); break;
    case kNoSlipInterp: bc_post_vjp<kNoSlipInterp>(p, qq, omega, g, out, ub); break;
    case kVelocityInterp: bc_post_vjp<kVelocityInterp>(p, qq, omega, g, out, ub); break;
    default: bc_post_vjp<kVelocityPressureInterp>(p, qq, omega, g, out, ub); break;
    }
    std::memcpy(fb, out, sizeof(out));
    std::memcpy(Ub, ub, sizeof(ub));
}

extern "C" void host_pressure_vjp(const float* fn, float rhoBC, const float* G, float* fnb, float* rb)
{
    float f[27], g[27], out[27];
    std::memcpy(f, fn, sizeof(f));
    std::memcpy(g, G, sizeof(g));
    pressure_vjp(f, rhoBC, g, out, *rb);
    std::memcpy(fnb, out, sizeof(out));
}

// the same with a Brinkman velocity scale per node; sb (n) receives its adjoint
extern "C" void host_k17_vjp_scaled(const float* fin, const float* g, long long n, float omega, const float* lim3,
                                    const float* vscale, float* fb, float* sb)
{
    const K17Rates r = rates(omega, lim3, nullptr);
    for (long long i = 0; i < n; ++i) {
        float f[27], gg[27], out[27];
        std::memcpy(f, fin + 27 * i, sizeof(f));
        std::memcpy(gg, g + 27 * i, sizeof(gg));
        k17_vjp(f, gg, r, out, nullptr, vscale[i], sb + i);
        std::memcpy(fb + 27 * i, out, sizeof(out));
    }
}

// ---- refinement interpolation of one cell (vf_scaling.h) and its reverse -----------------
// mode 0: CF (out: 8 children x 27), mode 1: FC (out: 27); src: 8 x 27 source populations
extern "C" void host_interp_cell(const float* src, float omegaSrc, float xo, float yo, float zo, int mode,
                                 float omegaDst, float* out)
{
    SrcMom m[8];
    for (int n = 0; n < 8; ++n) {
        float f[27];
        std::memcpy(f, src + 27 * n, sizeof(f));
        m[n] = source_moments_of(f, omegaSrc);
    }
    const Coef c = coefficients(m, xo, yo, zo);
    if (mode == 0) {
        for (int q = 0; q < 8; ++q) {
            float f[27];
            rebuild_populations(f, cf_node_moments(c, cf_child_offset(q, 0), cf_child_offset(q, 1),
                                                   cf_child_offset(q, 2), 0.5f, omegaDst));
            std::memcpy(out + 27 * q, f, sizeof(f));
        }
    } else {
        float f[27];
        rebuild_populations(f, fc_node_moments(c, omegaDst));
        std::memcpy(out, f, sizeof(f));
    }
}

// J^T g for one cell: g is 8 x 27 (CF) or 27 (FC); out 8 x 27
extern "C" void host_interp_cell_vjp(const float* src, float omegaSrc, float xo, float yo, float zo, int mode,
                                     float omegaDst, const float* g, float* out)
{
    float s[8][27], sb[8][27];
    std::memcpy(s, src, sizeof(s));
    if (mode == 0) {
        float gg[8][27];
        std::memcpy(gg, g, sizeof(gg));
        cf_cell_vjp(s, omegaSrc, omegaDst, xo, yo, zo, gg, sb);
    } else {
        float gg[27];
        std::memcpy(gg, g, sizeof(gg));
        fc_cell_vjp(s, omegaSrc, omegaDst, xo, yo, zo, gg, sb);
    }
    std::memcpy(out, sb, sizeof(sb));
}
