// Adjoints of the boundary-condition kernels (no counterpart in the reference: the reverse
// mode behind the differentiable time step).
//
// A post-collision BC kernel writes, for every active link d of a BC node, one cell W (the
// node's pre-collision inv(d) cell of the next step) with a function of the node's
// post-collision values. Its adjoint reads the adjoint at W, adds (d value / d post)^T to the
// adjoint of the node's post-collision cells, and clears W (its previous content was
// overwritten). The pressure BC (pre-collision, after the swap) writes the BC node's cells from
// the neighbour kN; its adjoint moves the written cells' adjoint to kN's cells. Linearisation
// points: for the post BCs the post-collision values the forward kernel stored (or, without
// them, each BC node's collision recomputed from the pre-collision state of the step), the state
// after the step for the pressure BC, with the parity each kernel used. Written cells are
// assumed to belong to nodes that no other BC node of the same list reads (links end outside the
// fluid; kN is not itself a pressure-BC node), which is how the grid builders lay boundaries out.
#include "vf_bc_adjoint.h"
#include "vf_lattice.h"

namespace vfa {

namespace {

struct BcAdjArgs {
    const int32_t* k;
    const float* q;
    uint32_t nbc;
    const float* vx;
    const float* vy;
    const float* vz;
    float omega;
    float* ubar; // 3*nbc (Vx block, Vy block, Vz block) or null
    K17Rates rates;
    const float* porosity;
    const float* post; // 27 * nbc post-collision values the forward kernel stored, or null
};

__device__ __forceinline__ float* lam_cell(float* lam, const Addr& a, const uint32_t (&n)[8], int d, bool even)
{
    return slot_ptr(lam, a.N, d, even) + n[node_sel(d)];
}

// One thread per node. SAVED: the node's post-collision values come from b.post (stored by the
// forward kernel; lists along x rows, whose loads of a direction are contiguous); otherwise they
// are recomputed from its pre-collision values in f
template <class S, int KIND, bool SAVED>
__global__ __launch_bounds__(256) void bc_post_adjoint_kernel(const S* __restrict__ f, float* __restrict__ lam, Addr a,
                                                              BcAdjArgs b, int even_flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.nbc) return;
    const bool even = even_flag != 0;
    uint32_t n[8];
    cell_nodes(a, (uint32_t)b.k[i], n);
    float post[27], q[27], G[27], fb[27], Ub[3];
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        post[d] = SAVED ? b.post[(uint64_t)d * b.nbc + i] : read_pre(f, a, n, d, even);
        q[d] = d == 0 ? -1.0f : b.q[(uint64_t)d * b.nbc + i];
    }
    if constexpr (!SAVED) {
        MacroOut mo;
        k17_collide(post, b.rates, mo, b.porosity ? 1.0f - 0.5f * b.porosity[b.k[i]] : 1.0f);
    }
    // adjoint of the written cells (incoming direction inv(d), next parity), then cleared; the
    // 27 candidate cells are loaded unconditionally (every one is a valid cell) so that the
    // loads overlap instead of waiting one by one behind the link tests
#pragma unroll
    for (int d = 0; d < 27; ++d) G[inv(d)] = *lam_cell(lam, a, n, inv(d), !even);
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        if (!(q[d] >= 0.0f && q[d] <= 1.0f)) {
            G[inv(d)] = 0.0f;
            continue;
        }
        *lam_cell(lam, a, n, inv(d), !even) = 0.0f;
    }
    bc_post_vjp<KIND>(post, q, b.omega, G, fb, Ub);
    // the post-collision value e sits in the cell of pre(inv(e)); the 27 cells are distinct, so
    // all of them are read before any is written (one memory latency instead of 27)
    float acc[27];
#pragma unroll
    for (int e = 0; e < 27; ++e) acc[e] = *lam_cell(lam, a, n, inv(e), even);
#pragma unroll
    for (int e = 0; e < 27; ++e) *lam_cell(lam, a, n, inv(e), even) = acc[e] + fb[e];
    if (b.ubar) {
        b.ubar[i] += Ub[0];
        b.ubar[b.nbc + i] += Ub[1];
        b.ubar[2 * b.nbc + i] += Ub[2];
    }
}

// ---- lane-per-direction forms (helpers in vf_bc_adjoint.h) ------------------------------------
// The post BC adjoint from stored post-collision values, a lane per direction: the arithmetic of
// bc_post_vjp, the node sums taken over the lane group. Lane d reads the written cell of link d
// (the node's incoming inv(d) of the next step) and the cell of the node's post-collision value
// d; both lie at the cell node moved by +1 where c_d > 0, in slots d and inv(d).
template <int KIND>
__global__ __launch_bounds__(kLaneBlock) void bc_post_adjoint_lanes_kernel(float* __restrict__ lam, Addr a, BcAdjArgs b,
                                                                           int even_flag)
{
    constexpr bool interp = KIND >= kNoSlipInterp;
    constexpr bool velocity = KIND == kVelocityBB || KIND == kVelocityInterp || KIND == kVelocityPressureInterp;
    const uint32_t i = blockIdx.x * kLaneNodes + (threadIdx.x / kLaneGroup);
    const bool node = i < b.nbc; // every lane of the group stays for the shuffles
    const LaneDir L = lane_dir();
    const bool on = node && L.valid;
    const bool even = even_flag != 0;
    float post = 0.0f, qd = -1.0f;
    float* wcell = nullptr;
    float* pcell = nullptr;
    if (on) {
        post = b.post[(uint64_t)L.d * b.nbc + i];
        if (L.d != 0) qd = b.q[(uint64_t)L.d * b.nbc + i];
        const uint32_t m = moved_node(a, (uint32_t)b.k[i], L.cx > 0, L.cy > 0, L.cz > 0);
        wcell = lam + (uint64_t)(even ? L.d : L.r) * a.N + m;
        pcell = lam + (uint64_t)(even ? L.r : L.d) * a.N + m;
    }
    const bool active = on && qd >= 0.0f && qd <= 1.0f;
    float g = 0.0f;
    if (active) {
        g = *wcell; // the adjoint of the written value; its old content never reached the output
        *wcell = 0.0f;
    }
    float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
    if (interp) {
        drho = group_sum(post);
        jx = group_sum(float(L.cx) * post);
        jy = group_sum(float(L.cy) * post);
        jz = group_sum(float(L.cz) * post);
    }
    float own = 0.0f, other = 0.0f, drhob = 0.0f, jxb = 0.0f, jyb = 0.0f, jzb = 0.0f;
    float wallFactor = 6.0f * L.w; // d value / d (c.U) = -wallFactor
    if (active) {
        if (!interp) {
            own = g;
        } else {
            const float ir = 1.0f / (1.0f + drho);
            const float j2 = jx * jx + jy * jy + jz * jz;
            const float om1 = 1.0f / (1.0f - b.omega);
            const float A = 0.5f * (1.0f - qd) / (1.0f + qd), Bq = 1.0f / (1.0f + qd);
            own = g * (A + A * om1 + Bq * qd);
            other = g * (-A + A * om1 + Bq * qd); // adjoint of the post-collision value inv(d)
            const float cj = float(L.cx) * jx + float(L.cy) * jy + float(L.cz) * jz;
            const float Q = 4.5f * cj * cj - 1.5f * j2;
            const float feqb = g * (-2.0f * A * b.omega * om1);
            drhob = feqb * L.w * (1.0f - Q * ir * ir);
            const float t = feqb * L.w * ir;
            jxb = t * (9.0f * cj * float(L.cx) - 3.0f * jx);
            jyb = t * (9.0f * cj * float(L.cy) - 3.0f * jy);
            jzb = t * (9.0f * cj * float(L.cz) - 3.0f * jz);
            if (KIND == kVelocityPressureInterp) drhob -= g * L.w;
            wallFactor *= Bq;
        }
    }
    float fb = own;
    if (interp) {
        fb += __shfl(other, L.valid ? L.r : L.d, kLaneGroup);
        fb += group_sum(drhob) + float(L.cx) * group_sum(jxb) + float(L.cy) * group_sum(jyb) +
              float(L.cz) * group_sum(jzb);
    }
    // the adjoint of the node's post-collision value d moves to its cell (bounce-back kinds only
    // touch their active links)
    if (on && (interp || active)) *pcell += fb;
    if (velocity && b.ubar) {
        const float gw = active ? -wallFactor * g : 0.0f;
        const float Ux = group_sum(gw * float(L.cx)), Uy = group_sum(gw * float(L.cy)), Uz = group_sum(gw * float(L.cz));
        if (node && L.d == 0) {
            b.ubar[i] += Ux;
            b.ubar[b.nbc + i] += Uy;
            b.ubar[2 * b.nbc + i] += Uz;
        }
    }
}

struct PressureAdjArgs {
    const int32_t* k;
    const int32_t* kN;
    const float* rhoBC;
    uint32_t nbc;
    int axis, sign;
    float* rhobar; // nbc or null
};

// DISTINCT: no two BC nodes share kN, so each neighbour cell has one writer (plain read-modify-
// write; atomics cost 35 against 20 us on a 96 x 96 face of FP32 cells on MI355X)
template <class S, bool DISTINCT>
__global__ __launch_bounds__(kBcBlock) void pressure_adjoint_kernel(const S* __restrict__ f, float* __restrict__ lam,
                                                                    Addr a, PressureAdjArgs b, int even_flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.nbc) return;
    const bool even = even_flag != 0;
    uint32_t nn[8], n[8];
    cell_nodes(a, (uint32_t)b.kN[i], nn);
    cell_nodes(a, (uint32_t)b.k[i], n);
    float fn[27], G[27], fnb[27], rb;
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        fn[d] = read_pre(f, a, nn, d, even);
        G[d] = *lam_cell(lam, a, n, d, even); // loaded unconditionally (valid cells), see above
    }
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const int c = b.axis == 0 ? cx(d) : (b.axis == 1 ? cy(d) : cz(d));
        if (c * b.sign > 0) {
            G[d] = 0.0f;
            continue;
        }
        *lam_cell(lam, a, n, d, even) = 0.0f;
    }
    pressure_vjp(fn, b.rhoBC[i], G, fnb, rb);
    if constexpr (DISTINCT) {
        // the 27 cells are distinct and this thread's alone: all reads before the writes
        float acc[27];
#pragma unroll
        for (int e = 0; e < 27; ++e) acc[e] = *lam_cell(lam, a, nn, e, even);
#pragma unroll
        for (int e = 0; e < 27; ++e) *lam_cell(lam, a, nn, e, even) = acc[e] + fnb[e];
    } else {
#pragma unroll
        for (int e = 0; e < 27; ++e) atomicAdd(lam_cell(lam, a, nn, e, even), fnb[e]);
    }
    if (b.rhobar) b.rhobar[i] += rb;
}

// The pressure BC adjoint with a lane per direction (the arithmetic of pressure_vjp) for faces
// normal to x: lane d reads the neighbour's pre-collision value d and, where the face writes
// direction d, the adjoint of the BC node's cell d, which it then clears. DISTINCT as above.
template <class S, bool DISTINCT>
__global__ __launch_bounds__(kLaneBlock) void pressure_adjoint_lanes_kernel(const S* __restrict__ f,
                                                                            float* __restrict__ lam, Addr a,
                                                                            PressureAdjArgs b, int even_flag)
{
    const uint32_t i = blockIdx.x * kLaneNodes + (threadIdx.x / kLaneGroup);
    const bool node = i < b.nbc;
    const LaneDir L = lane_dir();
    const bool on = node && L.valid;
    const bool even = even_flag != 0;
    const int c = b.axis == 0 ? L.cx : (b.axis == 1 ? L.cy : L.cz);
    const bool written = on && c * b.sign <= 0;
    float fn = 0.0f, G = 0.0f, rhoBC = 0.0f;
    float* ncell = nullptr;
    if (on) {
        const uint64_t slot = (uint64_t)(even ? L.d : L.r) * a.N;
        const bool sx = L.cx < 0, sy = L.cy < 0, sz = L.cz < 0;
        const uint32_t kn = moved_node(a, (uint32_t)b.kN[i], sx, sy, sz);
        fn = Codec<S>::load(f + slot + kn);
        ncell = lam + slot + kn;
        rhoBC = b.rhoBC[i];
        if (written) {
            float* gcell = lam + slot + moved_node(a, (uint32_t)b.k[i], sx, sy, sz);
            G = *gcell;
            *gcell = 0.0f;
        }
    }
    const float drho = group_sum(fn);
    const float jx = group_sum(float(L.cx) * fn), jy = group_sum(float(L.cy) * fn), jz = group_sum(float(L.cz) * fn);
    const float ir = 1.0f / (drho + 1.0f);
    const float ux = jx * ir, uy = jy * ir, uz = jz * ir;
    const float usq = ux * ux + uy * uy + uz * uz;
    const float cu = float(L.cx) * ux + float(L.cy) * uy + float(L.cz) * uz;
    const float E = 3.0f * cu + 4.5f * cu * cu - 1.5f * usq;
    const float t = G * L.w;
    const float A1 = group_sum(t * (1.0f + E));
    const float s = (rhoBC - drho) * ir;
    const float Vx = s * group_sum(t * (3.0f * float(L.cx) + 9.0f * cu * float(L.cx) - 3.0f * ux));
    const float Vy = s * group_sum(t * (3.0f * float(L.cy) + 9.0f * cu * float(L.cy) - 3.0f * uy));
    const float Vz = s * group_sum(t * (3.0f * float(L.cz) + 9.0f * cu * float(L.cz) - 3.0f * uz));
    if (on) {
        const float fnb = G - A1 + Vx * (float(L.cx) - ux) + Vy * (float(L.cy) - uy) + Vz * (float(L.cz) - uz);
        if constexpr (DISTINCT) *ncell += fnb;
        else atomicAdd(ncell, fnb);
        if (b.rhobar && L.d == 0) b.rhobar[i] += A1;
    }
}

template <template <class> class K, class... A> int by_storage(int storage, A&&... args)
{
    switch (storage) {
    case VF_STORAGE_FP32: K<float>::run(args...); return VF_OK;
    case VF_STORAGE_FP16: K<__half>::run(args...); return VF_OK;
    case VF_STORAGE_BF16: K<__hip_bfloat16>::run(args...); return VF_OK;
    default: set_last_error("unknown storage type"); return VF_ERR_INVALID_ARGUMENT;
    }
}

template <int KIND> struct PostAdjL {
    template <class S> struct L {
        static void run(const vf_lbm_parameter& p, const BcAdjArgs& b, bool rows, float* lam, hipStream_t s)
        {
            const dim3 grid((b.nbc + kBcBlock - 1) / kBcBlock);
            if (b.post && !rows)
                hipLaunchKernelGGL((bc_post_adjoint_lanes_kernel<KIND>), dim3((b.nbc + kLaneNodes - 1) / kLaneNodes),
                                   dim3(kLaneBlock), 0, s, lam, make_addr(p), b, p.isEvenTimestep);
            else if (b.post)
                hipLaunchKernelGGL((bc_post_adjoint_kernel<S, KIND, true>), grid, dim3(kBcBlock), 0, s,
                                   (const S*)p.distributions, lam, make_addr(p), b, p.isEvenTimestep);
            else
                hipLaunchKernelGGL((bc_post_adjoint_kernel<S, KIND, false>), grid, dim3(kBcBlock), 0, s,
                                   (const S*)p.distributions, lam, make_addr(p), b, p.isEvenTimestep);
        }
    };
};

template <class S> struct PressAdjL {
    static void run(const vf_lbm_parameter& p, const PressureAdjArgs& b, bool rows, float* lam, bool distinct, hipStream_t s)
    {
        const Addr a = make_addr(p);
        if (rows) {
            const dim3 grid((b.nbc + kBcBlock - 1) / kBcBlock);
            if (distinct)
                hipLaunchKernelGGL((pressure_adjoint_kernel<S, true>), grid, dim3(kBcBlock), 0, s, (const S*)p.distributions,
                                   lam, a, b, p.isEvenTimestep);
            else
                hipLaunchKernelGGL((pressure_adjoint_kernel<S, false>), grid, dim3(kBcBlock), 0, s,
                                   (const S*)p.distributions, lam, a, b, p.isEvenTimestep);
            return;
        }
        const dim3 grid((b.nbc + kLaneNodes - 1) / kLaneNodes);
        if (distinct)
            hipLaunchKernelGGL((pressure_adjoint_lanes_kernel<S, true>), grid, dim3(kLaneBlock), 0, s,
                               (const S*)p.distributions, lam, a, b, p.isEvenTimestep);
        else
            hipLaunchKernelGGL((pressure_adjoint_lanes_kernel<S, false>), grid, dim3(kLaneBlock), 0, s,
                               (const S*)p.distributions, lam, a, b, p.isEvenTimestep);
    }
};

int check_common(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, const float* lam)
{
    if (!p || !bc || !p->distributions || !lam) {
        set_last_error("adjoint BC: null parameter, boundary condition or adjoint buffer");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!p->neighborX && (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes) {
        set_last_error("adjoint BC: needs neighbour arrays or a dense box");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return VF_OK;
}

template <int KIND>
int launch_post_adjoint(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, float* lam, float* ubar,
                        void* stream, const char* name)
{
    int rc = check_common(p, bc, lam);
    if (rc) return rc;
    if (bc->numberOfBCnodes == 0) return VF_OK;
    if (!bc->k || !bc->q27) {
        set_last_error("adjoint BC: needs k and q27");
        return VF_ERR_INVALID_ARGUMENT;
    }
    constexpr bool velocity = KIND == kVelocityBB || KIND == kVelocityInterp || KIND == kVelocityPressureInterp;
    if (velocity && (!bc->Vx || !bc->Vy || !bc->Vz)) {
        set_last_error("adjoint velocity BC: needs Vx, Vy, Vz");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (KIND >= kNoSlipInterp && !(p->omega > 0.0f && p->omega < 2.0f && p->omega != 1.0f)) {
        set_last_error("adjoint interpolated BC: omega must lie in (0, 2) and differ from 1");
        return VF_ERR_INVALID_ARGUMENT;
    }
    const BcAdjArgs b{bc->k, bc->q27, bc->numberOfBCnodes, bc->Vx, bc->Vy, bc->Vz, p->omega, velocity ? ubar : nullptr,
                      make_k17_rates(*p), p->porosity, bc->postCollision};
    rc = by_storage<PostAdjL<KIND>::template L>(p->storage, *p, b, bc->nodesAlongRows != 0, lam, (hipStream_t)stream);
    return rc ? rc : check_launch(name);
}

} // namespace
} // namespace vfa

using namespace vfa;

extern "C" int vf_NoSlipBounceBackAdjoint(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, float* adjoint,
                                          float* velocityAdjoint, void* stream)
{
    return launch_post_adjoint<kNoSlipBB>(p, bc, adjoint, velocityAdjoint, stream, "NoSlipBounceBack adjoint");
}
extern "C" int vf_NoSlipInterpolatedCompressibleAdjoint(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc,
                                                        float* adjoint, float* velocityAdjoint, void* stream)
{
    return launch_post_adjoint<kNoSlipInterp>(p, bc, adjoint, velocityAdjoint, stream,
                                              "NoSlipInterpolatedCompressible adjoint");
}
extern "C" int vf_VelocityBounceBackAdjoint(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, float* adjoint,
                                            float* velocityAdjoint, void* stream)
{
    return launch_post_adjoint<kVelocityBB>(p, bc, adjoint, velocityAdjoint, stream, "VelocityBounceBack adjoint");
}
extern "C" int vf_VelocityInterpolatedCompressibleAdjoint(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc,
                                                          float* adjoint, float* velocityAdjoint, void* stream)
{
    return launch_post_adjoint<kVelocityInterp>(p, bc, adjoint, velocityAdjoint, stream,
                                                "VelocityInterpolatedCompressible adjoint");
}
extern "C" int vf_VelocityWithPressureInterpolatedCompressibleAdjoint(const vf_lbm_parameter* p,
                                                                      const vf_q_boundary_conditions* bc, float* adjoint,
                                                                      float* velocityAdjoint, void* stream)
{
    return launch_post_adjoint<kVelocityPressureInterp>(p, bc, adjoint, velocityAdjoint, stream,
                                                        "VelocityWithPressureInterpolatedCompressible adjoint");
}

extern "C" int vf_PressureNonEquilibriumCompressibleAdjoint(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc,
                                                            float* adjoint, float* rhoBCAdjoint, int32_t neighborsDistinct,
                                                            void* stream)
{
    int rc = check_common(p, bc, adjoint);
    if (rc) return rc;
    if (bc->numberOfBCnodes == 0) return VF_OK;
    if (!bc->k || !bc->kN || !bc->RhoBC) {
        set_last_error("adjoint pressure BC: needs k, kN and RhoBC");
        return VF_ERR_INVALID_ARGUMENT;
    }
    PressureAdjArgs b{bc->k, bc->kN, bc->RhoBC, bc->numberOfBCnodes, 0, 0, rhoBCAdjoint};
    switch (bc->direction) {
    case dP00: b.axis = 0; b.sign = 1; break;
    case dM00: b.axis = 0; b.sign = -1; break;
    case d0P0: b.axis = 1; b.sign = 1; break;
    case d0M0: b.axis = 1; b.sign = -1; break;
    case d00P: b.axis = 2; b.sign = 1; break;
    case d00M: b.axis = 2; b.sign = -1; break;
    default: return VF_OK; // the forward kernel writes nothing for other directions
    }
    rc = by_storage<PressAdjL>(p->storage, *p, b, bc->nodesAlongRows != 0, adjoint, neighborsDistinct != 0,
                               (hipStream_t)stream);
    return rc ? rc : check_launch("PressureNonEquilibriumCompressible adjoint");
}
