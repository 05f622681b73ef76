// ORACLE — test infrastructure only. Not part of the product; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// CPU restatement of VirtualFluids' single-GPU LBM time step: K17 collision with
// esoteric-twist in-place streaming, the post-collision no-slip/velocity BCs, the
// pre-collision pressure BC, equilibrium initialisation and macroscopic output.
// Every function is templated on the arithmetic type; the C ABI below exports the
// FP64 (and, where useful, FP32) instantiations. OpenMP parallelises over nodes.
//
// Reference semantics followed (file:line under the reference tree):
//   eso-twist pointer table     src/gpu/core/Utilities/KernelUtilities.h:66-125
//   neighbour chain (ListIndices) src/gpu/core/Utilities/KernelUtilities.h:221-245
//   pre/post gathers, scatters  src/gpu/core/Utilities/KernelUtilities.h:252-379
//   runCollision                src/gpu/core/Utilities/RunCollision.cuh:74-143
//   BC closed forms             src/gpu/core/Utilities/KernelUtilities.h:171-214
//   NoSlipBounceBack            src/gpu/core/BoundaryConditions/NoSlip/NoSlipBounceBack.cu:44-176
//   NoSlipInterpolatedCompr.    src/gpu/core/BoundaryConditions/NoSlip/NoSlipInterpolatedCompressible.cu:44-378
//   VelocityBounceBack          src/gpu/core/BoundaryConditions/Velocity/VelocityBounceBack.cu:44-186
//   VelocityInterpolatedCompr.  src/gpu/core/BoundaryConditions/Velocity/VelocityInterpolatedCompressible.cu:44-418
//   VelocityWithPressureInterp. src/gpu/core/BoundaryConditions/Velocity/VelocityWithPressureInterpolatedCompressible.cu
//   PressureNonEquilibriumCompr src/gpu/core/BoundaryConditions/Pressure/PressureNonEquilibriumCompressible.cu:45-311
//   InitNavierStokesCompressible src/gpu/core/PreProcessor/PreProcessorStrategy/InitNavierStokesCompressible/InitNavierStokesCompressible_Device.cu:44-204
//   macroscopic output          src/gpu/core/PostProcessor/MacroscopicQuantities.cu:194-237
//   step ordering               src/gpu/core/Calculation/UpdateGrid27.cpp:54-96 (driven from Python)
//   refinement cell gather      src/gpu/core/Utilities/ScalingUtilities.h:53-241
//   coarse -> fine kernel       src/gpu/core/GridScaling/scaleCoarseToFineCompressible.cu:46-528
//   fine -> coarse kernel       src/gpu/core/GridScaling/scaleFineToCoarseCompressible.cu:43-124
//   drag/lift momentum exchange src/gpu/core/PostProcessor/DragLift.cu:50-556, DragLift.cpp:67-72
//   kinetic energy analyzer     src/gpu/core/PostProcessor/KineticEnergyAnalyzer.cu:66-153
//   adjoint (no reference counterpart): forward-mode dual numbers through k17_collide
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_adjoint.h"
#include "oracle_k17.h"
#include "oracle_refine.h"

namespace orc {

using u32 = uint32_t;
using u64 = uint64_t;

// One population buffer of 27*N values, SoA: buf[d*N + k] (src/gpu/core/Cuda/CudaMemoryManager.cpp:169).
template <class T> struct Lattice {
    T* buf;
    u64 N;
    const u32* nbx; // +x neighbour
    const u32* nby; // +y neighbour
    const u32* nbz; // +z neighbour
};

// Eso-twist slot base per direction for a parity: even -> d, odd -> inv(d), d000 fixed.
template <class T> inline void slot_table(const Lattice<T>& L, bool even, T* P[27])
{
    for (int d = 0; d < 27; ++d) P[d] = L.buf + (u64)(even ? d : INV[d]) * L.N;
}

// The 8 nodes one update touches: bit0 = +x, bit1 = +y, bit2 = +z.
template <class T> inline void cell_nodes(const Lattice<T>& L, u32 k, u32 n[8])
{
    n[0] = k;
    n[1] = L.nbx[k];
    n[2] = L.nby[k];
    n[3] = L.nby[n[1]];
    n[4] = L.nbz[k];
    n[5] = L.nbz[n[1]];
    n[6] = L.nbz[n[2]];
    n[7] = L.nbz[n[3]];
}

// Node at which direction d of node k is stored: shifted along every axis where c_d < 0.
inline int node_sel(int d) { return (CX[d] < 0 ? 1 : 0) | (CY[d] < 0 ? 2 : 0) | (CZ[d] < 0 ? 4 : 0); }

template <class T> inline void read_pre(T* const P[27], const u32 n[8], T* f)
{
    for (int d = 0; d < 27; ++d) f[d] = P[d][n[node_sel(d)]];
}
template <class T> inline void write_post(T* const P[27], const u32 n[8], const T* f)
{
    for (int d = 0; d < 27; ++d) P[d][n[node_sel(d)]] = f[INV[d]];
}
template <class T> inline void read_post(T* const P[27], const u32 n[8], T* f)
{
    for (int d = 0; d < 27; ++d) f[INV[d]] = P[d][n[node_sel(d)]];
}
template <class T> inline void write_pre(T* const P[27], const u32 n[8], int d, T v)
{
    P[d][n[node_sel(d)]] = v;
}

// ---- collision over an index list (runCollision) -------------------------------------------
// Per-node body force of the ApplyBodyForce / AllFeatures / SubDomainBorder templates
// (RunCollision.cuh:87-99): the half-step shift is (forces + bodyForce[k]) * 1/2 * forceFactor,
// evaluated in that order, and the node's body force is reset to zero afterwards.
template <class T> struct BodyForce {
    T forces[3];   // global force (Parameter::getForcesDev), unscaled
    T forceFactor; // 2^-level
    T* bf[3];      // per-node body force, read and cleared
};

template <class T>
void collide_indexed(const Lattice<T>& L, const u32* fluid, u64 nfluid, const K17Param<T>& prm0, bool even, T* rho,
                     T* vx, T* vy, T* vz, const T* porosity = nullptr, const BodyForce<T>* body = nullptr)
{
    T* P[27];
    slot_table(L, even, P);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nfluid; ++i) {
        const u32 k = fluid[i];
        u32 n[8];
        cell_nodes(L, k, n);
        T f[27];
        read_pre(P, n, f);
        Macro<T> m;
        K17Param<T> prm = prm0;
        if (porosity) prm.vscale = T(1) - T(0.5) * porosity[k];
        if (body) {
            prm.fx = (body->forces[0] + body->bf[0][k]) * T(0.5) * body->forceFactor;
            prm.fy = (body->forces[1] + body->bf[1][k]) * T(0.5) * body->forceFactor;
            prm.fz = (body->forces[2] + body->bf[2][k]) * T(0.5) * body->forceFactor;
            body->bf[0][k] = body->bf[1][k] = body->bf[2][k] = T(0);
        }
        k17_collide(f, prm, &m);
        if (rho) {
            rho[k] = m.drho;
            vx[k] = m.vx;
            vy[k] = m.vy;
            vz[k] = m.vz;
        }
        write_post(P, n, f);
    }
}

// ---- collision on a dense periodic box with neighbours computed on the fly -----------------
// Node index k = x + nx*(y + ny*z). mask (optional, 1 byte per node): collide only where != 0.
template <class T>
void collide_box(T* buf, int nx, int ny, int nz, const uint8_t* mask, const K17Param<T>& prm, bool even)
{
    const u64 N = (u64)nx * ny * nz;
    T* P[27];
    for (int d = 0; d < 27; ++d) P[d] = buf + (u64)(even ? d : INV[d]) * N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            const int yp = (y + 1 == ny) ? 0 : y + 1;
            const int zp = (z + 1 == nz) ? 0 : z + 1;
            for (int x = 0; x < nx; ++x) {
                const int xp = (x + 1 == nx) ? 0 : x + 1;
                const u64 k = (u64)x + (u64)nx * ((u64)y + (u64)ny * z);
                if (mask && !mask[k]) continue;
                u32 n[8];
                const u64 r00 = (u64)nx * ((u64)y + (u64)ny * z), r10 = (u64)nx * ((u64)yp + (u64)ny * z);
                const u64 r01 = (u64)nx * ((u64)y + (u64)ny * zp), r11 = (u64)nx * ((u64)yp + (u64)ny * zp);
                n[0] = (u32)(r00 + x); n[1] = (u32)(r00 + xp);
                n[2] = (u32)(r10 + x); n[3] = (u32)(r10 + xp);
                n[4] = (u32)(r01 + x); n[5] = (u32)(r01 + xp);
                n[6] = (u32)(r11 + x); n[7] = (u32)(r11 + xp);
                T f[27];
                read_pre(P, n, f);
                k17_collide(f, prm, (Macro<T>*)nullptr);
                write_post(P, n, f);
            }
        }
}

// ---- boundary conditions -------------------------------------------------------------------
// Sub-grid distances are SoA: q[d*nbc + i] (src/gpu/core/Utilities/KernelUtilities.h:140-169);
// a link is active when 0 <= q <= 1.
template <class T> inline bool active(T q) { return q >= T(0) && q <= T(1); }

template <class T> inline T c_dot(int d, T ux, T uy, T uz)
{
    T s = T(0);
    if (CX[d]) s = s + T(CX[d]) * ux;
    if (CY[d]) s = s + T(CY[d]) * uy;
    if (CZ[d]) s = s + T(CZ[d]) * uz;
    return s;
}

// Macroscopic values of the post-collision state as the interpolated BC kernels form them
// (plain sums, src/gpu/core/BoundaryConditions/Velocity/VelocityInterpolatedCompressible.cu:170-186).
template <class T> inline void post_macro(const T* f, T& drho, T& ux, T& uy, T& uz)
{
    drho = T(0);
    T jx = T(0), jy = T(0), jz = T(0);
    for (int d = 0; d < 27; ++d) {
        drho += f[d];
        jx += T(CX[d]) * f[d];
        jy += T(CY[d]) * f[d];
        jz += T(CZ[d]) * f[d];
    }
    ux = jx / (T(1) + drho);
    uy = jy / (T(1) + drho);
    uz = jz / (T(1) + drho);
}

// getEquilibriumForBC (KernelUtilities.h:171-174): the even part of the shifted equilibrium.
template <class T> inline T feq_even(T drho, T cu, T cu_sq, T w)
{
    return w * (drho + T(4.5) * cu * cu * (T(1) + drho) - cu_sq);
}

// getInterpolatedDistributionFor{NoSlip,Velo,VeloWithPressure}BC (KernelUtilities.h:176-214)
template <class T> inline T interp_link(T q, T f, T finv, T feq, T omega, T wall_term, T pressure_term)
{
    return (T(1) - q) / (T(1) + q) * (f - finv + (f + finv - T(2) * feq * omega) / (T(1) - omega)) * T(0.5) +
           (q * (f + finv) - wall_term) / (T(1) + q) - pressure_term;
}

enum BcKind : int {
    BC_NOSLIP_BOUNCEBACK = 0,
    BC_VELOCITY_BOUNCEBACK = 1,
    BC_NOSLIP_INTERPOLATED = 2,
    BC_VELOCITY_INTERPOLATED = 3,
    BC_VELOCITY_PRESSURE_INTERPOLATED = 4,
};

// One BC node: from its post-collision values f, the reflected pre-collision value of the next
// step for every active link d (q_d in [0, 1]) goes to out[inv(d)]; act[inv(d)] marks it.
template <class T, class Q>
void bc_node(int kind, const T* f, const Q* q27, u64 nbc, u64 i, T Ux, T Uy, T Uz, T omega, T* out, bool* act)
{
    T drho = T(0), ux = T(0), uy = T(0), uz = T(0), cu_sq = T(0);
    if (kind >= BC_NOSLIP_INTERPOLATED) {
        post_macro(f, drho, ux, uy, uz);
        cu_sq = T(1.5) * (ux * ux + uy * uy + uz * uz) * (T(1) + drho);
    }
    for (int d = 0; d < 27; ++d) act[d] = false;
    for (int d = 1; d < 27; ++d) {
        const Q qraw = q27[(u64)d * nbc + i];
        if (!active(qraw)) continue;
        const T qd = T(qraw);
        const int r = INV[d];
        const T w = weight<T>(d);
        T v;
        switch (kind) {
        case BC_NOSLIP_BOUNCEBACK:
            v = f[d];
            break;
        case BC_VELOCITY_BOUNCEBACK:
            v = f[d] - T(6) * w * c_dot(d, Ux, Uy, Uz);
            break;
        default: {
            const T feq = feq_even(drho, c_dot(d, ux, uy, uz), cu_sq, w);
            const T wall = (kind == BC_NOSLIP_INTERPOLATED) ? T(0) : T(6) * w * c_dot(d, Ux, Uy, Uz);
            const T pres = (kind == BC_VELOCITY_PRESSURE_INTERPOLATED) ? w * drho : T(0);
            v = interp_link(qd, f[d], f[r], feq, omega, wall, pres);
        }
        }
        out[r] = v;
        act[r] = true;
    }
}

// Post-collision BC: read post-collision values with the current parity, write the
// reflected pre-collision values of the next step with the flipped parity.
template <class T>
void bc_post(int kind, const Lattice<T>& L, const int* k, const T* q, u64 nbc, const T* vxbc, const T* vybc,
             const T* vzbc, T omega, bool even)
{
    T* P[27];
    T* Pn[27];
    slot_table(L, even, P);
    slot_table(L, !even, Pn);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nbc; ++i) {
        u32 n[8];
        cell_nodes(L, (u32)k[i], n);
        T f[27], out[27];
        bool act[27];
        read_post(P, n, f);
        const T Ux = vxbc ? vxbc[i] : T(0), Uy = vybc ? vybc[i] : T(0), Uz = vzbc ? vzbc[i] : T(0);
        bc_node(kind, f, q, nbc, (u64)i, Ux, Uy, Uz, omega, out, act);
        for (int r = 0; r < 27; ++r)
            if (act[r]) write_pre(Pn, n, r, out[r]);
    }
}

// Full shifted equilibrium used by init and by the pressure BC (InitNavierStokesCompressible_Device.cu:173-201).
template <class T> inline T feq_full(int d, T drho, T ux, T uy, T uz, T cu_sq)
{
    if (d == d000) return T(8) / T(27) * (drho - cu_sq * (T(1) + drho));
    const T cu = c_dot(d, ux, uy, uz);
    return weight<T>(d) * (drho + (T(1) + drho) * (T(3) * cu + T(4.5) * cu * cu - cu_sq));
}

// One pressure-BC node: from the neighbour's pre-collision values fn, the values written to
// the BC node for the directions the face selects (act[d]).
template <class T> void pressure_node(const T* fn, T rbc, int ax, int sgn, T* out, bool* act)
{
    T drho, inv_rho, ux, uy, uz;
    compressible_macro(fn, drho, inv_rho, ux, uy, uz);
    const T cu_sq = T(1.5) * (ux * ux + uy * uy + uz * uz);
    for (int d = 0; d < 27; ++d) {
        const int c = ax == 0 ? CX[d] : (ax == 1 ? CY[d] : CZ[d]);
        act[d] = c * sgn <= 0;
        if (act[d]) out[d] = fn[d] - feq_full(d, drho, ux, uy, uz, cu_sq) + feq_full(d, rbc, ux, uy, uz, cu_sq);
    }
}

// Pre-collision pressure BC (runs after the parity swap). direction selects the face:
// dM00 writes the slots with c_x >= 0, dP00 those with c_x <= 0, and so on in y and z.
template <class T>
void bc_pressure_neq(const Lattice<T>& L, const int* k, const int* kN, const T* rhoBC, u64 nbc, bool even, int direction)
{
    T* P[27];
    slot_table(L, even, P);
    int dsel = -1;
    for (int d = 1; d <= 6; ++d)
        if (d == direction) dsel = d;
    if (dsel < 0) return;
    // the neighbour gathers happen before any BC-node write, as in one launch where every
    // thread reads first (the written BC-node slots and the neighbour slots are disjoint).
    T* fn = new T[27 * (nbc ? nbc : 1)];
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nbc; ++i) {
        u32 n[8];
        cell_nodes(L, (u32)kN[i], n);
        T* f = fn + 27 * i;
        read_pre(P, n, f);
        T drho, inv_rho, ux, uy, uz;
        compressible_macro(f, drho, inv_rho, ux, uy, uz);
        const T cu_sq = T(1.5) * (ux * ux + uy * uy + uz * uz);
        for (int d = 0; d < 27; ++d) f[d] -= feq_full(d, drho, ux, uy, uz, cu_sq);
        for (int d = 0; d < 27; ++d) f[d] += feq_full(d, rhoBC[i], ux, uy, uz, cu_sq);
    }
    const int ax = (dsel - 1) / 2;          // 0:x 1:y 2:z
    const int sgn = (dsel % 2 == 1) ? 1 : -1; // dP00 -> +1, dM00 -> -1
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nbc; ++i) {
        u32 n[8];
        cell_nodes(L, (u32)k[i], n);
        const T* f = fn + 27 * i;
        for (int d = 0; d < 27; ++d) {
            const int c = ax == 0 ? CX[d] : (ax == 1 ? CY[d] : CZ[d]);
            if (c * sgn <= 0) write_pre(P, n, d, f[d]);
        }
    }
    delete[] fn;
}

// ---- initialisation and output -------------------------------------------------------------
template <class T>
void init_eq(const Lattice<T>& L, const u32* geo, const T* rho, const T* ux, const T* uy, const T* uz, bool even)
{
    T* P[27];
    slot_table(L, even, P);
#pragma omp parallel for schedule(static)
    for (long long kk = 0; kk < (long long)L.N; ++kk) {
        const u32 k = (u32)kk;
        if (geo && (geo[k] == 15u || geo[k] == 16u)) continue; // GEO_SOLID, GEO_VOID
        u32 n[8];
        cell_nodes(L, k, n);
        const T cu_sq = T(1.5) * (ux[k] * ux[k] + uy[k] * uy[k] + uz[k] * uz[k]);
        for (int d = 0; d < 27; ++d) write_pre(P, n, d, feq_full(d, rho[k], ux[k], uy[k], uz[k], cu_sq));
    }
}

// Non-equilibrium initialisation (InitNavierStokesCompressibleNonEquilibrium_Device,
// InitNavierStokesCompressible_Device.cu:216-501): the equilibrium above plus (1 + drho) times a
// first-order non-equilibrium part built from central differences of the input velocity
// (neighbours +x/+y/+z and, through the (-1,-1,-1) neighbour nbinv, -x/-y/-z), with the
// relaxation rate omega for the deviatoric part and 1 for the bulk part. Solid and void
// nodes get 96 in their rest slot, as in the reference (a marker value).
template <class T> struct NeqCoefficients {
    T zero, e, n, t, ne, se, te, be, tn, bn, tne, tsw, tse, tnw;
};

// The closed forms of the reference (coefficients of the velocity-gradient entries
// a = d(ux)/d(x,y,z), b = d(uy)/..., c = d(uz)/...), with eps = 1 and bulk rate 1.
template <class T>
NeqCoefficients<T> neq_coefficients(T ax, T ay, T az, T bx, T by, T bz, T cx, T cy, T cz, T o)
{
    NeqCoefficients<T> q;
    const T div = ax + by + cz;
    q.e = (T(5) * div * o - T(8) * ax + T(4) * by + T(4) * cz) / (T(54) * o);
    q.n = q.e + T(2) * (ax - by) / (T(9) * o);
    q.t = q.e + T(2) * (ax - cz) / (T(9) * o);
    q.ne = -(T(5) * cz * o + T(3) * (ay + bx) - T(2) * cz + ax * (T(5) * o + T(1)) + by * (T(5) * o + T(1))) / (T(54) * o);
    q.se = q.ne + (ay + bx) / (T(9) * o);
    q.te = -(T(5) * cz * o + by * (T(5) * o - T(2)) + T(3) * (az + cx) + cz + ax * (T(5) * o + T(1))) / (T(54) * o);
    q.be = q.te + (az + cx) / (T(9) * o);
    q.tn = -(T(5) * div * o - T(2) * ax + by + T(3) * bz + T(3) * cy + cz) / (T(54) * o);
    q.bn = q.tn + (bz + cy) / (T(9) * o);
    q.zero = T(5) * div / T(9);
    q.tne = -(ay + az + bx + bz + cx + cy) / (T(72) * o);
    q.tsw = -q.tne - (ay + bx) / (T(36) * o);
    q.tse = -q.tne - (az + cx) / (T(36) * o);
    q.tnw = -q.tne - (bz + cy) / (T(36) * o);
    return q;
}

// per direction: which coefficient (the pairs d, inv(d) share one)
template <class T> inline T neq_of(const NeqCoefficients<T>& q, int d)
{
    switch (d) {
    case d000: return q.zero;
    case dP00: case dM00: return q.e;
    case d0P0: case d0M0: return q.n;
    case d00P: case d00M: return q.t;
    case dPP0: case dMM0: return q.ne;
    case dPM0: case dMP0: return q.se;
    case dP0P: case dM0M: return q.te;
    case dP0M: case dM0P: return q.be;
    case d0PP: case d0MM: return q.tn;
    case d0PM: case d0MP: return q.bn;
    case dPPP: case dMMM: return q.tne;
    case dPPM: case dMMP: return q.tsw;
    case dPMP: case dMPM: return q.tse;
    default: return q.tnw; // dPMM, dMPP
    }
}

template <class T>
void init_neq(const Lattice<T>& L, const u32* geo, const u32* nbinv, const T* rho, const T* ux, const T* uy,
              const T* uz, T omega, bool even)
{
    T* P[27];
    slot_table(L, even, P);
#pragma omp parallel for schedule(static)
    for (long long kk = 0; kk < (long long)L.N; ++kk) {
        const u32 k = (u32)kk;
        if (geo && (geo[k] == 15u || geo[k] == 16u)) {
            L.buf[(u64)d000 * L.N + k] = T(96);
            continue;
        }
        const u32 kPx = L.nbx[k], kPy = L.nby[k], kPz = L.nbz[k], kM = nbinv[k];
        const u32 kMx = L.nbz[L.nby[kM]], kMy = L.nbz[L.nbx[kM]], kMz = L.nby[L.nbx[kM]];
        const T h = T(2);
        const NeqCoefficients<T> q = neq_coefficients<T>(
            (ux[kPx] - ux[kMx]) / h, (ux[kPy] - ux[kMy]) / h, (ux[kPz] - ux[kMz]) / h, (uy[kPx] - uy[kMx]) / h,
            (uy[kPy] - uy[kMy]) / h, (uy[kPz] - uy[kMz]) / h, (uz[kPx] - uz[kMx]) / h, (uz[kPy] - uz[kMy]) / h,
            (uz[kPz] - uz[kMz]) / h, omega);
        u32 n[8];
        cell_nodes(L, k, n);
        const T cu_sq = T(1.5) * (ux[k] * ux[k] + uy[k] * uy[k] + uz[k] * uz[k]);
        for (int d = 0; d < 27; ++d) {
            T v = feq_full(d, rho[k], ux[k], uy[k], uz[k], cu_sq);
            v += (T(1) + rho[k]) * neq_of(q, d);
            write_pre(P, n, d, v);
        }
    }
}

template <class T>
void macroscopic(const Lattice<T>& L, const u32* geo, bool even, T* rho, T* vx, T* vy, T* vz, T* press)
{
    T* P[27];
    slot_table(L, even, P);
#pragma omp parallel for schedule(static)
    for (long long kk = 0; kk < (long long)L.N; ++kk) {
        const u32 k = (u32)kk;
        rho[k] = vx[k] = vy[k] = vz[k] = T(0);
        if (press) press[k] = T(0);
        // isValidFluidNode: GEO_FLUID or porous-media types (KernelUtilities.h:216-219)
        if (geo && !(geo[k] == 19u || geo[k] == 5u || geo[k] == 6u || geo[k] == 7u)) continue;
        u32 n[8];
        cell_nodes(L, k, n);
        T f[27];
        read_pre(P, n, f);
        const T r = density(f);
        rho[k] = r;
        vx[k] = momentum_x(f) / (r + T(1));
        vy[k] = momentum_y(f) / (r + T(1));
        vz[k] = momentum_z(f) / (r + T(1));
        if (press) press[k] = pressure(f, r, vx[k], vy[k], vz[k]);
    }
}

// ---- drag and lift by momentum exchange -------------------------------------------------------
// DragLiftPost27 / DragLiftPre27 (src/gpu/core/PostProcessor/DragLift.cu:50-302, 304-556): for a BC
// node and every link e with q_e in [0, 1],
//   post half (after the collision, before the BC kernels, collision parity): the population e
//     leaving towards the wall,                    drag_post = -sum_e c_e f_e^post
//   pre half (after the BC kernels and the parity swap): the population inv(e) coming back,
//                                                  drag_pre  = -sum_e c_inv(e) f_inv(e)^pre
// calcDragLift (DragLift.cpp:67-72) sums pre - post over the nodes in double.
template <class T>
void drag_lift(const Lattice<T>& L, const int* k, const T* q, u64 nbc, bool even, bool post, double* dx, double* dy,
               double* dz)
{
    T* P[27];
    slot_table(L, even, P);
    for (u64 i = 0; i < nbc; ++i) {
        u32 n[8];
        cell_nodes(L, (u32)k[i], n);
        T f[27];
        if (post) read_post(P, n, f);
        else read_pre(P, n, f);
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int e = 1; e < 27; ++e) {
            if (!active(q[(u64)e * nbc + i])) continue;
            const int d = post ? e : INV[e]; // the population counted for link e
            const double v = (double)f[d];
            sx += CX[d] * v;
            sy += CY[d] * v;
            sz += CZ[d] * v;
        }
        dx[i] = -sx;
        dy[i] = -sy;
        dz[i] = -sz;
    }
}

// KineticEnergyAnalyzer::run and kineticEnergyFunction
// (src/gpu/core/PostProcessor/KineticEnergyAnalyzer.cu:66-153): nodes of type fluid whose six face
// neighbours and (-1,-1,-1) neighbour (through which the -x/-y/-z neighbours are reached) are
// fluid contribute 1/2 |u|^2 (1 + drho); out[0] = sum, out[1] = count.
template <class T>
void kinetic_energy(const u32* nbx, const u32* nby, const u32* nbz, const u32* nbinv, const u32* geo, u64 N,
                    const T* vx, const T* vy, const T* vz, const T* rho, double* out)
{
    double e = 0.0, count = 0.0;
    auto fluid = [&](u32 j) { return geo[j] == 19u; };
    for (u64 k = 0; k < N; ++k) {
        if (!fluid((u32)k)) continue;
        const u32 m = nbinv[k];
        if (!fluid(nbx[k]) || !fluid(nby[k]) || !fluid(nbz[k]) || !fluid(m) || !fluid(nbz[nby[m]]) ||
            !fluid(nbz[nbx[m]]) || !fluid(nby[nbx[m]]))
            continue;
        count += 1.0;
        e += 0.5 * ((double)vx[k] * vx[k] + (double)vy[k] * vy[k] + (double)vz[k] * vz[k]) * ((double)rho[k] + 1.0);
    }
    out[0] = e;
    out[1] = count;
}

// ---- grid refinement ------------------------------------------------------------------------
// The 8 nodes of the cell whose lowest corner is `base`: n = ix + 2 iy + 4 iz.
template <class T> inline void cell_corner_nodes(const Lattice<T>& L, u32 base, u32 out[8])
{
    for (int n = 0; n < 8; ++n) {
        u32 k = base;
        if (n & 1) k = L.nbx[k];
        if (n & 2) k = L.nby[k];
        if (n & 4) k = L.nbz[k];
        out[n] = k;
    }
}

template <class T>
Coefficients<T> cell_coefficients(const Lattice<T>& L, bool even, u32 base, T omega, T xo, T yo, T zo)
{
    T* P[27];
    slot_table(L, even, P);
    u32 corners[8];
    cell_corner_nodes(L, base, corners);
    SourceMoments<T> m[8];
    for (int n = 0; n < 8; ++n) {
        u32 nn[8];
        cell_nodes(L, corners[n], nn);
        T f[27];
        read_pre(P, nn, f);
        m[n] = source_moments(f, omega);
    }
    return coefficients(m, xo, yo, zo);
}

// coarse -> fine: coarse read with its parity, the 8 fine nodes written with parity even
template <class T>
void scale_cf(const Lattice<T>& C, bool evenC, const Lattice<T>& F, const u32* cellsC, const u32* cellsF, u64 n,
              const T* ox, const T* oy, const T* oz, T omegaC, T omegaF)
{
    T* PF[27];
    slot_table(F, true, PF);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const Coefficients<T> c = cell_coefficients(C, evenC, cellsC[i], omegaC, ox[i], oy[i], oz[i]);
        u32 corners[8];
        cell_corner_nodes(F, cellsF[i], corners);
        for (int q = 0; q < 8; ++q) {
            const T x = (q & 1) ? T(0.25) : -T(0.25), y = (q & 2) ? T(0.25) : -T(0.25), z = (q & 4) ? T(0.25) : -T(0.25);
            T f[27];
            interpolate_cf(f, omegaF, T(0.5), c, x, y, z);
            u32 nn[8];
            cell_nodes(F, corners[q], nn);
            for (int d = 0; d < 27; ++d) write_pre(PF, nn, d, f[d]);
        }
    }
}

// fine -> coarse: fine read with parity even, the coarse node written with the coarse parity
template <class T>
void scale_fc(const Lattice<T>& C, bool evenC, const Lattice<T>& F, const u32* cellsC, const u32* cellsF, u64 n,
              const T* ox, const T* oy, const T* oz, T omegaC, T omegaF)
{
    T* PC[27];
    slot_table(C, evenC, PC);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const Coefficients<T> c = cell_coefficients(F, true, cellsF[i], omegaF, ox[i], oy[i], oz[i]);
        T f[27];
        interpolate_fc(f, T(2), omegaC, c);
        u32 nn[8];
        cell_nodes(C, cellsC[i], nn);
        for (int d = 0; d < 27; ++d) write_pre(PC, nn, d, f[d]);
    }
}

} // namespace orc

// =============================================================================================
// C ABI (ctypes). All buffers are caller-owned host arrays.
// =============================================================================================
using namespace orc;

namespace {
template <class T> K17Param<T> make_param(T omega, const T* lim3, const T* force3)
{
    K17Param<T> p;
    p.omega = omega;
    p.limP = lim3[0];
    p.limM = lim3[1];
    p.limD = lim3[2];
    p.fx = force3 ? force3[0] : T(0);
    p.fy = force3 ? force3[1] : T(0);
    p.fz = force3 ? force3[2] : T(0);
    return p;
}
} // namespace

#define ORC_INSTANTIATE(T, SFX)                                                                                  \
    extern "C" void orc_k17_nodes_##SFX(const T* fin, T* fout, long long n, T omega, const T* lim3,            \
                                        const T* force3, T* macro4)                                             \
    {                                                                                                          \
        const K17Param<T> p = make_param<T>(omega, lim3, force3);                                              \
        _Pragma("omp parallel for schedule(static)") for (long long i = 0; i < n; ++i)                         \
        {                                                                                                      \
            T f[27];                                                                                           \
            std::memcpy(f, fin + 27 * i, sizeof(f));                                                           \
            Macro<T> m;                                                                                        \
            k17_collide(f, p, &m);                                                                             \
            std::memcpy(fout + 27 * i, f, sizeof(f));                                                          \
            if (macro4) {                                                                                      \
                macro4[4 * i + 0] = m.drho;                                                                    \
                macro4[4 * i + 1] = m.vx;                                                                      \
                macro4[4 * i + 2] = m.vy;                                                                      \
                macro4[4 * i + 3] = m.vz;                                                                      \
            }                                                                                                  \
        }                                                                                                      \
    }                                                                                                          \
    extern "C" void orc_collide_indexed_##SFX(T* buf, unsigned long long N, const uint32_t* nbx,              \
                                              const uint32_t* nby, const uint32_t* nbz, const uint32_t* fluid, \
                                              unsigned long long nfluid, T omega, const T* lim3,             \
                                              const T* force3, int even, T* rho, T* vx, T* vy, T* vz)         \
    {                                                                                                          \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                   \
        collide_indexed(L, fluid, nfluid, make_param<T>(omega, lim3, force3), even != 0, rho, vx, vy, vz);    \
    }                                                                                                          \
    extern "C" void orc_collide_bodyforce_##SFX(T* buf, unsigned long long N, const uint32_t* nbx,            \
                                                const uint32_t* nby, const uint32_t* nbz, const uint32_t* fluid, \
                                                unsigned long long nfluid, T omega, const T* lim3,           \
                                                const T* forces3, T forceFactor, T* bfx, T* bfy, T* bfz,     \
                                                int even, T* rho, T* vx, T* vy, T* vz)                       \
    {                                                                                                          \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                   \
        const BodyForce<T> body{{forces3[0], forces3[1], forces3[2]}, forceFactor, {bfx, bfy, bfz}};           \
        collide_indexed(L, fluid, nfluid, make_param<T>(omega, lim3, nullptr), even != 0, rho, vx, vy, vz,     \
                        (const T*)nullptr, &body);                                                              \
    }                                                                                                          \
    extern "C" void orc_collide_porous_##SFX(T* buf, unsigned long long N, const uint32_t* nbx,               \
                                             const uint32_t* nby, const uint32_t* nbz, const uint32_t* fluid,  \
                                             unsigned long long nfluid, T omega, const T* lim3,              \
                                             const T* force3, const T* porosity, int even)                   \
    {                                                                                                          \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                   \
        collide_indexed(L, fluid, nfluid, make_param<T>(omega, lim3, force3), even != 0, (T*)nullptr,         \
                        (T*)nullptr, (T*)nullptr, (T*)nullptr, porosity);                                      \
    }                                                                                                          \
    extern "C" void orc_collide_box_##SFX(T* buf, int nx, int ny, int nz, const uint8_t* mask, T omega,        \
                                          const T* lim3, const T* force3, int even)                            \
    {                                                                                                          \
        collide_box(buf, nx, ny, nz, mask, make_param<T>(omega, lim3, force3), even != 0);                     \
    }                                                                                                          \
    extern "C" void orc_bc_post_##SFX(int kind, T* buf, unsigned long long N, const uint32_t* nbx,           \
                                      const uint32_t* nby, const uint32_t* nbz, const int* k, const T* q,     \
                                      unsigned long long nbc, const T* vx, const T* vy, const T* vz, T omega, \
                                      int even)                                                                \
    {                                                                                                          \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                   \
        bc_post(kind, L, k, q, nbc, vx, vy, vz, omega, even != 0);                                             \
    }                                                                                                          \
    extern "C" void orc_bc_pressure_neq_##SFX(T* buf, unsigned long long N, const uint32_t* nbx,             \
                                              const uint32_t* nby, const uint32_t* nbz, const int* k,         \
                                              const int* kN, const T* rhoBC, unsigned long long nbc, int even, \
                                              int direction)                                                   \
    {                                                                                                          \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                   \
        bc_pressure_neq(L, k, kN, rhoBC, nbc, even != 0, direction);                                           \
    }                                                                                                          \
    extern "C" void orc_init_eq_##SFX(T* buf, unsigned long long N, const uint32_t* nbx, const uint32_t* nby, \
                                      const uint32_t* nbz, const uint32_t* geo, const T* rho, const T* ux,    \
                                      const T* uy, const T* uz, int even)                                     \
    {                                                                                                          \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                   \
        init_eq(L, geo, rho, ux, uy, uz, even != 0);                                                           \
    }                                                                                                          \
    extern "C" void orc_init_neq_##SFX(T* buf, unsigned long long N, const uint32_t* nbx, const uint32_t* nby, \
                                       const uint32_t* nbz, const uint32_t* nbinv, const uint32_t* geo,        \
                                       const T* rho, const T* ux, const T* uy, const T* uz, T omega, int even)  \
    {                                                                                                          \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                   \
        init_neq(L, geo, nbinv, rho, ux, uy, uz, omega, even != 0);                                            \
    }                                                                                                          \
    extern "C" void orc_macroscopic_##SFX(T* buf, unsigned long long N, const uint32_t* nbx,                  \
                                          const uint32_t* nby, const uint32_t* nbz, const uint32_t* geo,      \
                                          int even, T* rho, T* vx, T* vy, T* vz, T* press)                     \
    {                                                                                                          \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                   \
        macroscopic(L, geo, even != 0, rho, vx, vy, vz, press);                                                \
    }

ORC_INSTANTIATE(double, f64)
ORC_INSTANTIATE(float, f32)

#define ORC_POSTPROC(T, SFX)                                                                                      \
    extern "C" void orc_drag_lift_##SFX(T* buf, unsigned long long N, const uint32_t* nbx, const uint32_t* nby,     \
                                        const uint32_t* nbz, const int* k, const T* q, unsigned long long nbc,    \
                                        int even, int post, double* dx, double* dy, double* dz)                   \
    {                                                                                                             \
        Lattice<T> L{buf, N, nbx, nby, nbz};                                                                      \
        drag_lift(L, k, q, nbc, even != 0, post != 0, dx, dy, dz);                                                \
    }                                                                                                             \
    extern "C" void orc_kinetic_energy_##SFX(unsigned long long N, const uint32_t* nbx, const uint32_t* nby,        \
                                             const uint32_t* nbz, const uint32_t* nbinv, const uint32_t* geo,     \
                                             const T* vx, const T* vy, const T* vz, const T* rho, double* out)    \
    {                                                                                                             \
        kinetic_energy(nbx, nby, nbz, nbinv, geo, N, vx, vy, vz, rho, out);                                       \
    }
ORC_POSTPROC(double, f64)
ORC_POSTPROC(float, f32)

#define ORC_SCALE(T, SFX)                                                                                         \
    extern "C" void orc_scale_##SFX(int fine_to_coarse, T* fc, unsigned long long Nc, const uint32_t* nbxc,        \
                                    const uint32_t* nbyc, const uint32_t* nbzc, T* ff, unsigned long long Nf,    \
                                    const uint32_t* nbxf, const uint32_t* nbyf, const uint32_t* nbzf, int evenC, \
                                    const uint32_t* cellsC, const uint32_t* cellsF, unsigned long long n,         \
                                    const T* ox, const T* oy, const T* oz, T omegaC, T omegaF)                    \
    {                                                                                                             \
        Lattice<T> C{fc, Nc, nbxc, nbyc, nbzc};                                                                   \
        Lattice<T> F{ff, Nf, nbxf, nbyf, nbzf};                                                                   \
        if (fine_to_coarse) scale_fc(C, evenC != 0, F, cellsC, cellsF, n, ox, oy, oz, omegaC, omegaF);          \
        else scale_cf(C, evenC != 0, F, cellsC, cellsF, n, ox, oy, oz, omegaC, omegaF);                          \
    }
ORC_SCALE(double, f64)
ORC_SCALE(float, f32)

// Reference adjoint of one interpolation cell by forward-mode dual numbers through the
// templated restatement above (source moments -> coefficients -> rebuilt nodes): one pass per
// source population (8 x 27) gives a column of the cell's Jacobian, contracted with the adjoint
// g of the written populations. mode 0: CF, g[8][27] of the eight fine children at
// (+-1/4, +-1/4, +-1/4) with eps 1/2 and omegaDst = omega_fine; mode 1: FC, g[27] of the coarse
// node at the centre with eps 2 and omegaDst = omega_coarse. out[8][27] = J^T g.
static void interp_cell_vjp(const double* f8, double omegaSrc, double xo, double yo, double zo, int mode,
                            double omegaDst, const double* g, double* out)
{
    for (int col = 0; col < 8 * 27; ++col) {
        SourceMoments<Dual> m[8];
        for (int n = 0; n < 8; ++n) {
            Dual f[27];
            for (int d = 0; d < 27; ++d) f[d] = Dual(f8[n * 27 + d], n * 27 + d == col ? 1.0 : 0.0);
            m[n] = source_moments(f, Dual(omegaSrc));
        }
        const Coefficients<Dual> c = coefficients(m, Dual(xo), Dual(yo), Dual(zo));
        double s = 0.0;
        if (mode == 0) {
            for (int q = 0; q < 8; ++q) {
                Dual fo[27];
                interpolate_cf(fo, Dual(omegaDst), Dual(0.5), c, Dual((q & 1) ? 0.25 : -0.25),
                               Dual((q & 2) ? 0.25 : -0.25), Dual((q & 4) ? 0.25 : -0.25));
                for (int d = 0; d < 27; ++d) s += fo[d].d * g[q * 27 + d];
            }
        } else {
            Dual fo[27];
            interpolate_fc(fo, Dual(2.0), Dual(omegaDst), c);
            for (int d = 0; d < 27; ++d) s += fo[d].d * g[d];
        }
        out[col] = s;
    }
}

extern "C" void orc_interp_cell_vjp_f64(const double* f8, double omegaSrc, double xo, double yo, double zo, int mode,
                                        double omegaDst, const double* g, double* out)
{
    interp_cell_vjp(f8, omegaSrc, xo, yo, zo, mode, omegaDst, g, out);
}

// Reference adjoint of scale_cf (fine_to_coarse 0) or scale_fc (1) over a lattice pair: per cell,
// the adjoint of the written cells is read and cleared and J^T of it (dual numbers) added to the
// source nodes' cells. Cells in order, one at a time.
extern "C" void orc_scale_adjoint_f64(int fine_to_coarse, double* fc, unsigned long long Nc, const uint32_t* nbxc,
                                      const uint32_t* nbyc, const uint32_t* nbzc, double* ff, unsigned long long Nf,
                                      const uint32_t* nbxf, const uint32_t* nbyf, const uint32_t* nbzf, int evenC,
                                      const uint32_t* cellsC, const uint32_t* cellsF, unsigned long long n,
                                      const double* ox, const double* oy, const double* oz, double omegaC,
                                      double omegaF, double* lamC, double* lamF)
{
    Lattice<double> C{fc, Nc, nbxc, nbyc, nbzc}, F{ff, Nf, nbxf, nbyf, nbzf};
    Lattice<double> LC{lamC, Nc, nbxc, nbyc, nbzc}, LF{lamF, Nf, nbxf, nbyf, nbzf};
    const bool cf = fine_to_coarse == 0;
    const Lattice<double>& src = cf ? C : F;
    const Lattice<double>& lsrc = cf ? LC : LF;
    const bool evenSrc = cf ? evenC != 0 : true;
    double *P[27], *PL[27], *PW[27];
    slot_table(src, evenSrc, P);
    slot_table(lsrc, evenSrc, PL);
    slot_table(cf ? LF : LC, cf ? true : evenC != 0, PW);
    for (unsigned long long i = 0; i < n; ++i) {
        u32 corners[8];
        cell_corner_nodes(src, cf ? cellsC[i] : cellsF[i], corners);
        std::vector<double> f8(8 * 27), g(cf ? 8 * 27 : 27), out(8 * 27);
        for (int q = 0; q < 8; ++q) {
            u32 nn[8];
            cell_nodes(src, corners[q], nn);
            read_pre(P, nn, f8.data() + 27 * q);
        }
        if (cf) {
            u32 fine[8];
            cell_corner_nodes(LF, cellsF[i], fine);
            for (int q = 0; q < 8; ++q) {
                u32 nn[8];
                cell_nodes(LF, fine[q], nn);
                read_pre(PW, nn, g.data() + 27 * q);
                for (int d = 0; d < 27; ++d) write_pre(PW, nn, d, 0.0);
            }
        } else {
            u32 nn[8];
            cell_nodes(LC, cellsC[i], nn);
            read_pre(PW, nn, g.data());
            for (int d = 0; d < 27; ++d) write_pre(PW, nn, d, 0.0);
        }
        interp_cell_vjp(f8.data(), cf ? omegaC : omegaF, ox ? ox[i] : 0.0, oy ? oy[i] : 0.0, oz ? oz[i] : 0.0, cf ? 0 : 1,
                        cf ? omegaF : omegaC, g.data(), out.data());
        for (int q = 0; q < 8; ++q) {
            u32 nn[8];
            cell_nodes(lsrc, corners[q], nn);
            for (int d = 0; d < 27; ++d) PL[d][nn[node_sel(d)]] += out[27 * q + d];
        }
    }
}

// One interpolation cell from 8 given source distributions (node-major [8][27], n = ix + 2 iy + 4 iz):
// coefficients (42 values, order of Coefficients<T>) and the rebuilt node (mode 0: coarse -> fine at
// (x, y, z) with eps and omegaDst = omega_fine; mode 1: fine -> coarse with eps and omegaDst = omega_coarse).
extern "C" void orc_interp_cell_f64(const double* f8, double omegaSrc, double xo, double yo, double zo, int mode,
                                    double omegaDst, double eps, double x, double y, double z, double* coeff42,
                                    double* fout27)
{
    SourceMoments<double> m[8];
    for (int n = 0; n < 8; ++n) m[n] = source_moments(f8 + 27 * n, omegaSrc);
    const Coefficients<double> c = coefficients(m, xo, yo, zo);
    static_assert(sizeof(Coefficients<double>) == 42 * sizeof(double), "coefficient layout");
    std::memcpy(coeff42, &c, sizeof(c));
    if (mode == 0) interpolate_cf(fout27, omegaDst, eps, c, x, y, z);
    else interpolate_fc(fout27, eps, omegaDst, c);
}

// Single-line chimera transforms and moments, for the reference's known-answer tests.
// kind: 0 forward-with-K, 1 backward-with-K, 2 forward, 3 backward.
extern "C" void orc_chimera_f64(int kind, double* abc, double u, double u2, double Kinv, double K)
{
    if (kind == 0) chimera_fwd(abc[0], abc[1], abc[2], u, u2, Kinv, K);
    if (kind == 1) chimera_bwd(abc[0], abc[1], abc[2], u, u2, Kinv, K);
    if (kind == 2) chimera_fwd(abc[0], abc[1], abc[2], u, u2, 0.0, 0.0);
    if (kind == 3) chimera_bwd(abc[0], abc[1], abc[2], u, u2, 0.0, 0.0);
}

// out: drho, jx, jy, jz (incompressible sums), 1/rho, ux, uy, uz (compressible)
extern "C" void orc_moments_f64(const double* f, double* out8)
{
    out8[0] = density(f);
    out8[1] = momentum_x(f);
    out8[2] = momentum_y(f);
    out8[3] = momentum_z(f);
    double d, ir, ux, uy, uz;
    compressible_macro(f, d, ir, ux, uy, uz);
    out8[4] = ir;
    out8[5] = ux;
    out8[6] = uy;
    out8[7] = uz;
}

// Eso-twist slot of direction d for a parity (KernelUtilities.h:66-125): returns the array index.
extern "C" int orc_eso_slot(int d, int even) { return even ? d : INV[d]; }

// ---- adjoint of the collide+stream step (forward-mode AD of k17_collide, oracle_adjoint.h) ----
// Per node: g_post[e] = adjoint at the cell where post[e] was written; the node's adjoint of its
// pre-collision values goes back to the cells they were read from (the step is block-diagonal
// in memory, so the adjoint is in place). ub3n (optional): per-node adjoint of the velocity shift.
extern "C" void orc_k17_vjp_f64(const double* fin, const double* g, long long n, double omega, const double* lim3,
                                const double* force3, double* fb, double* ub3n)
{
    const K17Param<double> p = make_param<double>(omega, lim3, force3);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i)
        k17_vjp_dual(fin + 27 * i, g + 27 * i, p, fb + 27 * i, ub3n ? ub3n + 3 * i : nullptr);
}

extern "C" void orc_adjoint_step_f64(const double* state, double* lam, unsigned long long N, const uint32_t* nbx,
                                     const uint32_t* nby, const uint32_t* nbz, const uint32_t* fluid,
                                     unsigned long long nfluid, double omega, const double* lim3,
                                     const double* force3, int even, double* ub3n, const double* porosity,
                                     double* porosityAdj)
{
    const K17Param<double> p = make_param<double>(omega, lim3, force3);
    Lattice<double> S{const_cast<double*>(state), N, nbx, nby, nbz};
    Lattice<double> A{lam, N, nbx, nby, nbz};
    double* PS[27];
    double* PA[27];
    slot_table(S, even != 0, PS);
    slot_table(A, even != 0, PA);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nfluid; ++i) {
        const u32 k = fluid[i];
        u32 nn[8];
        cell_nodes(S, k, nn);
        double f[27], g[27], fb[27], sb = 0.0;
        read_pre(PS, nn, f);
        read_post(PA, nn, g);
        K17Param<double> pk = p;
        if (porosity) pk.vscale = 1.0 - 0.5 * porosity[k];
        k17_vjp_dual(f, g, pk, fb, ub3n ? ub3n + 3 * (u64)k : nullptr, porosity ? &sb : nullptr);
        if (porosity && porosityAdj) porosityAdj[k] += -0.5 * sb;
        for (int d = 0; d < 27; ++d) write_pre(PA, nn, d, fb[d]);
    }
}

// ---- adjoints of the BC kernels (dual numbers through bc_node / pressure_node) -----------------
// Post-collision BC of one step: `state` holds the post-collision populations (the collision
// of the step done, parity `even` not yet swapped). The written cells (flipped parity) carry the
// adjoint of the values the BC wrote: it goes to the BC node's post-collision cells and the
// written cells' adjoint becomes zero (their previous content is overwritten). Ubar (optional,
// 3*nbc, added): adjoint of the wall velocity.
extern "C" void orc_bc_post_adjoint_f64(int kind, const double* state, double* lam, unsigned long long N,
                                        const uint32_t* nbx, const uint32_t* nby, const uint32_t* nbz, const int* k,
                                        const double* q, unsigned long long nbc, const double* vx, const double* vy,
                                        const double* vz, double omega, int even, double* Ubar)
{
    Lattice<double> S{const_cast<double*>(state), N, nbx, nby, nbz};
    Lattice<double> A{lam, N, nbx, nby, nbz};
    double *PS[27], *PA[27], *PAn[27];
    slot_table(S, even != 0, PS);
    slot_table(A, even != 0, PA);
    slot_table(A, even == 0, PAn);
    for (long long i = 0; i < (long long)nbc; ++i) {
        u32 n[8];
        cell_nodes(S, (u32)k[i], n);
        double f[27];
        read_post(PS, n, f);
        double g[27];
        bool act[27];
        {
            double out[27];
            bc_node<double, double>(kind, f, q, nbc, (u64)i, vx ? vx[i] : 0.0, vy ? vy[i] : 0.0, vz ? vz[i] : 0.0,
                                    omega, out, act);
        }
        for (int r = 0; r < 27; ++r) g[r] = act[r] ? PAn[r][n[node_sel(r)]] : 0.0;
        double fb[27] = {0};
        for (int col = 0; col < 30; ++col) {
            Dual fd[27], out[27];
            bool a2[27];
            for (int e = 0; e < 27; ++e) fd[e] = Dual(f[e], col == e ? 1.0 : 0.0);
            Dual Ux(vx ? vx[i] : 0.0, col == 27 ? 1.0 : 0.0), Uy(vy ? vy[i] : 0.0, col == 28 ? 1.0 : 0.0),
                Uz(vz ? vz[i] : 0.0, col == 29 ? 1.0 : 0.0);
            bc_node<Dual, double>(kind, fd, q, nbc, (u64)i, Ux, Uy, Uz, Dual(omega), out, a2);
            double sum = 0.0;
            for (int r = 0; r < 27; ++r)
                if (act[r]) sum += out[r].d * g[r];
            if (col < 27) fb[col] = sum;
            else if (Ubar) Ubar[3 * i + (col - 27)] += sum;
        }
        for (int r = 0; r < 27; ++r)
            if (act[r]) PAn[r][n[node_sel(r)]] = 0.0;
        for (int e = 0; e < 27; ++e) PA[INV[e]][n[node_sel(INV[e])]] += fb[e];
    }
}

// Pressure BC of one step: `state` holds the populations after the step (parity `even` is the
// one the BC used, after the swap). rhoBar (optional, nbc, added): adjoint of RhoBC.
extern "C" void orc_bc_pressure_adjoint_f64(const double* state, double* lam, unsigned long long N,
                                            const uint32_t* nbx, const uint32_t* nby, const uint32_t* nbz,
                                            const int* k, const int* kN, const double* rhoBC, unsigned long long nbc,
                                            int even, int direction, double* rhoBar)
{
    if (direction < 1 || direction > 6) return;
    const int ax = (direction - 1) / 2, sgn = (direction % 2 == 1) ? 1 : -1;
    Lattice<double> S{const_cast<double*>(state), N, nbx, nby, nbz};
    Lattice<double> A{lam, N, nbx, nby, nbz};
    double *PS[27], *PA[27];
    slot_table(S, even != 0, PS);
    slot_table(A, even != 0, PA);
    // gather every written cell's adjoint first, then clear, then scatter (a launch reads before it writes)
    std::vector<double> G(27 * nbc, 0.0);
    std::vector<double> FN(27 * nbc, 0.0);
    for (long long i = 0; i < (long long)nbc; ++i) {
        u32 n[8], nn[8];
        cell_nodes(S, (u32)k[i], n);
        cell_nodes(S, (u32)kN[i], nn);
        read_pre(PS, nn, &FN[27 * i]);
        for (int d = 0; d < 27; ++d) {
            const int c = ax == 0 ? CX[d] : (ax == 1 ? CY[d] : CZ[d]);
            if (c * sgn <= 0) G[27 * i + d] = PA[d][n[node_sel(d)]];
        }
    }
    for (long long i = 0; i < (long long)nbc; ++i) {
        u32 n[8];
        cell_nodes(S, (u32)k[i], n);
        for (int d = 0; d < 27; ++d) {
            const int c = ax == 0 ? CX[d] : (ax == 1 ? CY[d] : CZ[d]);
            if (c * sgn <= 0) PA[d][n[node_sel(d)]] = 0.0;
        }
    }
    for (long long i = 0; i < (long long)nbc; ++i) {
        u32 nn[8];
        cell_nodes(S, (u32)kN[i], nn);
        const double* f = &FN[27 * i];
        for (int col = 0; col < 28; ++col) {
            Dual fd[27], out[27];
            bool act[27];
            for (int e = 0; e < 27; ++e) fd[e] = Dual(f[e], col == e ? 1.0 : 0.0);
            pressure_node<Dual>(fd, Dual(rhoBC[i], col == 27 ? 1.0 : 0.0), ax, sgn, out, act);
            double sum = 0.0;
            for (int d = 0; d < 27; ++d)
                if (act[d]) sum += out[d].d * G[27 * i + d];
            if (col < 27) PA[col][nn[node_sel(col)]] += sum;
            else if (rhoBar) rhoBar[i] += sum;
        }
    }
}

// Per-node reverse mode of the BC formulas (dual numbers), for checking the GPU's hand-written
// node functions: G[r] is the adjoint of the value written for incoming direction r.
extern "C" void orc_bc_node_vjp_f64(int kind, const double* f, const double* q27, const double* U3, double omega,
                                    const double* G, double* fb, double* Ub)
{
    for (int col = 0; col < 30; ++col) {
        Dual fd[27], out[27];
        bool act[27];
        for (int e = 0; e < 27; ++e) fd[e] = Dual(f[e], col == e ? 1.0 : 0.0);
        const Dual Ux(U3[0], col == 27 ? 1.0 : 0.0), Uy(U3[1], col == 28 ? 1.0 : 0.0), Uz(U3[2], col == 29 ? 1.0 : 0.0);
        bc_node<Dual, double>(kind, fd, q27, 1, 0, Ux, Uy, Uz, Dual(omega), out, act);
        double s = 0.0;
        for (int r = 0; r < 27; ++r)
            if (act[r]) s += out[r].d * G[r];
        if (col < 27) fb[col] = s;
        else Ub[col - 27] = s;
    }
}

extern "C" void orc_pressure_node_vjp_f64(const double* fn, double rhoBC, int direction, const double* G, double* fnb,
                                          double* rb)
{
    const int ax = (direction - 1) / 2, sgn = (direction % 2 == 1) ? 1 : -1;
    for (int col = 0; col < 28; ++col) {
        Dual fd[27], out[27];
        bool act[27];
        for (int e = 0; e < 27; ++e) fd[e] = Dual(fn[e], col == e ? 1.0 : 0.0);
        pressure_node<Dual>(fd, Dual(rhoBC, col == 27 ? 1.0 : 0.0), ax, sgn, out, act);
        double s = 0.0;
        for (int d = 0; d < 27; ++d)
            if (act[d]) s += out[d].d * G[d];
        if (col < 27) fnb[col] = s;
        else *rb = s;
    }
}

extern "C" void orc_k17_vjp_scaled_f64(const double* fin, const double* g, long long n, double omega,
                                       const double* lim3, const double* vscale, double* fb, double* sb)
{
    const K17Param<double> p0 = make_param<double>(omega, lim3, nullptr);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        K17Param<double> p = p0;
        p.vscale = vscale[i];
        k17_vjp_dual(fin + 27 * i, g + 27 * i, p, fb + 27 * i, nullptr, sb + i);
    }
}
