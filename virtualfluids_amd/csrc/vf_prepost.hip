// Initialisation, macroscopic output and FP64 conservation reductions.
//
//   init         InitNavierStokesCompressible_Device
//                (src/gpu/core/PreProcessor/PreProcessorStrategy/InitNavierStokesCompressible/InitNavierStokesCompressible_Device.cu:44-204)
//   macroscopic  calculateMacroscopicQuantitiesCompressible_device
//                (src/gpu/core/PostProcessor/MacroscopicQuantities.cu:194-237)
//   reductions   sums of drho, momentum and kinetic energy accumulated in FP64: per-wave
//                shuffles, one LDS pass per workgroup, one FP64 atomic per workgroup.
#include "vf_lattice.h"

namespace vfa {

template <class S>
__global__ __launch_bounds__(256) void init_kernel(S* __restrict__ f, Addr a, const uint32_t* geo, const float* rho,
                                                   const float* ux, const float* uy, const float* uz, int even_flag)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.N) return;
    if (geo && (geo[k] == VF_GEO_SOLID || geo[k] == VF_GEO_VOID)) return;
    uint32_t n[8];
    cell_nodes(a, (uint32_t)k, n);
    const float r = rho[k], vx = ux[k], vy = uy[k], vz = uz[k];
    const float cu_sq = 1.5f * (vx * vx + vy * vy + vz * vz);
#pragma unroll
    for (int d = 0; d < 27; ++d) write_pre(f, a, n, d, even_flag != 0, feq_shifted(d, r, vx, vy, vz, cu_sq));
}

// Non-equilibrium initialisation: the equilibrium plus (1 + drho) times the first-order
// non-equilibrium part from central differences of the velocity field
// (InitNavierStokesCompressibleNonEquilibrium_Device, InitNavierStokesCompressible_Device.cu:216-501;
// eps = 1, bulk rate 1). The pairs d / inv(d) share one coefficient.
struct NeqCoef {
    float zero, e, n, t, ne, se, te, be, tn, bn, tne, tsw, tse, tnw;
};

__device__ __forceinline__ NeqCoef neq_coefficients(float ax, float ay, float az, float bx, float by, float bz, float cx_,
                                                    float cy_, float cz_, float o)
{
    NeqCoef q;
    const float div = ax + by + cz_;
    const float i54 = 1.0f / (54.0f * o), i9 = 1.0f / (9.0f * o), i36 = 1.0f / (36.0f * o), i72 = 1.0f / (72.0f * o);
    q.e = (5.0f * div * o - 8.0f * ax + 4.0f * by + 4.0f * cz_) * i54;
    q.n = q.e + 2.0f * (ax - by) * i9;
    q.t = q.e + 2.0f * (ax - cz_) * i9;
    q.ne = -(5.0f * cz_ * o + 3.0f * (ay + bx) - 2.0f * cz_ + (ax + by) * (5.0f * o + 1.0f)) * i54;
    q.se = q.ne + (ay + bx) * i9;
    q.te = -(5.0f * cz_ * o + by * (5.0f * o - 2.0f) + 3.0f * (az + cx_) + cz_ + ax * (5.0f * o + 1.0f)) * i54;
    q.be = q.te + (az + cx_) * i9;
    q.tn = -(5.0f * div * o - 2.0f * ax + by + 3.0f * bz + 3.0f * cy_ + cz_) * i54;
    q.bn = q.tn + (bz + cy_) * i9;
    q.zero = (5.0f / 9.0f) * div;
    q.tne = -(ay + az + bx + bz + cx_ + cy_) * i72;
    q.tsw = -q.tne - (ay + bx) * i36;
    q.tse = -q.tne - (az + cx_) * i36;
    q.tnw = -q.tne - (bz + cy_) * i36;
    return q;
}

__device__ __forceinline__ float neq_of(const NeqCoef& q, int d)
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
    default: return q.tnw;
    }
}

template <class S>
__global__ __launch_bounds__(256) void init_neq_kernel(S* __restrict__ f, Addr a, const uint32_t* geo,
                                                       const uint32_t* nbinv, const float* rho, const float* ux,
                                                       const float* uy, const float* uz, float omega, int even_flag)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.N) return;
    if (geo && (geo[k] == VF_GEO_SOLID || geo[k] == VF_GEO_VOID)) {
        Codec<S>::store(f + (uint64_t)d000 * a.N + k, 96.0f); // the reference's marker value
        return;
    }
    uint32_t kPx, kPy, kPz, kMx, kMy, kMz;
    if (a.nbx) {
        const uint32_t kM = nbinv[k];
        kPx = a.nbx[k];
        kPy = a.nby[k];
        kPz = a.nbz[k];
        kMx = a.nbz[a.nby[kM]];
        kMy = a.nbz[a.nbx[kM]];
        kMz = a.nby[a.nbx[kM]];
    } else {
        const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny, nz = (uint32_t)a.nz;
        const uint32_t x = (uint32_t)k % nx, t = (uint32_t)k / nx, y = t % ny, z = t / ny;
        const uint32_t row = nx * (y + ny * z);
        kPx = row + (x + 1 == nx ? 0 : x + 1);
        kMx = row + (x == 0 ? nx - 1 : x - 1);
        kPy = nx * ((y + 1 == ny ? 0 : y + 1) + ny * z) + x;
        kMy = nx * ((y == 0 ? ny - 1 : y - 1) + ny * z) + x;
        kPz = nx * (y + ny * (z + 1 == nz ? 0 : z + 1)) + x;
        kMz = nx * (y + ny * (z == 0 ? nz - 1 : z - 1)) + x;
    }
    const NeqCoef q = neq_coefficients(0.5f * (ux[kPx] - ux[kMx]), 0.5f * (ux[kPy] - ux[kMy]), 0.5f * (ux[kPz] - ux[kMz]),
                                       0.5f * (uy[kPx] - uy[kMx]), 0.5f * (uy[kPy] - uy[kMy]), 0.5f * (uy[kPz] - uy[kMz]),
                                       0.5f * (uz[kPx] - uz[kMx]), 0.5f * (uz[kPy] - uz[kMy]), 0.5f * (uz[kPz] - uz[kMz]),
                                       omega);
    uint32_t n[8];
    cell_nodes(a, (uint32_t)k, n);
    const float r = rho[k], vx = ux[k], vy = uy[k], vz = uz[k];
    const float cu_sq = 1.5f * (vx * vx + vy * vy + vz * vz);
#pragma unroll
    for (int d = 0; d < 27; ++d)
        write_pre(f, a, n, d, even_flag != 0, feq_shifted(d, r, vx, vy, vz, cu_sq) + (1.0f + r) * neq_of(q, d));
}

__device__ __forceinline__ bool is_valid_fluid(uint32_t t)
{
    return t == VF_GEO_FLUID || t == VF_GEO_PM_0 || t == VF_GEO_PM_1 || t == VF_GEO_PM_2;
}

// Moments of one node accumulated in FP64 (every population converted before the add).
template <class S> __device__ __forceinline__ void node_moments(const S* f, const Addr& a, uint32_t k, bool even,
                                                               double& drho, double& jx, double& jy, double& jz)
{
    uint32_t n[8];
    cell_nodes(a, k, n);
    drho = jx = jy = jz = 0.0;
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const double v = (double)read_pre(f, a, n, d, even);
        drho += v;
        if (cx(d)) jx += double(cx(d)) * v;
        if (cy(d)) jy += double(cy(d)) * v;
        if (cz(d)) jz += double(cz(d)) * v;
    }
}

template <class S>
__global__ __launch_bounds__(256) void macroscopic_kernel(const S* __restrict__ f, Addr a, const uint32_t* geo, float* rho,
                                                          float* vx, float* vy, float* vz, float* press, int even_flag)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.N) return;
    rho[k] = vx[k] = vy[k] = vz[k] = 0.0f;
    if (press) press[k] = 0.0f;
    if (geo && !is_valid_fluid(geo[k])) return;
    uint32_t n[8];
    cell_nodes(a, (uint32_t)k, n);
    float fv[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) fv[d] = read_pre(f, a, n, d, even_flag != 0);
    float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f, e1 = 0.0f, e2 = 0.0f, e3 = 0.0f;
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        drho += fv[d];
        if (cx(d)) jx += float(cx(d)) * fv[d];
        if (cy(d)) jy += float(cy(d)) * fv[d];
        if (cz(d)) jz += float(cz(d)) * fv[d];
        const int s = cx(d) * cx(d) + cy(d) * cy(d) + cz(d) * cz(d);
        if (s == 1) e1 += fv[d];
        if (s == 2) e2 += fv[d];
        if (s == 3) e3 += fv[d];
    }
    const float ir = 1.0f / (drho + 1.0f);
    const float ux = jx * ir, uy = jy * ir, uz = jz * ir;
    rho[k] = drho;
    vx[k] = ux;
    vy[k] = uy;
    vz[k] = uz;
    // getPressure (src/lbm/MacroscopicQuantities.h:139-150)
    if (press)
        press[k] = (e1 + 2.0f * e2 + 3.0f * e3 - drho - (ux * ux + uy * uy + uz * uz) * (1.0f + drho)) * 0.5f + drho;
}

__global__ void zero_doubles_kernel(double* out, int n)
{
    if ((int)threadIdx.x < n) out[threadIdx.x] = 0.0;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// out[0..3] += sum drho, jx, jy, jz; out[4] += sum 1/2 (1+drho) |u|^2 (when KE)
template <class S, bool KE>
__global__ __launch_bounds__(256) void moments_reduce_kernel(const S* __restrict__ f, Addr a, const uint32_t* geo,
                                                             int even_flag, double* out)
{
    constexpr int NV = KE ? 5 : 4;
    __shared__ double part[4][NV];
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.N; k += (uint64_t)gridDim.x * blockDim.x) {
        if (geo && geo[k] != VF_GEO_FLUID) continue;
        double drho, jx, jy, jz;
        node_moments(f, a, (uint32_t)k, even_flag != 0, drho, jx, jy, jz);
        acc[0] += drho;
        acc[1] += jx;
        acc[2] += jy;
        acc[3] += jz;
        if (KE) {
            acc[NV - 1] += 0.5 * (jx * jx + jy * jy + jz * jz) / (1.0 + drho);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const double s = wave_sum(acc[j]);
        if (lane == 0) part[wave][j] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const double s = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
        atomicAdd(out + threadIdx.x, s);
    }
}

template <template <class> class K, class... A> static int by_storage(int storage, A&&... args)
{
    switch (storage) {
    case VF_STORAGE_FP32: K<float>::run(args...); return VF_OK;
    case VF_STORAGE_FP16: K<__half>::run(args...); return VF_OK;
    case VF_STORAGE_BF16: K<__hip_bfloat16>::run(args...); return VF_OK;
    default: set_last_error("unknown storage type"); return VF_ERR_INVALID_ARGUMENT;
    }
}

template <class S> struct InitL {
    static void run(const vf_lbm_parameter& p, hipStream_t s)
    {
        const dim3 grid((unsigned)((p.numberOfNodes + 255) / 256));
        hipLaunchKernelGGL(init_kernel<S>, grid, dim3(256), 0, s, (S*)p.distributions, make_addr(p), p.typeOfGridNode,
                           p.rho, p.velocityX, p.velocityY, p.velocityZ, p.isEvenTimestep);
    }
};
template <class S> struct InitNeqL {
    static void run(const vf_lbm_parameter& p, const uint32_t* nbinv, hipStream_t s)
    {
        const dim3 grid((unsigned)((p.numberOfNodes + 255) / 256));
        hipLaunchKernelGGL(init_neq_kernel<S>, grid, dim3(256), 0, s, (S*)p.distributions, make_addr(p),
                           p.typeOfGridNode, nbinv, p.rho, p.velocityX, p.velocityY, p.velocityZ, p.omega,
                           p.isEvenTimestep);
    }
};
template <class S> struct MacroL {
    static void run(const vf_lbm_parameter& p, hipStream_t s)
    {
        const dim3 grid((unsigned)((p.numberOfNodes + 255) / 256));
        hipLaunchKernelGGL(macroscopic_kernel<S>, grid, dim3(256), 0, s, (const S*)p.distributions, make_addr(p),
                           p.typeOfGridNode, p.rho, p.velocityX, p.velocityY, p.velocityZ, p.pressure, p.isEvenTimestep);
    }
};
template <class S> struct ReduceL {
    static void run(const vf_lbm_parameter& p, double* out, bool ke, hipStream_t s)
    {
        const uint64_t blocks = (p.numberOfNodes + 255) / 256;
        const dim3 grid((unsigned)(blocks < 2048 ? blocks : 2048));
        hipLaunchKernelGGL(zero_doubles_kernel, dim3(1), dim3(64), 0, s, out, ke ? 5 : 4);
        if (ke)
            hipLaunchKernelGGL((moments_reduce_kernel<S, true>), grid, dim3(256), 0, s, (const S*)p.distributions,
                               make_addr(p), p.typeOfGridNode, p.isEvenTimestep, out);
        else
            hipLaunchKernelGGL((moments_reduce_kernel<S, false>), grid, dim3(256), 0, s, (const S*)p.distributions,
                               make_addr(p), p.typeOfGridNode, p.isEvenTimestep, out);
    }
};

static int check_lattice(const vf_lbm_parameter* p)
{
    if (!p || !p->distributions || p->numberOfNodes == 0 || p->numberOfNodes >= (1ull << 30)) {
        set_last_error("null parameter/distributions or numberOfNodes out of [1, 2^30)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!p->neighborX && (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes) {
        set_last_error("lattice needs neighbour arrays or a dense box");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return VF_OK;
}

} // namespace vfa

using namespace vfa;

extern "C" int vf_InitNavierStokesCompressible(const vf_lbm_parameter* p, void* stream)
{
    int rc = check_lattice(p);
    if (rc) return rc;
    if (!p->rho || !p->velocityX || !p->velocityY || !p->velocityZ) {
        set_last_error("init needs rho and velocityX/Y/Z");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = by_storage<InitL>(p->storage, *p, (hipStream_t)stream);
    return rc ? rc : check_launch("InitNavierStokesCompressible_Device execution failed");
}

extern "C" int vf_InitNavierStokesCompressibleNonEquilibrium(const vf_lbm_parameter* p, const uint32_t* neighborInverse,
                                                             void* stream)
{
    int rc = check_lattice(p);
    if (rc) return rc;
    if (!p->rho || !p->velocityX || !p->velocityY || !p->velocityZ) {
        set_last_error("init needs rho and velocityX/Y/Z");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->neighborX && !neighborInverse) {
        set_last_error("the non-equilibrium init on neighbour arrays needs neighborInverse");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!(p->omega > 0.0f && p->omega < 2.0f)) {
        set_last_error("omega must lie in (0, 2)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = by_storage<InitNeqL>(p->storage, *p, neighborInverse, (hipStream_t)stream);
    return rc ? rc : check_launch("InitNavierStokesCompressibleNonEquilibrium_Device execution failed");
}

extern "C" int vf_CalculateMacroscopicQuantitiesCompressible(const vf_lbm_parameter* p, void* stream)
{
    int rc = check_lattice(p);
    if (rc) return rc;
    if (!p->rho || !p->velocityX || !p->velocityY || !p->velocityZ) {
        set_last_error("macroscopic output needs rho and velocityX/Y/Z");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = by_storage<MacroL>(p->storage, *p, (hipStream_t)stream);
    return rc ? rc : check_launch("calculateMacroscopicQuantitiesCompressible execution failed");
}

extern "C" int vf_ConservationSums(const vf_lbm_parameter* p, double* out4, void* stream)
{
    int rc = check_lattice(p);
    if (rc) return rc;
    if (!out4) {
        set_last_error("null output");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = by_storage<ReduceL>(p->storage, *p, out4, false, (hipStream_t)stream);
    return rc ? rc : check_launch("conservation sums");
}

/* Sums drho, jx, jy, jz and kinetic energy in one pass (out5 device pointer, overwritten). */
extern "C" int vf_MomentSums(const vf_lbm_parameter* p, double* out5, void* stream)
{
    int rc = check_lattice(p);
    if (rc) return rc;
    if (!out5) {
        set_last_error("null output");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = by_storage<ReduceL>(p->storage, *p, out5, true, (hipStream_t)stream);
    return rc ? rc : check_launch("moment sums");
}
