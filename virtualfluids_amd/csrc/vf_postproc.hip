// Post-processing reductions in FP64: drag and lift by momentum exchange on a BC node list,
// and the kinetic-energy analyzer. These are the objectives of the flow cases (drag on a body,
// decay of the kinetic energy), evaluated on the device without copying fields to the host.
//
//  * vf_DragLiftPostD27 / vf_DragLiftPreD27: the two per-node halves of the reference
//    (src/gpu/core/PostProcessor/DragLift.cu:50-556), one double per node and axis.
//  * vf_DragLiftSum: their sum over the nodes, pre - post (DragLift.cpp:67-72), from the state
//    after the step alone. Both halves sit in cells no kernel touches between the two reference
//    calls: the population e a node sends towards the wall lands in the cell of the solid node
//    across the link, which nothing reads or writes, and the BC writes the returning population
//    inv(e) into the node's own next-step cell. With the parity after the swap both are the
//    slots e and inv(e) at the node that stores inv(e). One workgroup, fixed summation order:
//    the result does not change from run to run.
//  * vf_KineticEnergy: kineticEnergyKernel's node selection (the node, its six face neighbours
//    and its (-1,-1,-1) neighbour are fluid) and energy
//    (src/gpu/core/PostProcessor/KineticEnergyAnalyzer.cu:99-153), summed in double.
#include "vf_lattice.h"

namespace vfa {

namespace {

constexpr int kSumThreads = 256;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ bool link_active(float q) { return q >= 0.0f && q <= 1.0f; }

template <class S, bool POST>
__global__ __launch_bounds__(256) void drag_lift_kernel(const S* __restrict__ f, Addr a, const int32_t* __restrict__ k,
                                                        const float* __restrict__ q, uint32_t nbc, int even_flag,
                                                        double* dx, double* dy, double* dz)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbc) return;
    const bool even = even_flag != 0;
    uint32_t n[8];
    cell_nodes(a, (uint32_t)k[i], n);
    double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
    for (int e = 1; e < 27; ++e) {
        if (!link_active(q[(uint64_t)e * nbc + i])) continue;
        // post half: the population e leaving towards the wall; pre half: inv(e) coming back
        const int d = POST ? e : inv(e);
        const double v = POST ? (double)read_post(f, a, n, d, even) : (double)read_pre(f, a, n, d, even);
        sx += cx(d) * v;
        sy += cy(d) * v;
        sz += cz(d) * v;
    }
    dx[i] = -sx;
    dy[i] = -sy;
    dz[i] = -sz;
}

// force[0..2] = sum over links c_e (f_e leaving + f_inv(e) returning), state after the swap
template <class S>
__global__ __launch_bounds__(kSumThreads) void drag_sum_kernel(const S* __restrict__ f, Addr a,
                                                               const int32_t* __restrict__ k,
                                                               const float* __restrict__ q, uint32_t nbc,
                                                               int even_flag, double* force)
{
    __shared__ double part[kSumThreads / 64][3];
    const bool even = even_flag != 0;
    double acc[3] = {0.0, 0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < nbc; i += kSumThreads) {
        uint32_t n[8];
        cell_nodes(a, (uint32_t)k[i], n);
#pragma unroll
        for (int e = 1; e < 27; ++e) {
            if (!link_active(q[(uint64_t)e * nbc + i])) continue;
            const uint32_t m = n[node_sel(inv(e))];
            const double out = (double)Codec<S>::load(slot_ptr(const_cast<S*>(f), a.N, e, even) + m);
            const double back = (double)Codec<S>::load(slot_ptr(const_cast<S*>(f), a.N, inv(e), even) + m);
            const double s = out + back;
            acc[0] += cx(e) * s;
            acc[1] += cy(e) * s;
            acc[2] += cz(e) * s;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double s = wave_sum(acc[j]);
        if (lane == 0) part[wave][j] = s;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0.0;
        for (int w = 0; w < kSumThreads / 64; ++w) s += part[w][threadIdx.x];
        force[threadIdx.x] = s;
    }
}

__global__ void zero2_kernel(double* out)
{
    if (threadIdx.x < 2) out[threadIdx.x] = 0.0;
}

// the node k, its six face neighbours and its (-1,-1,-1) neighbour are fluid
// (kineticEnergyFunction checks the diagonal node it reaches the -x/-y/-z neighbours through)
__device__ __forceinline__ bool interior_fluid(const Addr& a, const uint32_t* nbinv, const uint32_t* geo, uint32_t k)
{
    auto fl = [&](uint32_t j) { return geo[j] == VF_GEO_FLUID; };
    if (!fl(k)) return false;
    if (a.nbx) {
        const uint32_t m = nbinv[k];
        return fl(a.nbx[k]) && fl(a.nby[k]) && fl(a.nbz[k]) && fl(m) && fl(a.nbz[a.nby[m]]) && fl(a.nbz[a.nbx[m]]) &&
               fl(a.nby[a.nbx[m]]);
    }
    const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny, nz = (uint32_t)a.nz;
    const uint32_t x = k % nx, t = k / nx, y = t % ny, z = t / ny;
    const uint32_t xp = x + 1 == nx ? 0 : x + 1, xm = x == 0 ? nx - 1 : x - 1;
    const uint32_t yp = y + 1 == ny ? 0 : y + 1, ym = y == 0 ? ny - 1 : y - 1;
    const uint32_t zp = z + 1 == nz ? 0 : z + 1, zm = z == 0 ? nz - 1 : z - 1;
    auto at = [&](uint32_t xx, uint32_t yy, uint32_t zz) { return xx + nx * (yy + ny * zz); };
    return fl(at(xp, y, z)) && fl(at(xm, y, z)) && fl(at(x, yp, z)) && fl(at(x, ym, z)) && fl(at(x, y, zp)) &&
           fl(at(x, y, zm)) && fl(at(xm, ym, zm));
}

__global__ __launch_bounds__(256) void kinetic_energy_kernel(Addr a, const uint32_t* nbinv, const uint32_t* geo,
                                                             const float* vx, const float* vy, const float* vz,
                                                             const float* rho, double* out)
{
    __shared__ double part[4][2];
    double e = 0.0, c = 0.0;
    for (uint64_t kk = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < a.N; kk += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)kk;
        if (!interior_fluid(a, nbinv, geo, k)) continue;
        const double ux = vx[k], uy = vy[k], uz = vz[k];
        e += 0.5 * (ux * ux + uy * uy + uz * uz) * ((double)rho[k] + 1.0);
        c += 1.0;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    e = wave_sum(e);
    c = wave_sum(c);
    if (lane == 0) {
        part[wave][0] = e;
        part[wave][1] = c;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const double s = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
        atomicAdd(out + threadIdx.x, s);
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

template <bool POST> struct DragL {
    template <class S> struct L {
        static void run(const vf_lbm_parameter& p, const vf_q_boundary_conditions& b, double* dx, double* dy,
                        double* dz, hipStream_t s)
        {
            const dim3 grid((b.numberOfBCnodes + 255) / 256);
            hipLaunchKernelGGL((drag_lift_kernel<S, POST>), grid, dim3(256), 0, s, (const S*)p.distributions,
                               make_addr(p), b.k, b.q27, b.numberOfBCnodes, p.isEvenTimestep, dx, dy, dz);
        }
    };
};

template <class S> struct DragSumL {
    static void run(const vf_lbm_parameter& p, const vf_q_boundary_conditions& b, double* force, hipStream_t s)
    {
        hipLaunchKernelGGL(drag_sum_kernel<S>, dim3(1), dim3(kSumThreads), 0, s, (const S*)p.distributions,
                           make_addr(p), b.k, b.q27, b.numberOfBCnodes, p.isEvenTimestep, force);
    }
};

int check_drag(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc)
{
    if (!p || !bc || !p->distributions) {
        set_last_error("drag/lift: null parameter, boundary condition or distributions");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->numberOfNodes == 0 || p->numberOfNodes >= (1ull << 30)) {
        set_last_error("drag/lift: numberOfNodes must lie in [1, 2^30)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!p->neighborX && (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes) {
        set_last_error("drag/lift: needs neighbour arrays or a dense box");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (bc->numberOfBCnodes > 0 && (!bc->k || !bc->q27)) {
        set_last_error("drag/lift: BC nodes without k or q27");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return VF_OK;
}

template <bool POST>
int drag_half(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, double* dx, double* dy, double* dz,
              void* stream, const char* name)
{
    int rc = check_drag(p, bc);
    if (rc) return rc;
    if (bc->numberOfBCnodes == 0) return VF_OK;
    if (!dx || !dy || !dz) {
        set_last_error("drag/lift: null output array");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = by_storage<DragL<POST>::template L>(p->storage, *p, *bc, dx, dy, dz, (hipStream_t)stream);
    return rc ? rc : check_launch(name);
}

} // namespace
} // namespace vfa

using namespace vfa;

extern "C" int vf_DragLiftPostD27(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, double* dragX,
                                  double* dragY, double* dragZ, void* stream)
{
    return drag_half<true>(p, bc, dragX, dragY, dragZ, stream, "DragLiftPost27 execution failed");
}

extern "C" int vf_DragLiftPreD27(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, double* dragX,
                                 double* dragY, double* dragZ, void* stream)
{
    return drag_half<false>(p, bc, dragX, dragY, dragZ, stream, "DragLiftPre27 execution failed");
}

extern "C" int vf_DragLiftSum(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, double* force3,
                              void* stream)
{
    int rc = check_drag(p, bc);
    if (rc) return rc;
    if (!force3) {
        set_last_error("drag/lift: null output");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = by_storage<DragSumL>(p->storage, *p, *bc, force3, (hipStream_t)stream);
    return rc ? rc : check_launch("DragLiftSum");
}

extern "C" int vf_KineticEnergy(const vf_lbm_parameter* p, const uint32_t* neighborInverse, double* out2, void* stream)
{
    if (!p || !out2 || !p->typeOfGridNode || !p->rho || !p->velocityX || !p->velocityY || !p->velocityZ) {
        set_last_error("kinetic energy: needs typeOfGridNode, rho, velocityX/Y/Z and an output");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->numberOfNodes == 0 || p->numberOfNodes >= (1ull << 30)) {
        set_last_error("kinetic energy: numberOfNodes must lie in [1, 2^30)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->neighborX && !neighborInverse) {
        set_last_error("kinetic energy on neighbour arrays needs neighborInverse");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!p->neighborX && (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes) {
        set_last_error("kinetic energy: needs neighbour arrays or a dense box");
        return VF_ERR_INVALID_ARGUMENT;
    }
    hipStream_t s = (hipStream_t)stream;
    const uint64_t blocks = (p->numberOfNodes + 255) / 256;
    hipLaunchKernelGGL(zero2_kernel, dim3(1), dim3(64), 0, s, out2);
    hipLaunchKernelGGL(kinetic_energy_kernel, dim3((unsigned)(blocks < 2048 ? blocks : 2048)), dim3(256), 0, s,
                       make_addr(*p), neighborInverse, p->typeOfGridNode, p->velocityX, p->velocityY, p->velocityZ,
                       p->rho, out2);
    return check_launch("kineticEnergyKernel execution failed");
}
