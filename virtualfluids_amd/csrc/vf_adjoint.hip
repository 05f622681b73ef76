// Adjoint (reverse mode) of the K17 collide + eso-twist stream step.
//
// In memory the forward step is block-diagonal: node k reads its 27 pre-collision values from
// 27 cells and writes its post-collision values back into the same cells (post[inv(d)] into the
// cell of pre[d]); nodes that do not collide leave their cells alone. The adjoint is therefore
// the same sweep with the transposed node Jacobian: node k reads the adjoint of its
// post-collision values from its cells (g[e] from the cell of pre[inv(e)]), forms
// (d post / d pre)^T g with k17_vjp, and writes the result into the cells of pre[d]. Streaming
// needs nothing extra: the parity the forward step used selects the cells. The population
// state before the forward step is the linearisation point (checkpointing and recomputation
// are the caller's: the Python layer keeps a checkpoint every K steps and replays a segment).
//
// No counterpart in the reference.
#include <type_traits>

#include "vf_k17_adjoint.h"

namespace vfa {

namespace {

constexpr int kAdjBlock = 256;

template <class S, bool EVEN>
__device__ __forceinline__ void adj_gather(const S* __restrict__ f, const float* __restrict__ lam, uint64_t N,
                                           const uint32_t (&n)[8], float (&v)[27], float (&g)[27])
{
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const uint64_t cell = (uint64_t)slot<EVEN>(d) * N + n[node_sel(d)];
        v[d] = Codec<S>::load(f + cell);
        g[inv(d)] = lam[cell];
    }
}

template <bool EVEN>
__device__ __forceinline__ void adj_scatter(float* __restrict__ lam, uint64_t N, const uint32_t (&n)[8],
                                            const float (&fb)[27])
{
#pragma unroll
    for (int d = 0; d < 27; ++d) lam[(uint64_t)slot<EVEN>(d) * N + n[node_sel(d)]] = fb[d];
}

struct AdjBoxArgs {
    uint64_t N;
    int nx, ny, nz;
    const uint8_t* mask;
    float* shiftAdj;       // 3*N or null
    const float* porosity; // Brinkman porosity or null
    float* porosityAdj;    // N or null
};

// adjoint of the velocity scale s = 1 - porosity/2 goes to the porosity with the factor -1/2
__device__ __forceinline__ void brinkman_adjoint(const float* porosity, float* porosityAdj, uint32_t k, float sb)
{
    if (porosity && porosityAdj) porosityAdj[k] += -0.5f * sb;
}

template <class S, bool EVEN, bool MASK, bool SHIFT>
__global__ __launch_bounds__(kAdjBlock) void k17_adjoint_box_kernel(const S* __restrict__ f, float* __restrict__ lam,
                                                                    AdjBoxArgs a, K17Rates r)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int z = blockIdx.z;
    if (x >= a.nx) return;
    const int xp = (x + 1 == a.nx) ? 0 : x + 1;
    const int yp = (y + 1 == a.ny) ? 0 : y + 1;
    const int zp = (z + 1 == a.nz) ? 0 : z + 1;
    const uint32_t row00 = (uint32_t)a.nx * ((uint32_t)y + (uint32_t)a.ny * (uint32_t)z);
    const uint32_t row10 = (uint32_t)a.nx * ((uint32_t)yp + (uint32_t)a.ny * (uint32_t)z);
    const uint32_t row01 = (uint32_t)a.nx * ((uint32_t)y + (uint32_t)a.ny * (uint32_t)zp);
    const uint32_t row11 = (uint32_t)a.nx * ((uint32_t)yp + (uint32_t)a.ny * (uint32_t)zp);
    const uint32_t n[8] = {row00 + x, row00 + xp, row10 + x, row10 + xp, row01 + x, row01 + xp, row11 + x, row11 + xp};
    if (MASK && a.mask[n[0]] == 0) return;
    float v[27], g[27], fb[27], ub[3], sb = 0.0f;
    adj_gather<S, EVEN>(f, lam, a.N, n, v, g);
    const float vs = a.porosity ? 1.0f - 0.5f * a.porosity[n[0]] : 1.0f;
    k17_vjp(v, g, r, fb, SHIFT ? ub : nullptr, vs, a.porosity ? &sb : nullptr);
    adj_scatter<EVEN>(lam, a.N, n, fb);
    if (SHIFT) {
        a.shiftAdj[n[0]] += ub[0];
        a.shiftAdj[a.N + n[0]] += ub[1];
        a.shiftAdj[2 * a.N + n[0]] += ub[2];
    }
    brinkman_adjoint(a.porosity, a.porosityAdj, n[0], sb);
}

struct AdjIndexedArgs {
    uint64_t N;
    const uint32_t* nbx;
    const uint32_t* nby;
    const uint32_t* nbz;
    const uint32_t* indices;
    uint32_t count;
    float* shiftAdj;
    const float* porosity;
    float* porosityAdj;
};

template <class S, bool EVEN, bool SHIFT>
__global__ __launch_bounds__(kAdjBlock) void k17_adjoint_indexed_kernel(const S* __restrict__ f,
                                                                        float* __restrict__ lam, AdjIndexedArgs a,
                                                                        K17Rates r)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const uint32_t k = a.indices[i];
    uint32_t n[8];
    n[0] = k;
    n[1] = a.nbx[k];
    n[2] = a.nby[k];
    n[4] = a.nbz[k];
    n[3] = a.nby[n[1]];
    n[5] = a.nbz[n[1]];
    n[6] = a.nbz[n[2]];
    n[7] = a.nbz[n[3]];
    float v[27], g[27], fb[27], ub[3], sb = 0.0f;
    adj_gather<S, EVEN>(f, lam, a.N, n, v, g);
    const float vs = a.porosity ? 1.0f - 0.5f * a.porosity[k] : 1.0f;
    k17_vjp(v, g, r, fb, SHIFT ? ub : nullptr, vs, a.porosity ? &sb : nullptr);
    adj_scatter<EVEN>(lam, a.N, n, fb);
    if (SHIFT) {
        a.shiftAdj[k] += ub[0];
        a.shiftAdj[a.N + k] += ub[1];
        a.shiftAdj[2 * a.N + k] += ub[2];
    }
    brinkman_adjoint(a.porosity, a.porosityAdj, k, sb);
}

int check_adjoint_args(const vf_lbm_parameter* p, const float* adjoint)
{
    if (!p || !p->distributions || !adjoint) {
        set_last_error("adjoint: null level, populations or adjoint buffer");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->numberOfNodes == 0 || p->numberOfNodes >= (1ull << 30)) {
        set_last_error("adjoint: numberOfNodes must lie in [1, 2^30)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!(p->omega > 0.0f && p->omega < 2.0f)) {
        set_last_error("adjoint: omega must lie in (0, 2)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->storage != VF_STORAGE_FP32 && p->storage != VF_STORAGE_FP16 && p->storage != VF_STORAGE_BF16) {
        set_last_error("adjoint: unknown storage type");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return VF_OK;
}

template <class S, bool EVEN>
void launch_box(const vf_lbm_parameter& p, float* lam, float* shift, float* padj, hipStream_t s)
{
    AdjBoxArgs a{p.numberOfNodes, p.nx, p.ny, p.nz, p.fluidMask, shift, p.porosity, padj};
    const K17Rates r = make_k17_rates(p);
    const int bx = p.nx >= kAdjBlock ? kAdjBlock : ((p.nx + 63) / 64) * 64;
    const dim3 grid((p.nx + bx - 1) / bx, p.ny, p.nz);
    const S* f = static_cast<const S*>(p.distributions);
    const bool mask = p.fluidMask != nullptr;
    if (mask && shift) hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, true, true>), grid, dim3(bx), 0, s, f, lam, a, r);
    else if (mask) hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, true, false>), grid, dim3(bx), 0, s, f, lam, a, r);
    else if (shift) hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, false, true>), grid, dim3(bx), 0, s, f, lam, a, r);
    else hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, false, false>), grid, dim3(bx), 0, s, f, lam, a, r);
}

template <class S, bool EVEN>
void launch_indexed(const vf_lbm_parameter& p, float* lam, float* shift, float* padj, hipStream_t s)
{
    AdjIndexedArgs a{p.numberOfNodes, p.neighborX, p.neighborY, p.neighborZ, p.fluidNodeIndices,
                     p.numberOfFluidNodes, shift, p.porosity, padj};
    const K17Rates r = make_k17_rates(p);
    const dim3 grid((p.numberOfFluidNodes + kAdjBlock - 1) / kAdjBlock);
    const S* f = static_cast<const S*>(p.distributions);
    if (shift) hipLaunchKernelGGL((k17_adjoint_indexed_kernel<S, EVEN, true>), grid, dim3(kAdjBlock), 0, s, f, lam, a, r);
    else hipLaunchKernelGGL((k17_adjoint_indexed_kernel<S, EVEN, false>), grid, dim3(kAdjBlock), 0, s, f, lam, a, r);
}

template <class F>
void dispatch(const vf_lbm_parameter& p, F&& launch)
{
    const bool even = p.isEvenTimestep != 0;
    switch (p.storage) {
    case VF_STORAGE_FP32: even ? launch((float*)nullptr, std::true_type{}) : launch((float*)nullptr, std::false_type{}); break;
    case VF_STORAGE_FP16: even ? launch((__half*)nullptr, std::true_type{}) : launch((__half*)nullptr, std::false_type{}); break;
    default: even ? launch((__hip_bfloat16*)nullptr, std::true_type{}) : launch((__hip_bfloat16*)nullptr, std::false_type{}); break;
    }
}

} // namespace
} // namespace vfa

using namespace vfa;

extern "C" int vf_k17_adjoint_box(const vf_lbm_parameter* p, float* adjoint, float* velocityShiftAdjoint,
                                  float* porosityAdjoint, void* stream)
{
    const int st = check_adjoint_args(p, adjoint);
    if (st != VF_OK) return st;
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0 || (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes ||
        p->ny > 65535 || p->nz > 65535) {
        set_last_error("adjoint box: nx*ny*nz must equal numberOfNodes (ny, nz <= 65535)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    hipStream_t s = (hipStream_t)stream;
    dispatch(*p, [&](auto tag, auto even) {
        using S = std::remove_pointer_t<decltype(tag)>;
        launch_box<S, decltype(even)::value>(*p, adjoint, velocityShiftAdjoint, porosityAdjoint, s);
    });
    return check_launch("k17_adjoint_box_kernel execution failed");
}

extern "C" int vf_k17_adjoint(const vf_lbm_parameter* p, float* adjoint, float* velocityShiftAdjoint,
                              float* porosityAdjoint, void* stream)
{
    const int st = check_adjoint_args(p, adjoint);
    if (st != VF_OK) return st;
    if (!p->neighborX || !p->neighborY || !p->neighborZ || !p->fluidNodeIndices) {
        set_last_error("adjoint: neighbour arrays and fluidNodeIndices are required");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->numberOfFluidNodes == 0) return VF_OK;
    hipStream_t s = (hipStream_t)stream;
    dispatch(*p, [&](auto tag, auto even) {
        using S = std::remove_pointer_t<decltype(tag)>;
        launch_indexed<S, decltype(even)::value>(*p, adjoint, velocityShiftAdjoint, porosityAdjoint, s);
    });
    return check_launch("k17_adjoint_indexed_kernel execution failed");
}
