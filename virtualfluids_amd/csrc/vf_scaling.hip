// Grid-refinement interpolation kernels (compressible, K17 chimera moments).
//
//   coarse -> fine  scaleCoarseToFineCompressible_Device (src/gpu/core/GridScaling/scaleCoarseToFineCompressible.cu:305-528)
//   fine -> coarse  scaleFineToCoarseCompressible_Device (src/gpu/core/GridScaling/scaleFineToCoarseCompressible.cu:76-129)
//   arithmetic      src/lbm/interpolation/InterpolationCoefficients.h, src/lbm/refinement/Interpolation{CF,FC}.h
//
// Eight lanes per interface cell. The 8 source nodes of a cell are the corners of the cell
// whose lowest corner is the listed index, n = ix + 2 iy + 4 iz (M = 0, P = 1); each source
// node's pre-collision state gives drho, u and the non-equilibrium second moments; every
// coefficient of the interpolation polynomials is a short combination of the signed sums
// S_p(q) = sum_n p(n) q(n), p in {1, sx, sy, sz, sxsy, sxsz, sysz, sxsysz}. CF writes the 8
// fine nodes at (+-1/4, +-1/4, +-1/4) of the coarse cell with the fine grid's even parity
// (the fine grid has taken two steps); FC writes the one coarse node at the centre of the
// fine cell with the coarse parity, reading the fine grid with the even parity.
#include <mutex>

#include "vf_lattice.h"
#include "vf_scaling.h"

namespace vfa {

// the 27 pre-collision populations of node k
template <class S>
__device__ __forceinline__ void node_populations(const S* f, const Addr& a, uint32_t k, bool even, float (&v)[27])
{
    uint32_t n[8];
    cell_nodes(a, k, n);
#pragma unroll
    for (int d = 0; d < 27; ++d) v[d] = read_pre(f, a, n, d, even);
}

__device__ __forceinline__ uint32_t corner(const Addr& a, uint32_t base, int n)
{
    uint32_t k = base;
    if (n & 1) k = a.nbx ? a.nbx[k] : (k % a.nx + 1 == (uint32_t)a.nx ? k + 1 - a.nx : k + 1);
    if (n & 2) {
        if (a.nbx) k = a.nby[k];
        else {
            const uint32_t y = (k / a.nx) % a.ny;
            k = (y + 1 == (uint32_t)a.ny) ? k - (uint32_t)a.nx * (a.ny - 1) : k + a.nx;
        }
    }
    if (n & 4) {
        if (a.nbx) k = a.nbz[k];
        else {
            const uint32_t z = k / ((uint32_t)a.nx * a.ny);
            k = (z + 1 == (uint32_t)a.nz) ? k - (uint32_t)a.nx * a.ny * (a.nz - 1) : k + (uint32_t)a.nx * a.ny;
        }
    }
    return k;
}

struct ScaleArgs {
    const uint32_t* cellsC;
    const uint32_t* cellsF;
    uint32_t n;
    const float* ox;
    const float* oy;
    const float* oz;
    float omegaC, omegaF;
};

// Eight lanes per interface cell (a 64-lane workgroup holds 8 cells): lane q loads and reduces
// source node q, the 8 lanes trade their source moments through LDS, each lane forms the cell's
// coefficients (the same arithmetic in every lane), and lane q then rebuilds fine child q (CF)
// or an eighth of the coarse node's 27 populations (FC). One cell per thread would leave most
// of the chip idle: an interface has tens of thousands of cells, each a long serial chain.
constexpr int kScaleBlock = 64;
constexpr int kScaleCells = kScaleBlock / 8;

struct CellLane {
    uint32_t i;  // interface cell
    int q;       // lane within the cell: source corner, fine child, or slice of the 27 directions
    bool valid;  // cell < number of cells (the other lanes still reach every barrier)
};

__device__ __forceinline__ CellLane cell_lane(uint32_t n)
{
    CellLane c;
    c.i = blockIdx.x * kScaleCells + (threadIdx.x >> 3);
    c.q = threadIdx.x & 7;
    c.valid = c.i < n;
    return c;
}

// the 8 source moments of this lane's cell: lane q computes corner q, LDS hands them round;
// own (optional) receives corner q's population sums, all its reverse needs later
template <class S>
__device__ __forceinline__ void cell_source_moments(const S* f, const Addr& a, uint32_t base, const CellLane& cl,
                                                    bool even, float omega, SrcMom (&m)[8], SrcSums* own = nullptr)
{
    __shared__ SrcMom sm[kScaleBlock];
    if (cl.valid) {
        float v[27];
        node_populations(f, a, corner(a, base, cl.q), even, v);
        const SrcSums sums = source_sums_of(v);
        sm[threadIdx.x] = source_moments_of_sums(sums, omega);
        if (own) *own = sums;
    }
    __syncthreads();
    const int g0 = threadIdx.x & ~7;
#pragma unroll
    for (int n = 0; n < 8; ++n) m[n] = sm[g0 + n];
}

__device__ __forceinline__ void cell_offsets(const ScaleArgs& s, const CellLane& cl, float& xo, float& yo, float& zo)
{
    xo = (s.ox && cl.valid) ? s.ox[cl.i] : 0.f;
    yo = (s.oy && cl.valid) ? s.oy[cl.i] : 0.f;
    zo = (s.oz && cl.valid) ? s.oz[cl.i] : 0.f;
}

template <class S>
__global__ __launch_bounds__(kScaleBlock) void scale_cf_kernel(const S* __restrict__ fc, S* __restrict__ ff, Addr ac,
                                                               Addr af, ScaleArgs s, int evenC)
{
    const CellLane cl = cell_lane(s.n);
    SrcMom m[8];
    cell_source_moments(fc, ac, cl.valid ? s.cellsC[cl.i] : 0u, cl, evenC != 0, s.omegaC, m);
    if (!cl.valid) return;
    float xo, yo, zo;
    cell_offsets(s, cl, xo, yo, zo);
    const Coef c = coefficients(m, xo, yo, zo);
    const int q = cl.q;
    const NodeMoments nm = cf_node_moments(c, cf_child_offset(q, 0), cf_child_offset(q, 1), cf_child_offset(q, 2),
                                           0.5f, s.omegaF);
    float f[27];
    rebuild_populations(f, nm);
    uint32_t nn[8];
    cell_nodes(af, corner(af, s.cellsF[cl.i], q), nn);
#pragma unroll
    for (int d = 0; d < 27; ++d) write_pre(ff, af, nn, d, true, f[d]);
}

template <class S>
__global__ __launch_bounds__(kScaleBlock) void scale_fc_kernel(S* __restrict__ fc, const S* __restrict__ ff, Addr ac,
                                                               Addr af, ScaleArgs s, int evenC)
{
    const CellLane cl = cell_lane(s.n);
    SrcMom m[8];
    cell_source_moments(ff, af, cl.valid ? s.cellsF[cl.i] : 0u, cl, true, s.omegaF, m);
    if (!cl.valid) return;
    float xo, yo, zo;
    cell_offsets(s, cl, xo, yo, zo);
    const Coef c = coefficients(m, xo, yo, zo);
    float f[27];
    rebuild_populations(f, fc_node_moments(c, s.omegaC));
    uint32_t nn[8];
    cell_nodes(ac, s.cellsC[cl.i], nn);
#pragma unroll
    for (int d = 0; d < 27; ++d)
        if ((d & 7) == cl.q) write_pre(fc, ac, nn, d, evenC != 0, f[d]);
}

// ---- reverse mode ------------------------------------------------------------------------
// The adjoint buffers have the population layout of their level (FP32). The cells a forward
// kernel wrote are read and cleared (their old content did not reach the output); the
// transposed interpolation of their adjoint is added to the source nodes' cells, which
// neighbouring interface cells share (atomic adds). Linearisation point: the source state the
// forward kernel read. Per cell the chain is recomputed from that state (nothing is stored).
// Lanes as in the forward kernels: lane q takes fine child q (CF) and source node q.

__device__ __forceinline__ float* adj_cell(float* lam, const Addr& a, const uint32_t (&n)[8], int d, bool even)
{
    return slot_ptr(lam, a.N, d, even) + n[node_sel(d)];
}

// sum of a Coef over the 8 lanes of a cell (butterfly within the lane group)
__device__ __forceinline__ void coef_group_sum(Coef& c)
{
    float* p = &c.a000;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(Coef) / sizeof(float)); ++k) {
        float v = p[k];
        v += __shfl_xor(v, 1, 8);
        v += __shfl_xor(v, 2, 8);
        v += __shfl_xor(v, 4, 8);
        p[k] = v;
    }
}

// adjoint of source node k's moments back to its 27 cells (atomic adds: cells share source
// nodes), from the node's population sums
__device__ __forceinline__ void source_node_adjoint(float* lam, const Addr& a, uint32_t k, bool even, float omega,
                                                    const SrcSums& sums, const SrcMom& mb)
{
    float fb[27];
    source_moments_vjp_sums(sums, omega, mb, fb);
    uint32_t nn[8];
    cell_nodes(a, k, nn);
#pragma unroll
    for (int d = 0; d < 27; ++d) atomicAdd(adj_cell(lam, a, nn, d, even), fb[d]);
}

// GATHER: lane l = 8 cell + corner stores the adjoint of its source node's moments in mbOut
// (field-major, 9 x 8 n floats) instead of adding it to the node's cells; source_gather_kernel
// then sums the lanes of every source node and applies it once per node
template <class S, bool GATHER = false>
__global__ __launch_bounds__(kScaleBlock) void scale_cf_adjoint_kernel(const S* __restrict__ fc,
                                                                       float* __restrict__ lamC,
                                                                       float* __restrict__ lamF, Addr ac, Addr af,
                                                                       ScaleArgs s, int evenC, float* mbOut = nullptr)
{
    const CellLane cl = cell_lane(s.n);
    SrcMom m[8];
    SrcSums own;
    cell_source_moments(fc, ac, cl.valid ? s.cellsC[cl.i] : 0u, cl, evenC != 0, s.omegaC, m, GATHER ? nullptr : &own);
    float xo, yo, zo;
    cell_offsets(s, cl, xo, yo, zo);
    const Coef c = coefficients(m, xo, yo, zo);
    Coef cb = coef_zero();
    const int q = cl.q;
    if (cl.valid) {
        uint32_t nn[8];
        cell_nodes(af, corner(af, s.cellsF[cl.i], q), nn);
        // the 27 cells are distinct: read them all, then clear them (a clear between the reads
        // would make every read wait for the previous one)
        float g[27];
#pragma unroll
        for (int d = 0; d < 27; ++d) g[d] = *adj_cell(lamF, af, nn, d, true);
#pragma unroll
        for (int d = 0; d < 27; ++d) *adj_cell(lamF, af, nn, d, true) = 0.f;
        const float x = cf_child_offset(q, 0), y = cf_child_offset(q, 1), z = cf_child_offset(q, 2);
        const NodeMoments nb = rebuild_vjp(cf_node_moments(c, x, y, z, 0.5f, s.omegaF), g);
        cf_node_moments_vjp(c, x, y, z, 0.5f, s.omegaF, nb, cb);
    }
    coef_group_sum(cb);  // every lane of the group takes part (invalid cells add zeros)
    if (!cl.valid) return;
    const SrcMom mb = node_of_sums(coefficients_vjp_sums(m, xo, yo, zo, cb), q);
    if constexpr (GATHER) {
        const uint64_t L = 8ull * s.n, l = 8ull * cl.i + (uint64_t)q;
        const float* src = &mb.drho;
#pragma unroll
        for (int k = 0; k < 9; ++k) mbOut[k * L + l] = src[k];
    } else {
        source_node_adjoint(lamC, ac, corner(ac, s.cellsC[cl.i], q), evenC != 0, s.omegaC, own, mb);
    }
}

// The distinct source nodes of a set of interface cells: node[u], and first[u] .. first[u+1]
// indexing `lanes`, the lanes 8 cell + corner that read node u.
struct SourceArgs {
    const uint32_t* node;
    const uint32_t* first;
    const uint32_t* lanes;
    uint32_t count;
    const float* mb; // 9 x L partial adjoints of the source moments, field-major
    uint32_t L;      // 8 x cells
    float omega;
};

// One thread per source node: the sum of its lanes' moment adjoints, taken back to the node's 27
// cells once (source_node_adjoint without atomics, and one read of the node's populations
// instead of one per cell that shares it). Lane and node indices outside the arrays are skipped.
template <class S>
__global__ __launch_bounds__(64) void source_gather_kernel(const S* __restrict__ f, float* __restrict__ lam, Addr a,
                                                           SourceArgs s, int even_flag)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= s.count) return;
    const uint32_t k = s.node[u];
    if ((uint64_t)k >= a.N) return;
    float acc9[9] = {};
    for (uint32_t j = s.first[u], e = s.first[u + 1]; j < e; ++j) {
        const uint32_t l = s.lanes[j];
        if (l >= s.L) continue;
#pragma unroll
        for (int c = 0; c < 9; ++c) acc9[c] += s.mb[(uint64_t)c * s.L + l];
    }
    SrcMom mb;
    float* dst = &mb.drho;
#pragma unroll
    for (int c = 0; c < 9; ++c) dst[c] = acc9[c];
    const bool even = even_flag != 0;
    float v[27], fb[27], old[27];
    node_populations(f, a, k, even, v);
    source_moments_vjp(v, s.omega, mb, fb);
    uint32_t nn[8];
    cell_nodes(a, k, nn);
#pragma unroll
    for (int d = 0; d < 27; ++d) old[d] = *adj_cell(lam, a, nn, d, even);
#pragma unroll
    for (int d = 0; d < 27; ++d) *adj_cell(lam, a, nn, d, even) = old[d] + fb[d];
}

template <class S>
__global__ __launch_bounds__(kScaleBlock) void scale_fc_adjoint_kernel(const S* __restrict__ ff,
                                                                       float* __restrict__ lamC,
                                                                       float* __restrict__ lamF, Addr ac, Addr af,
                                                                       ScaleArgs s, int evenC)
{
    const CellLane cl = cell_lane(s.n);
    SrcMom m[8];
    SrcSums own;
    cell_source_moments(ff, af, cl.valid ? s.cellsF[cl.i] : 0u, cl, true, s.omegaF, m, &own);
    // every lane reads the written coarse node's adjoint; after the barrier lane q clears its
    // slice of the 27 cells
    uint32_t nc[8];
    float g[27];
    if (cl.valid) {
        cell_nodes(ac, s.cellsC[cl.i], nc);
#pragma unroll
        for (int d = 0; d < 27; ++d) g[d] = *adj_cell(lamC, ac, nc, d, evenC != 0);
    }
    __syncthreads();
    if (!cl.valid) return;
#pragma unroll
    for (int d = 0; d < 27; ++d)
        if ((d & 7) == cl.q) *adj_cell(lamC, ac, nc, d, evenC != 0) = 0.f;
    float xo, yo, zo;
    cell_offsets(s, cl, xo, yo, zo);
    const Coef c = coefficients(m, xo, yo, zo);
    Coef cb = coef_zero();
    fc_node_moments_vjp(c, s.omegaC, rebuild_vjp(fc_node_moments(c, s.omegaC), g), cb);
    const SrcMom mb = node_of_sums(coefficients_vjp_sums(m, xo, yo, zo, cb), cl.q);
    source_node_adjoint(lamF, af, corner(af, s.cellsF[cl.i], cl.q), true, s.omegaF, own, mb);
}

static int check_level(const vf_lbm_parameter* p)
{
    if (!p || !p->distributions || p->numberOfNodes == 0 || p->numberOfNodes >= (1ull << 30)) return 0;
    if (!p->neighborX && (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes) return 0;
    return 1;
}

static_assert(sizeof(SrcMom) == 9 * sizeof(float), "the gather workspace holds SrcMom as 9 floats");

// The library-owned workspace of the gathered coarse-to-fine adjoint for callers that cannot
// allocate (vf_ScaleCoarseToFineCompressibleAdjointGatherNoAlloc): 9 floats per cell corner for at
// most kNoAllocCells cells (18 MiB), one copy per GPU, held until the process ends. Host code
// never passes this symbol itself to a kernel: it passes the device address that
// hipGetSymbolAddress returns for the current device.
constexpr uint32_t kNoAllocCells = 1u << 16;
static __device__ __align__(256) float scratch[9u * 8u * kNoAllocCells];

// The device address of `scratch` on the current device, looked up once per device.
static int scratch_address(float** out)
{
    static std::mutex lock;
    static float* address[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_last_error("no-alloc grid scaling adjoint: no current device (or a device index above 63)");
        return VF_ERR_HIP;
    }
    std::lock_guard<std::mutex> guard(lock);
    if (!address[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_last_error("no-alloc grid scaling adjoint: hipGetSymbolAddress of the workspace failed");
            return VF_ERR_HIP;
        }
        address[dev] = static_cast<float*>(ptr);
    }
    *out = address[dev];
    return VF_OK;
}

// lamC/lamF null: the forward kernels; otherwise their adjoints. sources/workspace (the
// coarse-to-fine adjoint only): per-source-node gather instead of atomic adds
static int launch_scale(bool fine_to_coarse, const vf_lbm_parameter* c, const vf_lbm_parameter* f,
                        const vf_interpolation_cells* cells, const vf_interpolation_cell_neighbor* off, void* stream,
                        float* lamC = nullptr, float* lamF = nullptr, const vf_interpolation_sources* sources = nullptr,
                        float* workspace = nullptr)
{
    const bool adjoint = lamC || lamF;
    if (adjoint && (!lamC || !lamF)) {
        set_last_error("grid scaling adjoint needs the adjoint buffers of both levels");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!check_level(c) || !check_level(f) || !cells) {
        set_last_error("grid scaling needs two valid levels and the interface cells");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (cells->numberOfCells == 0) return VF_OK;
    if (!cells->coarseCellIndices || !cells->fineCellIndices) {
        set_last_error("interface cells need coarse and fine cell indices");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (c->storage != f->storage) {
        set_last_error("both levels must use the same population storage type");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!(c->omega > 0.f && c->omega < 2.f && f->omega > 0.f && f->omega < 2.f)) {
        set_last_error("omega must lie in (0, 2) on both levels");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (sources && (!sources->node || !sources->first || !sources->lanes || !workspace)) {
        set_last_error("grid scaling adjoint gather needs node, first and lanes of the sources and a workspace");
        return VF_ERR_INVALID_ARGUMENT;
    }
    ScaleArgs s{cells->coarseCellIndices, cells->fineCellIndices, cells->numberOfCells,
                off ? off->x : nullptr, off ? off->y : nullptr, off ? off->z : nullptr, c->omega, f->omega};
    const Addr ac = make_addr(*c), af = make_addr(*f);
    const dim3 grid((cells->numberOfCells + kScaleCells - 1) / kScaleCells), block(kScaleBlock);
    hipStream_t st = (hipStream_t)stream;
    const SourceArgs src{sources ? sources->node : nullptr, sources ? sources->first : nullptr,
                         sources ? sources->lanes : nullptr, sources ? sources->numberOfNodes : 0u, workspace,
                         8u * cells->numberOfCells, c->omega};
    const dim3 sgrid((src.count + 63) / 64);
#define VF_LAUNCH(S)                                                                                                 \
    if (sources) {                                                                                                   \
        hipLaunchKernelGGL((scale_cf_adjoint_kernel<S, true>), grid, block, 0, st, (const S*)c->distributions, lamC, \
                           lamF, ac, af, s, c->isEvenTimestep, workspace);                                           \
        if (src.count)                                                                                               \
            hipLaunchKernelGGL(source_gather_kernel<S>, sgrid, dim3(64), 0, st, (const S*)c->distributions, lamC, ac, \
                               src, c->isEvenTimestep);                                                              \
    } else if (adjoint && fine_to_coarse)                                                                            \
        hipLaunchKernelGGL(scale_fc_adjoint_kernel<S>, grid, block, 0, st, (const S*)f->distributions, lamC, lamF,  \
                           ac, af, s, c->isEvenTimestep);                                                            \
    else if (adjoint)                                                                                                \
        hipLaunchKernelGGL((scale_cf_adjoint_kernel<S, false>), grid, block, 0, st, (const S*)c->distributions, lamC, \
                           lamF, ac, af, s, c->isEvenTimestep, nullptr);                                             \
    else if (fine_to_coarse)                                                                                         \
        hipLaunchKernelGGL(scale_fc_kernel<S>, grid, block, 0, st, (S*)c->distributions, (const S*)f->distributions, \
                           ac, af, s, c->isEvenTimestep);                                                            \
    else                                                                                                             \
        hipLaunchKernelGGL(scale_cf_kernel<S>, grid, block, 0, st, (const S*)c->distributions, (S*)f->distributions, \
                           ac, af, s, c->isEvenTimestep);
    switch (c->storage) {
    case VF_STORAGE_FP32: VF_LAUNCH(float) break;
    case VF_STORAGE_FP16: VF_LAUNCH(__half) break;
    case VF_STORAGE_BF16: VF_LAUNCH(__hip_bfloat16) break;
    default: set_last_error("unknown storage type"); return VF_ERR_INVALID_ARGUMENT;
    }
#undef VF_LAUNCH
    if (adjoint) return check_launch(fine_to_coarse ? "scaleFineToCoarseCompressible adjoint" : "scaleCoarseToFineCompressible adjoint");
    return check_launch(fine_to_coarse ? "scaleFineToCoarseCompressible_Device execution failed"
                                       : "scaleCoarseToFineCompressible_Device execution failed");
}

} // namespace vfa

using namespace vfa;

extern "C" int vf_ScaleCoarseToFineCompressible(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                                const vf_interpolation_cells* coarseToFine,
                                                const vf_interpolation_cell_neighbor* offsets, void* stream)
{
    return launch_scale(false, coarse, fine, coarseToFine, offsets, stream);
}

extern "C" int vf_ScaleFineToCoarseCompressible(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                                const vf_interpolation_cells* fineToCoarse,
                                                const vf_interpolation_cell_neighbor* offsets, void* stream)
{
    return launch_scale(true, coarse, fine, fineToCoarse, offsets, stream);
}

extern "C" int vf_ScaleCoarseToFineCompressibleAdjoint(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                                       const vf_interpolation_cells* coarseToFine,
                                                       const vf_interpolation_cell_neighbor* offsets,
                                                       float* coarseAdjoint, float* fineAdjoint, void* stream)
{
    return launch_scale(false, coarse, fine, coarseToFine, offsets, stream, coarseAdjoint, fineAdjoint);
}

extern "C" int vf_ScaleCoarseToFineCompressibleAdjointGather(const vf_lbm_parameter* coarse,
                                                             const vf_lbm_parameter* fine,
                                                             const vf_interpolation_cells* coarseToFine,
                                                             const vf_interpolation_cell_neighbor* offsets,
                                                             const vf_interpolation_sources* sources, float* workspace,
                                                             float* coarseAdjoint, float* fineAdjoint, void* stream)
{
    if (!sources) {
        set_last_error("grid scaling adjoint gather needs the source nodes of the cells");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!coarseAdjoint || !fineAdjoint) {
        set_last_error("grid scaling adjoint needs the adjoint buffers of both levels");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return launch_scale(false, coarse, fine, coarseToFine, offsets, stream, coarseAdjoint, fineAdjoint, sources, workspace);
}

extern "C" int vf_ScaleCoarseToFineCompressibleAdjointGatherNoAlloc(const vf_lbm_parameter* coarse,
                                                                    const vf_lbm_parameter* fine,
                                                                    const vf_interpolation_cells* coarseToFine,
                                                                    const vf_interpolation_cell_neighbor* offsets,
                                                                    const vf_interpolation_sources* sources,
                                                                    float* coarseAdjoint, float* fineAdjoint,
                                                                    void* stream)
{
    if (!sources || !coarseToFine) {
        set_last_error("grid scaling adjoint gather needs the interface cells and their source nodes");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (coarseToFine->numberOfCells > kNoAllocCells) {
        set_last_error("no-alloc grid scaling adjoint: at most 65536 interface cells (the library's workspace)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!coarseAdjoint || !fineAdjoint) {
        set_last_error("grid scaling adjoint needs the adjoint buffers of both levels");
        return VF_ERR_INVALID_ARGUMENT;
    }
    float* work = nullptr;
    const int rc = scratch_address(&work);
    if (rc != VF_OK) return rc;
    // every lane of every cell overwrites its 9 workspace floats before the gather reads them:
    // nothing needs clearing between calls
    return launch_scale(false, coarse, fine, coarseToFine, offsets, stream, coarseAdjoint, fineAdjoint, sources, work);
}

extern "C" int vf_ScaleFineToCoarseCompressibleAdjoint(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                                       const vf_interpolation_cells* fineToCoarse,
                                                       const vf_interpolation_cell_neighbor* offsets,
                                                       float* coarseAdjoint, float* fineAdjoint, void* stream)
{
    return launch_scale(true, coarse, fine, fineToCoarse, offsets, stream, coarseAdjoint, fineAdjoint);
}
