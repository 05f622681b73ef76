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
#include <cstdlib>
#include <type_traits>

#include "vf_k17_adjoint.h"
#include "vf_lattice.h"
#include "vf_rowio.h"

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
    uint32_t rowEnd = 0;   // row kernel: columns [0, rowEnd) (0: the whole row)
    uint32_t x0 = 0;       // scalar kernel: 0 for every node, else the columns [x0, nx) only
    PlaneDiv plane{};      // scalar kernel over every node: plane-flat numbering
};

// (x, y, z) of this thread's node in the scalar kernel; false for threads beyond the nodes.
// XCD: the plane-flat workgroups taken in XCD block ranges (as the forward node kernel).
template <bool XCD = false>
__device__ __forceinline__ bool adj_coords(const AdjBoxArgs& a, int& x, int& y, int& z)
{
    if (a.x0 == 0) {
        uint32_t bx = blockIdx.x, bz = blockIdx.y;
        if constexpr (XCD) {
            const uint32_t b = xcd_block(bx + gridDim.x * bz, gridDim.x * gridDim.y);
            bx = b % gridDim.x;
            bz = b / gridDim.x;
        }
        const uint32_t t = bx * blockDim.x + threadIdx.x;
        if (t >= a.plane.planeNodes) return false;
        const uint32_t row = plane_row(t, a.plane);
        x = (int)(t - row * a.plane.nx);
        y = (int)row;
        z = (int)bz;
        return true;
    }
    const uint32_t tw = (uint32_t)a.nx - a.x0;
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= tw * (uint32_t)a.ny * (uint32_t)a.nz) return false;
    const uint32_t r = id / tw;
    x = (int)(a.x0 + id % tw);
    y = (int)(r % (uint32_t)a.ny);
    z = (int)(r / (uint32_t)a.ny);
    return true;
}

// adjoint of the velocity scale s = 1 - porosity/2 goes to the porosity with the factor -1/2
__device__ __forceinline__ void brinkman_adjoint(const float* porosity, float* porosityAdj, uint32_t k, float sb)
{
    if (porosity && porosityAdj) porosityAdj[k] += -0.5f * sb;
}

template <class S, bool EVEN, bool MASK, bool SHIFT, bool XCD = false>
__global__ __launch_bounds__(kAdjBlock) void k17_adjoint_box_kernel(const S* __restrict__ f, float* __restrict__ lam,
                                                                    AdjBoxArgs a, K17Rates r)
{
    int x, y, z;
    if (!adj_coords<XCD>(a, x, y, z)) return;
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

// ---- row kernel: one workgroup per x-row, one node per lane -------------------------------
// The access pattern of the forward row kernel: every global access of a wave is aligned and
// contiguous. The state and the adjoint of the nine c_x = -1 directions are read one node
// further in +x, so a lane takes them from lane+1 (__shfl) and gives its result for those
// cells to lane+1; across waves and around the periodic wrap the boundary lanes trade through
// LDS. Lanes beyond the row idle. Solid nodes write their adjoint cells back unchanged.
constexpr int kAdjRowMaxThreads = 512;

template <class S, bool EVEN, bool MASK, bool SHIFT, bool BUF>
__global__ __launch_bounds__(kAdjRowMaxThreads) void k17_adjoint_row_kernel(const S* f, float* lam, AdjBoxArgs a,
                                                                           K17Rates r)
{
    using IO = RowIO<S, 1>;
    using LO = RowIO<float, 1>;
    __shared__ unsigned xinS[kAdjRowMaxThreads / 64][9];
    __shared__ float xinL[kAdjRowMaxThreads / 64][9];
    __shared__ float xout[kAdjRowMaxThreads / 64][9];
    __shared__ float xlast[9];
    __shared__ bool xoutFluid[kAdjRowMaxThreads / 64];
    __shared__ bool xlastFluid;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny;
    // split: the kernel covers the columns [0, rowEnd) and a tail launch of the scalar kernel
    // the rest; the cells either side of that seam are read and written one by one
    const uint32_t rowEnd = a.rowEnd ? a.rowEnd : nx;
    const bool split = rowEnd < nx;
    const bool active = threadIdx.x < rowEnd;
    const bool lastLane = threadIdx.x + 1 == rowEnd;
    const uint32_t y = blockIdx.x, z = blockIdx.y;
    const uint32_t yp = (y + 1 == ny) ? 0 : y + 1;
    const uint32_t zp = (z + 1 == (uint32_t)a.nz) ? 0 : z + 1;
    // idle lanes read the row's last node (so that every load is unconditional and in bounds)
    // and drop their results
    const uint32_t x = active ? threadIdx.x : rowEnd - 1;
    // this lane's node in the 4 rows its cells touch, indexed by the y/z bits of node_sel
    const uint32_t rows[4] = {nx * (y + ny * z) + x, nx * (yp + ny * z) + x, nx * (y + ny * zp) + x,
                              nx * (yp + ny * zp) + x};
    const uint64_t slotS = a.N * sizeof(S), slotL = a.N * sizeof(float);
    __amdgpu_buffer_rsrc_t sLo, sHi, lLo, lHi;
    if constexpr (BUF) {
        void* fb0 = const_cast<S*>(f);
        sLo = __builtin_amdgcn_make_buffer_rsrc(fb0, (short)0, (int)0xffffffff, 0x00020000);
        sHi = __builtin_amdgcn_make_buffer_rsrc(static_cast<char*>(fb0) + kRowSplit * slotS, (short)0, (int)0xffffffff,
                                                0x00020000);
        lLo = __builtin_amdgcn_make_buffer_rsrc(lam, (short)0, (int)0xffffffff, 0x00020000);
        lHi = __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<char*>(lam) + kRowSplit * slotL, (short)0,
                                                (int)0xffffffff, 0x00020000);
    }
    auto soff = [&](int d, uint64_t slotBytes) -> uint32_t {
        const int sl = slot<EVEN>(d);
        return (uint32_t)(sl >= kRowSplit ? sl - kRowSplit : sl) * (uint32_t)slotBytes;
    };
    auto loadS = [&](int d) {
        const uint32_t off = rows[node_sel(d) >> 1] * (uint32_t)sizeof(S);
        if constexpr (BUF) return IO::load(slot<EVEN>(d) >= kRowSplit ? sHi : sLo, off, soff(d, slotS));
        else return IO::load(reinterpret_cast<const S*>(reinterpret_cast<const char*>(f) + slot<EVEN>(d) * slotS + off));
    };
    auto loadL = [&](int d) {
        const uint32_t off = rows[node_sel(d) >> 1] * (uint32_t)sizeof(float);
        if constexpr (BUF) return LO::load(slot<EVEN>(d) >= kRowSplit ? lHi : lLo, off, soff(d, slotL));
        else return LO::load(reinterpret_cast<const float*>(reinterpret_cast<const char*>(lam) + slot<EVEN>(d) * slotL + off));
    };
    // direction d's cell at column cx of this lane's row (split rows only)
    auto cellS = [&](int d, uint32_t cx) {
        return f + (uint64_t)slot<EVEN>(d) * a.N + (rows[node_sel(d) >> 1] - x + cx);
    };
    auto cellL = [&](int d, uint32_t cx) {
        return lam + (uint64_t)slot<EVEN>(d) * a.N + (rows[node_sel(d) >> 1] - x + cx);
    };
    auto storeL = [&](int d, float val) {
        const uint32_t off = rows[node_sel(d) >> 1] * (uint32_t)sizeof(float);
        const float vals[1] = {val};
        if constexpr (BUF) LO::store(slot<EVEN>(d) >= kRowSplit ? lHi : lLo, off, soff(d, slotL), vals);
        else LO::store(reinterpret_cast<float*>(reinterpret_cast<char*>(lam) + slot<EVEN>(d) * slotL + off), vals);
    };

    typename IO::Raw rawS[27];
    typename LO::Raw rawL[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        rawS[d] = loadS(d);
        rawL[d] = loadL(d);
    }

    // state and adjoint of the cells one node further in +x, for the c_x = -1 directions
    unsigned nxtS[9];
    float nxtL[9];
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            xinS[wave][j] = IO::word0(rawS[kShiftedDirs[j]]);
            xinL[wave][j] = LO::elem(rawL[kShiftedDirs[j]], 0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        nxtS[j] = __shfl(IO::word0(rawS[kShiftedDirs[j]]), (lane + 1) & 63, 64);
        nxtL[j] = __shfl(LO::elem(rawL[kShiftedDirs[j]], 0), (lane + 1) & 63, 64);
    }
    if (active && (lane == 63 || lastLane)) {
        if (lastLane && split) {
            // the first tail column
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                nxtS[j] = IO::elem_bits(cellS(kShiftedDirs[j], rowEnd));
                nxtL[j] = *cellL(kShiftedDirs[j], rowEnd);
            }
        } else {
            // the next wave's first node, or the row's first node around the periodic wrap
            const int src = lastLane ? 0 : wave + 1;
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                nxtS[j] = xinS[src][j];
                nxtL[j] = xinL[src][j];
            }
        }
    }

    // state v[d] and adjoint of the post-collision values g[inv(d)], both from the cell of pre[d]
    float v[27], g[27], fb[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        if (node_sel(d) & 1) {
            v[d] = IO::of_word(nxtS[shifted_index(d)], 0);
            g[inv(d)] = nxtL[shifted_index(d)];
        } else {
            v[d] = IO::elem(rawS[d], 0);
            g[inv(d)] = LO::elem(rawL[d], 0);
        }
    }
    const uint32_t k = rows[0];
    bool fluid = active;
    if (MASK && active) fluid = a.mask[k] != 0;
    if (fluid) {
        float ub[3], sb = 0.0f;
        const float vs = a.porosity ? 1.0f - 0.5f * a.porosity[k] : 1.0f;
        k17_vjp(v, g, r, fb, SHIFT ? ub : nullptr, vs, a.porosity ? &sb : nullptr);
        if (SHIFT) {
            a.shiftAdj[k] += ub[0];
            a.shiftAdj[a.N + k] += ub[1];
            a.shiftAdj[2 * a.N + k] += ub[2];
        }
        brinkman_adjoint(a.porosity, a.porosityAdj, k, sb);
    } else {
        // the forward step leaves a solid node's cells alone: their adjoint passes through
#pragma unroll
        for (int d = 0; d < 27; ++d) fb[d] = g[inv(d)];
    }

    // A solid node's cells keep their adjoint (the forward step leaves them alone): their stores
    // are skipped. Cells without an x-shift are this lane's own
    if (active && fluid) {
#pragma unroll
        for (int d = 0; d < 27; ++d) {
            if (node_sel(d) & 1) continue;
            storeL(d, fb[d]);
        }
    }
    // the shifted cells belong to lane+1: take lane-1's, the previous wave's last lane's, or
    // (first lane of the row) the row's last lane's; the owner's fluid flag travels with them
    if (active && lane == 63) {
#pragma unroll
        for (int j = 0; j < 9; ++j) xout[wave][j] = fb[kShiftedDirs[j]];
        xoutFluid[wave] = fluid;
    }
    if (lastLane) {
        if (split) {
            // into the first tail column, which the tail launch does not write for these slots
            if (fluid) {
#pragma unroll
                for (int j = 0; j < 9; ++j) *cellL(kShiftedDirs[j], rowEnd) = fb[kShiftedDirs[j]];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 9; ++j) xlast[j] = fb[kShiftedDirs[j]];
            xlastFluid = fluid;
        }
    }
    __syncthreads();
    float s0[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        s0[j] = __shfl(fb[kShiftedDirs[j]], (lane + 63) & 63, 64);
        if (lane == 0) s0[j] = (wave == 0) ? xlast[j] : xout[wave - 1][j];
    }
    bool ownerFluid = __shfl((int)fluid, (lane + 63) & 63, 64) != 0;
    if (lane == 0) ownerFluid = (wave == 0) ? xlastFluid : xoutFluid[wave - 1];
    // split: column 0 of these slots belongs to the last tail node
    if (active && ownerFluid && !(split && threadIdx.x == 0)) {
#pragma unroll
        for (int j = 0; j < 9; ++j) storeL(kShiftedDirs[j], s0[j]);
    }
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

template <class S, bool EVEN, bool BUF>
void launch_row_b(const AdjBoxArgs& a, const K17Rates& r, const S* f, float* lam, bool mask, bool shift, hipStream_t s)
{
    const dim3 grid(a.ny, a.nz), block(((a.rowEnd + 63) / 64) * 64);
    if (mask && shift) hipLaunchKernelGGL((k17_adjoint_row_kernel<S, EVEN, true, true, BUF>), grid, block, 0, s, f, lam, a, r);
    else if (mask) hipLaunchKernelGGL((k17_adjoint_row_kernel<S, EVEN, true, false, BUF>), grid, block, 0, s, f, lam, a, r);
    else if (shift) hipLaunchKernelGGL((k17_adjoint_row_kernel<S, EVEN, false, true, BUF>), grid, block, 0, s, f, lam, a, r);
    else hipLaunchKernelGGL((k17_adjoint_row_kernel<S, EVEN, false, false, BUF>), grid, block, 0, s, f, lam, a, r);
}

template <class S, bool EVEN>
void launch_scalar(const AdjBoxArgs& a, const K17Rates& r, const S* f, float* lam, bool mask, bool shift, dim3 grid,
                   hipStream_t s, bool xcd = false)
{
    const dim3 block(kAdjBlock);
    if (xcd && mask && !shift) hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, true, false, true>), grid, block, 0, s, f, lam, a, r);
    else if (xcd && !mask && !shift) hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, false, false, true>), grid, block, 0, s, f, lam, a, r);
    else if (mask && shift) hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, true, true>), grid, block, 0, s, f, lam, a, r);
    else if (mask) hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, true, false>), grid, block, 0, s, f, lam, a, r);
    else if (shift) hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, false, true>), grid, block, 0, s, f, lam, a, r);
    else hipLaunchKernelGGL((k17_adjoint_box_kernel<S, EVEN, false, false>), grid, block, 0, s, f, lam, a, r);
}

template <class S, bool EVEN>
void launch_box(const vf_lbm_parameter& p, float* lam, float* shift, float* padj, hipStream_t s)
{
    AdjBoxArgs a{p.numberOfNodes, p.nx, p.ny, p.nz, p.fluidMask, shift, p.porosity, padj};
    a.plane = make_plane_div(p.nx, p.ny);
    const K17Rates r = make_k17_rates(p);
    const bool mask = p.fluidMask != nullptr;
    const S* f = static_cast<const S*>(p.distributions);
    // A row is one workgroup of ceil(nx / 64) waves. At ~245 VGPRs a CU holds 8 waves of this
    // kernel, so only rows of 1, 2, 4 or 8 whole waves fill it; those take the row kernel (256^3
    // FP32: 0.99 against 1.09 ms). Every other width takes the scalar kernel, whose plane-flat
    // workgroups keep every lane busy (258 x 130 x 130 FP16: 247 us, against 371 us for 256
    // columns in the row kernel and a tail launch for the last two; measured on MI355X).
    // kernelChoice.adjointKernel pins the scalar kernel, whole rows (ROW_WHOLE), or the row kernel
    // over the whole waves of any row with the scalar kernel on the columns beyond (ROW: a row
    // whose last wave would carry fewer than 32 lanes is split there, 258 = 256 + 2).
    const vf_kernel_choice& kc = p.kernelChoice;
    const bool pinnedRow = kc.adjointKernel == VF_ADJOINT_KERNEL_ROW || kc.adjointKernel == VF_ADJOINT_KERNEL_ROW_WHOLE;
    uint32_t rowEnd = (uint32_t)p.nx;
    if (p.nx > 64 && p.nx % 64 != 0 && p.nx % 64 < 32 && kc.adjointKernel == VF_ADJOINT_KERNEL_ROW)
        rowEnd = p.nx - p.nx % 64;
    const int waves = (int)((rowEnd + 63) / 64);
    const bool wholeWaves = p.nx % 64 == 0 && 8 % waves == 0;
    if (rowEnd <= (uint32_t)kAdjRowMaxThreads && (wholeWaves || pinnedRow) && kc.adjointKernel != VF_ADJOINT_KERNEL_NODE) {
        // the buffer form needs kRowSplit slots of the adjoint (the larger element) within
        // 32-bit offsets
        AdjBoxArgs ar = a;
        ar.rowEnd = rowEnd;
        if ((uint64_t)kRowSplit * a.N * sizeof(float) <= 0xffffffffull && kc.rowAddressing != VF_ROW_ADDRESSING_GLOBAL)
            launch_row_b<S, EVEN, true>(ar, r, f, lam, mask, shift != nullptr, s);
        else
            launch_row_b<S, EVEN, false>(ar, r, f, lam, mask, shift != nullptr, s);
        if (rowEnd == (uint32_t)p.nx) return;
        // the tail columns, one thread per node over all rows
        AdjBoxArgs at = a;
        at.x0 = rowEnd;
        const uint64_t n = (uint64_t)(p.nx - rowEnd) * p.ny * p.nz;
        launch_scalar<S, EVEN>(at, r, f, lam, mask, shift != nullptr, dim3((unsigned)((n + kAdjBlock - 1) / kAdjBlock)), s);
        return;
    }
    static_assert(kAdjBlock == kPlaneBlock, "the scalar adjoint kernel takes plane-flat workgroups");
    // XCD block ranges only where pinned: the adjoint's three streams per direction ran slower
    // with them (sphere channel 125 -> 151 us, 258 x 130 x 130 FP16 227 -> 284 us)
    launch_scalar<S, EVEN>(a, r, f, lam, mask, shift != nullptr, dim3(plane_blocks(p.nx, p.ny), p.nz), s,
                           kc.rowXcdRanges == 1);
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
    return check_launch("k17 adjoint box kernel execution failed");
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
