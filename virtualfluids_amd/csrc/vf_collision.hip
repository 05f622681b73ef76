// K17 collide + esoteric-twist stream kernels for gfx950 and their C entry points.
//
// One thread updates one node in place: it gathers the 27 pre-collision populations
// from the 8 nodes {k, +x, +y, +z, ...} of its cell, collides, and writes the
// post-collision populations back to exactly the addresses it read, in the opposite
// slots (src/gpu/core/Utilities/KernelUtilities.h:252-281, 350-379). Because every
// address is read and written by one thread only, the update needs no second buffer
// and no synchronisation; the parity flag swaps the slot table between steps.
//
// Two addressings:
//  * box:     dense periodic nx*ny*nz block, neighbours recomputed from (x,y,z);
//             a workgroup covers whole x-rows, so every cache line it touches in a
//             direction array belongs to rows it owns (no line is shared between
//             workgroups, hence none between the XCDs' L2 caches).
//  * indexed: the reference's sparse layout (fluid-node list + neighbour arrays,
//             src/gpu/core/Utilities/RunCollision.cuh:74-143).
#include <atomic>
#include <cmath>
#include <cstdio>

#include "vf_k17.h"
#include "vf_lattice.h"
#include "vf_rowio.h"

namespace vfa {

// ---- device: gather / scatter over the 8 cell nodes -------------------------------------
template <class S, bool EVEN>
__device__ __forceinline__ void gather(const S* __restrict__ f, uint64_t N, const uint32_t (&n)[8], float (&v)[27])
{
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const char* base = reinterpret_cast<const char*>(f) + (uint64_t)slot<EVEN>(d) * N * sizeof(S);
        v[d] = Codec<S>::load(reinterpret_cast<const S*>(base + n[node_sel(d)] * (uint32_t)sizeof(S)));
    }
}

template <class S, bool EVEN>
__device__ __forceinline__ void scatter(S* __restrict__ f, uint64_t N, const uint32_t (&n)[8], const float (&v)[27])
{
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        char* base = reinterpret_cast<char*>(f) + (uint64_t)slot<EVEN>(d) * N * sizeof(S);
        Codec<S>::store(reinterpret_cast<S*>(base + n[node_sel(d)] * (uint32_t)sizeof(S)), v[inv(d)]);
    }
}

struct BoxArgs {
    uint64_t N;
    int nx, ny, nz;
    const uint8_t* mask;
    float* rho;
    float* vx;
    float* vy;
    float* vz;
    const float* porosity; // Brinkman porosity per node, or null
    PlaneDiv plane;        // plane-flat numbering of the one-node-per-thread kernels
};

// Brinkman drag a = -porosity u through the half-step velocity shift: u (1 - porosity/2)
__device__ __forceinline__ float brinkman_scale(const float* porosity, uint32_t k)
{
    return porosity ? 1.0f - 0.5f * porosity[k] : 1.0f;
}

constexpr int kBoxBlock = kPlaneBlock; // one-node-per-thread kernels: plane-flat workgroups

// Bounce-back lists applied by the collision itself (vf_post_bounce_back): for a node with an
// entry, list l bounces back the links of entryLinks[3 e + l] (velocity lists subtract
// 6 w_d c_d.U of the node's wall velocity).
struct FusedBB {
    const int32_t* nodeEntry;
    const int32_t* entryIndex;
    const uint32_t* entryLinks;
    const float* vx[3];
    const float* vy[3];
    const float* vz[3];
};

// The post-collision bounce-back of one node, from its post-collision values v (as the
// collision left them in registers): exactly what NoSlipBounceBack / VelocityBounceBack compute
// from the stored values, list by list, writing the incoming population inv(d) of every wall
// link d into the cell the next step reads (slot d with this step's parity, at the node of
// inv(d)'s cell).
template <class S, bool EVEN>
__device__ __forceinline__ void bounce_back_epilogue(S* __restrict__ f, uint64_t N, const uint32_t (&n)[8],
                                                     const float (&v)[27], const FusedBB& bb)
{
    const int32_t e = bb.nodeEntry[n[0]];
    if (e < 0) return;
    float post[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) post[d] = Codec<S>::round(v[d]); // the stored post-collision values
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const uint32_t links = bb.entryLinks[3 * e + l];
        if (!links) continue;
        float Ux = 0.0f, Uy = 0.0f, Uz = 0.0f;
        const bool velocity = bb.vx[l] != nullptr;
        if (velocity) {
            const int32_t i = bb.entryIndex[3 * e + l];
            Ux = bb.vx[l][i];
            Uy = bb.vy[l][i];
            Uz = bb.vz[l][i];
        }
#pragma unroll
        for (int d = 1; d < 27; ++d) {
            if (!(links >> d & 1u)) continue;
            const float val = velocity ? velocity_bounce_back(post[d], d, Ux, Uy, Uz) : post[d];
            Codec<S>::store(f + (uint64_t)slot<EVEN>(d) * N + n[node_sel(inv(d))], val);
        }
    }
}

// The node of this thread in the plane-flat numbering; XCD: the (plane_blocks, nz) workgroups
// taken in XCD block ranges, so that consecutive spans of a plane, which share the cache lines at
// their ends, meet in one L2
template <bool XCD> __device__ __forceinline__ bool plane_node(const PlaneDiv& p, int& x, int& y, int& z)
{
    uint32_t bx = blockIdx.x, bz = blockIdx.y;
    if constexpr (XCD) {
        const uint32_t b = xcd_block(bx + gridDim.x * bz, gridDim.x * gridDim.y);
        bx = b % gridDim.x;
        bz = b / gridDim.x;
    }
    const uint32_t t = bx * kBoxBlock + threadIdx.x;
    if (t >= p.planeNodes) return false;
    const uint32_t row = plane_row(t, p);
    x = (int)(t - row * p.nx);
    y = (int)row;
    z = (int)bz;
    return true;
}

template <class S, bool EVEN, bool MASK, bool MACRO, bool BB = false, bool XCD = false>
__global__ __launch_bounds__(kBoxBlock) void k17_box_kernel(S* __restrict__ f, BoxArgs a, K17Rates r, FusedBB bb)
{
    int x, y, z;
    if (!plane_node<XCD>(a.plane, x, y, z)) return;
    const int xp = (x + 1 == a.nx) ? 0 : x + 1;
    const int yp = (y + 1 == a.ny) ? 0 : y + 1;
    const int zp = (z + 1 == a.nz) ? 0 : z + 1;
    const uint32_t row00 = (uint32_t)a.nx * ((uint32_t)y + (uint32_t)a.ny * (uint32_t)z);
    const uint32_t row10 = (uint32_t)a.nx * ((uint32_t)yp + (uint32_t)a.ny * (uint32_t)z);
    const uint32_t row01 = (uint32_t)a.nx * ((uint32_t)y + (uint32_t)a.ny * (uint32_t)zp);
    const uint32_t row11 = (uint32_t)a.nx * ((uint32_t)yp + (uint32_t)a.ny * (uint32_t)zp);
    const uint32_t n[8] = {row00 + x, row00 + xp, row10 + x, row10 + xp, row01 + x, row01 + xp, row11 + x, row11 + xp};
    if (MASK && a.mask[n[0]] == 0) return;

    float v[27];
    gather<S, EVEN>(f, a.N, n, v);
    MacroOut mo;
    k17_collide(v, r, mo, brinkman_scale(a.porosity, n[0]));
    if (MACRO) {
        a.rho[n[0]] = mo.drho;
        a.vx[n[0]] = mo.vx;
        a.vy[n[0]] = mo.vy;
        a.vz[n[0]] = mo.vz;
    }
    scatter<S, EVEN>(f, a.N, n, v);
    if (BB) bounce_back_epilogue<S, EVEN>(f, a.N, n, v, bb);
}

// Out-of-place variant: reads `in`, writes the stepped state to `out` (same layout). Masked
// nodes copy their 27 cells unchanged, so `out` is the complete state the in-place step
// would leave. Used to replay a segment of steps while keeping every intermediate state.
template <class S, bool EVEN, bool MASK, bool XCD = false>
__global__ __launch_bounds__(kBoxBlock) void k17_box_into_kernel(const S* __restrict__ in, S* __restrict__ out,
                                                                 BoxArgs a, K17Rates r)
{
    int x, y, z;
    if (!plane_node<XCD>(a.plane, x, y, z)) return;
    const int xp = (x + 1 == a.nx) ? 0 : x + 1;
    const int yp = (y + 1 == a.ny) ? 0 : y + 1;
    const int zp = (z + 1 == a.nz) ? 0 : z + 1;
    const uint32_t row00 = (uint32_t)a.nx * ((uint32_t)y + (uint32_t)a.ny * (uint32_t)z);
    const uint32_t row10 = (uint32_t)a.nx * ((uint32_t)yp + (uint32_t)a.ny * (uint32_t)z);
    const uint32_t row01 = (uint32_t)a.nx * ((uint32_t)y + (uint32_t)a.ny * (uint32_t)zp);
    const uint32_t row11 = (uint32_t)a.nx * ((uint32_t)yp + (uint32_t)a.ny * (uint32_t)zp);
    const uint32_t n[8] = {row00 + x, row00 + xp, row10 + x, row10 + xp, row01 + x, row01 + xp, row11 + x, row11 + xp};
    float v[27];
    gather<S, EVEN>(in, a.N, n, v);
    if (MASK && a.mask[n[0]] == 0) {
        // unchanged cells: value of pre(d) back into its own cell (scatter writes v[inv(d)])
        float w[27];
#pragma unroll
        for (int d = 0; d < 27; ++d) w[inv(d)] = v[d];
        scatter<S, EVEN>(out, a.N, n, w);
        return;
    }
    MacroOut mo;
    k17_collide(v, r, mo, brinkman_scale(a.porosity, n[0]));
    scatter<S, EVEN>(out, a.N, n, v);
}

// ---- row kernel: one workgroup per x-row, V consecutive x-nodes per lane ----------------
// Every global access of a wave is an aligned, contiguous, non-temporal vector access.
// The nine directions with c_x = -1 live one node further in +x (node_sel bit 0); a lane
// loads its own aligned elements and takes the element after its last node from lane+1
// (__shfl), and stores the element before its first node from lane-1. Across the waves of
// the row, and around the periodic wrap, the boundary lanes trade those values through LDS.
// A row need not fill its last wave (the lanes beyond it idle), and a Brinkman porosity field
// scales each node's collision velocity.
constexpr int kRowMaxThreads = 256;

// BUF: the direction arrays are addressed through two buffer descriptors (kRowSplit), which
// needs kRowSplit slots to span less than 4 GiB; otherwise each direction takes a 64-bit base
// and a global access.

// INTO: the stepped state goes to `fout` (same addresses, another buffer) and `f` is only read;
// solid nodes then copy their cells, so `fout` is the complete state the in-place step leaves.
// XCD: the rows are taken in the XCD block ranges of xcd_block (rows that do not fill whole
// cache lines)
template <class S, int V, bool EVEN, bool MASK, bool MACRO, bool BUF, bool INTO = false, bool XCD = false>
__global__ __launch_bounds__(kRowMaxThreads) void k17_row_kernel(S* f, S* fout, BoxArgs a, K17Rates r)
{
    using IO = RowIO<S, V>;
    __shared__ unsigned xin[kRowMaxThreads / 64][9];
    __shared__ float xout[kRowMaxThreads / 64][9];
    __shared__ float xlast[9];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // a row is nx / V lanes; the lanes of the last wave beyond it idle (loads, stores and the
    // collision are skipped; they still reach every barrier)
    const uint32_t lanes = (uint32_t)a.nx / V;
    const bool active = threadIdx.x < lanes;
    const bool lastLane = threadIdx.x + 1 == lanes;
    uint32_t y = blockIdx.x, z = blockIdx.y;
    if constexpr (XCD) {
        const uint32_t row = xcd_block(y + gridDim.x * z, gridDim.x * gridDim.y);
        y = row % (uint32_t)a.ny;
        z = row / (uint32_t)a.ny;
    }
    const uint32_t yp = (y + 1 == (uint32_t)a.ny) ? 0 : y + 1;
    const uint32_t zp = (z + 1 == (uint32_t)a.nz) ? 0 : z + 1;
    const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny;
    const uint32_t x0 = threadIdx.x * V;
    // node index of this lane's first node in the 4 rows its cells touch, indexed by the
    // y/z bits of node_sel; each direction array has its own (uniform) 64-bit base and the
    // per-lane byte offsets stay 32-bit (N * sizeof(S) < 4 GiB)
    const uint32_t rows[4] = {nx * (y + ny * z) + x0, nx * (yp + ny * z) + x0, nx * (y + ny * zp) + x0,
                              nx * (yp + ny * zp) + x0};
    const uint32_t boff[4] = {rows[0] * (uint32_t)sizeof(S), rows[1] * (uint32_t)sizeof(S),
                              rows[2] * (uint32_t)sizeof(S), rows[3] * (uint32_t)sizeof(S)};
    const uint64_t slotBytes = a.N * sizeof(S);
    S* const dst = INTO ? fout : f;
    __amdgpu_buffer_rsrc_t rsLo, rsHi, wsLo, wsHi;
    if constexpr (BUF) {
        rsLo = __builtin_amdgcn_make_buffer_rsrc(f, (short)0, (int)0xffffffff, 0x00020000);
        rsHi = __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<char*>(f) + kRowSplit * slotBytes, (short)0,
                                                 (int)0xffffffff, 0x00020000);
        if constexpr (INTO) {
            wsLo = __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, (int)0xffffffff, 0x00020000);
            wsHi = __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<char*>(dst) + kRowSplit * slotBytes, (short)0,
                                                     (int)0xffffffff, 0x00020000);
        } else {
            wsLo = rsLo;
            wsHi = rsHi;
        }
    }
    // byte offset of direction d's slot from its descriptor's base
    auto soff = [&](int d) -> uint32_t {
        const int sl = slot<EVEN>(d);
        return (uint32_t)(sl >= kRowSplit ? sl - kRowSplit : sl) * (uint32_t)slotBytes;
    };
    auto load = [&](int d) {
        if constexpr (BUF) return IO::load(slot<EVEN>(d) >= kRowSplit ? rsHi : rsLo, boff[node_sel(d) >> 1], soff(d));
        else {
            const char* base = reinterpret_cast<const char*>(f) + (uint64_t)slot<EVEN>(d) * slotBytes;
            return IO::load(reinterpret_cast<const S*>(base + boff[node_sel(d) >> 1]));
        }
    };
    auto store = [&](int d, const float (&vals)[V]) {
        if constexpr (BUF) IO::store(slot<EVEN>(d) >= kRowSplit ? wsHi : wsLo, boff[node_sel(d) >> 1], soff(d), vals);
        else {
            char* base = reinterpret_cast<char*>(dst) + (uint64_t)slot<EVEN>(d) * slotBytes;
            IO::store(reinterpret_cast<S*>(base + boff[node_sel(d) >> 1]), vals);
        }
    };

    typename IO::Raw raw[27];
    if (active) {
#pragma unroll
        for (int d = 0; d < 27; ++d) raw[d] = load(d);
    } else {
#pragma unroll
        for (int d = 0; d < 27; ++d) raw[d] = typename IO::Raw{};
    }

    // word holding the element after this lane's last node, for the c_x = -1 directions
    unsigned nxt[9];
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < 9; ++j) xin[wave][j] = IO::word0(raw[kShiftedDirs[j]]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 9; ++j) nxt[j] = __shfl(IO::word0(raw[kShiftedDirs[j]]), (lane + 1) & 63, 64);
    if (active && (lane == 63 || lastLane)) {
        // the next wave's first element, or the row's first element around the periodic wrap
        const int src = lastLane ? 0 : wave + 1;
#pragma unroll
        for (int j = 0; j < 9; ++j) nxt[j] = xin[src][j];
    }

    float out[27][V];
    float carry[9];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        float fl[27];
#pragma unroll
        for (int d = 0; d < 27; ++d) {
            if (node_sel(d) & 1) fl[d] = (i + 1 < V) ? IO::elem(raw[d], i + 1) : IO::of_word(nxt[shifted_index(d)], 0);
            else fl[d] = IO::elem(raw[d], i);
        }
        bool fluid = active;
        if (MASK && active) fluid = a.mask[rows[0] + i] != 0;
        if (fluid) {
            MacroOut mo;
            // Brinkman porosity: the velocity scale 1 - porosity / 2 (1 without a field)
            const float vs = a.porosity ? 1.0f - 0.5f * a.porosity[rows[0] + i] : 1.0f;
            k17_collide(fl, r, mo, vs);
            if (MACRO) {
                a.rho[rows[0] + i] = mo.drho;
                a.vx[rows[0] + i] = mo.vx;
                a.vy[rows[0] + i] = mo.vy;
                a.vz[rows[0] + i] = mo.vz;
            }
        } else {
            // solid node: write back exactly what was read (post[inv d] := pre[d])
#pragma unroll
            for (int d = 1; d < 27; ++d) {
                if (d < inv(d)) {
                    const float t = fl[d];
                    fl[d] = fl[inv(d)];
                    fl[inv(d)] = t;
                }
            }
        }
#pragma unroll
        for (int d = 0; d < 27; ++d) {
            const float val = fl[inv(d)];
            if (node_sel(d) & 1) {
                if (i + 1 < V) out[d][i + 1] = val;
                else carry[shifted_index(d)] = val;
            } else {
                out[d][i] = val;
            }
        }
    }

    // directions without an x-shift are complete: store them before the exchange
    if (active) {
#pragma unroll
        for (int d = 0; d < 27; ++d) {
            if (node_sel(d) & 1) continue;
            store(d, out[d]);
        }
    }

    // element before this lane's first node, for the c_x = -1 directions: from the previous
    // lane, the previous wave's last lane, or (first lane) the row's last lane
    if (active && lane == 63) {
#pragma unroll
        for (int j = 0; j < 9; ++j) xout[wave][j] = carry[j];
    }
    if (lastLane) {
#pragma unroll
        for (int j = 0; j < 9; ++j) xlast[j] = carry[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        float s0 = __shfl(carry[j], (lane + 63) & 63, 64);
        if (lane == 0) s0 = (wave == 0) ? xlast[j] : xout[wave - 1][j];
        const int d = kShiftedDirs[j];
        out[d][0] = s0;
        if (active) store(d, out[d]);
    }
}

struct IndexedArgs {
    uint64_t N;
    const uint32_t* nbx;
    const uint32_t* nby;
    const uint32_t* nbz;
    const uint32_t* indices;
    uint32_t count;
    float* rho;
    float* vx;
    float* vy;
    float* vz;
    float* bfx;
    float* bfy;
    float* bfz;
    float forces[3];  // global force, unscaled (Parameter::getForcesDev)
    float forceFactor; // 2^-level
    const float* porosity;
};

constexpr int kIndexedBlock = 256;

template <class S, bool EVEN, bool MACRO, bool BODYFORCE>
__global__ __launch_bounds__(kIndexedBlock) void k17_indexed_kernel(S* __restrict__ f, IndexedArgs a, K17Rates r)
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
    K17Rates rr = r;
    if (BODYFORCE) {
        // RunCollision.cuh:87-94: (forces + bodyForce[k]) * 1/2 * forceFactor in that order,
        // then the node's body force is reset
        rr.fx = (a.forces[0] + a.bfx[k]) * 0.5f * a.forceFactor;
        rr.fy = (a.forces[1] + a.bfy[k]) * 0.5f * a.forceFactor;
        rr.fz = (a.forces[2] + a.bfz[k]) * 0.5f * a.forceFactor;
        a.bfx[k] = 0.0f;
        a.bfy[k] = 0.0f;
        a.bfz[k] = 0.0f;
    }
    float v[27];
    gather<S, EVEN>(f, a.N, n, v);
    MacroOut mo;
    k17_collide(v, rr, mo, brinkman_scale(a.porosity, k));
    if (MACRO) {
        a.rho[k] = mo.drho;
        a.vx[k] = mo.vx;
        a.vy[k] = mo.vy;
        a.vz[k] = mo.vz;
    }
    scatter<S, EVEN>(f, a.N, n, v);
}

// ---- host launchers ----------------------------------------------------------------------
// Width (nodes per lane) of the row kernel per storage type; 0 selects the scalar kernel.
// Measured on MI355X at 256^3: FP32 1 node/lane (5.9 TB/s; 2 nodes/lane 5.8, scalar 4.9);
// FP16 and BF16 4 nodes/lane where rows are whole waves of it (1-2 % above 2 nodes/lane; both
// near the VALU issue limit), else 2. Rows longer than one workgroup's lanes take the next width
// that fits; rows no width fits take the scalar kernel. kernelChoice.rowWidth pins a width.
template <class S> static int row_width(int nx, const vf_kernel_choice& kc)
{
    if (kc.rowWidth != 0) return kc.rowWidth;
    auto fits = [nx](int v) { return nx % v == 0 && nx / v <= kRowMaxThreads; };
    auto waves = [nx](int v) { return (nx / v + 63) / 64; };
    if (Codec<S>::storage == VF_STORAGE_FP32) {
        // two nodes per lane where one node per lane leaves a last wave of a few lanes (a row
        // of 130 nodes: one full wave and one of 33 lanes instead of two full waves and one of 2)
        const int rem = (nx % 128) % 64;
        if (fits(1) && fits(2) && rem != 0 && rem < 32 && waves(2) < waves(1)) return 2;
        return fits(1) ? 1 : fits(2) ? 2 : fits(4) ? 4 : 0;
    }
    if (nx % 256 == 0 && fits(4)) return 4;
    return fits(2) ? 2 : fits(4) ? 4 : 0;
}

// Resident workgroups per CU for the row kernel (0: as many as registers allow). The FP32
// kernel moves the most bytes per instruction and runs best with few resident rows when it
// streams from HBM: at 256^3 on MI355X, 8 resident 4-wave workgroups per CU give 5.72 TB/s, 3-4
// give 5.84-5.86 and 2 give 5.92-5.93 (fewer concurrent direction streams per memory channel);
// one gives 3.8. A working set that fits the 256 MB infinity cache is latency-bound instead
// and takes every wave. The 16-bit kernels are bound by VALU issue and need every wave they can
// get. The cap is imposed by reserving LDS at launch; kernelChoice.rowWorkgroupsPerCU pins it.
constexpr uint64_t kCacheResidentBytes = 256ull << 20;
template <class S, int V, bool BUF> static int row_wgs_per_cu(uint64_t workingSet, const vf_kernel_choice& kc)
{
    if (kc.rowWorkgroupsPerCU != 0) return kc.rowWorkgroupsPerCU < 0 ? 0 : kc.rowWorkgroupsPerCU;
    const bool streams = workingSet > kCacheResidentBytes;
    return (Codec<S>::storage == VF_STORAGE_FP32 && V == 1 && BUF && streams) ? 2 : 0;
}

// out: null for the in-place step, else the buffer the out-of-place step writes
template <class S, int V, bool EVEN, bool BUF>
static void launch_row_b(const BoxArgs& a, const K17Rates& r, bool mask, bool macro, S* f, S* out, uint64_t ws,
                         const vf_kernel_choice& kc, hipStream_t s)
{
    dim3 grid(a.ny, a.nz);
    dim3 block(((a.nx / V + 63) / 64) * 64); // whole waves; lanes beyond the row idle
    const size_t lds = lds_reserve_for(row_wgs_per_cu<S, V, BUF>(ws, kc));
    if (out) {
        if (mask) hipLaunchKernelGGL((k17_row_kernel<S, V, EVEN, true, false, BUF, true>), grid, block, lds, s, f, out, a, r);
        else hipLaunchKernelGGL((k17_row_kernel<S, V, EVEN, false, false, BUF, true>), grid, block, lds, s, f, out, a, r);
        return;
    }
    if (mask && macro) hipLaunchKernelGGL((k17_row_kernel<S, V, EVEN, true, true, BUF>), grid, block, lds, s, f, f, a, r);
    else if (macro) hipLaunchKernelGGL((k17_row_kernel<S, V, EVEN, false, true, BUF>), grid, block, lds, s, f, f, a, r);
    else if (BUF && xcd_rows(a.nx, sizeof(S), kc.rowXcdRanges)) {
        if (mask) hipLaunchKernelGGL((k17_row_kernel<S, V, EVEN, true, false, BUF, false, true>), grid, block, lds, s, f, f, a, r);
        else hipLaunchKernelGGL((k17_row_kernel<S, V, EVEN, false, false, BUF, false, true>), grid, block, lds, s, f, f, a, r);
    }
    else if (mask) hipLaunchKernelGGL((k17_row_kernel<S, V, EVEN, true, false, BUF>), grid, block, lds, s, f, f, a, r);
    else hipLaunchKernelGGL((k17_row_kernel<S, V, EVEN, false, false, BUF>), grid, block, lds, s, f, f, a, r);
}

template <class S, int V, bool EVEN>
static void launch_row_t(const BoxArgs& a, const K17Rates& r, bool mask, bool macro, S* f, S* out, uint64_t ws,
                         const vf_kernel_choice& kc, hipStream_t s)
{
    // the larger descriptor spans kRowSplit slots; buffer offsets are 32-bit
    if ((uint64_t)kRowSplit * a.N * sizeof(S) <= 0xffffffffull && kc.rowAddressing != VF_ROW_ADDRESSING_GLOBAL)
        launch_row_b<S, V, EVEN, true>(a, r, mask, macro, f, out, ws, kc, s);
    else
        launch_row_b<S, V, EVEN, false>(a, r, mask, macro, f, out, ws, kc, s);
}

// Nodes per lane of the row kernel the dense-box collision takes for this box, or 0 when it
// takes the one-node-per-thread kernel: a row of nx nodes is nx/V lanes, at most one workgroup.
// A working set that stays in the 256 MB infinity cache is latency-bound: the node kernel (no
// shuffles or barriers, every lane busy, more resident waves) serves it best. Beyond the cache
// the row kernel leads where its rows are whole waves (256^3 FP32 620 against 682-736 us per
// step); rows of other lengths leave lanes idle and start inside cache lines, and there the
// node kernel, whose workgroups take consecutive nodes of a plane, leads (258 x 130 x 130 FP16
// 87 against 153 us, FP32 203 against 264 us; measured on MI355X, one lattice stepped alone).
// A pinned row kernel or row width keeps the row kernel.
template <class S> static int box_row_width(const vf_lbm_parameter& p)
{
    const vf_kernel_choice& kc = p.kernelChoice;
    if (kc.boxKernel == VF_BOX_KERNEL_NODE) return 0;
    const int V = row_width<S>(p.nx, kc);
    if (V <= 0 || p.nx % V != 0 || p.nx / V > kRowMaxThreads) return 0;
    if (V == 1 && sizeof(S) != 4) return 0;
    if (V != 1 && V != 2 && V != 4) return 0;
    const uint64_t ws = p.workingSetBytes ? p.workingSetBytes : 27 * p.numberOfNodes * sizeof(S);
    const bool pinned = kc.boxKernel == VF_BOX_KERNEL_ROW || kc.rowWidth != 0;
    if (pinned) return V;
    return (ws <= kCacheResidentBytes || (p.nx / V) % 64 != 0) ? 0 : V;
}

// The row kernel when box_row_width selects it. Returns false when it does not apply.
template <class S, bool EVEN>
static bool try_launch_row(const vf_lbm_parameter& p, const BoxArgs& a, const K17Rates& r, bool macro, S* out,
                           hipStream_t s)
{
    const int V = box_row_width<S>(p);
    if (V == 0) return false;
    const vf_kernel_choice& kc = p.kernelChoice;
    S* f = static_cast<S*>(p.distributions);
    const bool mask = p.fluidMask != nullptr;
    const uint64_t ws = p.workingSetBytes ? p.workingSetBytes : 27 * a.N * sizeof(S);
    if constexpr (sizeof(S) == 4) {
        if (V == 1) { launch_row_t<S, 1, EVEN>(a, r, mask, macro, f, out, ws, kc, s); return true; }
    }
    if (V == 2) { launch_row_t<S, 2, EVEN>(a, r, mask, macro, f, out, ws, kc, s); return true; }
    launch_row_t<S, 4, EVEN>(a, r, mask, macro, f, out, ws, kc, s);
    return true;
}

template <class S, bool EVEN>
static void launch_box_t(const vf_lbm_parameter& p, bool macro, hipStream_t s)
{
    BoxArgs a{p.numberOfNodes, p.nx, p.ny, p.nz, p.fluidMask, p.rho, p.velocityX, p.velocityY, p.velocityZ,
              p.porosity, make_plane_div(p.nx, p.ny)};
    const K17Rates r = make_k17_rates(p);
    S* f = static_cast<S*>(p.distributions);
    const bool mask = p.fluidMask != nullptr;
    if (try_launch_row<S, EVEN>(p, a, r, macro, nullptr, s)) return;
    const dim3 grid(plane_blocks(p.nx, p.ny), p.nz), bx(kBoxBlock);
    const FusedBB none{};
    // XCD block ranges where the working set streams and planes share cache lines at their ends
    // (or where pinned): inside the two-level step they take the 258 x 130 x 130 FP16 coarse
    // level's fetched bytes from 1.21 to 0.99 x algorithmic at the same time (107 against
    // 104-109 us); on a cache-resident lattice (the sphere channel) they cost 59 -> 61 us
    const uint64_t ws = p.workingSetBytes ? p.workingSetBytes : 27 * p.numberOfNodes * sizeof(S);
    const bool xcd = p.kernelChoice.rowXcdRanges == 1 ||
                     (p.kernelChoice.rowXcdRanges == 0 && ws > kCacheResidentBytes &&
                      xcd_planes(p.nx, p.ny, sizeof(S), 0));
    if (mask && macro) hipLaunchKernelGGL((k17_box_kernel<S, EVEN, true, true>), grid, dim3(bx), 0, s, f, a, r, none);
    else if (mask && xcd) hipLaunchKernelGGL((k17_box_kernel<S, EVEN, true, false, false, true>), grid, dim3(bx), 0, s, f, a, r, none);
    else if (mask) hipLaunchKernelGGL((k17_box_kernel<S, EVEN, true, false>), grid, dim3(bx), 0, s, f, a, r, none);
    else if (macro) hipLaunchKernelGGL((k17_box_kernel<S, EVEN, false, true>), grid, dim3(bx), 0, s, f, a, r, none);
    else if (xcd) hipLaunchKernelGGL((k17_box_kernel<S, EVEN, false, false, false, true>), grid, dim3(bx), 0, s, f, a, r, none);
    else hipLaunchKernelGGL((k17_box_kernel<S, EVEN, false, false>), grid, dim3(bx), 0, s, f, a, r, none);
}

// The node kernel with the bounce-back epilogue. Returns false where the dense-box collision
// takes the row kernel (the caller then launches the collision and the bounce-back kernels).
template <class S, bool EVEN>
static bool launch_box_bb_t(const vf_lbm_parameter& p, bool macro, const FusedBB& bb, hipStream_t s)
{
    if (box_row_width<S>(p) != 0) return false;
    BoxArgs a{p.numberOfNodes, p.nx, p.ny, p.nz, p.fluidMask, p.rho, p.velocityX, p.velocityY, p.velocityZ,
              p.porosity, make_plane_div(p.nx, p.ny)};
    const K17Rates r = make_k17_rates(p);
    S* f = static_cast<S*>(p.distributions);
    const bool mask = p.fluidMask != nullptr;
    const dim3 grid(plane_blocks(p.nx, p.ny), p.nz), bx(kBoxBlock);
    if (mask && macro) hipLaunchKernelGGL((k17_box_kernel<S, EVEN, true, true, true>), grid, dim3(bx), 0, s, f, a, r, bb);
    else if (mask) hipLaunchKernelGGL((k17_box_kernel<S, EVEN, true, false, true>), grid, dim3(bx), 0, s, f, a, r, bb);
    else if (macro) hipLaunchKernelGGL((k17_box_kernel<S, EVEN, false, true, true>), grid, dim3(bx), 0, s, f, a, r, bb);
    else hipLaunchKernelGGL((k17_box_kernel<S, EVEN, false, false, true>), grid, dim3(bx), 0, s, f, a, r, bb);
    return true;
}

template <class S, bool EVEN>
static void launch_indexed_t(const vf_lbm_parameter& p, const uint32_t* idx, uint32_t n, bool macro, bool force,
                             hipStream_t s)
{
    IndexedArgs a{p.numberOfNodes, p.neighborX, p.neighborY, p.neighborZ, idx, n,
                  p.rho, p.velocityX, p.velocityY, p.velocityZ, p.bodyForceX, p.bodyForceY, p.bodyForceZ,
                  {p.forces[0], p.forces[1], p.forces[2]}, p.forceFactor, p.porosity};
    const K17Rates r = make_k17_rates(p);
    dim3 grid((n + kIndexedBlock - 1) / kIndexedBlock);
    S* f = static_cast<S*>(p.distributions);
    if (macro && force) hipLaunchKernelGGL((k17_indexed_kernel<S, EVEN, true, true>), grid, dim3(kIndexedBlock), 0, s, f, a, r);
    else if (macro) hipLaunchKernelGGL((k17_indexed_kernel<S, EVEN, true, false>), grid, dim3(kIndexedBlock), 0, s, f, a, r);
    else if (force) hipLaunchKernelGGL((k17_indexed_kernel<S, EVEN, false, true>), grid, dim3(kIndexedBlock), 0, s, f, a, r);
    else hipLaunchKernelGGL((k17_indexed_kernel<S, EVEN, false, false>), grid, dim3(kIndexedBlock), 0, s, f, a, r);
}

template <template <class, bool> class L, class... A>
static int dispatch_storage(const vf_lbm_parameter& p, A&&... args)
{
    const bool even = p.isEvenTimestep != 0;
    switch (p.storage) {
    case VF_STORAGE_FP32:
        even ? L<float, true>::run(p, args...) : L<float, false>::run(p, args...);
        return VF_OK;
    case VF_STORAGE_FP16:
        even ? L<__half, true>::run(p, args...) : L<__half, false>::run(p, args...);
        return VF_OK;
    case VF_STORAGE_BF16:
        even ? L<__hip_bfloat16, true>::run(p, args...) : L<__hip_bfloat16, false>::run(p, args...);
        return VF_OK;
    default:
        set_last_error("unknown storage type");
        return VF_ERR_INVALID_ARGUMENT;
    }
}

template <class S, bool EVEN> struct BoxL {
    static void run(const vf_lbm_parameter& p, bool macro, hipStream_t s) { launch_box_t<S, EVEN>(p, macro, s); }
};
template <class S, bool EVEN> struct BoxIntoL {
    static void run(const vf_lbm_parameter& p, void* out, hipStream_t s)
    {
        BoxArgs a{p.numberOfNodes, p.nx, p.ny, p.nz, p.fluidMask, nullptr, nullptr, nullptr, nullptr, p.porosity,
                  make_plane_div(p.nx, p.ny)};
        const K17Rates r = make_k17_rates(p);
        if (try_launch_row<S, EVEN>(p, a, r, false, static_cast<S*>(out), s)) return;
        const dim3 grid(plane_blocks(p.nx, p.ny), p.nz), bx(kBoxBlock);
        const S* in = static_cast<const S*>(p.distributions);
        const bool xcd = xcd_planes(p.nx, p.ny, sizeof(S), p.kernelChoice.rowXcdRanges);
        if (p.fluidMask && xcd)
            hipLaunchKernelGGL((k17_box_into_kernel<S, EVEN, true, true>), grid, dim3(bx), 0, s, in, (S*)out, a, r);
        else if (p.fluidMask)
            hipLaunchKernelGGL((k17_box_into_kernel<S, EVEN, true>), grid, dim3(bx), 0, s, in, (S*)out, a, r);
        else if (xcd)
            hipLaunchKernelGGL((k17_box_into_kernel<S, EVEN, false, true>), grid, dim3(bx), 0, s, in, (S*)out, a, r);
        else
            hipLaunchKernelGGL((k17_box_into_kernel<S, EVEN, false>), grid, dim3(bx), 0, s, in, (S*)out, a, r);
    }
};
template <class S, bool EVEN> struct IndexedL {
    static void run(const vf_lbm_parameter& p, const uint32_t* idx, uint32_t n, bool macro, bool force, hipStream_t s)
    {
        launch_indexed_t<S, EVEN>(p, idx, n, macro, force, s);
    }
};

static int check_common(const vf_lbm_parameter* p)
{
    if (!p || !p->distributions) {
        set_last_error("null parameter or distribution buffer");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->numberOfNodes == 0 || p->numberOfNodes >= (1ull << 30)) {
        set_last_error("numberOfNodes must be in [1, 2^30)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!(p->omega > 0.0f && p->omega < 2.0f)) {
        set_last_error("omega must lie in (0, 2)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return VF_OK;
}

} // namespace vfa

using namespace vfa;

extern "C" int vf_k17_run_on_indices(const vf_lbm_parameter* p, const uint32_t* indices, uint32_t n,
                                     int32_t collisionTemplate, void* stream)
{
    int rc = check_common(p);
    if (rc) return rc;
    if (n == 0) return VF_OK;
    if (!indices || !p->neighborX || !p->neighborY || !p->neighborZ) {
        set_last_error("runOnIndices needs indices and neighborX/Y/Z");
        return VF_ERR_INVALID_ARGUMENT;
    }
    bool macro = false, force = false;
    switch (collisionTemplate) {
    case VF_COLLISION_DEFAULT: break;
    case VF_COLLISION_WRITE_MACRO_VARS: macro = true; break;
    case VF_COLLISION_APPLY_BODY_FORCE: force = true; break;
    case VF_COLLISION_ALL_FEATURES:
    case VF_COLLISION_SUBDOMAIN_BORDER: macro = force = true; break;
    default:
        // K17CompressibleNavierStokes.cu:139-141 throws on an unknown template
        set_last_error("Invalid CollisionTemplate in CumulantK17::runOnIndices()");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (macro && (!p->rho || !p->velocityX || !p->velocityY || !p->velocityZ)) {
        set_last_error("WriteMacroVars needs rho and velocityX/Y/Z");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (force && (!p->bodyForceX || !p->bodyForceY || !p->bodyForceZ)) {
        set_last_error("ApplyBodyForce needs bodyForceX/Y/Z");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = dispatch_storage<IndexedL>(*p, indices, n, macro, force, (hipStream_t)stream);
    if (rc) return rc;
    return check_launch("K17CompressibleNavierStokes (indexed)");
}

extern "C" int vf_k17_run(const vf_lbm_parameter* p, void* stream)
{
    if (!p) {
        set_last_error("null parameter");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return vf_k17_run_on_indices(p, p->fluidNodeIndices, p->numberOfFluidNodes, VF_COLLISION_DEFAULT, stream);
}

extern "C" int vf_k17_run_box_into(const vf_lbm_parameter* p, void* out, void* stream)
{
    int rc = check_common(p);
    if (rc) return rc;
    if (!out || out == p->distributions) {
        set_last_error("out-of-place step needs a separate output buffer");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0 ||
        (uint64_t)p->nx * (uint64_t)p->ny * (uint64_t)p->nz != p->numberOfNodes || p->ny > 65535 || p->nz > 65535) {
        set_last_error("box path needs nx*ny*nz == numberOfNodes (ny, nz <= 65535)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    rc = dispatch_storage<BoxIntoL>(*p, out, (hipStream_t)stream);
    if (rc) return rc;
    return check_launch("K17CompressibleNavierStokes (box, out of place)");
}

// argument checks of the dense-box collision; macro: the template writes the macroscopic values
static int check_box(const vf_lbm_parameter* p, int32_t collisionTemplate, bool& macro)
{
    int rc = check_common(p);
    if (rc) return rc;
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0 ||
        (uint64_t)p->nx * (uint64_t)p->ny * (uint64_t)p->nz != p->numberOfNodes || p->ny > 65535 || p->nz > 65535) {
        set_last_error("box path needs nx*ny*nz == numberOfNodes (ny, nz <= 65535)");
        return VF_ERR_INVALID_ARGUMENT;
    }
    macro = false;
    if (collisionTemplate == VF_COLLISION_WRITE_MACRO_VARS) {
        macro = true;
        if (!p->rho || !p->velocityX || !p->velocityY || !p->velocityZ) {
            set_last_error("WriteMacroVars needs rho and velocityX/Y/Z");
            return VF_ERR_INVALID_ARGUMENT;
        }
    } else if (collisionTemplate != VF_COLLISION_DEFAULT) {
        set_last_error("box path supports the Default and WriteMacroVars templates");
        return VF_ERR_UNSUPPORTED;
    }
    return VF_OK;
}

extern "C" int vf_k17_run_box(const vf_lbm_parameter* p, int32_t collisionTemplate, void* stream)
{
    bool macro = false;
    int rc = check_box(p, collisionTemplate, macro);
    if (rc) return rc;
    rc = dispatch_storage<BoxL>(*p, macro, (hipStream_t)stream);
    if (rc) return rc;
    return check_launch("K17CompressibleNavierStokes (box)");
}

template <class S, bool EVEN> struct BoxBBL {
    static void run(const vf_lbm_parameter& p, bool macro, const FusedBB& bb, bool& fused, hipStream_t s)
    {
        fused = launch_box_bb_t<S, EVEN>(p, macro, bb, s);
    }
};

extern "C" int vf_k17_run_box_bounce_back(const vf_lbm_parameter* p, int32_t collisionTemplate,
                                          const vf_post_bounce_back* bb, void* stream)
{
    bool macro = false;
    int rc = check_box(p, collisionTemplate, macro);
    if (rc) return rc;
    if (!bb) {
        set_last_error("run_box_bounce_back: null bounce-back lists");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (bb->numberOfEntries > 0 && (!bb->nodeEntry || !bb->entryIndex || !bb->entryLinks)) {
        set_last_error("run_box_bounce_back: nodeEntry, entryIndex and entryLinks are required");
        return VF_ERR_INVALID_ARGUMENT;
    }
    FusedBB fb{bb->nodeEntry, bb->entryIndex, bb->entryLinks, {}, {}, {}};
    for (int l = 0; l < 3; ++l) {
        const vf_q_boundary_conditions* q = bb->lists[l];
        if (!q || q->numberOfBCnodes == 0) continue;
        if (!q->k || !q->q27) {
            set_last_error("run_box_bounce_back: a list needs k and q27");
            return VF_ERR_INVALID_ARGUMENT;
        }
        if (!bb->velocity[l]) continue;
        if (!q->Vx || !q->Vy || !q->Vz) {
            set_last_error("run_box_bounce_back: a velocity list needs Vx, Vy, Vz");
            return VF_ERR_INVALID_ARGUMENT;
        }
        fb.vx[l] = q->Vx;
        fb.vy[l] = q->Vy;
        fb.vz[l] = q->Vz;
    }
    if (bb->numberOfEntries > 0) {
        bool fused = false;
        rc = dispatch_storage<BoxBBL>(*p, macro, fb, fused, (hipStream_t)stream);
        if (rc) return rc;
        if (fused) return check_launch("K17CompressibleNavierStokes (box, bounce-back epilogue)");
    }
    // the row kernel serves this box (or no node has a wall link): the collision, then each list
    rc = vf_k17_run_box(p, collisionTemplate, stream);
    for (int l = 0; l < 3 && rc == VF_OK; ++l) {
        const vf_q_boundary_conditions* q = bb->lists[l];
        if (!q || q->numberOfBCnodes == 0) continue;
        rc = bb->velocity[l] ? vf_VelocityBounceBack(p, q, stream) : vf_NoSlipBounceBack(p, q, stream);
    }
    return rc;
}
