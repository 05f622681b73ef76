// Boundary-condition kernels of the K17 time step and their C entry points.
//
// Post-collision BCs run after the collision, before the parity swap
// (src/gpu/core/Calculation/UpdateGrid27.cpp:69-77): one thread per BC node (a lane per
// direction for interpolated lists not along x rows, vf_bc_adjoint.h) reads the
// node's post-collision populations with the current parity and, for every link d with
// 0 <= q_d <= 1, writes the incoming population inv(d) of the next step into the slot the
// next collision reads (flipped parity). Formulas follow the closed forms of
// src/gpu/core/Utilities/KernelUtilities.h:171-214:
//   NoSlipBounceBack          f_inv(d) = f_d                                   (NoSlipBounceBack.cu:141-174)
//   VelocityBounceBack        f_inv(d) = f_d - 6 w_d c_d.U                      (VelocityBounceBack.cu:150-185)
//   *InterpolatedCompressible q-interpolated bounce-back with pre-collision reconstruction
// The pressure BC is a pre-collision kernel run after the swap (PressureNonEquilibriumCompressible.cu).
#include "vf_bc_adjoint.h" // BcKind
#include "vf_lattice.h"

namespace vfa {

struct BcArgs {
    const int32_t* k;
    const float* q;
    uint32_t nbc;
    const float* vx;
    const float* vy;
    const float* vz;
    float omega;
    float* postOut; // 27 * nbc post-collision values stored for the adjoint, or null
};

template <class S, int KIND>
__global__ __launch_bounds__(256) void bc_post_kernel(S* __restrict__ f, Addr a, BcArgs b, int even_flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.nbc) return;
    const bool even = even_flag != 0;
    uint32_t n[8];
    cell_nodes(a, (uint32_t)b.k[i], n);
    // every load issued before the first use: the link distances decide the branches below,
    // and loading each inside its branch would serialise 26 memory latencies
    float qd[27], post[27];
#pragma unroll
    for (int d = 1; d < 27; ++d) qd[d] = b.q[(uint64_t)d * b.nbc + i];
#pragma unroll
    for (int d = 0; d < 27; ++d) post[d] = read_post(f, a, n, d, even);
    if (b.postOut) {
        // the linearisation point of this node's adjoint (SoA: coalesced over the BC nodes)
#pragma unroll
        for (int d = 0; d < 27; ++d) b.postOut[(uint64_t)d * b.nbc + i] = post[d];
    }

    float drho = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f, cu_sq = 0.0f;
    if (KIND >= kNoSlipInterp) {
        float jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
        for (int d = 0; d < 27; ++d) {
            drho += post[d];
            if (cx(d)) jx += float(cx(d)) * post[d];
            if (cy(d)) jy += float(cy(d)) * post[d];
            if (cz(d)) jz += float(cz(d)) * post[d];
        }
        const float ir = 1.0f / (1.0f + drho);
        ux = jx * ir;
        uy = jy * ir;
        uz = jz * ir;
        cu_sq = 1.5f * (ux * ux + uy * uy + uz * uz) * (1.0f + drho);
    }
    float Ux = 0.0f, Uy = 0.0f, Uz = 0.0f;
    if (KIND == kVelocityBB || KIND == kVelocityInterp || KIND == kVelocityPressureInterp) {
        Ux = b.vx[i];
        Uy = b.vy[i];
        Uz = b.vz[i];
    }
    const float omega = b.omega;
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        const float q = qd[d];
        if (!(q >= 0.0f && q <= 1.0f)) continue;
        const int r = inv(d);
        const float w = weight(d);
        float v;
        if (KIND == kNoSlipBB) {
            v = post[d];
        } else if (KIND == kVelocityBB) {
            v = velocity_bounce_back(post[d], d, Ux, Uy, Uz);
        } else {
            const float cu = c_dot(d, ux, uy, uz);
            const float feq = w * (drho + 4.5f * cu * cu * (1.0f + drho) - cu_sq); // getEquilibriumForBC
            const float fd = post[d], fr = post[r];
            const float wall = (KIND == kNoSlipInterp) ? 0.0f : 6.0f * w * c_dot(d, Ux, Uy, Uz);
            v = (1.0f - q) / (1.0f + q) * (fd - fr + (fd + fr - 2.0f * feq * omega) / (1.0f - omega)) * 0.5f +
                (q * (fd + fr) - wall) / (1.0f + q);
            if (KIND == kVelocityPressureInterp) v -= w * drho;
        }
        write_pre(f, a, n, r, !even, v);
    }
}

// The interpolated post-collision BCs with a lane per direction (lists not along x rows; the
// arithmetic of bc_post_kernel, the node sums over the lane group). Lane d reads the node's
// post-collision value d and writes the incoming population inv(d) of its link; both cells lie at
// the cell node moved by +1 where c_d > 0, in slots inv(d) and d (even parity; swapped for odd),
// so no lane writes a cell another lane of the list reads (links end outside the fluid).
template <class S, int KIND>
__global__ __launch_bounds__(kLaneBlock) void bc_post_lanes_kernel(S* __restrict__ f, Addr a, BcArgs b, int even_flag)
{
    static_assert(KIND >= kNoSlipInterp, "bounce-back kinds keep one thread per node");
    const uint32_t i = blockIdx.x * kLaneNodes + (threadIdx.x / kLaneGroup);
    const bool node = i < b.nbc; // every lane of the group stays for the shuffles
    const LaneDir L = lane_dir();
    const bool on = node && L.valid;
    const bool even = even_flag != 0;
    float post = 0.0f, q = -1.0f;
    S* wcell = nullptr;
    if (on) {
        const uint32_t m = moved_node(a, (uint32_t)b.k[i], L.cx > 0, L.cy > 0, L.cz > 0);
        post = Codec<S>::load(f + (uint64_t)(even ? L.r : L.d) * a.N + m);
        if (L.d != 0) q = b.q[(uint64_t)L.d * b.nbc + i];
        wcell = f + (uint64_t)(even ? L.d : L.r) * a.N + m;
        if (b.postOut) b.postOut[(uint64_t)L.d * b.nbc + i] = post;
    }
    const float drho = group_sum(post);
    const float ir = 1.0f / (1.0f + drho);
    const float ux = group_sum(float(L.cx) * post) * ir, uy = group_sum(float(L.cy) * post) * ir,
                uz = group_sum(float(L.cz) * post) * ir;
    const float cu_sq = 1.5f * (ux * ux + uy * uy + uz * uz) * (1.0f + drho);
    const float fr = __shfl(post, L.valid ? L.r : L.d, kLaneGroup);
    if (!(on && q >= 0.0f && q <= 1.0f)) return;
    const float omega = b.omega, w = L.w, fd = post;
    const float cu = float(L.cx) * ux + float(L.cy) * uy + float(L.cz) * uz;
    const float feq = w * (drho + 4.5f * cu * cu * (1.0f + drho) - cu_sq); // getEquilibriumForBC
    float wall = 0.0f;
    if (KIND != kNoSlipInterp)
        wall = 6.0f * w * (float(L.cx) * b.vx[i] + float(L.cy) * b.vy[i] + float(L.cz) * b.vz[i]);
    float v = (1.0f - q) / (1.0f + q) * (fd - fr + (fd + fr - 2.0f * feq * omega) / (1.0f - omega)) * 0.5f +
              (q * (fd + fr) - wall) / (1.0f + q);
    if (KIND == kVelocityPressureInterp) v -= w * drho;
    Codec<S>::store(wcell, v);
}

struct PressureArgs {
    const int32_t* k;
    const int32_t* kN;
    const float* rhoBC;
    uint32_t nbc;
    int axis; // 0 x, 1 y, 2 z
    int sign; // +1 for the P side (write c <= 0), -1 for the M side (write c >= 0)
};

template <class S>
__global__ __launch_bounds__(256) void pressure_neq_kernel(S* __restrict__ f, Addr a, PressureArgs b, int even_flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.nbc) return;
    const bool even = even_flag != 0;
    uint32_t nn[8];
    cell_nodes(a, (uint32_t)b.kN[i], nn);
    float fn[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) fn[d] = read_pre(f, a, nn, d, even);
    float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        drho += fn[d];
        if (cx(d)) jx += float(cx(d)) * fn[d];
        if (cy(d)) jy += float(cy(d)) * fn[d];
        if (cz(d)) jz += float(cz(d)) * fn[d];
    }
    const float ir = 1.0f / (drho + 1.0f);
    const float ux = jx * ir, uy = jy * ir, uz = jz * ir;
    const float cu_sq = 1.5f * (ux * ux + uy * uy + uz * uz);
    const float rbc = b.rhoBC[i];
    uint32_t n[8];
    cell_nodes(a, (uint32_t)b.k[i], n);
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const int c = b.axis == 0 ? cx(d) : (b.axis == 1 ? cy(d) : cz(d));
        if (c * b.sign > 0) continue;
        // non-equilibrium part of the neighbour plus the equilibrium at the prescribed density
        const float v = fn[d] - feq_shifted(d, drho, ux, uy, uz, cu_sq) + feq_shifted(d, rbc, ux, uy, uz, cu_sq);
        write_pre(f, a, n, d, even, v);
    }
}

// The pressure BC with a lane per direction (faces normal to x; the arithmetic of
// pressure_neq_kernel): lane d reads the neighbour's pre-collision value d and, where the face
// writes direction d, writes the BC node's cell d (kN is not itself a pressure BC node).
template <class S>
__global__ __launch_bounds__(kLaneBlock) void pressure_neq_lanes_kernel(S* __restrict__ f, Addr a, PressureArgs b,
                                                                        int even_flag)
{
    const uint32_t i = blockIdx.x * kLaneNodes + (threadIdx.x / kLaneGroup);
    const bool node = i < b.nbc;
    const LaneDir L = lane_dir();
    const bool on = node && L.valid;
    const bool even = even_flag != 0;
    float fn = 0.0f, rbc = 0.0f;
    S* wcell = nullptr;
    if (on) {
        const uint64_t slot = (uint64_t)(even ? L.d : L.r) * a.N;
        const bool sx = L.cx < 0, sy = L.cy < 0, sz = L.cz < 0;
        fn = Codec<S>::load(f + slot + moved_node(a, (uint32_t)b.kN[i], sx, sy, sz));
        rbc = b.rhoBC[i];
        const int c = b.axis == 0 ? L.cx : (b.axis == 1 ? L.cy : L.cz);
        if (c * b.sign <= 0) wcell = f + slot + moved_node(a, (uint32_t)b.k[i], sx, sy, sz);
    }
    const float drho = group_sum(fn);
    const float ir = 1.0f / (drho + 1.0f);
    const float ux = group_sum(float(L.cx) * fn) * ir, uy = group_sum(float(L.cy) * fn) * ir,
                uz = group_sum(float(L.cz) * fn) * ir;
    const float cu_sq = 1.5f * (ux * ux + uy * uy + uz * uz);
    if (!wcell) return;
    // non-equilibrium part of the neighbour plus the equilibrium at the prescribed density
    const float cu = float(L.cx) * ux + float(L.cy) * uy + float(L.cz) * uz;
    const float e = 3.0f * cu + 4.5f * cu * cu - cu_sq;
    const float v = fn - L.w * (drho + (1.0f + drho) * e) + L.w * (rbc + (1.0f + rbc) * e);
    Codec<S>::store(wcell, v);
}

template <int KIND> static int launch_bc_post(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, void* stream,
                                              const char* name)
{
    if (!p || !bc || !p->distributions) {
        set_last_error("null parameter or boundary condition");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (bc->numberOfBCnodes == 0) return VF_OK;
    if (!bc->k || !bc->q27) {
        set_last_error("boundary condition needs k and q27");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if ((KIND == kVelocityBB || KIND == kVelocityInterp || KIND == kVelocityPressureInterp) &&
        (!bc->Vx || !bc->Vy || !bc->Vz)) {
        set_last_error("velocity boundary condition needs Vx, Vy, Vz");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!p->neighborX && (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes) {
        set_last_error("boundary condition needs neighbour arrays or a dense box");
        return VF_ERR_INVALID_ARGUMENT;
    }
    const Addr a = make_addr(*p);
    BcArgs b{bc->k, bc->q27, bc->numberOfBCnodes, bc->Vx, bc->Vy, bc->Vz, p->omega, bc->postCollision};
    const dim3 grid((bc->numberOfBCnodes + kBcBlock - 1) / kBcBlock);
    hipStream_t s = (hipStream_t)stream;
    if constexpr (KIND >= kNoSlipInterp) {
        if (!bc->nodesAlongRows) {
            const dim3 lanes((bc->numberOfBCnodes + kLaneNodes - 1) / kLaneNodes);
            switch (p->storage) {
            case VF_STORAGE_FP32:
                hipLaunchKernelGGL((bc_post_lanes_kernel<float, KIND>), lanes, dim3(kLaneBlock), 0, s,
                                   (float*)p->distributions, a, b, p->isEvenTimestep);
                break;
            case VF_STORAGE_FP16:
                hipLaunchKernelGGL((bc_post_lanes_kernel<__half, KIND>), lanes, dim3(kLaneBlock), 0, s,
                                   (__half*)p->distributions, a, b, p->isEvenTimestep);
                break;
            case VF_STORAGE_BF16:
                hipLaunchKernelGGL((bc_post_lanes_kernel<__hip_bfloat16, KIND>), lanes, dim3(kLaneBlock), 0, s,
                                   (__hip_bfloat16*)p->distributions, a, b, p->isEvenTimestep);
                break;
            default:
                set_last_error("unknown storage type");
                return VF_ERR_INVALID_ARGUMENT;
            }
            return check_launch(name);
        }
    }
    switch (p->storage) {
    case VF_STORAGE_FP32:
        hipLaunchKernelGGL((bc_post_kernel<float, KIND>), grid, dim3(kBcBlock), 0, s, (float*)p->distributions, a, b, p->isEvenTimestep);
        break;
    case VF_STORAGE_FP16:
        hipLaunchKernelGGL((bc_post_kernel<__half, KIND>), grid, dim3(kBcBlock), 0, s, (__half*)p->distributions, a, b, p->isEvenTimestep);
        break;
    case VF_STORAGE_BF16:
        hipLaunchKernelGGL((bc_post_kernel<__hip_bfloat16, KIND>), grid, dim3(kBcBlock), 0, s, (__hip_bfloat16*)p->distributions, a, b,
                           p->isEvenTimestep);
        break;
    default:
        set_last_error("unknown storage type");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return check_launch(name);
}

} // namespace vfa

using namespace vfa;

extern "C" int vf_NoSlipBounceBack(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, void* stream)
{
    return launch_bc_post<kNoSlipBB>(p, bc, stream, "NoSlipBounceBack_Device execution failed");
}
extern "C" int vf_NoSlipInterpolatedCompressible(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, void* stream)
{
    return launch_bc_post<kNoSlipInterp>(p, bc, stream, "NoSlipInterpolatedCompressible_Device execution failed");
}
extern "C" int vf_VelocityBounceBack(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, void* stream)
{
    return launch_bc_post<kVelocityBB>(p, bc, stream, "VelocityBounceBack_Device execution failed");
}
extern "C" int vf_VelocityInterpolatedCompressible(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, void* stream)
{
    return launch_bc_post<kVelocityInterp>(p, bc, stream, "VelocityInterpolatedCompressible_Device execution failed");
}
extern "C" int vf_VelocityWithPressureInterpolatedCompressible(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc,
                                                               void* stream)
{
    return launch_bc_post<kVelocityPressureInterp>(p, bc, stream,
                                                   "VelocityWithPressureInterpolatedCompressible_Device execution failed");
}

extern "C" int vf_PressureNonEquilibriumCompressible(const vf_lbm_parameter* p, const vf_q_boundary_conditions* bc, void* stream)
{
    if (!p || !bc || !p->distributions) {
        set_last_error("null parameter or boundary condition");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (bc->numberOfBCnodes == 0) return VF_OK;
    if (!bc->k || !bc->kN || !bc->RhoBC) {
        set_last_error("pressure boundary condition needs k, kN and RhoBC");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!p->neighborX && (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes) {
        set_last_error("boundary condition needs neighbour arrays or a dense box");
        return VF_ERR_INVALID_ARGUMENT;
    }
    PressureArgs b{bc->k, bc->kN, bc->RhoBC, bc->numberOfBCnodes, 0, 0};
    // directions of src/gpu/core/BoundaryConditions/Pressure/PressureNonEquilibriumCompressible.cu:179-309
    switch (bc->direction) {
    case dP00: b.axis = 0; b.sign = 1; break;
    case dM00: b.axis = 0; b.sign = -1; break;
    case d0P0: b.axis = 1; b.sign = 1; break;
    case d0M0: b.axis = 1; b.sign = -1; break;
    case d00P: b.axis = 2; b.sign = 1; break;
    case d00M: b.axis = 2; b.sign = -1; break;
    default: return VF_OK; // the reference kernel writes nothing for other directions
    }
    const Addr a = make_addr(*p);
    const dim3 grid((bc->numberOfBCnodes + kBcBlock - 1) / kBcBlock);
    hipStream_t s = (hipStream_t)stream;
    if (!bc->nodesAlongRows) {
        const dim3 lanes((bc->numberOfBCnodes + kLaneNodes - 1) / kLaneNodes);
        switch (p->storage) {
        case VF_STORAGE_FP32:
            hipLaunchKernelGGL((pressure_neq_lanes_kernel<float>), lanes, dim3(kLaneBlock), 0, s,
                               (float*)p->distributions, a, b, p->isEvenTimestep);
            break;
        case VF_STORAGE_FP16:
            hipLaunchKernelGGL((pressure_neq_lanes_kernel<__half>), lanes, dim3(kLaneBlock), 0, s,
                               (__half*)p->distributions, a, b, p->isEvenTimestep);
            break;
        case VF_STORAGE_BF16:
            hipLaunchKernelGGL((pressure_neq_lanes_kernel<__hip_bfloat16>), lanes, dim3(kLaneBlock), 0, s,
                               (__hip_bfloat16*)p->distributions, a, b, p->isEvenTimestep);
            break;
        default:
            set_last_error("unknown storage type");
            return VF_ERR_INVALID_ARGUMENT;
        }
        return check_launch("PressureNonEquilibriumCompressible_Device execution failed");
    }
    switch (p->storage) {
    case VF_STORAGE_FP32:
        hipLaunchKernelGGL((pressure_neq_kernel<float>), grid, dim3(kBcBlock), 0, s, (float*)p->distributions, a, b, p->isEvenTimestep);
        break;
    case VF_STORAGE_FP16:
        hipLaunchKernelGGL((pressure_neq_kernel<__half>), grid, dim3(kBcBlock), 0, s, (__half*)p->distributions, a, b, p->isEvenTimestep);
        break;
    case VF_STORAGE_BF16:
        hipLaunchKernelGGL((pressure_neq_kernel<__hip_bfloat16>), grid, dim3(kBcBlock), 0, s, (__hip_bfloat16*)p->distributions, a, b,
                           p->isEvenTimestep);
        break;
    default:
        set_last_error("unknown storage type");
        return VF_ERR_INVALID_ARGUMENT;
    }
    return check_launch("PressureNonEquilibriumCompressible_Device execution failed");
}
