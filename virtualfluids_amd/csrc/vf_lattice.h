// Device helpers shared by the boundary, initialisation and output kernels:
// cell-node lookup (neighbour arrays or dense periodic box) and esoteric-twist
// reads/writes with a runtime parity flag.
#pragma once

#include "vf_common.h"

namespace vfa {

#define VF_PLANE_HD __host__ __device__ __forceinline__

struct Addr {
    uint64_t N;
    const uint32_t* nbx; // +x/+y/+z neighbour arrays (sparse grids)
    const uint32_t* nby;
    const uint32_t* nbz;
    int nx, ny, nz; // dense periodic box when nbx == nullptr
};

inline Addr make_addr(const vf_lbm_parameter& p)
{
    return Addr{p.numberOfNodes, p.neighborX, p.neighborY, p.neighborZ, p.nx, p.ny, p.nz};
}

// The 8 nodes of the cell at k: bit0 = +x, bit1 = +y, bit2 = +z
// (ListIndices, src/gpu/core/Utilities/KernelUtilities.h:221-245).
__device__ __forceinline__ void cell_nodes(const Addr& a, uint32_t k, uint32_t (&n)[8])
{
    if (a.nbx) {
        n[0] = k;
        n[1] = a.nbx[k];
        n[2] = a.nby[k];
        n[4] = a.nbz[k];
        n[3] = a.nby[n[1]];
        n[5] = a.nbz[n[1]];
        n[6] = a.nbz[n[2]];
        n[7] = a.nbz[n[3]];
    } else {
        const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny, nz = (uint32_t)a.nz;
        const uint32_t x = k % nx, t = k / nx, y = t % ny, z = t / ny;
        const uint32_t xp = x + 1 == nx ? 0 : x + 1, yp = y + 1 == ny ? 0 : y + 1, zp = z + 1 == nz ? 0 : z + 1;
        const uint32_t r00 = nx * (y + ny * z), r10 = nx * (yp + ny * z), r01 = nx * (y + ny * zp),
                       r11 = nx * (yp + ny * zp);
        n[0] = r00 + x; n[1] = r00 + xp; n[2] = r10 + x; n[3] = r10 + xp;
        n[4] = r01 + x; n[5] = r01 + xp; n[6] = r11 + x; n[7] = r11 + xp;
    }
}

// Plane-flat numbering of the one-node-per-thread box kernels: a launch of (plane_blocks, nz)
// workgroups of kPlaneBlock threads gives thread t of plane z the node t of the plane's nx * ny
// nodes, so every lane has a node whatever nx is. y = t / nx comes from a double reciprocal:
// v = (t + 1/2) / nx lies at least 1/(2 nx) from an integer, and the two roundings (of 1/nx and
// of the product) err by at most v 2^-52 < 1/(2 nx) for every t < 2^51, so the truncation is
// exact.
constexpr int kPlaneBlock = 256;
struct PlaneDiv {
    uint32_t nx;
    uint32_t planeNodes; // nx * ny
    double invNx;        // 1 / nx
};
inline PlaneDiv make_plane_div(int nx, int ny)
{
    return PlaneDiv{(uint32_t)nx, (uint32_t)nx * (uint32_t)ny, 1.0 / (double)nx};
}
inline unsigned plane_blocks(int nx, int ny)
{
    return (unsigned)(((uint64_t)nx * (uint64_t)ny + kPlaneBlock - 1) / kPlaneBlock);
}
VF_PLANE_HD uint32_t plane_row(uint32_t t, const PlaneDiv& p) { return (uint32_t)(((double)t + 0.5) * p.invNx); }
// (x, y) of this thread's node in its plane (blockIdx.y); false for threads beyond the plane
__device__ __forceinline__ bool plane_xy(const PlaneDiv& p, int& x, int& y)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.planeNodes) return false;
    const uint32_t r = plane_row(t, p);
    y = (int)r;
    x = (int)(t - r * p.nx);
    return true;
}

template <class S> __device__ __forceinline__ S* slot_ptr(S* f, uint64_t N, int d, bool even)
{
    return f + (uint64_t)(even ? d : inv(d)) * N;
}

// pre-collision value of direction d (the value a collision at this cell would read)
template <class S> __device__ __forceinline__ float read_pre(const S* f, const Addr& a, const uint32_t (&n)[8], int d, bool even)
{
    return Codec<S>::load(slot_ptr(const_cast<S*>(f), a.N, d, even) + n[node_sel(d)]);
}
template <class S> __device__ __forceinline__ void write_pre(S* f, const Addr& a, const uint32_t (&n)[8], int d, bool even, float v)
{
    Codec<S>::store(slot_ptr(f, a.N, d, even) + n[node_sel(d)], v);
}
// post-collision value of direction d (stored in the slot of inv(d) at inv(d)'s node)
template <class S> __device__ __forceinline__ float read_post(const S* f, const Addr& a, const uint32_t (&n)[8], int d, bool even)
{
    return read_pre(f, a, n, inv(d), even);
}

__device__ __forceinline__ float c_dot(int d, float ux, float uy, float uz)
{
    float s = 0.0f;
    if (cx(d)) s += float(cx(d)) * ux;
    if (cy(d)) s += float(cy(d)) * uy;
    if (cz(d)) s += float(cz(d)) * uz;
    return s;
}

// VelocityBounceBack: the incoming population inv(d) from the post-collision value of the wall
// link d, f_inv(d) = f_d - 6 w_d c_d.U (getBounceBackDistributionForVeloBC,
// src/gpu/core/Utilities/KernelUtilities.h:186-190); NoSlipBounceBack returns f_d itself
__device__ __forceinline__ float velocity_bounce_back(float post_d, int d, float Ux, float Uy, float Uz)
{
    return post_d - 6.0f * weight(d) * c_dot(d, Ux, Uy, Uz);
}

// Full shifted equilibrium w_d (drho + (1+drho)(3 c.u + 9/2 (c.u)^2 - 3/2 u^2))
// (InitNavierStokesCompressible_Device.cu:173-201).
__device__ __forceinline__ float feq_shifted(int d, float drho, float ux, float uy, float uz, float cu_sq)
{
    if (d == 0) return (8.0f / 27.0f) * (drho - cu_sq * (1.0f + drho));
    const float cu = c_dot(d, ux, uy, uz);
    return weight(d) * (drho + (1.0f + drho) * (3.0f * cu + 4.5f * cu * cu - cu_sq));
}

} // namespace vfa
