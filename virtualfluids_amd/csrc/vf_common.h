// Shared device/host definitions of the gfx950 LBM time step.
//
// D3Q27 direction order and opposites follow src/lbm/constants/D3Q27.h:46-110;
// lattice weights follow src/basics/constants/NumericConstants.h:79-85.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/virtualfluids_amd.h"

namespace vfa {

enum : int {
    d000 = 0, dP00, dM00, d0P0, d0M0, d00P, d00M,
    dPP0, dMM0, dPM0, dMP0, dP0P, dM0M, dP0M, dM0P, d0PP, d0MM, d0PM, d0MP,
    dPPP, dMPP, dPMP, dMMP, dPPM, dMPM, dPMM, dMMM
};

__host__ __device__ constexpr int cx(int d)
{
    return (d == dP00 || d == dPP0 || d == dPM0 || d == dP0P || d == dP0M || d == dPPP || d == dPMP || d == dPPM ||
            d == dPMM)
               ? 1
               : ((d == dM00 || d == dMM0 || d == dMP0 || d == dM0M || d == dM0P || d == dMPP || d == dMMP ||
                   d == dMPM || d == dMMM)
                      ? -1
                      : 0);
}
__host__ __device__ constexpr int cy(int d)
{
    return (d == d0P0 || d == dPP0 || d == dMP0 || d == d0PP || d == d0PM || d == dPPP || d == dMPP || d == dPPM ||
            d == dMPM)
               ? 1
               : ((d == d0M0 || d == dMM0 || d == dPM0 || d == d0MM || d == d0MP || d == dPMP || d == dMMP ||
                   d == dPMM || d == dMMM)
                      ? -1
                      : 0);
}
__host__ __device__ constexpr int cz(int d)
{
    return (d == d00P || d == dP0P || d == dM0P || d == d0PP || d == d0MP || d == dPPP || d == dMPP || d == dPMP ||
            d == dMMP)
               ? 1
               : ((d == d00M || d == dM0M || d == dP0M || d == d0MM || d == d0PM || d == dPPM || d == dMPM ||
                   d == dPMM || d == dMMM)
                      ? -1
                      : 0);
}
__host__ __device__ constexpr int inv(int d)
{
    return d == 0 ? 0 : (d <= 18 ? ((d & 1) ? d + 1 : d - 1) : 45 - d);
}
// Eso-twist slot of direction d: even parity -> d, odd parity -> inv(d).
template <bool EVEN> __host__ __device__ constexpr int slot(int d) { return EVEN ? d : inv(d); }
// Which of the 8 cell nodes holds direction d: shifted (+1) along each axis with c < 0.
__host__ __device__ constexpr int node_sel(int d)
{
    return (cx(d) < 0 ? 1 : 0) | (cy(d) < 0 ? 2 : 0) | (cz(d) < 0 ? 4 : 0);
}
__host__ __device__ constexpr float weight(int d)
{
    return d == 0 ? 8.0f / 27.0f
                  : (d <= 6 ? 2.0f / 27.0f : (d <= 18 ? 1.0f / 54.0f : 1.0f / 216.0f));
}

// ---- population storage codecs -------------------------------------------------------------
// Populations are stored shifted (f_i - w_i) in FP32, FP16 or BF16; all arithmetic is FP32.
template <class S> struct Codec;

// round(v): the value a store followed by a load gives back
template <> struct Codec<float> {
    static constexpr int storage = VF_STORAGE_FP32;
    __device__ static float load(const float* p) { return *p; }
    __device__ static void store(float* p, float v) { *p = v; }
    __device__ static float round(float v) { return v; }
};

template <> struct Codec<__half> {
    static constexpr int storage = VF_STORAGE_FP16;
    __device__ static float load(const __half* p) { return __half2float(*p); }
    __device__ static void store(__half* p, float v) { *p = __float2half_rn(v); }
    __device__ static float round(float v) { return __half2float(__float2half_rn(v)); }
};

// bfloat16 is the upper half of an FP32 word; stores round to nearest even.
template <> struct Codec<__hip_bfloat16> {
    static constexpr int storage = VF_STORAGE_BF16;
    __device__ static float load(const __hip_bfloat16* p)
    {
        return __uint_as_float(((unsigned)*reinterpret_cast<const unsigned short*>(p)) << 16);
    }
    __device__ static void store(__hip_bfloat16* p, float v)
    {
        *reinterpret_cast<unsigned short*>(p) = __bfloat16_as_ushort(__float2bfloat16(v));
    }
    __device__ static float round(float v)
    {
        return __uint_as_float(((unsigned)__bfloat16_as_ushort(__float2bfloat16(v))) << 16);
    }
};

// Launch error bookkeeping shared by every entry point.
void set_last_error(const char* msg);
int check_launch(const char* what);

// Dynamic LDS bytes a launch reserves (and never touches) so that at most k of its workgroups
// are resident on one CU of the current device (0 for k <= 0: no cap).
size_t lds_reserve_for(int k);

} // namespace vfa
