// Aligned, non-temporal vector loads and stores of V consecutive populations of one
// direction array, decoded to / encoded from FP32. Used by the row kernel, where every
// lane owns V consecutive x-nodes and every global access of a wave is aligned and
// contiguous (the x-shift of the eso-twist layout is resolved with lane shuffles).
#pragma once

#include "vf_common.h"

namespace vfa {

template <int BYTES> struct RawT;
template <> struct RawT<2> { using T = unsigned short; };
template <> struct RawT<4> { using T = unsigned int; };
template <> struct RawT<8> { using T = unsigned int __attribute__((ext_vector_type(2))); };
template <> struct RawT<16> { using T = unsigned int __attribute__((ext_vector_type(4))); };

template <int BYTES> __device__ __forceinline__ unsigned word(const typename RawT<BYTES>::T& r, int i)
{
    if constexpr (BYTES <= 4) return (unsigned)r;
    else return r[i];
}
template <int BYTES> __device__ __forceinline__ void set_word(typename RawT<BYTES>::T& r, int i, unsigned w)
{
    if constexpr (BYTES <= 4) r = (typename RawT<BYTES>::T)w;
    else r[i] = w;
}

template <class S, int V> struct RowIO {
    static constexpr int BYTES = V * (int)sizeof(S);
    using Raw = typename RawT<BYTES>::T;

    __device__ static void load(const S* p, float (&out)[V])
    {
        const Raw r = __builtin_nontemporal_load(reinterpret_cast<const Raw*>(p));
        if constexpr (sizeof(S) == 4) {
#pragma unroll
            for (int i = 0; i < V; ++i) out[i] = __uint_as_float(word<BYTES>(r, i));
        } else if constexpr (Codec<S>::storage == VF_STORAGE_BF16) {
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const unsigned w = word<BYTES>(r, i / 2);
                out[i] = __uint_as_float((i & 1) ? (w & 0xffff0000u) : (w << 16));
            }
        } else {
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const unsigned w = word<BYTES>(r, i / 2);
                out[i] = __half2float(__ushort_as_half((unsigned short)((i & 1) ? (w >> 16) : (w & 0xffffu))));
            }
        }
    }

    __device__ static void store(S* p, const float (&in)[V])
    {
        Raw r;
        if constexpr (sizeof(S) == 4) {
#pragma unroll
            for (int i = 0; i < V; ++i) set_word<BYTES>(r, i, __float_as_uint(in[i]));
        } else {
#pragma unroll
            for (int i = 0; i < V; i += 2) {
                unsigned lo, hi = 0;
                if constexpr (Codec<S>::storage == VF_STORAGE_BF16) {
                    lo = __bfloat16_as_ushort(__float2bfloat16(in[i]));
                    if (i + 1 < V) hi = __bfloat16_as_ushort(__float2bfloat16(in[i + 1]));
                } else {
                    lo = __half_as_ushort(__float2half_rn(in[i]));
                    if (i + 1 < V) hi = __half_as_ushort(__float2half_rn(in[i + 1]));
                }
                if constexpr (BYTES == 2) r = (Raw)lo;
                else set_word<BYTES>(r, i / 2, lo | (hi << 16));
            }
        }
        __builtin_nontemporal_store(r, reinterpret_cast<Raw*>(p));
    }
};

} // namespace vfa
