// Aligned, non-temporal vector loads and stores of V consecutive populations of one
// direction array. Used by the row kernels, where every lane owns V consecutive x-nodes and
// every global access of a wave is aligned and contiguous (the x-shift of the eso-twist
// layout is resolved with lane shuffles).
//
// Loads keep the raw lane word (the word, not a widened value, is what lanes trade for the
// shifted directions); elements are widened to FP32 where they are used. Stores narrow two
// values per instruction (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32, round to nearest even, the
// rounding of Codec<S>::store).
#pragma once

#include <cstdlib>

#include "vf_common.h"

namespace vfa {

// The nine directions with c_x = -1: their cells lie one node further in +x (node_sel bit 0).
__device__ constexpr int kShiftedDirs[9] = {dM00, dMM0, dMP0, dM0M, dM0P, dMPP, dMMP, dMPM, dMMM};
__host__ __device__ constexpr int shifted_index(int d)
{
    return d == dM00 ? 0 : d == dMM0 ? 1 : d == dMP0 ? 2 : d == dM0M ? 3 : d == dM0P ? 4 : d == dMPP ? 5
         : d == dMMP ? 6 : d == dMPM ? 7 : d == dMMM ? 8 : -1;
}

// Buffer-descriptor addressing of a row kernel: slots below kRowSplit from the buffer's base,
// the rest from base + kRowSplit slots, so 32-bit offsets reach every slot while kRowSplit
// slots span less than 4 GiB.
constexpr int kRowSplit = 14;

// Workgroup b of a launch runs on XCD b % 8; each of the eight XCDs has its own L2. With
// xcd_block the XCDs take contiguous ranges of R blocks instead (XCD j: [j q + min(j, R % 8),
// ...) with q = R / 8), so that neighbouring rows meet in one L2. Rows whose bytes are not a
// whole number of 128-byte lines share a line at either end with the rows beside them in y and
// z, which then meet in one L2 instead of two (on MI355X 258-node FP16 rows run 12 % faster
// with it, 130-node FP32 rows 13 %, while aligned rows run 2-5 % slower).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t R)
{
    const uint32_t q = R >> 3, rem = R & 7, j = b & 7;
    return j * q + (j < rem ? j : rem) + (b >> 3);
}

// host: whether a launch over rows of nx elements of elemBytes takes the XCD block ranges
// (choice: 0 automatic, 1 always, 2 never; vf_kernel_choice.rowXcdRanges)
inline uint32_t xcd_rows(int nx, size_t elemBytes, int choice)
{
    if (choice == 1 || choice == 2) return choice == 1;
    return ((size_t)nx * elemBytes) % 128 != 0;
}

// host: the same for the plane-flat out-of-place kernel, whose workgroups take consecutive spans
// of a plane: spans share their end lines with the next one where a plane's bytes are not a
// whole number of lines (on MI355X, 258 x 130 x 130 FP16 out of place 140 -> 125 us alone and
// 121-130 -> 118 us inside the two-level step; the sphere channel's 194 x 98 x 98 FP32 92 -> 87)
inline uint32_t xcd_planes(int nx, int ny, size_t elemBytes, int choice)
{
    if (choice == 1 || choice == 2) return choice == 1;
    return ((size_t)nx * (size_t)ny * elemBytes) % 128 != 0;
}

template <int BYTES> struct RawT;
template <> struct RawT<2> { using T = unsigned short; };
template <> struct RawT<4> { using T = unsigned int; };
template <> struct RawT<8> { using T = unsigned int __attribute__((ext_vector_type(2))); };
template <> struct RawT<16> { using T = unsigned int __attribute__((ext_vector_type(4))); };

typedef float rowio_f2 __attribute__((ext_vector_type(2)));
typedef _Float16 rowio_h2 __attribute__((ext_vector_type(2)));
typedef __bf16 rowio_b2 __attribute__((ext_vector_type(2)));

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
    static constexpr int PER_WORD = 4 / (int)sizeof(S); // elements per 32-bit word
    using Raw = typename RawT<BYTES>::T;

    __device__ static Raw load(const S* p) { return __builtin_nontemporal_load(reinterpret_cast<const Raw*>(p)); }

    // element j (0 or 1) of a 32-bit word, widened to FP32
    __device__ static float of_word(unsigned w, int j)
    {
        if constexpr (sizeof(S) == 4) return __uint_as_float(w);
        else if constexpr (Codec<S>::storage == VF_STORAGE_BF16) return __uint_as_float(j ? (w & 0xffff0000u) : (w << 16));
        else return (float)__builtin_bit_cast(rowio_h2, w)[j];
    }

    // element i of the lane's V values
    __device__ static float elem(const Raw& r, int i) { return of_word(word<BYTES>(r, i / PER_WORD), i % PER_WORD); }

    // the word holding element 0 (what the previous lane needs for its shifted directions)
    __device__ static unsigned word0(const Raw& r) { return word<BYTES>(r, 0); }

    // one element's bits in the low part of a word (of_word(w, 0) reads it back)
    __device__ static unsigned elem_bits(const S* p)
    {
        if constexpr (sizeof(S) == 4) return *reinterpret_cast<const unsigned*>(p);
        else return *reinterpret_cast<const unsigned short*>(p);
    }

    // cache policy of the buffer forms: non-temporal (each population is touched once per step)
    static constexpr int kNT = 2;

    // buffer forms: address = descriptor base + soff (uniform, SGPR) + voff (per lane), so no
    // per-lane 64-bit address arithmetic is needed
    __device__ static Raw load(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff)
    {
        if constexpr (BYTES == 2) return __builtin_amdgcn_raw_buffer_load_b16(rs, voff, soff, kNT);
        else if constexpr (BYTES == 4) return __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, kNT);
        else if constexpr (BYTES == 8) return __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, kNT);
        else return __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, kNT);
    }
    __device__ static void store(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff, const float (&in)[V])
    {
        const Raw r = pack(in);
        if constexpr (BYTES == 4) __builtin_amdgcn_raw_buffer_store_b32(r, rs, voff, soff, kNT);
        else if constexpr (BYTES == 8) __builtin_amdgcn_raw_buffer_store_b64(r, rs, voff, soff, kNT);
        else __builtin_amdgcn_raw_buffer_store_b128(r, rs, voff, soff, kNT);
    }

    __device__ static void store(S* p, const float (&in)[V])
    {
        __builtin_nontemporal_store(pack(in), reinterpret_cast<Raw*>(p));
    }

    __device__ static Raw pack(const float (&in)[V])
    {
        Raw r;
        if constexpr (sizeof(S) == 4) {
#pragma unroll
            for (int i = 0; i < V; ++i) set_word<BYTES>(r, i, __float_as_uint(in[i]));
        } else {
            static_assert(V % 2 == 0, "16-bit rows store whole 32-bit words");
#pragma unroll
            for (int i = 0; i < V; i += 2) {
                const rowio_f2 x = {in[i], in[i + 1]};
                unsigned w;
                if constexpr (Codec<S>::storage == VF_STORAGE_BF16) w = __builtin_bit_cast(unsigned, __builtin_convertvector(x, rowio_b2));
                else w = __builtin_bit_cast(unsigned, __builtin_convertvector(x, rowio_h2));
                set_word<BYTES>(r, i / 2, w);
            }
        }
        return r;
    }
};

} // namespace vfa
