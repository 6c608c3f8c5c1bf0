// gcs_device.hpp -- wave64 helpers and small device functions for gfx950.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gcs_common.hpp"

namespace gcs {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Lanes below the calling lane that have their bit set in `mask`.
__device__ __forceinline__ int mbcnt(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename T>
__device__ __forceinline__ T wave_min(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o);
        v = u < v ? u : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Inclusive prefix sum of an integer over the wave from DPP moves (row
// shifts by 1, 2, 4, 8 within each row of 16 lanes, then the broadcasts of
// lanes 15 and 31 into the later rows): six adds, no LDS traffic.  Called
// with all 64 lanes active.
__device__ __forceinline__ int wave_scan_incl_dpp(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2, 3
    return x;
}

// 32-bit words at base, as a buffer resource: a store at a byte offset of
// kDrop (beyond any such range) is dropped by the range check, which makes a
// per-lane conditional store a plain store.
constexpr int kDrop = (int)0x80000000;
constexpr int kBufferDword3 = 0x00020000;  // gfx9 raw buffer: 32-bit data, no swizzle

__device__ __forceinline__ __amdgpu_buffer_rsrc_t word_rsrc(int* base, int words)
{
    return __builtin_amdgcn_make_buffer_rsrc(base, 0, words * 4, kBufferDword3);
}

__device__ __forceinline__ void store_word(__amdgpu_buffer_rsrc_t r, int byte_off, int v)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, byte_off, 0, 0);
}

// Inclusive prefix sum over the wave.
// Exclusive prefix sum over the wave of a value below 2^bits (bits uniform),
// from one ballot per bit: no cross-lane data movement.  *total = wave sum.
__device__ __forceinline__ int wave_scan_excl_bits(int v, int bits, int* total)
{
    int ex = 0, tot = 0;
    for (int b = 0; b < bits; ++b) {
        const uint64_t m = ballot((v >> b) & 1);
        ex += mbcnt(m) << b;
        tot += __popcll(m) << b;
    }
    *total = uniform(tot);
    return ex;
}

__device__ __forceinline__ int wave_scan_incl(int v)
{
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(v, o);
        if (l >= o) v += u;
    }
    return v;
}

// ---- Morton (morton.h:11-28) -------------------------------------------------
__device__ __forceinline__ uint64_t expand_bits(uint32_t v)
{
    uint64_t w = (uint64_t)(v & 0x1FFFFFu);
    w = (w | (w << 32)) & 0x001F00000000FFFFull;
    w = (w | (w << 16)) & 0x001F0000FF0000FFull;
    w = (w | (w << 8)) & 0x100F00F00F00F00Full;
    w = (w | (w << 4)) & 0x10C30C30C30C30C3ull;
    w = (w | (w << 2)) & 0x1249249249249249ull;
    return w;
}

// Saturating round-toward-zero float -> uint32 (NaN and values below 1 give 0,
// values at or above 2^32 give 0xFFFFFFFF): the conversion the reference's
// morton.h:24-26 cast performs on its GPU.
__device__ __forceinline__ uint32_t f2u_sat(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

__device__ __forceinline__ uint64_t morton_key(float x, float y, float z, float mn, float mx)
{
#pragma clang fp contract(off)
    float scale = 2097151.0f / (mx - mn);
    uint32_t ix = f2u_sat((x - mn) * scale);
    uint32_t iy = f2u_sat((y - mn) * scale);
    uint32_t iz = f2u_sat((z - mn) * scale);
    return (expand_bits(ix) << 2) | (expand_bits(iy) << 1) | expand_bits(iz);
}

// Order-preserving float <-> uint32 encoding for atomic min/max.
__device__ __forceinline__ uint32_t fenc(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float fdec(uint32_t e)
{
    uint32_t u = (e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ int clz32s(uint32_t v) { return v ? __builtin_clz(v) : 32; }
__device__ __forceinline__ int clz64s(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

__device__ __forceinline__ uint32_t pack_rgbt(uint32_t type, uint32_t r, uint32_t g, uint32_t b)
{
    return type | (r << 8) | (g << 16) | (b << 24);
}

// Block-wide exclusive scan of one value per thread (256 threads; wsum is a
// 4-entry LDS scratch).  *total receives the block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total,
                                                         uint32_t* wsum)
{
    const int l = lane_id(), w = threadIdx.x >> 6;
    uint32_t incl = (uint32_t)wave_scan_incl((int)v);
    if (l == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t s = wsum[k];
        if (k < w) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

}  // namespace gcs
