// device.h -- device helpers shared by the kernel translation units: wave
// scans and sums, the record types' payload widths and kMinNum, payload loads
// and the decode of one record into its values.
//
// Record decode follows ProcessArrayFunctor (src/AdaBool/inverted_list.cu.h:
// 61-94) and the bitset emit (src/AdaBool/key_transformer.cu.h:100-111):
// header value always; payload bytes [0, kMinNum) always, later bytes until the
// first zero; a bitset's 8 payload words are the bits of its 256-value window.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "adabool.h"

namespace ab {

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// Inclusive wave scan by DPP: row shifts by 1, 2, 4, 8 inside rows of 16
// lanes (lanes shifted in from outside the row add 0), then lane 15 of each
// row added to the next row and lane 31 to the upper half.
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);
    return x;
}

__device__ __forceinline__ unsigned long long wave_incl_scan64(unsigned long long x)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        unsigned long long y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

template <int T> struct TypeTraits;
template <> struct TypeTraits<AB_ARRAY1>  { static constexpr int P = 0;  static constexpr int kMin = 0; };
template <> struct TypeTraits<AB_ARRAY4>  { static constexpr int P = 4;  static constexpr int kMin = 1; };
template <> struct TypeTraits<AB_ARRAY8>  { static constexpr int P = 8;  static constexpr int kMin = 5; };
template <> struct TypeTraits<AB_ARRAY16> { static constexpr int P = 16; static constexpr int kMin = 9; };
template <> struct TypeTraits<AB_ARRAY32> { static constexpr int P = 32; static constexpr int kMin = 17; };
template <> struct TypeTraits<AB_BITSET>  { static constexpr int P = 32; static constexpr int kMin = 0; };

// Loads a record's payload as little-endian words (one 4/8/16-byte load, or two
// 16-byte loads for 32-byte payloads).
template <int T>
__device__ __forceinline__ void load_payload(const uint8_t *__restrict__ pay, uint32_t idx, uint32_t (&w)[8])
{
    if constexpr (T == AB_ARRAY4) {
        w[0] = reinterpret_cast<const uint32_t *>(pay)[idx];
    } else if constexpr (T == AB_ARRAY8) {
        uint2 v = reinterpret_cast<const uint2 *>(pay)[idx];
        w[0] = v.x; w[1] = v.y;
    } else if constexpr (T == AB_ARRAY16) {
        uint4 v = reinterpret_cast<const uint4 *>(pay)[idx];
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else if constexpr (T == AB_ARRAY32 || T == AB_BITSET) {
        const uint4 *p = reinterpret_cast<const uint4 *>(pay) + 2ull * idx;
        uint4 a = p[0], b = p[1];
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
        w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
    }
}

// Calls flush(word, bits) once per result word the record touches.
template <int T, class F>
__device__ __forceinline__ void emit_record(uint32_t h, const uint32_t (&w)[8], F &&flush)
{
    if constexpr (T == AB_BITSET) {
        const uint32_t base = h >> 5;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (w[j]) flush(base + j, w[j]);
    } else if constexpr (T == AB_ARRAY1) {
        flush(h >> 5, 1u << (h & 31));
    } else {
        constexpr int P = TypeTraits<T>::P, kMin = TypeTraits<T>::kMin;
        const uint32_t hi = h & 0xffffff00u;
        uint32_t cw = h >> 5, bits = 1u << (h & 31);
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const uint32_t b = (w[i >> 2] >> (8 * (i & 3))) & 0xffu;
            if (i >= kMin && b == 0) break;
            const uint32_t v = hi | b;
            if ((v >> 5) != cw) {
                flush(cw, bits);
                cw = v >> 5;
                bits = 0;
            }
            bits |= 1u << (v & 31);
        }
        flush(cw, bits);
    }
}

// Calls f(value) for every value of a record, in ascending order.
template <int T, class F>
__device__ __forceinline__ void record_values(uint32_t h, const uint32_t (&w)[8], F &&f)
{
    if constexpr (T == AB_BITSET) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint32_t x = w[j];
            while (x) {
                f(h + 32u * j + (uint32_t)__builtin_ctz(x));
                x &= x - 1;
            }
        }
    } else if constexpr (T == AB_ARRAY1) {
        f(h);
    } else {
        constexpr int P = TypeTraits<T>::P, kMin = TypeTraits<T>::kMin;
        const uint32_t hi = h & 0xffffff00u;
        f(h);
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const uint32_t b = (w[i >> 2] >> (8 * (i & 3))) & 0xffu;
            if (i >= kMin && b == 0) break;
            f(hi | b);
        }
    }
}

// Number of values of a record.
template <int T>
__device__ __forceinline__ uint32_t record_count(const uint32_t (&w)[8])
{
    if constexpr (T == AB_BITSET) {
        uint32_t n = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) n += (uint32_t)__popc(w[j]);
        return n;
    } else if constexpr (T == AB_ARRAY1) {
        return 1;
    } else {
        constexpr int P = TypeTraits<T>::P, kMin = TypeTraits<T>::kMin;
        uint32_t n = 1 + kMin;
#pragma unroll
        for (int i = kMin; i < P; ++i) {
            if (((w[i >> 2] >> (8 * (i & 3))) & 0xffu) == 0) break;
            ++n;
        }
        return n;
    }
}

// Largest j in [0, K) with koff[j] <= x (koff non-decreasing, koff[0] <= x <
// koff[K]); (NT+1)-ary search by the whole work-group.
template <int NT>
__device__ uint64_t wg_find(const uint64_t *__restrict__ koff, uint64_t K, uint64_t x, unsigned long long *s_lo,
                            unsigned long long *s_hi)
{
    uint64_t lo = 0, hi = K;
    while (hi - lo > 1) {
        if (threadIdx.x == 0) {
            *s_lo = lo;
            *s_hi = hi;
        }
        __syncthreads();
        const uint64_t span = hi - lo;
        const uint64_t p = lo + (span * (threadIdx.x + 1)) / (NT + 1);
        if (p > lo && p < hi) {
            if (koff[p] <= x)
                atomicMax(s_lo, (unsigned long long)p);
            else
                atomicMin(s_hi, (unsigned long long)p);
        }
        __syncthreads();
        lo = *s_lo;
        hi = *s_hi;
        __syncthreads();
    }
    return lo;
}

}  // namespace ab
