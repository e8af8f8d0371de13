// Device helpers shared by the kernel translation units: orderable f32 keys,
// LDS bitonic sort and the block-level streaming top-k.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ovx {
namespace {

constexpr uint64_t kKeyMax = ~0ull;

__device__ __forceinline__ uint32_t f32_ord(float x) {
    uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float ord_f32(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}

// Ascending key = (score desc under f32::total_cmp, id asc).
__device__ __forceinline__ uint64_t score_key(float s, uint32_t id) {
    return ((uint64_t)(~f32_ord(s)) << 32) | id;
}

__device__ __forceinline__ float key_score(uint64_t k) { return ord_f32(~(uint32_t)(k >> 32)); }

__device__ __forceinline__ uint32_t next_pow2(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

template <typename T>
__device__ void bitonic_sort(T* a, uint32_t P) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < P / 2; i += nt) {
                uint32_t lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                uint32_t hi = lo + j;
                bool asc = (lo & k) == 0;
                T x = a[lo], y = a[hi];
                if ((x > y) == asc) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Block-level streaming top-k on 64-bit keys (ascending = better).
// keys[0, k) hold the best seen so far (sorted after each flush, padded with
// kKeyMax); offers that beat the current k-th key are appended at
// keys[k + pos].  A flush sorts only the next power of two >= k + buffered.
// ---------------------------------------------------------------------------
constexpr uint32_t TK_CAP = 2048;

// CAP (a power of two) bounds k + buffered offers.
template <uint32_t CAP>
struct TopKS {
    static constexpr uint32_t kCap = CAP;
    uint64_t keys[CAP];
    uint64_t thresh;
    uint32_t nbuf;
};
using TopKShared = TopKS<TK_CAP>;

template <class T>
__device__ void tk_init(T& s, uint32_t k) {
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) s.keys[i] = kKeyMax;
    if (threadIdx.x == 0) {
        s.thresh = kKeyMax;
        s.nbuf = 0;
    }
    __syncthreads();
}

template <class T>
__device__ void tk_flush(T& s, uint32_t k) {
    const uint32_t total = k + s.nbuf;
    const uint32_t P = next_pow2(total);
    for (uint32_t i = total + threadIdx.x; i < P; i += blockDim.x) s.keys[i] = kKeyMax;
    __syncthreads();
    bitonic_sort(s.keys, P);
    if (threadIdx.x == 0) {
        s.nbuf = 0;
        s.thresh = s.keys[k - 1];
    }
    __syncthreads();
}

// Called by all threads between rounds: make room for up to `r` offers.
template <class T>
__device__ void tk_reserve(T& s, uint32_t k, uint32_t r) {
    __syncthreads();
    if (k + s.nbuf + r > T::kCap) tk_flush(s, k);
}

template <class T>
__device__ __forceinline__ void tk_offer(T& s, uint32_t k, uint64_t key) {
    if (key < s.thresh) {
        uint32_t pos = atomicAdd(&s.nbuf, 1u);
        s.keys[k + pos] = key;
    }
}

}  // namespace
}  // namespace ovx
