// CDNA4 (gfx950) kernels for the ordinal retrieval hot path.
//
// Exactness rules (the engine is bit-exact with the reference CPU path):
//  * every f32 add/mul that the reference performs separately is written with
//    __fadd_rn / __fmul_rn, every _mm*_fmadd_ps with __builtin_fmaf, and the
//    file is compiled with -ffp-contract=off;
//  * summation orders follow the reference tier that the same (dim, bits)
//    selects on an AVX-512 host (src/quant.rs:300-329);
//  * selections are total orders on integer keys, so ties resolve exactly as
//    the reference's (score desc, id asc) / (hamming asc, id asc) contracts.
#include "common.h"
#include "device_common.h"

#include <hip/hip_runtime.h>

#include <atomic>

#include <mutex>

#include <algorithm>
#include <cmath>

namespace ovx {

namespace {


// ---------------------------------------------------------------------------
// Query preparation (src/util.rs:93-115, src/quant.rs:788-791)
// ---------------------------------------------------------------------------

__global__ void k_check_finite(const float* __restrict__ x, uint64_t count, uint32_t* flag) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    bool bad = false;
    for (; i < count; i += stride) bad |= !isfinite(x[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

// Sequential f32 sum of squares / of values over an LDS chunk by one thread
// (the reference folds left to right starting at -0.0).  The next 32 values
// are read from LDS while the current 32 are folded, so the fold runs at the
// add chain's latency rather than LDS latency plus the chain.
template <bool SQ>
__device__ __forceinline__ float seq_fold(const float* xs, uint32_t len, float s) {
    uint32_t i = 0;
    if (len >= 32) {
        float4 cur[8], nxt[8];
#pragma unroll
        for (int j = 0; j < 8; j++) cur[j] = *reinterpret_cast<const float4*>(xs + 4 * j);
        for (; i + 32 <= len; i += 32) {
            const uint32_t in = i + 64 <= len ? i + 32 : i;  // in bounds; unused after the last block
#pragma unroll
            for (int j = 0; j < 8; j++) nxt[j] = *reinterpret_cast<const float4*>(xs + in + 4 * j);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (SQ) {
                    s = __fadd_rn(s, __fmul_rn(cur[j].x, cur[j].x));
                    s = __fadd_rn(s, __fmul_rn(cur[j].y, cur[j].y));
                    s = __fadd_rn(s, __fmul_rn(cur[j].z, cur[j].z));
                    s = __fadd_rn(s, __fmul_rn(cur[j].w, cur[j].w));
                } else {
                    s = __fadd_rn(__fadd_rn(__fadd_rn(__fadd_rn(s, cur[j].x), cur[j].y), cur[j].z), cur[j].w);
                }
            }
#pragma unroll
            for (int j = 0; j < 8; j++) cur[j] = nxt[j];
        }
    }
    for (; i < len; i++) s = SQ ? __fadd_rn(s, __fmul_rn(xs[i], xs[i])) : __fadd_rn(s, xs[i]);
    return s;
}
__device__ __forceinline__ float seq_sum_sq(const float* xs, uint32_t len, float s) { return seq_fold<true>(xs, len, s); }
__device__ __forceinline__ float seq_sum(const float* xs, uint32_t len, float s) { return seq_fold<false>(xs, len, s); }

// One workgroup per query: q_hat = x * (1/||x||) (zeros when ||x|| <= 1e-12)
// and the SIMD tiers' TopK offset -centre * sum(q_hat) / R.  The two sums are
// sequential like the reference's; chunks are staged in LDS (xs, kPrepChunk
// floats) so the single summing thread reads on-chip memory.  qstat bit 0
// records non-finite input.
constexpr uint32_t kPrepChunk = 2048;

__device__ __forceinline__ void prepare_row(const float* __restrict__ x, uint32_t d, int tier, float centre,
                                            float inv_norm, float* __restrict__ out, float* __restrict__ offset,
                                            uint32_t* __restrict__ qstat, float* xs) {
    __shared__ float s_inv;
    __shared__ int s_zero;
    float sq = -0.0f;
    bool bad = false;
    for (uint32_t c0 = 0; c0 < d; c0 += kPrepChunk) {
        const uint32_t len = min(kPrepChunk, d - c0);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
            float v = x[c0 + i];
            bad |= !isfinite(v);
            xs[i] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) sq = seq_sum_sq(xs, len, sq);
    }
    const int any_bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) {
        float norm = __fsqrt_rn(sq);
        s_zero = norm <= 1e-12f;
        s_inv = __fdiv_rn(1.0f, norm);
        if (qstat) *qstat = any_bad ? 1u : 0u;
    }
    __syncthreads();
    const float inv = s_inv;
    const int zero = s_zero;
    float qsum = -0.0f;
    for (uint32_t c0 = 0; c0 < d; c0 += kPrepChunk) {
        const uint32_t len = min(kPrepChunk, d - c0);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
            float v = zero ? 0.0f : __fmul_rn(x[c0 + i], inv);
            out[c0 + i] = v;
            xs[i] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0 && tier != 0) qsum = seq_sum(xs, len, qsum);
    }
    if (threadIdx.x == 0) *offset = tier != 0 ? __fmul_rn(__fmul_rn(-centre, qsum), inv_norm) : 0.0f;
}

__global__ __launch_bounds__(256) void k_prepare(const float* __restrict__ q, uint64_t nq, uint32_t d, int tier,
                                                 float centre, float inv_norm, float* __restrict__ qhat,
                                                 float* __restrict__ offset, uint32_t* __restrict__ qstat) {
    __shared__ __align__(16) float xs[kPrepChunk];
    const uint64_t qi = blockIdx.x;
    prepare_row(q + qi * d, d, tier, centre, inv_norm, qhat + qi * d, offset + qi, qstat ? qstat + qi : nullptr, xs);
}

// ---------------------------------------------------------------------------
// Encode (src/sign_bitmap.rs:180-200, src/rank.rs:38-156)
// ---------------------------------------------------------------------------

// One wave per row: word w of the row is the 64-lane ballot of v[w*64+lane] > 0.
__global__ __launch_bounds__(256) void k_sign_encode(const float* __restrict__ x, uint64_t n, uint32_t d,
                                                     uint64_t* __restrict__ out, uint64_t first_doc,
                                                     int interleaved) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const uint32_t qpv = d / 64;
    const float* v = x + r * d;
    const uint64_t doc = first_doc + r;
    for (uint32_t w = 0; w < qpv; w++) {
        uint64_t mask = __ballot(v[w * 64 + lane] > 0.0f);
        if (lane == 0) {
            uint64_t pos = interleaved ? (((doc >> 6) * qpv + w) << 6) + (doc & 63) : doc * qpv + w;
            out[pos] = mask;
        }
    }
}

// One workgroup per row: LDS bitonic sort of (orderable value, index) keys,
// scatter rank -> bucket = floor(rank * 2^bits / d), pack MSB-first.
// -0.0 is canonicalised to +0.0 first: the reference compares with `==`
// before total_cmp, so the two zeros tie and fall back to index order.
// RANKS: write the u16 rank vector (Rank::add, src/rank.rs:329-350) instead.
// list != null: rows list[0 .. *list_count) only (rows the one-pass encode
// handed over).
template <bool RANKS>
__global__ __launch_bounds__(256) void k_rank_encode(const float* __restrict__ x, uint64_t n, uint32_t d,
                                                     uint32_t bits, uint8_t* __restrict__ codes,
                                                     uint16_t* __restrict__ ranks,
                                                     const uint32_t* __restrict__ list = nullptr,
                                                     const uint32_t* __restrict__ list_count = nullptr) {
    extern __shared__ __align__(16) uint64_t smem[];
    const uint32_t P = next_pow2(d);
    uint64_t* keys = smem;
    uint8_t* buck = reinterpret_cast<uint8_t*>(smem + P);
    const uint32_t cpb = RANKS ? 1 : 8 / bits, bpv = d / cpb, nb = 1u << bits;
    const uint64_t total = list ? min((uint64_t)*list_count, n) : n;
    for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint64_t r = list ? list[t] : t;
        const float* v = x + r * d;
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
            if (i < d) {
                float f = v[i];
                if (f == 0.0f) f = 0.0f;
                keys[i] = ((uint64_t)f32_ord(f) << 16) | i;
            } else {
                keys[i] = kKeyMax;
            }
        }
        __syncthreads();
        bitonic_sort(keys, P);
        if (RANKS) {
            for (uint32_t p = threadIdx.x; p < d; p += blockDim.x) ranks[r * d + (keys[p] & 0xffffu)] = (uint16_t)p;
            __syncthreads();
            continue;
        }
        for (uint32_t p = threadIdx.x; p < d; p += blockDim.x) {
            uint32_t idx = (uint32_t)(keys[p] & 0xffffu);
            buck[idx] = (uint8_t)(((uint64_t)p * nb) / d);
        }
        __syncthreads();
        uint8_t* out = codes + r * bpv;
        for (uint32_t t = threadIdx.x; t < bpv; t += blockDim.x) {
            uint32_t byte = 0;
            for (uint32_t j = 0; j < cpb; j++) byte |= (uint32_t)buck[t * cpb + j] << ((cpb - 1 - j) * bits);
            out[t] = (uint8_t)byte;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// One-pass row encode (src/rank.rs:38-156 with src/quant.rs:587-614, and
// src/sign_bitmap.rs:180-200): one wave per row, lane l holding the EPL
// consecutive coordinates l*EPL .. l*EPL+EPL-1 in registers.  In the same
// pass over the f32 row it records non-finite values, writes the sign bits
// and the RankQuant codes.
//
// The codes need each coordinate's bucket floor(rank * 2^b / d), not the
// full ranking.  The row's values are binned into d equal-width bins over
// [min, max] (the map v -> bin is monotone, and equal values share a bin),
// a histogram gives every bin's rank range [cum, cum + cnt), and a bin whose
// whole range falls in one bucket gives its elements that bucket whatever
// their order inside the bin.  Only bins whose range straddles a bucket
// boundary (at most 2^b - 1 of them) need the exact order: their elements
// are ranked inside the bin by (value, index), the reference's order
// (cmp_finite_f32_then_index, src/util.rs:19-30; -0.0 and +0.0 tie).  A row
// whose straddling bins hold more than kEncSpecial elements, or whose range
// does not bin (overflowing max - min), is appended to a list that
// k_rank_encode (the whole-row sort) finishes in a second launch, so every
// row is exact.  Each wave requests its next row before working on the
// current one.
// ---------------------------------------------------------------------------
constexpr uint32_t kEncSpecial = 128;

template <int EPL>
__device__ __forceinline__ void load_row(const float* __restrict__ p, float (&f)[EPL]) {
    if constexpr (EPL % 4 == 0) {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int i = 0; i < EPL / 4; i++) {
            const f32x4 v = *(reinterpret_cast<const f32x4*>(p) + i);
            f[4 * i] = v.x;
            f[4 * i + 1] = v.y;
            f[4 * i + 2] = v.z;
            f[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < EPL; i++) f[i] = p[i];
    }
}

// Wave-wide min / max: DPP steps inside each 16-lane row (quad swaps, half
// and full row mirrors), then the four row results read as scalars.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float a) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(a), __float_as_int(a), CTRL, 0xF, 0xF, false));
}

template <bool MAX>
__device__ __forceinline__ float wave_minmax(float v) {
    auto op = [](float a, float b) { return MAX ? fmaxf(a, b) : fminf(a, b); };
    v = op(v, dpp_f32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = op(v, dpp_f32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = op(v, dpp_f32<0x141>(v));  // row_half_mirror
    v = op(v, dpp_f32<0x140>(v));  // row_mirror
    const float a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float b = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return op(op(a, b), op(c, d));
}

// Wave minimum of a 64-bit value by DPP (both halves moved with the same
// control; no LDS round trips), then the four row results as scalars.
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t a) {
    const uint32_t lo = (uint32_t)a, hi = (uint32_t)(a >> 32);
    const uint32_t l2 = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, CTRL, 0xF, 0xF, false);
    const uint32_t h2 = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, CTRL, 0xF, 0xF, false);
    return ((uint64_t)h2 << 32) | l2;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
    v = min(v, dpp_u64<0xB1>(v));   // quad_perm [1,0,3,2]
    v = min(v, dpp_u64<0x4E>(v));   // quad_perm [2,3,0,1]
    v = min(v, dpp_u64<0x141>(v));  // row_half_mirror
    v = min(v, dpp_u64<0x140>(v));  // row_mirror
    uint64_t m = ~0ull;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 16 * r);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 16 * r);
        m = min(m, ((uint64_t)hi << 32) | lo);
    }
    return m;
}

// Exclusive prefix sum over the wave's lanes: in-row inclusive scan by DPP
// row shifts (lanes shifted in from outside the row read 0), then the row
// totals added as scalars.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t lane) {
    uint32_t x = v;
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);  // row_shr:8
    const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)x, 15);
    const uint32_t t1 = (uint32_t)__builtin_amdgcn_readlane((int)x, 31);
    const uint32_t t2 = (uint32_t)__builtin_amdgcn_readlane((int)x, 47);
    const uint32_t row = lane >> 4;
    const uint32_t base = (row > 0 ? t0 : 0u) + (row > 1 ? t1 : 0u) + (row > 2 ? t2 : 0u);
    return base + x - v;
}

// Writes NB bytes held little-endian in w[] at p.
template <int NB>
__device__ __forceinline__ void store_bytes(uint8_t* p, const uint32_t (&w)[(NB + 3) / 4]) {
    if constexpr (NB == 16) {
        *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
    } else if constexpr (NB == 8) {
        *reinterpret_cast<uint2*>(p) = make_uint2(w[0], w[1]);
    } else if constexpr (NB == 4) {
        *reinterpret_cast<uint32_t*>(p) = w[0];
    } else if constexpr (NB == 2) {
        *reinterpret_cast<uint16_t*>(p) = (uint16_t)w[0];
    } else {
#pragma unroll
        for (int i = 0; i < NB; i++) p[i] = (uint8_t)(w[i / 4] >> ((i % 4) * 8));
    }
}

// The encode workgroup is a single wave, and LDS instructions of a wave run
// in order: its lanes share LDS data with no hardware barrier, only a code
// motion barrier (so the next row's global loads stay in flight).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
}

// LDS per workgroup (one wave): the histogram and the straddling-bin list.
template <int EPL>
__host__ __device__ constexpr size_t enc_lds_bytes() {
    return (size_t)64 * EPL * 4 + kEncSpecial * (8 + 4 + 4);
}

template <int EPL, int BITS, bool CODES, bool SIGNS>
__global__ __launch_bounds__(64) void k_encode_rows(const float* __restrict__ x, uint64_t n,
                                                    uint8_t* __restrict__ codes, uint64_t* __restrict__ words,
                                                    uint64_t first_doc, uint32_t* __restrict__ flag,
                                                    uint32_t* __restrict__ fb_rows, uint32_t* __restrict__ fb_count) {
    constexpr uint32_t D = 64 * EPL, NB = 1u << BITS, CPB = 8 / BITS, LANE_BYTES = EPL * BITS / 8;
    constexpr uint32_t NW = (LANE_BYTES + 3) / 4;
    static_assert(!CODES || (EPL * BITS) % 8 == 0, "a lane's codes must fill whole bytes");
    static_assert(!SIGNS || EPL % 8 == 0, "a lane's sign bits must fill whole bytes");
    extern __shared__ __align__(16) uint8_t lds[];
    uint32_t* hist = reinterpret_cast<uint32_t*>(lds);                       // [D] counts, then cum << 16 | cnt
    uint64_t* spec_key = reinterpret_cast<uint64_t*>(lds + (size_t)D * 4);   // [kEncSpecial]
    uint32_t* spec_bin = reinterpret_cast<uint32_t*>(spec_key + kEncSpecial);  // [kEncSpecial]
    uint32_t* spec_out = spec_bin + kEncSpecial;                             // [kEncSpecial] exact buckets
    const uint32_t lane = threadIdx.x;
    bool bad = false;
    float nxt[EPL];
    if (blockIdx.x < n) load_row<EPL>(x + (uint64_t)blockIdx.x * D + (uint64_t)lane * EPL, nxt);
    for (uint64_t r = blockIdx.x; r < n; r += gridDim.x) {
        float f[EPL];
#pragma unroll
        for (int i = 0; i < EPL; i++) f[i] = nxt[i];
        if (r + gridDim.x < n) load_row<EPL>(x + (r + gridDim.x) * D + (uint64_t)lane * EPL, nxt);
#pragma unroll
        for (int i = 0; i < EPL; i++) bad |= !isfinite(f[i]);
        if constexpr (SIGNS) {
            // bit j of word j / 64 is (v_j > 0); the lane's EPL bits sit at
            // bit (lane * EPL) % 64 of word lane * EPL / 64
            uint32_t sw[(EPL / 8 + 3) / 4] = {};
#pragma unroll
            for (int i = 0; i < EPL; i++) sw[i / 32] |= (f[i] > 0.0f ? 1u : 0u) << (i % 32);
            const uint64_t doc = first_doc + r;
            const uint32_t w = lane * EPL / 64, qpv = D / 64;
            const uint64_t widx = (((doc >> 6) * qpv + w) << 6) + (doc & 63);
            store_bytes<EPL / 8>(reinterpret_cast<uint8_t*>(words + widx) + (lane * EPL % 64) / 8, sw);
        }
        if constexpr (CODES) {
            float lo = INFINITY, hi = -INFINITY;
#pragma unroll
            for (int i = 0; i < EPL; i++) {
                f[i] = __fadd_rn(f[i], 0.0f);  // -0.0 + 0.0 = +0.0: the two zeros tie
                lo = fminf(lo, f[i]);
                hi = fmaxf(hi, f[i]);
            }
            lo = wave_minmax<false>(lo);
            hi = wave_minmax<true>(hi);
            const float scale = (float)D / (hi - lo);
            const bool all_equal = !(hi > lo);  // every element equal: rank = index order
            uint32_t cw[NW] = {};
            // code of element i lands in byte i / CPB, MSB-first inside the byte
            auto put = [&](int i, uint32_t b) { cw[(i / CPB) / 4] |= b << (((i / CPB) % 4) * 8 + (CPB - 1 - i % CPB) * BITS); };
            bool listed = false;
            if (all_equal) {
#pragma unroll
                for (int i = 0; i < EPL; i++) put(i, ((lane * EPL + i) * NB) / D);
            } else if (!(isfinite(scale) && scale > 0.0f)) {
                listed = true;
            } else {
                uint32_t bin[EPL];
#pragma unroll
                for (int i = 0; i < EPL; i++) {
                    const float t = __fmul_rn(__fsub_rn(f[i], lo), scale);
                    bin[i] = t >= (float)D ? D - 1 : (t > 0.0f ? (uint32_t)t : 0u);
                }
#pragma unroll
                for (int i = 0; i < EPL; i += 4)
                    *reinterpret_cast<uint4*>(hist + lane * EPL + i) = make_uint4(0, 0, 0, 0);
                wave_lds_sync();
#pragma unroll
                for (int i = 0; i < EPL; i++) atomicAdd(&hist[bin[i]], 1u);
                wave_lds_sync();
                // exclusive prefix over bins (lane owns bins lane*EPL ..)
                uint32_t own[EPL], tot = 0;
#pragma unroll
                for (int i = 0; i < EPL; i += 4) {
                    const uint4 c = *reinterpret_cast<const uint4*>(hist + lane * EPL + i);
                    own[i] = c.x;
                    own[i + 1] = c.y;
                    own[i + 2] = c.z;
                    own[i + 3] = c.w;
                    tot += c.x + c.y + c.z + c.w;
                }
                uint32_t cum = wave_excl_scan(tot, lane);
#pragma unroll
                for (int i = 0; i < EPL; i++) {
                    const uint32_t c = own[i];
                    own[i] = (cum << 16) | c;
                    cum += c;
                }
#pragma unroll
                for (int i = 0; i < EPL; i += 4)
                    *reinterpret_cast<uint4*>(hist + lane * EPL + i) = make_uint4(own[i], own[i + 1], own[i + 2], own[i + 3]);
                wave_lds_sync();
                // a bin whose rank range stays in one bucket gives it to all its
                // elements; the others' elements are listed for exact ranking
                uint32_t nspec = 0, spec_mask = 0;
#pragma unroll
                for (int i = 0; i < EPL; i++) {
                    const uint32_t h = hist[bin[i]];
                    const uint32_t c0 = h >> 16, c1 = c0 + (h & 0xffffu) - 1;
                    const uint32_t b0 = (c0 * NB) / D, b1 = (c1 * NB) / D;
                    put(i, b0);
                    if (b0 != b1) {
                        spec_mask |= 1u << i;
                        nspec++;
                    }
                }
                const uint32_t sbase = wave_excl_scan(nspec, lane);
                const uint32_t stotal = (uint32_t)__builtin_amdgcn_readlane((int)(sbase + nspec), 63);
                if (stotal > kEncSpecial) {
                    listed = true;
                } else if (stotal > 0) {
                    uint32_t pos = sbase;
#pragma unroll
                    for (int i = 0; i < EPL; i++)
                        if (spec_mask & (1u << i)) {
                            spec_key[pos] = ((uint64_t)f32_ord(f[i]) << 16) | (lane * EPL + i);
                            spec_bin[pos] = bin[i];
                            pos++;
                        }
                    wave_lds_sync();
                    // rank inside the bin by (value, index), from the bin's first rank
                    for (uint32_t e = lane; e < stotal; e += 64) {
                        const uint64_t ke = spec_key[e];
                        const uint32_t be = spec_bin[e];
                        uint32_t below = 0;
                        for (uint32_t g = 0; g < stotal; g++)
                            below += (spec_bin[g] == be && spec_key[g] < ke) ? 1u : 0u;
                        spec_out[e] = (((hist[be] >> 16) + below) * NB) / D;
                    }
                    wave_lds_sync();
                    pos = sbase;
#pragma unroll
                    for (int i = 0; i < EPL; i++)
                        if (spec_mask & (1u << i)) {
                            const int sh = ((i / CPB) % 4) * 8 + (CPB - 1 - i % CPB) * BITS;
                            cw[(i / CPB) / 4] = (cw[(i / CPB) / 4] & ~((NB - 1) << sh)) | (spec_out[pos] << sh);
                            pos++;
                        }
                }
                wave_lds_sync();  // the histogram is reused by the next row
            }
            if (listed) {  // wave-uniform: the whole-row sort finishes this row
                if (lane == 0) fb_rows[atomicAdd(fb_count, 1u)] = (uint32_t)r;
            } else {
                store_bytes<LANE_BYTES>(codes + r * (D * BITS / 8) + lane * LANE_BYTES, cw);
            }
        }
    }
    if (__any(bad) && lane == 0) atomicOr(flag, 1u);
}

// Uniform row offsets q * row_m for the segmented sort (large k, long rows).
__global__ void k_row_offsets(uint64_t* __restrict__ offs, uint64_t nq, uint64_t row_m) {
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= nq; q += (uint64_t)gridDim.x * blockDim.x)
        offs[q] = q * row_m;
}

// Rank transform of long rows (d above the LDS sort's capacity, up to
// 65535): keys (orderable value << 16 | index) per element, a per-row sort,
// then each sorted position scattered to its element (rank[idx] = position);
// -0.0 ties +0.0 as in the one-pass kernels.
__global__ void k_rank_keys(const float* __restrict__ x, uint64_t count, uint32_t d, uint64_t* __restrict__ keys) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        keys[i] = ((uint64_t)f32_ord(__fadd_rn(x[i], 0.0f)) << 16) | (uint32_t)(i % d);
}

// ranks != null: ranks[row][idx] = position; else codes (packed MSB-first,
// bucket = position * 2^bits / d) built through an LDS row of buckets.
__global__ __launch_bounds__(256) void k_rank_scatter(const uint64_t* __restrict__ sorted, uint64_t n, uint32_t d,
                                                      uint16_t* __restrict__ ranks, uint32_t bits,
                                                      uint8_t* __restrict__ codes) {
    extern __shared__ __align__(16) uint8_t rowb[];
    for (uint64_t r = blockIdx.x; r < n; r += gridDim.x) {
        const uint64_t* sr = sorted + r * d;
        if (ranks) {
            for (uint32_t p = threadIdx.x; p < d; p += blockDim.x) ranks[r * d + (sr[p] & 0xffffu)] = (uint16_t)p;
            continue;
        }
        const uint32_t nb = 1u << bits, cpb = 8 / bits, bpv = d / cpb;
        for (uint32_t p = threadIdx.x; p < d; p += blockDim.x)
            rowb[sr[p] & 0xffffu] = (uint8_t)(((uint64_t)p * nb) / d);
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < bpv; t += blockDim.x) {
            uint32_t byte = 0;
            for (uint32_t j = 0; j < cpb; j++) byte |= (uint32_t)rowb[t * cpb + j] << ((cpb - 1 - j) * bits);
            codes[r * bpv + t] = (uint8_t)byte;
        }
        __syncthreads();
    }
}

// rankquant_eval_search (src/quant.rs:1541-1610): rank buckets of any width
// 1..8 computed on the fly.  Buckets from u16 ranks; query centres; scores
// (sum over coordinates, in order, of q_centre * centre(doc bucket)) times
// 1 / norm^2 as score keys for every document.
__global__ void k_eval_buckets(const uint16_t* __restrict__ ranks, uint64_t count, uint32_t d, uint32_t bits,
                               uint8_t* __restrict__ out, float* __restrict__ centres) {
    const float half = (float)((1u << bits) - 1) / 2.0f;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = (uint32_t)(((uint64_t)ranks[i] << bits) / d);
        if (out) out[i] = (uint8_t)b;
        if (centres) centres[i] = __fsub_rn((float)b, half);
    }
}

__global__ __launch_bounds__(256) void k_eval_keys(const uint8_t* __restrict__ buckets, uint64_t n, uint32_t d,
                                                   uint32_t bits, const float* __restrict__ qcent, float inv_norm_sq,
                                                   uint64_t* __restrict__ keys) {
    extern __shared__ __align__(16) float eval_qc[];
    float* qc = eval_qc;
    const uint64_t q = blockIdx.y;
    const float half = (float)((1u << bits) - 1) / 2.0f;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) qc[j] = qcent[q * d + j];
    __syncthreads();
    for (uint64_t doc = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; doc < n; doc += (uint64_t)gridDim.x * blockDim.x) {
        const uint8_t* row = buckets + doc * d;
        float acc = -0.0f;
        for (uint32_t j = 0; j < d; j++) acc = __fadd_rn(acc, __fmul_rn(qc[j], __fsub_rn((float)row[j], half)));
        keys[q * n + doc] = score_key(__fmul_rn(acc, inv_norm_sq), (uint32_t)doc);
    }
}

// SignBitmap::score_all / score_all_batched_flat (src/sign_bitmap.rs:
// 492-548): out[q][doc] = dim - popcount(doc XOR q), the sign agreement,
// for every document.  Grid (64-document blocks / 4, query tiles of 8); a
// wave holds one block's words (coalesced 512-byte loads) and scores them
// against the tile's queries (wave-uniform words).
__global__ __launch_bounds__(256) void k_sign_score_all(const uint64_t* __restrict__ words, uint64_t n, uint32_t qpv,
                                                        const uint64_t* __restrict__ qbits, uint32_t nq,
                                                        uint32_t* __restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t blk = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t q0 = blockIdx.y * 8, qn = min(8u, nq - q0);
    const uint64_t doc = blk * 64 + lane;
    if (blk * 64 >= n) return;
    uint32_t acc[8] = {};
    const uint64_t* base = words + blk * (uint64_t)qpv * 64 + lane;
    for (uint32_t w = 0; w < qpv; w++) {
        const uint64_t dw = base[(uint64_t)w * 64];
#pragma unroll
        for (uint32_t t = 0; t < 8; t++)
            if (t < qn) acc[t] += (uint32_t)__popcll(dw ^ qbits[(uint64_t)(q0 + t) * qpv + w]);
    }
    if (doc < n) {
#pragma unroll
        for (uint32_t t = 0; t < 8; t++)
            if (t < qn) out[(uint64_t)(q0 + t) * n + doc] = qpv * 64 - acc[t];
    }
}

// src/fastscan.rs:87-110: out[b*(d/2)*32 + p*32 + lane] = code[2p]<<2 | code[2p+1]
// from packed b=2 rows (tail lanes of the last block stay zero).
__global__ void k_fastscan_pack(const uint8_t* __restrict__ codes, uint64_t n, uint32_t d, uint8_t* __restrict__ out) {
    const uint32_t pairs = d / 2, bpv = d / 4;
    uint64_t total = n * pairs;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t doc = t / pairs;
        uint32_t p = (uint32_t)(t % pairs);
        uint8_t byte = codes[doc * bpv + p / 2];
        uint8_t nib = (p & 1) ? (byte & 0xf) : (byte >> 4);
        out[(doc / 32) * pairs * 32 + (uint64_t)p * 32 + (doc & 31)] = nib;
    }
}

// ---------------------------------------------------------------------------
// Stage 1: XOR-popcount scan + exact top-m (src/sign_bitmap.rs:231-483)
// ---------------------------------------------------------------------------

// Per-word statistic of the candidate scans.  OP 0: popcount(d XOR q), the
// sign-bitmap Hamming distance (src/util.rs:192-238).  OP 1: popcount(d AND q),
// the top-bucket bitmap overlap (src/bitmap.rs:607-1000), reported as the
// distance n_top - overlap so both kinds select ascending with ties by id.
template <int OP>
__device__ __forceinline__ uint32_t word_stat(uint64_t d, uint64_t q) {
    return (uint32_t)__popcll(OP ? (d & q) : (d ^ q));
}
template <int OP>
__device__ __forceinline__ uint32_t to_dist(uint32_t acc, uint32_t ntop) {
    return OP ? ntop - acc : acc;
}

// K1: grid (query tiles, doc segments).  Each wave scores one 64-doc block per
// step against QT queries: doc words are 512-byte coalesced loads, query
// words are wave-uniform (scalar loads).  Writes H[q][doc] (u16) and a
// per-(query, segment) Hamming histogram kept in LDS as packed 16-bit bins.
template <int QT, int WCH, int OP>
__global__ __launch_bounds__(256) void k_sign_scan(const uint64_t* __restrict__ words, uint32_t nblk, uint32_t qpv,
                                                   uint64_t n, const uint64_t* __restrict__ qbits, uint32_t nq,
                                                   uint32_t blocks_per_seg, uint64_t n_pad, uint16_t* __restrict__ H,
                                                   uint16_t* __restrict__ hist_part, uint32_t nseg, uint32_t nbins,
                                                   uint32_t hwords, uint32_t ntop) {
    extern __shared__ __align__(16) uint32_t lds_hist[];
    const uint32_t q0 = blockIdx.x * QT;
    const uint32_t tq = min((uint32_t)QT, nq - q0);
    const uint32_t seg = blockIdx.y;
    for (uint32_t i = threadIdx.x; i < QT * hwords; i += blockDim.x) lds_hist[i] = 0;
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t b_begin = seg * blocks_per_seg;
    const uint32_t b_end = min(nblk, b_begin + blocks_per_seg);
    const uint64_t* qb = qbits + (uint64_t)q0 * qpv;
    for (uint32_t b = b_begin + wave; b < b_end; b += 4) {
        uint32_t acc[QT];
#pragma unroll
        for (int i = 0; i < QT; i++) acc[i] = 0;
        const uint64_t* base = words + (uint64_t)b * qpv * 64 + lane;
        for (uint32_t w0 = 0; w0 < qpv; w0 += WCH) {
            uint64_t dw[WCH];
#pragma unroll
            for (int j = 0; j < WCH; j++) dw[j] = base[(uint64_t)(w0 + j) * 64];
#pragma unroll
            for (int i = 0; i < QT; i++) {
                if (i < (int)tq) {
                    const uint64_t* qp = qb + (uint64_t)i * qpv + w0;
#pragma unroll
                    for (int j = 0; j < WCH; j++) acc[i] += word_stat<OP>(dw[j], qp[j]);
                }
            }
        }
        const uint64_t doc = (uint64_t)b * 64 + lane;
        const bool valid = doc < n;
#pragma unroll
        for (int i = 0; i < QT; i++) {
            acc[i] = to_dist<OP>(acc[i], ntop);
            if (i < (int)tq) {
                H[(uint64_t)(q0 + i) * n_pad + doc] = (uint16_t)acc[i];
                if (valid) atomicAdd(&lds_hist[i * hwords + (acc[i] >> 1)], 1u << ((acc[i] & 1) * 16));
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tq * nbins; i += blockDim.x) {
        uint32_t qi = i / nbins, bin = i - qi * nbins;
        uint32_t w = lds_hist[qi * hwords + (bin >> 1)];
        hist_part[((uint64_t)(q0 + qi) * nseg + seg) * nbins + bin] = (uint16_t)((bin & 1) ? (w >> 16) : (w & 0xffffu));
    }
}

// K2: per query, reduce segment histograms, find h* = smallest h with
// count(hamming <= h) >= m.  thr = {h*, below = count(< h*), need, tie_total}.
__global__ __launch_bounds__(256) void k_stage1_threshold(const uint16_t* __restrict__ hist_part, uint32_t nseg,
                                                          uint32_t nbins, uint64_t m_eff, uint32_t* __restrict__ thr,
                                                          uint32_t* __restrict__ counters) {
    extern __shared__ __align__(16) uint32_t cnt[];
    __shared__ uint32_t part[256];
    const uint64_t q = blockIdx.x;
    const uint16_t* hp = hist_part + q * nseg * nbins;
    // segment sums, eight segments' loads in flight per bin
    for (uint32_t bin = threadIdx.x; bin < nbins; bin += blockDim.x) {
        uint32_t s = 0, g = 0;
        for (; g + 8 <= nseg; g += 8) {
            uint32_t v[8];
#pragma unroll
            for (uint32_t k = 0; k < 8; k++) v[k] = hp[(uint64_t)(g + k) * nbins + bin];
#pragma unroll
            for (uint32_t k = 0; k < 8; k++) s += v[k];
        }
        for (; g < nseg; g++) s += hp[(uint64_t)g * nbins + bin];
        cnt[bin] = s;
    }
    __syncthreads();
    const uint32_t per = (nbins + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = min(nbins, threadIdx.x * per), hi = min(nbins, lo + per);
    uint32_t local = 0;
    for (uint32_t b = lo; b < hi; b++) local += cnt[b];
    // exclusive prefix over the 256 threads: wave scans, then the 4 wave totals
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t wex = wave_excl_scan(local, lane);
    if (lane == 63) part[wave] = wex + local;
    __syncthreads();
    uint32_t wbase = 0;
    for (uint32_t w = 0; w < wave; w++) wbase += part[w];
    uint64_t cum = wbase + wex;
    for (uint32_t b = lo; b < hi; b++) {
        uint64_t c = cnt[b];
        if (cum < m_eff && cum + c >= m_eff) {
            thr[q * 4 + 0] = b;
            thr[q * 4 + 1] = (uint32_t)cum;
            thr[q * 4 + 2] = (uint32_t)(m_eff - cum);
            thr[q * 4 + 3] = (uint32_t)c;
        }
        cum += c;
    }
    if (threadIdx.x == 0) {
        counters[q * 2 + 0] = 0;
        counters[q * 2 + 1] = 0;
    }
}

// K3: stream H[q] in segments; keys below h* go to `below` (order fixed later
// by the sort), boundary-tier ids to `ties` (capacity kTieCap).  Appends are
// wave-aggregated over a thread's 8 elements: at most one global atomic per
// wave and list per 2048 documents (one per element slot had serialised the
// atomics' round trips).
__global__ __launch_bounds__(256) void k_stage1_compact(const uint16_t* __restrict__ H, uint64_t n, uint64_t n_pad,
                                                        uint64_t per_seg, const uint32_t* __restrict__ thr,
                                                        uint32_t* __restrict__ counters, uint64_t m_eff,
                                                        uint64_t* __restrict__ below, uint32_t* __restrict__ ties) {
    const uint64_t q = blockIdx.y;
    const uint32_t hstar = thr[q * 4 + 0];
    const uint64_t begin = (uint64_t)blockIdx.x * per_seg;
    const uint64_t end = min(n, begin + per_seg);
    const uint16_t* row = H + q * n_pad;
    uint64_t* bq = below + q * m_eff;
    uint32_t* tq = ties + q * kTieCap;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint64_t base = begin + (uint64_t)threadIdx.x * 8; base < end + 0; base += (uint64_t)blockDim.x * 8) {
        uint4 v = *reinterpret_cast<const uint4*>(row + base);
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
        uint64_t mb[8], mt[8];
        uint32_t nb = 0, nt = 0;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const uint32_t h = (w[e >> 1] >> ((e & 1) * 16)) & 0xffffu;
            const bool ok = base + e < end;
            mb[e] = __ballot(ok && h < hstar);
            mt[e] = __ballot(ok && h == hstar);
            nb += (uint32_t)__popcll(mb[e]);
            nt += (uint32_t)__popcll(mt[e]);
        }
        uint32_t pb = 0, pt = 0;  // this wave's first positions in the two lists
        if (lane == 0) {
            if (nb) pb = atomicAdd(&counters[q * 2 + 0], nb);
            if (nt) pt = atomicAdd(&counters[q * 2 + 1], nt);
        }
        pb = __shfl(pb, 0);
        pt = __shfl(pt, 0);
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const uint64_t idx = base + e;
            const uint32_t h = (w[e >> 1] >> ((e & 1) * 16)) & 0xffffu;
            if ((mb[e] >> lane) & 1) bq[pb + __popcll(mb[e] & lt_mask)] = ((uint64_t)h << 32) | idx;
            if ((mt[e] >> lane) & 1) {
                const uint32_t pos = pt + (uint32_t)__popcll(mt[e] & lt_mask);
                if (pos < kTieCap) tq[pos] = (uint32_t)idx;
            }
            pb += (uint32_t)__popcll(mb[e]);
            pt += (uint32_t)__popcll(mt[e]);
        }
    }
}

// K4: per query, pick the `need` lowest boundary-tier ids (sorted buffer, or
// an ordered scan of H when the tier overflowed kTieCap), merge with the
// below-threshold keys and sort by (hamming, id).
__global__ __launch_bounds__(256) void k_stage1_finalize(const uint16_t* __restrict__ H, uint64_t n, uint64_t n_pad,
                                                         const uint32_t* __restrict__ thr,
                                                         const uint64_t* __restrict__ below,
                                                         const uint32_t* __restrict__ ties, uint64_t m_eff,
                                                         uint32_t* __restrict__ cand, uint32_t* __restrict__ cand_h) {
    __shared__ uint64_t keys[kStage1MaxM];
    __shared__ uint32_t tid_buf[kTieCap];
    __shared__ uint32_t wave_cnt[4];
    __shared__ uint32_t s_taken;
    const uint64_t q = blockIdx.x;
    const uint32_t hstar = thr[q * 4 + 0], nbelow = thr[q * 4 + 1], need = thr[q * 4 + 2], tie_total = thr[q * 4 + 3];
    const uint64_t* bq = below + q * m_eff;
    for (uint32_t i = threadIdx.x; i < nbelow; i += blockDim.x) keys[i] = bq[i];
    if (need > 0) {
        if (tie_total <= kTieCap) {
            const uint32_t* tq = ties + q * kTieCap;
            const uint32_t P = next_pow2(tie_total);
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) tid_buf[i] = i < tie_total ? tq[i] : 0xffffffffu;
            __syncthreads();
            if (tie_total > need) bitonic_sort(tid_buf, P);
        } else {
            // Ordered scan: tie ids in ascending id order until `need` are taken.
            const uint16_t* row = H + q * n_pad;
            const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
            if (threadIdx.x == 0) s_taken = 0;
            __syncthreads();
            for (uint64_t base = 0; base < n; base += blockDim.x) {
                uint64_t idx = base + threadIdx.x;
                bool is_tie = idx < n && row[idx] == hstar;
                uint64_t mt = __ballot(is_tie);
                if (lane == 0) wave_cnt[wave] = (uint32_t)__popcll(mt);
                __syncthreads();
                uint32_t before = s_taken;
                for (uint32_t w = 0; w < wave; w++) before += wave_cnt[w];
                uint32_t ord = before + (uint32_t)__popcll(mt & lt_mask);
                if (is_tie && ord < need) tid_buf[ord] = (uint32_t)idx;
                __syncthreads();
                if (threadIdx.x == 0) s_taken += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
                __syncthreads();
                if (s_taken >= need) break;
            }
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < need; j += blockDim.x)
            keys[nbelow + j] = ((uint64_t)hstar << 32) | tid_buf[j];
    }
    const uint32_t P = next_pow2((uint32_t)m_eff);
    for (uint32_t i = (uint32_t)m_eff + threadIdx.x; i < P; i += blockDim.x) keys[i] = kKeyMax;
    __syncthreads();
    bitonic_sort(keys, P);
    for (uint32_t j = threadIdx.x; j < m_eff; j += blockDim.x) {
        cand[q * m_eff + j] = (uint32_t)(keys[j] & 0xffffffffu);
        if (cand_h) cand_h[q * m_eff + j] = (uint32_t)(keys[j] >> 32);
    }
}

// Large m (m_eff > kStage1MaxM): the exact top-m by an ordered scatter
// instead of compact + sort.  The documents with hamming == h of segment s
// take the output positions starting at
//   run(s, h) = sum_{h' < h} count(h') + sum_{s' < s} hist_part[s'][h],
// in id order, so every document with hamming <= h* and position < m_eff
// lands at its place in the (hamming asc, id asc) order with no sort.
// K5: per query, run(s, h) for h <= h* into off[q][s][h] (u32).
__global__ __launch_bounds__(256) void k_stage1_run_offsets(const uint16_t* __restrict__ hist_part, uint32_t nseg,
                                                            uint32_t nbins, const uint32_t* __restrict__ thr,
                                                            uint32_t* __restrict__ off) {
    extern __shared__ __align__(16) uint32_t start[];
    __shared__ uint32_t part[256];
    const uint64_t q = blockIdx.x;
    const uint16_t* hp = hist_part + q * nseg * nbins;
    const uint32_t hstar = thr[q * 4 + 0];
    for (uint32_t bin = threadIdx.x; bin < nbins; bin += blockDim.x) {
        uint32_t c = 0;
        if (bin <= hstar)
            for (uint32_t g = 0; g < nseg; g++) c += hp[(uint64_t)g * nbins + bin];
        start[bin] = c;
    }
    __syncthreads();
    // exclusive prefix over bins: per-thread runs, then one pass over the runs
    const uint32_t per = (nbins + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = min(nbins, threadIdx.x * per), hi = min(nbins, lo + per);
    uint32_t local = 0;
    for (uint32_t b = lo; b < hi; b++) local += start[b];
    part[threadIdx.x] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t t = 0; t < blockDim.x; t++) {
            const uint32_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    uint32_t cum = part[threadIdx.x];
    for (uint32_t b = lo; b < hi; b++) {
        const uint32_t c = start[b];
        start[b] = cum;
        cum += c;
    }
    __syncthreads();
    uint32_t* oq = off + q * (uint64_t)nseg * nbins;
    for (uint32_t bin = threadIdx.x; bin <= hstar && bin < nbins; bin += blockDim.x) {
        uint32_t run = start[bin];
        for (uint32_t g = 0; g < nseg; g++) {
            oq[(uint64_t)g * nbins + bin] = run;
            run += hp[(uint64_t)g * nbins + bin];
        }
    }
}

// K6: grid (segments, queries), one wave per workgroup walking its segment's
// documents in id order, 64 at a time.  Lanes with hamming <= h* are placed
// one distinct hamming value at a time: a ballot of the lanes sharing the
// leading lane's value gives each its order, the run counter in LDS its base.
__global__ __launch_bounds__(64) void k_stage1_scatter(const uint16_t* __restrict__ H, uint64_t n, uint64_t n_pad,
                                                       uint32_t blocks_per_seg, uint32_t nseg, uint32_t nbins,
                                                       const uint32_t* __restrict__ thr,
                                                       const uint32_t* __restrict__ off, uint64_t m_eff,
                                                       uint32_t* __restrict__ cand, uint32_t* __restrict__ cand_h) {
    extern __shared__ __align__(16) uint32_t run[];
    const uint64_t q = blockIdx.y;
    const uint32_t seg = blockIdx.x;
    const uint32_t hstar = thr[q * 4 + 0];
    const uint32_t lane = threadIdx.x;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t* oq = off + (q * nseg + seg) * (uint64_t)nbins;
    for (uint32_t b = lane; b <= hstar && b < nbins; b += 64) run[b] = oq[b];
    __syncthreads();
    const uint64_t begin = (uint64_t)seg * blocks_per_seg * 64;
    const uint64_t end = min(n, begin + (uint64_t)blocks_per_seg * 64);
    const uint16_t* row = H + q * n_pad;
    uint32_t* cq = cand + q * m_eff;
    uint32_t* hq = cand_h ? cand_h + q * m_eff : nullptr;
    for (uint64_t base = begin; base < end; base += 64) {
        const uint64_t idx = base + lane;
        const uint32_t h = idx < end ? (uint32_t)row[idx] : 0xffffffffu;
        bool active = h <= hstar;
        uint64_t mask = __ballot(active);
        while (mask) {
            const int lead = __ffsll((long long)mask) - 1;
            const uint32_t hv = __shfl(h, lead);
            const bool mine = active && h == hv;
            const uint64_t same = __ballot(mine);
            const uint32_t basepos = run[hv];
            __syncthreads();
            const uint64_t pos = (uint64_t)basepos + (uint32_t)__popcll(same & lt_mask);
            if (mine && pos < m_eff) {
                cq[pos] = (uint32_t)idx;
                if (hq) hq[pos] = h;
            }
            if (lane == (uint32_t)lead) run[hv] = basepos + (uint32_t)__popcll(same);
            __syncthreads();
            active = active && !mine;
            mask &= ~same;
        }
    }
}

// ---------------------------------------------------------------------------
// Stage 1, fast path: sampled threshold + survivor compaction.
//
// A strided sample of the corpus gives each query a Hamming threshold T
// expected to admit ~2.5m documents.  The full scan then appends every
// document with hamming <= T as a key (hamming << 32 | id).  If at least m
// keys arrived and none were dropped, the buffer holds *every* document with
// hamming <= T, hence every member of the exact top-m by (hamming, id), and a
// sort of the buffer yields the reference's result.  Otherwise the query is
// flagged and the exact histogram pipeline above recomputes it.
// ---------------------------------------------------------------------------

// Per query: sign bits (wave ballot), zero the sample histogram and the
// survivor counter, record non-finite input in qstat bit 0.
// One workgroup (4 waves) per query; wave w writes sign words w, w+4, ...
__global__ __launch_bounds__(256) void k_s1_prep(const float* __restrict__ q, uint32_t d, uint64_t* __restrict__ qbits,
                                                 uint32_t* __restrict__ hist, uint32_t nbins, uint32_t* __restrict__ cnt,
                                                 uint32_t* __restrict__ qstat) {
    const uint64_t qi = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6, qpv = d / 64;
    const float* x = q + qi * d;
    bool bad = false;
    for (uint32_t w = wave; w < qpv; w += nwave) {
        const float v = x[w * 64 + lane];
        bad |= !isfinite(v);
        const uint64_t mask = __ballot(v > 0.0f);
        if (lane == 0) qbits[qi * qpv + w] = mask;
    }
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) hist[qi * nbins + i] = 0;
    const int any_bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) {
        cnt[qi] = 0;
        if (qstat) qstat[qi] = any_bad ? 1u : 0u;
    }
}

// k_prepare and k_s1_prep in one workgroup per query (sign-bitmap stage 1
// behind the matrix-core scan, where the rerank's query prep is a launch of
// its own): the waves first write the query's sign words, zero its histogram
// and counter and record non-finite input in qstat, then prepare q_hat.
__global__ __launch_bounds__(256) void k_prepare_s1(const float* __restrict__ q, uint32_t d, int tier, float centre,
                                                    float inv_norm, float* __restrict__ qhat,
                                                    float* __restrict__ offset, uint64_t* __restrict__ qbits,
                                                    uint32_t* __restrict__ hist, uint32_t nbins,
                                                    uint32_t* __restrict__ cnt, uint32_t* __restrict__ qstat) {
    __shared__ __align__(16) float xs[kPrepChunk];
    const uint64_t qi = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6, qpv = d / 64;
    const float* x = q + qi * d;
    bool bad = false;
    for (uint32_t w = wave; w < qpv; w += nwave) {
        const float v = x[w * 64 + lane];
        bad |= !isfinite(v);
        const uint64_t mask = __ballot(v > 0.0f);
        if (lane == 0) qbits[qi * qpv + w] = mask;
    }
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) hist[qi * nbins + i] = 0;
    const int any_bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) {
        cnt[qi] = 0;
        if (qstat) qstat[qi] = any_bad ? 1u : 0u;
    }
    prepare_row(x, d, tier, centre, inv_norm, qhat + qi * d, offset + qi, nullptr, xs);
}

// Top-bucket bitmaps (src/bitmap.rs:140-187): one workgroup per row, LDS
// bitonic sort of 64-bit keys, bit j set for the chosen n_top coordinates.
//  * document (QUERY == false): rank_transform's order (value ascending, -0 ==
//    +0, ties by index ascending); bits for ranks >= d - n_top.
//  * query (QUERY == true): build_query_bitmap_fp32's stable sort by value
//    descending (ties keep index order); bits for the first n_top.
// Documents go to the interleaved block layout, queries to qbits[q][w]; the
// query form also zeroes the fast path's histogram and counter and records
// non-finite input in qstat bit 0.
template <bool QUERY>
__global__ __launch_bounds__(256) void k_topbits(const float* __restrict__ x, uint64_t n, uint32_t d, uint32_t ntop,
                                                 uint64_t* __restrict__ out, uint64_t first_doc,
                                                 uint32_t* __restrict__ hist, uint32_t nbins,
                                                 uint32_t* __restrict__ cnt, uint32_t* __restrict__ qstat) {
    extern __shared__ __align__(16) uint64_t smem[];
    const uint32_t P = next_pow2(d), qpv = d / 64;
    uint64_t* keys = smem;
    uint32_t* bits = reinterpret_cast<uint32_t*>(smem + P);
    for (uint64_t r = blockIdx.x; r < n; r += gridDim.x) {
        const float* v = x + r * d;
        bool bad = false;
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
            if (i < d) {
                float f = v[i];
                bad |= !isfinite(f);
                if (f == 0.0f) f = 0.0f;
                const uint32_t o = QUERY ? ~f32_ord(f) : f32_ord(f);
                keys[i] = ((uint64_t)o << 16) | i;
            } else {
                keys[i] = kKeyMax;
            }
        }
        for (uint32_t i = threadIdx.x; i < 2 * qpv; i += blockDim.x) bits[i] = 0;
        const int any_bad = __syncthreads_or(bad);
        bitonic_sort(keys, P);
        const uint32_t lo = QUERY ? 0 : d - ntop, hi = QUERY ? ntop : d;
        for (uint32_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
            const uint32_t idx = (uint32_t)(keys[p] & 0xffffu);
            atomicOr(&bits[idx >> 5], 1u << (idx & 31));
        }
        __syncthreads();
        for (uint32_t w = threadIdx.x; w < qpv; w += blockDim.x) {
            const uint64_t word = (uint64_t)bits[2 * w] | ((uint64_t)bits[2 * w + 1] << 32);
            if (QUERY) {
                out[r * qpv + w] = word;
            } else {
                const uint64_t doc = first_doc + r;
                out[(((doc >> 6) * qpv + w) << 6) + (doc & 63)] = word;
            }
        }
        if (QUERY) {
            for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) hist[r * nbins + i] = 0;
            if (threadIdx.x == 0) {
                if (cnt) cnt[r] = 0;
                if (qstat) qstat[r] = any_bad ? 1u : 0u;
            }
        }
        __syncthreads();
    }
}

// Bitmap subset search (src/bitmap.rs:425-505): per query, overlap of each
// listed row (duplicates scored independently), top-k by (overlap desc, row
// asc); scores are the overlaps as f32.
__global__ __launch_bounds__(256) void k_bm_subset(const uint64_t* __restrict__ words, uint32_t qpv, uint64_t n_docs,
                                                   const uint64_t* __restrict__ qbits, const uint64_t* __restrict__ offsets,
                                                   const uint32_t* __restrict__ rows, uint32_t ntop, uint32_t out_k,
                                                   float* __restrict__ out_s, int64_t* __restrict__ out_i) {
    __shared__ TopKShared tk;
    const uint64_t q = blockIdx.x;
    const uint64_t b = offsets[q], e = offsets[q + 1];
    const uint64_t* qb = qbits + q * qpv;
    tk_init(tk, out_k);
    for (uint64_t r0 = b; r0 < e; r0 += blockDim.x) {
        tk_reserve(tk, out_k, blockDim.x);
        const uint64_t j = r0 + threadIdx.x;
        const uint32_t row = j < e ? rows[j] : 0u;
        if (j < e && row < n_docs) {
            const uint64_t* base = words + ((uint64_t)(row >> 6) * qpv << 6) + (row & 63);
            uint32_t ov = 0;
            for (uint32_t w = 0; w < qpv; w++) ov += word_stat<1>(base[(uint64_t)w << 6], qb[w]);
            tk_offer(tk, out_k, ((uint64_t)(ntop - ov) << 32) | row);
        }
    }
    __syncthreads();
    tk_flush(tk, out_k);
    for (uint32_t j = threadIdx.x; j < out_k; j += blockDim.x) {
        const uint64_t key = tk.keys[j];
        out_s[q * out_k + j] = key == kKeyMax ? -INFINITY : (float)(ntop - (uint32_t)(key >> 32));
        out_i[q * out_k + j] = key == kKeyMax ? -1 : (int64_t)(key & 0xffffffffu);
    }
}

// body_overlap_scores_subset (src/bitmap.rs:425-460): out[i] = popcount(doc
// rows[i] AND q) for a caller-supplied query bitmap.
__global__ void k_bm_overlap_rows(const uint64_t* __restrict__ words, uint32_t qpv, const uint64_t* __restrict__ q,
                                  const uint32_t* __restrict__ rows, uint64_t count, uint32_t* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t row = rows[i];
        const uint64_t* base = words + ((uint64_t)(row >> 6) * qpv << 6) + (row & 63);
        uint32_t ov = 0;
        for (uint32_t w = 0; w < qpv; w++) ov += word_stat<1>(base[(uint64_t)w << 6], q[w]);
        out[i] = ov;
    }
}

// Full Bitmap search output (src/bitmap.rs:202-250): stage-1 rows of (distance,
// id) -> overlap scores as f32 and i64 ids.
__global__ void k_bm_finish(const uint32_t* __restrict__ cand, const uint32_t* __restrict__ cand_h, uint64_t count,
                            uint32_t ntop, float* __restrict__ out_s, int64_t* __restrict__ out_i) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        out_s[i] = (float)(ntop - cand_h[i]);
        out_i[i] = (int64_t)cand[i];
    }
}

// Hamming histogram of the sample blocks (every `stride`-th 64-doc block),
// LDS packed 16-bit bins per query, merged into hist with global atomics.
template <int QT, int WCH, int OP>
__global__ __launch_bounds__(256) void k_s1_sample(const uint64_t* __restrict__ words, uint32_t nblk, uint32_t qpv,
                                                   uint64_t n, const uint64_t* __restrict__ qbits, uint32_t nq,
                                                   uint32_t stride, uint32_t nsample, uint32_t blocks_per_seg,
                                                   uint32_t* __restrict__ hist, uint32_t nbins, uint32_t hwords,
                                                   uint32_t ntop) {
    extern __shared__ __align__(16) uint32_t lds_hist[];
    const uint32_t q0 = blockIdx.x * QT;
    const uint32_t tq = min((uint32_t)QT, nq - q0);
    for (uint32_t i = threadIdx.x; i < QT * hwords; i += blockDim.x) lds_hist[i] = 0;
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t j0 = blockIdx.y * blocks_per_seg, j1 = min(nsample, j0 + blocks_per_seg);
    const uint64_t* qb = qbits + (uint64_t)q0 * qpv;
    for (uint32_t j = j0 + wave; j < j1; j += 4) {
        const uint32_t b = j * stride;
        if (b >= nblk) break;
        uint32_t acc[QT];
#pragma unroll
        for (int i = 0; i < QT; i++) acc[i] = 0;
        const uint64_t* base = words + (uint64_t)b * qpv * 64 + lane;
        for (uint32_t w0 = 0; w0 < qpv; w0 += WCH) {
            uint64_t dw[WCH];
#pragma unroll
            for (int jj = 0; jj < WCH; jj++) dw[jj] = base[(uint64_t)(w0 + jj) * 64];
#pragma unroll
            for (int i = 0; i < QT; i++) {
                if (i < (int)tq) {
                    const uint64_t* qp = qb + (uint64_t)i * qpv + w0;
#pragma unroll
                    for (int jj = 0; jj < WCH; jj++) acc[i] += word_stat<OP>(dw[jj], qp[jj]);
                }
            }
        }
        if ((uint64_t)b * 64 + lane < n) {
#pragma unroll
            for (int i = 0; i < QT; i++) {
                const uint32_t h = to_dist<OP>(acc[i], ntop);
                if (i < (int)tq) atomicAdd(&lds_hist[i * hwords + (h >> 1)], 1u << ((h & 1) * 16));
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tq * nbins; i += blockDim.x) {
        uint32_t qi = i / nbins, bin = i - qi * nbins;
        uint32_t w = lds_hist[qi * hwords + (bin >> 1)];
        uint32_t c = (bin & 1) ? (w >> 16) : (w & 0xffffu);
        if (c) atomicAdd(&hist[(uint64_t)(q0 + qi) * nbins + bin], c);
    }
}

// T = smallest h with sample count(hamming <= h) >= s_target, or the largest
// Hamming value (admit everything) when s_target is ~0u or unreachable.
// The smallest distance t whose sampled count (hist[0..t]) reaches s_target,
// or nbins - 1 when the sample holds fewer (or every document is admitted).
// One wave; every lane gets the value.
__device__ __forceinline__ uint32_t sample_threshold(const uint32_t* __restrict__ h, uint32_t nbins,
                                                     uint32_t s_target, uint32_t lane) {
    if (s_target == 0xffffffffu) return nbins - 1;
    const uint32_t per = (nbins + 63) / 64;
    const uint32_t lo = min(nbins, lane * per), hi = min(nbins, lo + per);
    uint32_t local = 0;
    for (uint32_t b = lo; b < hi; b++) local += h[b];
    uint32_t incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t v = __shfl_up(incl, off);
        if ((int)lane >= off) incl += v;
    }
    const uint32_t excl = incl - local;
    const bool here = excl < s_target && incl >= s_target;  // one lane at most
    uint32_t t = nbins - 1;
    if (here) {
        uint32_t cum = excl;
        for (uint32_t b = lo; b < hi; b++) {
            cum += h[b];
            if (cum >= s_target) {
                t = b;
                break;
            }
        }
    }
    const uint64_t m = __ballot(here);
    return m ? (uint32_t)__shfl((int)t, __ffsll((long long)m) - 1) : nbins - 1;
}

__global__ __launch_bounds__(64) void k_s1_threshold(const uint32_t* __restrict__ hist, uint32_t nbins,
                                                     uint32_t s_target, uint32_t* __restrict__ thr) {
    const uint64_t q = blockIdx.x;
    const uint32_t t = sample_threshold(hist + q * nbins, nbins, s_target, threadIdx.x);
    if (threadIdx.x == 0) thr[q] = t;
}

// Stage-1 head of a one-query sign-bitmap call: the query's sign words
// (k_s1_prep: ballots of q > 0, non-finite check in qstat bit 0), the sampled
// Hamming histogram (k_s1_sample's blocks and packed LDS counting), and, in
// the workgroup that arrives last, the survivor threshold (k_s1_threshold's
// rule).  That workgroup also leaves what the next call expects at zero: the
// histogram `hist` and the arrival counter `done`, both zero when the kernel
// starts; and it clears this call's survivor counter.  Dynamic LDS:
// max(hwords, nbins) words.
template <int WCH>
__global__ __launch_bounds__(256) void k_s1_head(const uint64_t* __restrict__ words, uint32_t nblk, uint32_t qpv,
                                                 uint64_t n, const float* __restrict__ q, uint32_t stride,
                                                 uint32_t nsample, uint32_t blocks_per_seg, uint32_t* __restrict__ hist,
                                                 uint32_t nbins, uint32_t hwords, uint32_t s_target,
                                                 uint32_t* __restrict__ thr, uint32_t* __restrict__ cnt,
                                                 uint64_t* __restrict__ qbits, uint32_t* __restrict__ qstat,
                                                 uint32_t* __restrict__ done) {
    extern __shared__ __align__(16) uint32_t lds_h[];
    __shared__ uint64_t s_qb[kSignMaxDim / 64];
    __shared__ int s_last;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    bool bad = false;
    for (uint32_t w = wave; w < qpv; w += 4) {
        const float v = q[w * 64 + lane];
        bad |= !isfinite(v);
        const uint64_t m = __ballot(v > 0.0f);
        if (lane == 0) s_qb[w] = m;
    }
    for (uint32_t i = threadIdx.x; i < hwords; i += blockDim.x) lds_h[i] = 0;
    const int any_bad = __syncthreads_or(bad);
    if (blockIdx.x == 0) {
        for (uint32_t w = threadIdx.x; w < qpv; w += blockDim.x) qbits[w] = s_qb[w];
        if (threadIdx.x == 0) *qstat = any_bad ? 1u : 0u;
    }
    const uint32_t j0 = blockIdx.x * blocks_per_seg, j1 = min(nsample, j0 + blocks_per_seg);
    for (uint32_t j = j0 + wave; j < j1; j += 4) {
        const uint32_t b = j * stride;
        if (b >= nblk) break;
        const uint64_t* base = words + (uint64_t)b * qpv * 64 + lane;
        uint32_t acc = 0;
        if (WCH > 0) {
            uint64_t dw[WCH > 0 ? WCH : 1];
#pragma unroll
            for (int jj = 0; jj < WCH; jj++) dw[jj] = base[(uint64_t)jj * 64];
#pragma unroll
            for (int jj = 0; jj < WCH; jj++) acc += (uint32_t)__popcll(dw[jj] ^ s_qb[jj]);
        } else {
            for (uint32_t w = 0; w < qpv; w++) acc += (uint32_t)__popcll(base[(uint64_t)w * 64] ^ s_qb[w]);
        }
        if ((uint64_t)b * 64 + lane < n) atomicAdd(&lds_h[acc >> 1], 1u << ((acc & 1) * 16));
    }
    __syncthreads();
    // Cross-workgroup hand-off without device-scope fences (each would write
    // back the XCD's L2): the histogram and the counter change only through
    // device-scope atomics, and the barrier below waits for this
    // workgroup's atomics to complete before its arrival is counted; the
    // last workgroup then reads and clears them with device-scope atomics.
    for (uint32_t bin = threadIdx.x; bin < nbins; bin += blockDim.x) {
        const uint32_t w = lds_h[bin >> 1];
        const uint32_t c = (bin & 1) ? (w >> 16) : (w & 0xffffu);
        if (c) __hip_atomic_fetch_add(&hist[bin], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
        s_last = __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;
    for (uint32_t bin = threadIdx.x; bin < nbins; bin += blockDim.x)
        lds_h[bin] = __hip_atomic_exchange(&hist[bin], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (wave == 0) {
        const uint32_t t = sample_threshold(lds_h, nbins, s_target, lane);
        if (lane == 0) *thr = t;
    }
    if (threadIdx.x == 0) {
        *cnt = 0;
        __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Full scan: every document with hamming <= T[q] is appended (wave-aggregated
// atomics) to the query's key buffer; counts past `cap` are counted, not stored.
// Each wave holds one 64-doc block's sign words in VGPRs (WCH = words per
// doc, 512-byte coalesced loads) and loops over the tile's queries, whose
// words are wave-uniform scalar loads: 2*WCH VALU XOR + 2*WCH popcount per
// query and doc.  WCH == 0 is the generic path for word counts without a
// specialisation (doc words re-read per query from L1).
template <int WCH, int OP>
__device__ __forceinline__ uint32_t hamming_regs(const uint64_t* dw, const uint64_t* __restrict__ qp, uint32_t ntop) {
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < WCH; j++) acc += word_stat<OP>(dw[j], qp[j]);
    return to_dist<OP>(acc, ntop);
}

__device__ __forceinline__ void s1_append(bool pass, uint32_t acc, uint64_t doc, uint32_t lane, uint64_t lt_mask,
                                          uint32_t* cnt_q, uint64_t* keys_q, uint32_t cap) {
    const uint64_t mask = __ballot(pass);
    if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        uint32_t basepos = 0;
        if ((int)lane == leader) basepos = atomicAdd(cnt_q, (uint32_t)__popcll(mask));
        basepos = __shfl(basepos, leader);
        const uint32_t pos = basepos + (uint32_t)__popcll(mask & lt_mask);
        if (pass && pos < cap) keys_q[pos] = ((uint64_t)acc << 32) | doc;
    }
}

// Workgroup-list form of s1_append: the wave's passing lanes take consecutive
// entries (query slot << 56 | hamming << 32 | doc) of the workgroup's LDS list
// through one LDS atomic; the list is drained once at the end of the
// workgroup, with one global atomic per query.  (One global atomic per
// appending wave on a single query's counter serialises: at 1M documents and
// one query, about 3000 of them doubled the scan time.)
__device__ __forceinline__ void s1_append_lds(bool pass, uint32_t acc, uint64_t doc, uint32_t lane, uint64_t lt_mask,
                                              uint32_t slot, uint64_t* list, uint32_t* count) {
    const uint64_t mask = __ballot(pass);
    if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        uint32_t basepos = 0;
        if ((int)lane == leader) basepos = atomicAdd(count, (uint32_t)__popcll(mask));
        basepos = __shfl(basepos, leader);
        const uint32_t pos = basepos + (uint32_t)__popcll(mask & lt_mask);
        if (pass) list[pos] = ((uint64_t)slot << 56) | ((uint64_t)acc << 32) | doc;
    }
}

template <int WCH, int OP>
__global__ __launch_bounds__(256) void k_s1_scan(const uint64_t* __restrict__ words, uint32_t nblk, uint32_t qpv,
                                                 uint64_t n, const uint64_t* __restrict__ qbits, uint32_t nq,
                                                 uint32_t qt, uint32_t blocks_per_seg, uint32_t* __restrict__ thr,
                                                 uint32_t cap, uint64_t* __restrict__ keys, uint32_t* __restrict__ cnt,
                                                 uint32_t ntop, uint32_t wgl, PrepFuse pf, uint32_t prep_rows,
                                                 const uint32_t* __restrict__ hist, uint32_t nbins,
                                                 uint32_t s_target) {
    extern __shared__ __align__(16) uint64_t s_wl[];
    __shared__ uint32_t s_wn, s_slot[4], s_thr[4];
    // The first prep_rows rows of the grid prepare query y * gridDim.x + x
    // for the rerank (dispatched first, so they run beside the scan); the
    // dynamic LDS holds at least kPrepChunk floats when prep_rows != 0.
    if (blockIdx.y < prep_rows) {
        const uint64_t qi = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
        if (qi < nq)
            prepare_row(pf.q + qi * pf.d, pf.d, pf.tier, ((float)(1u << pf.bits) - 1.0f) / 2.0f, pf.inv_norm,
                        pf.qhat + qi * pf.d, pf.offset + qi, pf.qstat ? pf.qstat + qi : nullptr,
                        reinterpret_cast<float*>(s_wl));
        return;
    }
    const uint32_t q0 = blockIdx.x * qt;
    const uint32_t tq = min(qt, nq - q0);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t b_begin = (blockIdx.y - prep_rows) * blocks_per_seg;
    const uint32_t b_end = min(nblk, b_begin + blocks_per_seg);
    const uint64_t* qb = qbits + (uint64_t)q0 * qpv;
    const uint32_t* tb = thr + q0;
    uint32_t* cb = cnt + q0;
    uint64_t* kb = keys + (uint64_t)q0 * cap;
    // wgl != 0: survivors go to the workgroup's LDS list (wgl entries, at
    // least qt x 64 x blocks_per_seg, so it cannot overflow) and are drained
    // at the end; otherwise each appending wave takes global positions.
    // hist != null (tiles of at most 4 queries): wave w derives query w's
    // threshold from the sample histogram (k_s1_threshold's rule) instead of
    // reading thr, and the first segment's workgroup stores it there for the
    // select that follows.
    if (hist && wave < tq) {
        const uint32_t t = sample_threshold(hist + (uint64_t)(q0 + wave) * nbins, nbins, s_target, lane);
        if (lane == 0) {
            s_thr[wave] = t;
            if (blockIdx.y == prep_rows) thr[q0 + wave] = t;
        }
    }
    if (wgl) {
        if (threadIdx.x == 0) s_wn = 0;
        if (threadIdx.x < 4) s_slot[threadIdx.x] = 0;
    }
    if (wgl || hist) __syncthreads();
    uint32_t t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    if (hist) {
        t0 = s_thr[0];
        t1 = tq > 1 ? s_thr[1] : 0u;
        t2 = tq > 2 ? s_thr[2] : 0u;
        t3 = tq > 3 ? s_thr[3] : 0u;
    }
    auto thr_of = [&](uint32_t i) -> uint32_t {
        return hist ? (i == 0 ? t0 : (i == 1 ? t1 : (i == 2 ? t2 : t3))) : tb[i];
    };
    auto append = [&](bool pass, uint32_t a, uint64_t doc, uint32_t i) {
        if (wgl) s1_append_lds(pass, a, doc, lane, lt_mask, i, s_wl, &s_wn);
        else s1_append(pass, a, doc, lane, lt_mask, cb + i, kb + (uint64_t)i * cap, cap);
    };
    if (WCH > 0) {
        // Each wave's next block is requested before the current one is
        // scored (the last block requests itself again: an in-bounds load
        // whose value is not used keeps the loop free of branches).
        uint64_t dw[WCH > 0 ? WCH : 1], nx[WCH > 0 ? WCH : 1];
        if (b_begin + wave < b_end) {
            const uint64_t* base = words + (uint64_t)(b_begin + wave) * qpv * 64 + lane;
#pragma unroll
            for (int j = 0; j < WCH; j++) dw[j] = base[(uint64_t)j * 64];
        }
        for (uint32_t b = b_begin + wave; b < b_end; b += 4) {
            const uint64_t doc = (uint64_t)b * 64 + lane;
            const bool valid = doc < n;
            const uint64_t* nb = words + (uint64_t)(b + 4 < b_end ? b + 4 : b) * qpv * 64 + lane;
#pragma unroll
            for (int j = 0; j < WCH; j++) nx[j] = nb[(uint64_t)j * 64];
            uint32_t i = 0;
            // Two queries per step: their scalar loads share one wait.
            for (; i + 1 < tq; i += 2) {
                const uint32_t a0 = hamming_regs<WCH, OP>(dw, qb + (uint64_t)i * WCH, ntop);
                const uint32_t a1 = hamming_regs<WCH, OP>(dw, qb + (uint64_t)(i + 1) * WCH, ntop);
                const uint32_t u0 = thr_of(i), u1 = thr_of(i + 1);
                append(valid && a0 <= u0, a0, doc, i);
                append(valid && a1 <= u1, a1, doc, i + 1);
            }
            if (i < tq) {
                const uint32_t a0 = hamming_regs<WCH, OP>(dw, qb + (uint64_t)i * WCH, ntop);
                append(valid && a0 <= thr_of(i), a0, doc, i);
            }
#pragma unroll
            for (int j = 0; j < WCH; j++) dw[j] = nx[j];
        }
    } else {
        for (uint32_t b = b_begin + wave; b < b_end; b += 4) {
            const uint64_t* base = words + (uint64_t)b * qpv * 64 + lane;
            const uint64_t doc = (uint64_t)b * 64 + lane;
            const bool valid = doc < n;
            for (uint32_t i = 0; i < tq; i++) {
                const uint64_t* qp = qb + (uint64_t)i * qpv;
                uint32_t acc = 0;
                for (uint32_t j = 0; j < qpv; j++) acc += word_stat<OP>(base[(uint64_t)j * 64], qp[j]);
                acc = to_dist<OP>(acc, ntop);
                append(valid && acc <= thr_of(i), acc, doc, i);
            }
        }
    }
    if (wgl) {
        // Drain: count per query slot, reserve each slot's run with one
        // global atomic, hand out positions inside the runs from LDS.
        __syncthreads();
        const uint32_t total = s_wn;
        for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) atomicAdd(&s_slot[(uint32_t)(s_wl[e] >> 56)], 1u);
        __syncthreads();
        if (threadIdx.x < tq && threadIdx.x < 4) {
            const uint32_t c = s_slot[threadIdx.x];
            s_slot[threadIdx.x] = c ? atomicAdd(&cb[threadIdx.x], c) : 0u;
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
            const uint64_t x = s_wl[e];
            const uint32_t slot = (uint32_t)(x >> 56);
            const uint32_t pos = atomicAdd(&s_slot[slot], 1u);
            if (pos < cap) kb[(uint64_t)slot * cap + pos] = x & ((1ull << 56) - 1);
        }
    }
}

// Full scan on the matrix cores.  Documents expand to w_k d_k with d_k in
// {0, 1} and queries to s_k / w_k with s_k = +1 / -1 for bit 1 / 0 (sign
// bitmaps) or s_k in {0, 1} (top-bucket bitmaps), both as FP4 (e2m1); a
// 32 x 32 tile of (document, query) dot products is then
// v_mfma_scale_f32_32x32x64_f8f6f4 (FP4 x FP4, unit scales) over one 64-bit
// code word per instruction.  The weight w_k in {0.5, 1, 2} is whatever the
// bit's position in its nibble gives for free: a document dword masked with
// 0x11111111, 0x22222222 or 0x44444444 is already a vector of FP4 {0, 0.5},
// {0, 1.0} or {0, 2.0} (one AND each; only the fourth bit of each nibble,
// the sign position, needs a shift), and the query carries 1 / w_k (2.0, 1.0,
// 0.5).  Every product w_k d_k * s_k / w_k is 0 or +-1 exactly and every
// partial sum an integer below 2^24, so the f32 accumulation is exact:
//   sign:   sum_k d_k s_k = 2|d & q| - |d|,  so  popcount(d ^ q) = |q| - dot
//   bitmap: sum_k d_k q_k = |d & q|,         so  n_top - overlap = n_top - dot
// A (documents): lane l holds row l%32 and the 32 coordinates of half l/32 of
// the word as 4 dwords of 8 nibbles, coordinate 4i + j in nibble i of dword j
// (five mask-and-shift ops, no table); B (queries) uses the same assignment,
// so the k order inside the instruction never matters.  D: lane l holds
// column (query) l%32 and rows 8*(v/4) + 4*(l/32) + v%4, v = 0..15.
// A workgroup (8 waves) holds NQT query tiles of 32 whose nibbles are
// expanded once into LDS (one 16-byte fragment per lane per word); each wave
// scores 64-document blocks (two row groups), expanding each document half
// word once and reusing it for all NQT tiles.  Code dwords stream through a
// ring of RING words loaded ahead.  Survivors (distance <= T[q], i.e.
// dot >= base - T[q]) are appended as in k_s1_scan.
typedef int ovx_v8i __attribute__((ext_vector_type(8)));
typedef float ovx_v16f __attribute__((ext_vector_type(16)));

// Document side: 32 bits -> 32 weighted FP4 values, nibble i of dword j is
// bit 4i + j times w_j, (w_0, w_1, w_2, w_3) = (0.5, 1.0, 2.0, 0.5)
// (e2m1 0x1 = 0.5, 0x2 = 1.0, 0x4 = 2.0).
__device__ __forceinline__ ovx_v8i fp4_doc(uint32_t x) {
    ovx_v8i r;
    r[0] = (int)(x & 0x11111111u);
    r[1] = (int)(x & 0x22222222u);
    r[2] = (int)(x & 0x44444444u);
    r[3] = (int)((x >> 3) & 0x11111111u);
    r[4] = r[5] = r[6] = r[7] = 0;  // FP4 operands use the first four dwords
    return r;
}

// Query side, the same assignment with the inverse weights 1 / w_j = (2.0,
// 1.0, 0.5, 2.0) (e2m1 0x4, 0x2, 0x1; bit 3 of a nibble is the sign):
// PM: +1 / w_j for a set bit, -1 / w_j otherwise (sign bitmaps);
// otherwise 1 / w_j for a set bit, 0 otherwise (top-bucket bitmaps).
template <bool PM>
__device__ __forceinline__ uint4 fp4_query(uint32_t x) {
    const uint32_t mag[4] = {0x44444444u, 0x22222222u, 0x11111111u, 0x44444444u};
    uint32_t o[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t b = (x >> j) & 0x11111111u;  // bit 4i + j at bit 4i
        o[j] = PM ? (mag[j] | ((~b & 0x11111111u) << 3)) : (b * (mag[j] & 0xfu));
    }
    return make_uint4(o[0], o[1], o[2], o[3]);
}

constexpr uint32_t kMfmaThreads = 512;
constexpr int kFp4Fmt = 4;                  // e2m1 operand format
constexpr int kUnitScale = 0x7f7f7f7f;      // E8M0 127 = 2^0 in every byte

// One code word (64 coordinates) of documents col and col + 32 (this lane's
// 32-bit halves x0, x1) against every query tile.
// FIRST: the accumulators start from zero (inline constant), no clearing pass.
// pf: with two or more tiles, every tile's fragment for this slot, read
// during the previous slot; this slot reads the next slot's (slot 0 after the
// last) so that their LDS latency is covered by this slot's MFMAs.
template <int NQT, int QPV, bool FIRST>
__device__ __forceinline__ void mfma_word(uint32_t x0, uint32_t x1, int w, uint32_t lane,
                                          const uint4* __restrict__ sB, ovx_v16f (&acc)[2][NQT], uint4 (&pf)[4]) {
    // All of the word's query fragments are requested before the document
    // expansion, so the VALU work covers part of the LDS latency.
    uint4 bq[NQT];
    if constexpr (NQT >= 2) {
#pragma unroll
        for (int t = 0; t < NQT; t++) bq[t] = pf[t];
        const int wn = (w + 1) % QPV;
#pragma unroll
        for (int t = 0; t < NQT; t++) pf[t] = sB[(t * QPV + wn) * 64 + lane];
    } else {
#pragma unroll
        for (int t = 0; t < NQT; t++) bq[t] = sB[(t * QPV + w) * 64 + lane];
    }
    const ovx_v8i a0 = fp4_doc(x0), a1 = fp4_doc(x1);
#pragma unroll
    for (int t = 0; t < NQT; t++) {
        const ovx_v8i bf = {(int)bq[t].x, (int)bq[t].y, (int)bq[t].z, (int)bq[t].w, 0, 0, 0, 0};
        if (FIRST) {
            const ovx_v16f z = {};
            acc[0][t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a0, bf, z, kFp4Fmt, kFp4Fmt, 0, kUnitScale,
                                                                        0, kUnitScale);
            acc[1][t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a1, bf, z, kFp4Fmt, kFp4Fmt, 0, kUnitScale,
                                                                        0, kUnitScale);
        } else {
            acc[0][t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a0, bf, acc[0][t], kFp4Fmt, kFp4Fmt, 0,
                                                                        kUnitScale, 0, kUnitScale);
            acc[1][t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a1, bf, acc[1][t], kFp4Fmt, kFp4Fmt, 0,
                                                                        kUnitScale, 0, kUnitScale);
        }
    }
}

// Survivor lists: 320 entries per 32-query tile (1280 at four tiles) of
// (query slot << 56 | hamming << 32 | doc)
// per wave (plus one spare slot per lane), drained to the per-query buffers
// when all but 64 entries per tile are used, and at the end.  A drain takes one global atomic per query
// slot rather than one per entry: the wave counts its entries per slot in
// LDS (wcnt, QW counters), reserves that many positions with one atomicAdd
// per slot, leaves the reserved base in wcnt, and then hands out positions
// with LDS atomics.  Positions at or past cap are dropped; cnt[q] still
// counts them, which sends query q to the exact path.
template <uint32_t QW>
__device__ __forceinline__ void s1_flush_list(const uint64_t* list, uint32_t count, uint32_t lane, uint32_t q0,
                                              uint64_t n, uint32_t cap, uint64_t* __restrict__ keys,
                                              uint32_t* __restrict__ cnt, uint32_t* wcnt) {
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t s = lane; s < QW; s += 64) wcnt[s] = 0;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < count; i += 64) {
        const uint64_t e = list[i];
        if ((e & 0xffffffffull) < n) atomicAdd(&wcnt[(uint32_t)(e >> 56)], 1u);  // padding rows of the last block skip
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t c[(QW + 63) / 64], r[(QW + 63) / 64];
#pragma unroll
    for (uint32_t j = 0; j < (QW + 63) / 64; j++) {
        const uint32_t s = lane + 64 * j;
        c[j] = s < QW ? wcnt[s] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < (QW + 63) / 64; j++) r[j] = c[j] ? atomicAdd(&cnt[q0 + lane + 64 * j], c[j]) : 0u;
#pragma unroll
    for (uint32_t j = 0; j < (QW + 63) / 64; j++)
        if (lane + 64 * j < QW) wcnt[lane + 64 * j] = r[j];
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < count; i += 64) {
        const uint64_t e = list[i];
        if ((e & 0xffffffffull) >= n) continue;
        const uint32_t slot = (uint32_t)(e >> 56);
        const uint32_t pos = atomicAdd(&wcnt[slot], 1u);
        if (pos < cap) keys[(uint64_t)(q0 + slot) * cap + pos] = e & ((1ull << 56) - 1);
    }
    __builtin_amdgcn_wave_barrier();
}

template <int NQT, int OP, int QPV>
__global__ __launch_bounds__(kMfmaThreads, 1) void k_s1_scan_mfma(const uint64_t* __restrict__ words, uint32_t nblk,
                                                                  uint64_t n, const uint64_t* __restrict__ qbits,
                                                                  uint32_t nq, uint32_t blocks_per_seg,
                                                                  const uint32_t* __restrict__ thr, uint32_t cap,
                                                                  uint64_t* __restrict__ keys,
                                                                  uint32_t* __restrict__ cnt, uint32_t ntop,
                                                                  uint32_t ntiles, uint32_t nsegs) {
    constexpr uint32_t QW = 32 * NQT, NWAVE = kMfmaThreads / 64;
    constexpr int RING = QPV % 8 == 0 ? 8 : (QPV % 4 == 0 ? 4 : QPV);  // words in flight
    constexpr bool PAIR = QPV % 2 == 0;  // whole 64-bit words per lane (half the load instructions)
    extern __shared__ __align__(16) uint4 sB[];  // [NQT][QPV][64]
    __shared__ int s_base[QW];
    __shared__ uint32_t s_thr[QW];
    // Per-wave survivor lists (+ a spare slot per lane), sized by the tile
    // count so one- and two-tile workgroups leave LDS for a second workgroup
    // (four tiles: 84 KiB of lists + 64 KiB of fragments + counters fill the
    // 160 KiB).  A list is drained once it holds kListCap - kListHeadroom
    // entries, about 8 per query slot, so each drain writes runs of about 64
    // bytes per query; one block adds about 18 entries on average and never
    // more than the headroom on ordinary data (more is caught and sends the
    // tile to the exact path).
    constexpr uint32_t kListCap = 320 * NQT, kListHeadroom = 64 * NQT;
    __shared__ uint64_t s_list[NWAVE * (kListCap + 64)];
    __shared__ uint32_t s_wcnt[NWAVE * QW];  // per-wave slot counters of the list drain
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t col = lane & 31, g = lane >> 5;
    // A workgroup takes (document segment, query tile) pairs p0, p0 + pstep,
    // ... .  Workgroups are dealt to the 8 XCDs round-robin by id; with
    // nsegs % 8 == 0 the pairs of XCD x are its segments x, x + 8, ... times
    // every tile, so the tiles of a segment run on one XCD (its codes come
    // from HBM once per XCD L2, not once per tile), and when the workgroups
    // per XCD are a multiple of ntiles a workgroup keeps one tile, whose
    // query fragments it expands into LDS once for all its segments.
    const uint32_t f = blockIdx.x;
    const bool grouped = nsegs % 8 == 0;
    const uint32_t xcd = grouped ? f % 8 : 0u;
    const uint32_t p0 = grouped ? f / 8 : f, pstep = grouped ? gridDim.x / 8 : gridDim.x;
    const uint32_t pend = grouped ? ntiles * (nsegs / 8) : ntiles * nsegs;
    uint32_t cur_tile = 0xffffffffu, q0 = 0;
    uint32_t wbase = 0;  // entries in this wave's list (wave-uniform)
    uint64_t* const wlist = s_list + wave * (kListCap + 64);
    // pass <=> dot >= base - T; columns past nq never pass.
    float need[NQT];
    int basev[NQT];
    const uint32_t* words32 = reinterpret_cast<const uint32_t*>(words);
    const uint2* words2 = reinterpret_cast<const uint2*>(words);
    for (uint32_t pr = p0; pr < pend; pr += pstep) {
        const uint32_t tile = pr % ntiles;
        const uint32_t seg = grouped ? xcd + 8 * (pr / ntiles) : pr / ntiles;
        if (tile != cur_tile) {
            if (cur_tile != 0xffffffffu) {
                // the list's query slots belong to the old tile
                s1_flush_list<QW>(wlist, wbase, lane, q0, n, cap, keys, cnt, s_wcnt + wave * QW);
                wbase = 0;
                __syncthreads();  // every wave is done with the old tile's fragments
            }
            cur_tile = tile;
            q0 = tile * QW;
            for (uint32_t e = threadIdx.x; e < NQT * QPV * 64; e += blockDim.x) {
                const uint32_t ln = e & 63, w = (e >> 6) % QPV, t = (e >> 6) / QPV;
                const uint32_t q = q0 + t * 32 + (ln & 31);
                if (q < nq) {
                    // slot w: K half h = ln / 32 is dword h of word w (one word per
                    // slot), or with paired loads dword w % 2 of word 2 (w / 2) + h
                    const uint32_t word = PAIR ? 2 * (w / 2) + (ln >> 5) : w;
                    const uint32_t half = PAIR ? (w & 1) : (ln >> 5);
                    const uint32_t bits = (uint32_t)(qbits[(uint64_t)q * QPV + word] >> (32 * half));
                    sB[e] = fp4_query<!OP>(bits);
                } else {
                    sB[e] = make_uint4(0, 0, 0, 0);
                }
            }
            for (uint32_t j = threadIdx.x; j < QW; j += blockDim.x) {
                const uint32_t q = q0 + j;
                uint32_t pc = 0;
                if (q < nq && !OP)
                    for (uint32_t w = 0; w < QPV; w++) pc += (uint32_t)__popcll(qbits[(uint64_t)q * QPV + w]);
                s_base[j] = (int)(OP ? ntop : pc);
                s_thr[j] = q < nq ? thr[q] : 0u;
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < NQT; t++) {
                const uint32_t j = t * 32 + col;
                basev[t] = s_base[j];
                need[t] = q0 + j < nq ? (float)(basev[t] - (int)s_thr[j]) : 3.0e38f;
            }
        }
        uint32_t* const cnt_col = cnt + q0 + col;
        const uint32_t b_begin = seg * blocks_per_seg;
        const uint32_t b_end = min(nblk, b_begin + blocks_per_seg);
        // dword g of documents b*64 + col (+32): word w at [w * 128] (+64).  The
        // ring runs across blocks: the last RING words of a block prefetch the
        // first RING words of the wave's next block.  With paired loads (QPV
        // even) a lane loads whole words instead, word 2p + g of documents col
        // and col + 32 for pair p, feeding slots 2p (low dwords) and 2p + 1.
        uint32_t r0[RING], r1[RING];
        uint2 p0[PAIR ? RING / 2 : 1], p1[PAIR ? RING / 2 : 1];
        if (b_begin + wave < b_end) {
            if constexpr (PAIR) {
                const uint2* first = words2 + (uint64_t)(b_begin + wave) * QPV * 64 + (uint64_t)g * 64 + col;
#pragma unroll
                for (int i = 0; i < RING / 2; i++) {
                    p0[i] = first[2 * i * 64];
                    p1[i] = first[2 * i * 64 + 32];
                }
            } else {
                const uint32_t* first = words32 + ((uint64_t)(b_begin + wave) * QPV * 64 + col) * 2 + g;
#pragma unroll
                for (int i = 0; i < RING; i++) {
                    r0[i] = first[i * 128];
                    r1[i] = first[i * 128 + 64];
                }
            }
        }
        uint4 pf[4] = {};  // every tile's fragment of the next slot (mfma_word reads them one slot ahead)
        if constexpr (NQT >= 2) {
#pragma unroll
            for (int t = 0; t < NQT; t++) pf[t] = sB[t * QPV * 64 + lane];
        }
        for (uint32_t b = b_begin + wave; b < b_end; b += NWAVE) {
            ovx_v16f acc[2][NQT];
            // The query fragments do not depend on b; an opaque lane index keeps
            // the compiler from hoisting all of them out of the loop (they would
            // not fit in registers).
            uint32_t slane = lane;
            asm volatile("" : "+v"(slane));
            if constexpr (PAIR) {
                constexpr int RP = RING / 2, NP = QPV / 2;  // pairs in flight, pairs per block
                const uint2* base = words2 + (uint64_t)b * QPV * 64 + (uint64_t)g * 64 + col;
                // the wave's next block, or (after its last block) this block
                // again: an in-bounds load whose value is never used
                const uint2* next = b + NWAVE < b_end ? base + (uint64_t)NWAVE * QPV * 64 : base;
#pragma unroll
                for (int j0 = 0; j0 < NP; j0 += RP) {
#pragma unroll
                    for (int i = 0; i < RP; i++) {
                        const int p = j0 + i;
                        const uint2 u0 = p0[i], u1 = p1[i];
                        const int pn = p + RP;
                        if (pn < NP) {
                            p0[i] = base[2 * pn * 64];
                            p1[i] = base[2 * pn * 64 + 32];
                        } else {
                            p0[i] = next[2 * (pn - NP) * 64];
                            p1[i] = next[2 * (pn - NP) * 64 + 32];
                        }
                        if (p == 0) mfma_word<NQT, QPV, true>(u0.x, u1.x, 0, slane, sB, acc, pf);
                        else mfma_word<NQT, QPV, false>(u0.x, u1.x, 2 * p, slane, sB, acc, pf);
                        __builtin_amdgcn_sched_barrier(0);
                        mfma_word<NQT, QPV, false>(u0.y, u1.y, 2 * p + 1, slane, sB, acc, pf);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            } else {
                const uint32_t* base = words32 + ((uint64_t)b * QPV * 64 + col) * 2 + g;
                // the wave's next block, or (after its last block) this block again:
                // an in-bounds load whose value is never used, which keeps the
                // unrolled word loop free of branches
                const uint32_t* next = b + NWAVE < b_end ? base + (uint64_t)NWAVE * QPV * 128 : base;
#pragma unroll
                for (int w0 = 0; w0 < QPV; w0 += RING) {
#pragma unroll
                    for (int i = 0; i < RING; i++) {
                        const int w = w0 + i;
                        const uint32_t x0 = r0[i], x1 = r1[i];
                        const int wn = w + RING;
                        if (wn < QPV) {
                            r0[i] = base[wn * 128];
                            r1[i] = base[wn * 128 + 64];
                        } else {
                            r0[i] = next[(wn - QPV) * 128];
                            r1[i] = next[(wn - QPV) * 128 + 64];
                        }
                        if (w == 0) mfma_word<NQT, QPV, true>(x0, x1, 0, slane, sB, acc, pf);
                        else mfma_word<NQT, QPV, false>(x0, x1, w, slane, sB, acc, pf);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }
            // Every accumulator is consumed here (an empty asm reading one element
            // of each), so the compiler cannot sink MFMA chains into the
            // conditional code below: MFMAs read every lane's operands and must
            // run with the full exec mask.
#pragma unroll
            for (int t = 0; t < NQT; t++) asm volatile("" ::"v"(acc[0][t][0]), "v"(acc[1][t][0]));
            // Survivors: for each 16-slot row group, 16 compares give the wave
            // masks (scalar registers); each non-empty mask appends (query, key)
            // entries to this wave's LDS list at ballot offsets.  The append has no
            // branch or exec change inside: lanes that do not pass, and entries
            // past the list's capacity, write a spare slot of their own.  Rows past
            // n (last block only) are dropped when the list is drained.
            const uint32_t dbase = (uint32_t)b * 64 + 4 * g;
#pragma unroll
            for (int t = 0; t < NQT; t++) {
                const uint32_t hbase = ((uint32_t)(t * 32 + col) << 24) | (uint32_t)basev[t];
#pragma unroll
                for (int G = 0; G < 2; G++) {
                    uint64_t m[16];
#pragma unroll
                    for (int v = 0; v < 16; v++) m[v] = __ballot(acc[G][t][v] >= need[t]);
#pragma unroll
                    for (int v = 0; v < 16; v++) {
                        if (__builtin_expect(m[v] != 0, 0)) {  // out of line: the common path falls through
                            // list position = wbase + rank of this lane among the passing lanes
                            const uint32_t pos = __builtin_amdgcn_mbcnt_hi(
                                (uint32_t)(m[v] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m[v], wbase));
                            const bool p = acc[G][t][v] >= need[t];
                            const uint32_t slot = min(p ? pos : 0xffffffffu, kListCap + lane);
                            // (query slot << 24 | base) - dot = query slot << 24 | hamming
                            const uint32_t hi = hbase - (uint32_t)(int)acc[G][t][v];
                            wlist[slot] = ((uint64_t)hi << 32) | (dbase + (G * 32 + 8 * (v / 4) + (v % 4)));
                            // wave-uniform: kept in a scalar register
                            wbase = __builtin_amdgcn_readfirstlane(wbase + (uint32_t)__popcll(m[v]));
                        }
                    }
                }
            }
            if (wbase > kListCap) {  // entries were dropped: send every query of the tile to the exact path
#pragma unroll
                for (int t = 0; t < NQT; t++)
                    if (q0 + t * 32 + col < nq) atomicAdd(cnt_col + t * 32, cap + 1);
                wbase = kListCap;
            }
            if (wbase >= kListCap - kListHeadroom) {
                s1_flush_list<QW>(wlist, wbase, lane, q0, n, cap, keys, cnt, s_wcnt + wave * QW);
                wbase = 0;
            }
        }
    }
    if (cur_tile != 0xffffffffu) s1_flush_list<QW>(wlist, wbase, lane, q0, n, cap, keys, cnt, s_wcnt + wave * QW);
}

// Sampled Hamming histogram on the matrix cores (the k_s1_sample contract):
// 32 queries per workgroup, their FP4 fragments in LDS as in k_s1_scan_mfma;
// each wave scores sampled 64-document blocks and adds every distance to the
// workgroup's packed 16-bit LDS histogram.  Only bins up to the workgroup's own
// threshold (smallest h whose local count reaches s_target) are added to the
// global histogram: the global threshold can only be lower, so the bins the
// threshold pass reads are exact.
template <int OP, int QPV>
__global__ __launch_bounds__(kMfmaThreads, 1) void k_s1_sample_mfma(
    const uint64_t* __restrict__ words, uint32_t nblk, uint64_t n, const uint64_t* __restrict__ qbits, uint32_t nq,
    uint32_t stride, uint32_t nsample, uint32_t blocks_per_seg, uint32_t* __restrict__ hist, uint32_t nbins,
    uint32_t hwords, uint32_t s_target, uint32_t ntop) {
    constexpr uint32_t NWAVE = kMfmaThreads / 64;
    extern __shared__ __align__(16) uint4 sB[];  // [QPV][64] query fragments, then [32][hwords] packed bins
    uint32_t* lds_hist = reinterpret_cast<uint32_t*>(sB + QPV * 64);
    __shared__ int s_base[32];
    const uint32_t q0 = blockIdx.x * 32;
    for (uint32_t e = threadIdx.x; e < QPV * 64; e += blockDim.x) {
        const uint32_t ln = e & 63, w = e >> 6, q = q0 + (ln & 31);
        if (q < nq) {
            const uint32_t bits = (uint32_t)(qbits[(uint64_t)q * QPV + w] >> (32 * (ln >> 5)));
            sB[e] = fp4_query<!OP>(bits);
        } else {
            sB[e] = make_uint4(0, 0, 0, 0);
        }
    }
    for (uint32_t i = threadIdx.x; i < 32 * hwords; i += blockDim.x) lds_hist[i] = 0;
    for (uint32_t j = threadIdx.x; j < 32; j += blockDim.x) {
        const uint32_t q = q0 + j;
        uint32_t pc = 0;
        if (q < nq && !OP)
            for (uint32_t w = 0; w < QPV; w++) pc += (uint32_t)__popcll(qbits[(uint64_t)q * QPV + w]);
        s_base[j] = (int)(OP ? ntop : pc);
    }
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t col = lane & 31, g = lane >> 5;
    const int basev = s_base[col];
    const bool qvalid = q0 + col < nq;
    const uint32_t* words32 = reinterpret_cast<const uint32_t*>(words);
    const uint32_t j0 = blockIdx.y * blocks_per_seg, j1 = min(nsample, j0 + blocks_per_seg);
    for (uint32_t j = j0 + wave; j < j1; j += NWAVE) {
        const uint32_t b = j * stride;
        if (b >= nblk) break;
        uint32_t slane = lane;
        asm volatile("" : "+v"(slane));
        ovx_v16f acc[2][1];
        uint4 pf[4] = {};  // unused with one tile
        const uint32_t* base = words32 + ((uint64_t)b * QPV * 64 + col) * 2 + g;
        uint32_t x0[QPV], x1[QPV];
#pragma unroll
        for (int w = 0; w < QPV; w++) {
            x0[w] = base[w * 128];
            x1[w] = base[w * 128 + 64];
        }
#pragma unroll
        for (int w = 0; w < QPV; w++) {
            if (w == 0) mfma_word<1, QPV, true>(x0[w], x1[w], 0, slane, sB, acc, pf);
            else mfma_word<1, QPV, false>(x0[w], x1[w], w, slane, sB, acc, pf);
        }
        // The MFMAs read every lane's operands, so they must run with the full
        // exec mask: an empty asm reading the results keeps the compiler from
        // sinking them into the per-lane branch below.
        asm volatile("" ::"v"(acc[0][0][0]), "v"(acc[1][0][0]));
        if (qvalid) {
            const uint64_t doc0 = (uint64_t)b * 64 + 4 * g;
#pragma unroll
            for (int G = 0; G < 2; G++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const uint32_t row = G * 32 + 8 * (v / 4) + (v % 4);
                    if (doc0 + row < n) {
                        const uint32_t h = (uint32_t)(basev - (int)acc[G][0][v]);
                        atomicAdd(&lds_hist[col * hwords + (h >> 1)], 1u << ((h & 1) * 16));
                    }
                }
        }
    }
    __syncthreads();
    // One wave per query: the packed bins in chunks of 128 (one word, two
    // bins, per lane) from distance 0 up, each chunk's bins going to the
    // global histogram, until the chunk where the local count reaches
    // s_target; there only the bins up to that local threshold go.
    for (uint32_t qi = wave; qi < 32 && q0 + qi < nq; qi += NWAVE) {
        const uint32_t* hq = lds_hist + qi * hwords;
        uint32_t* hg = hist + (uint64_t)(q0 + qi) * nbins;
        uint32_t run = 0;  // local count below this chunk
        for (uint32_t w0 = 0; w0 < hwords; w0 += 64) {
            const uint32_t w = w0 + lane;
            const uint32_t word = w < hwords ? hq[w] : 0u;
            const uint32_t c0 = word & 0xffffu, c1 = word >> 16;  // bins 2w, 2w + 1
            const uint32_t ex = wave_excl_scan(c0 + c1, lane);
            const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)(ex + c0 + c1), 63);
            uint32_t last = 0xffffffffu;  // every bin of the chunk
            if (run + tot >= s_target) {
                // the local threshold is in this chunk: the first lane whose
                // count reaches s_target, then its first or second bin
                const uint64_t m = __ballot(run + ex + c0 + c1 >= s_target);
                const int l0 = (int)__builtin_ctzll(m);
                const uint32_t cut = run + ex + c0 >= s_target ? 2 * w : 2 * w + 1;
                last = (uint32_t)__shfl((int)cut, l0);
            }
            if (c0 && 2 * w <= last) atomicAdd(&hg[2 * w], c0);
            if (c1 && 2 * w + 1 <= last) atomicAdd(&hg[2 * w + 1], c1);
            if (last != 0xffffffffu) break;
            run += tot;
        }
    }
}

// Sort each query's survivors by (hamming, id) and keep the first m; flag the
// query (qstat bit 1) when the buffer cannot prove exactness.  A histogram of
// min(hamming - hmin, kSelBins-1) (hmin: the smallest survivor distance)
// first narrows the sort to the keys at or below the
// bin that reaches m (a superset of the answer that keeps every tie).
constexpr uint32_t kSelBins = 2048, kSelBuf = 4096;

__global__ __launch_bounds__(1024) void k_s1_select(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ cnt,
                                                    const uint32_t* __restrict__ thr, uint32_t idbits, uint32_t cap,
                                                    uint64_t m_eff, uint32_t* __restrict__ cand,
                                                    uint32_t* __restrict__ cand_h, uint32_t* __restrict__ qstat) {
    constexpr uint32_t KPT = kFastCap / 1024;  // keys per thread (buffer capacity / threads)
    __shared__ uint32_t hist[kSelBins];        // keys per bin
    __shared__ uint32_t start[kSelBins];       // first slot of each bin in sk
    __shared__ uint32_t fill[kSelBins];
    __shared__ uint64_t sk[kSelBuf];           // selected keys grouped by bin
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t s_cut, s_big, s_hmin;
    const uint64_t q = blockIdx.x;
    const uint32_t count = cnt[q];
    if (count < m_eff || count > cap) {
        if (threadIdx.x == 0) qstat[q] |= 2u;
        return;
    }
    const uint64_t* kq = keys + q * cap;
    // One read of the survivors into registers.
    uint64_t kr[KPT];
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++) {
        const uint32_t i = threadIdx.x + j * 1024;
        kr[j] = i < count ? kq[i] : kKeyMax;
    }
    for (uint32_t i = threadIdx.x; i < kSelBins; i += blockDim.x) {
        hist[i] = 0;
        fill[i] = 0;
    }
    // bins count from the smallest survivor distance, so that wide rows
    // (distances above kSelBins) keep distinct bins
    uint64_t kmin = kKeyMax;
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++) kmin = min(kmin, kr[j]);
    uint32_t hmin = (uint32_t)(kmin >> 32);
    for (int off = 32; off >= 1; off >>= 1) hmin = min(hmin, (uint32_t)__shfl_xor((int)hmin, off));
    if (threadIdx.x == 0) {
        s_big = 0;
        s_hmin = 0xffffffffu;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) atomicMin(&s_hmin, hmin);
    __syncthreads();
    const uint32_t hbase = count ? s_hmin : 0u;
    // bins follow the key order and split each distance by its leading id
    // bits, as far as the distance range hbase..thr[q] leaves room (s1_cut's
    // binning), so that a bin holds few keys to rank
    const uint32_t hcap = thr[q], span = hcap >= hbase ? hcap - hbase + 1 : 1u;
    uint32_t sbits = 0;
    while (sbits < idbits && (span << (sbits + 1)) <= kSelBins) sbits++;
    const uint32_t ishift = idbits - sbits;
    auto bin_of = [&](uint64_t key) {
        const uint32_t dh = (uint32_t)(key >> 32) - hbase;
        return dh >= (kSelBins >> sbits) ? kSelBins - 1 : (dh << sbits | (uint32_t)((key & 0xffffffffull) >> ishift));
    };
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++)
        if (kr[j] != kKeyMax) atomicAdd(&hist[bin_of(kr[j])], 1u);
    __syncthreads();
    // Exclusive bin starts (2 bins per thread): wave scans, then the 16 wave totals.
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lo = threadIdx.x * 2;
    const uint32_t h0 = hist[lo], h1 = hist[lo + 1], local = h0 + h1;
    uint32_t incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off);
        if ((int)lane >= off) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t w = threadIdx.x < 16 ? wsum[threadIdx.x] : 0u;
        uint32_t wi = w;
        for (int off = 1; off < 16; off <<= 1) {
            const uint32_t v = __shfl_up(wi, off);
            if ((int)threadIdx.x >= off) wi += v;
        }
        if (threadIdx.x < 16) wsum[threadIdx.x] = wi - w;  // exclusive wave offsets
    }
    __syncthreads();
    const uint32_t excl = wsum[wave] + incl - local;
    start[lo] = excl;
    start[lo + 1] = excl + h0;
    // cut = smallest bin whose cumulative count reaches m
    if (excl < m_eff && excl + local >= m_eff) s_cut = (excl + h0 >= m_eff) ? lo : lo + 1;
    __syncthreads();
    const uint32_t cut = s_cut;
    const uint32_t nsel = start[cut] + hist[cut];
    if (nsel > kSelBuf) {
        if (threadIdx.x == 0) qstat[q] |= 2u;
        return;
    }
    if (lo <= cut) atomicMax(&s_big, max(h0, lo + 1 <= cut ? h1 : 0u));
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++) {
        const uint32_t bin = bin_of(kr[j]);
        if (kr[j] != kKeyMax && bin <= cut) sk[start[bin] + atomicAdd(&fill[bin], 1u)] = kr[j];
    }
    __syncthreads();
    if (s_big <= 512) {
        // Position of a key = its bin's start + the keys of its bin that order
        // before it ((hamming, id) order; one bin holds one hamming value
        // below the clamp and ids are unique).
        for (uint32_t i = threadIdx.x; i < nsel; i += blockDim.x) {
            const uint64_t k = sk[i];
            const uint32_t bin = bin_of(k), b0 = start[bin], sz = hist[bin];
            uint32_t rank = 0;
            for (uint32_t t = 0; t < sz; t++) rank += sk[b0 + t] < k ? 1u : 0u;
            const uint32_t pos = b0 + rank;
            if (pos < m_eff) {
                cand[q * m_eff + pos] = (uint32_t)(k & 0xffffffffu);
                if (cand_h) cand_h[q * m_eff + pos] = (uint32_t)(k >> 32);
            }
        }
        return;
    }
    // Large tie groups: sort the selection.
    const uint32_t P = next_pow2(nsel);
    for (uint32_t i = nsel + threadIdx.x; i < P; i += blockDim.x) sk[i] = kKeyMax;
    __syncthreads();
    bitonic_sort(sk, P);
    for (uint32_t j = threadIdx.x; j < m_eff; j += blockDim.x) {
        cand[q * m_eff + j] = (uint32_t)(sk[j] & 0xffffffffu);
        if (cand_h) cand_h[q * m_eff + j] = (uint32_t)(sk[j] >> 32);
    }
}

// ---------------------------------------------------------------------------
// Asymmetric scoring (src/quant_kernels.rs): one thread per document, with the
// reference tier's accumulator layout held in registers.  A SIMD lane l of
// accumulator i in the reference becomes register acc[16*i + l]; since chunk c
// of 16 codes feeds accumulator c % 4, coordinate p always lands in
// acc[p % 64] (AVX-512), acc[p % 16] (AVX2 b=2) or acc[p % 8] (AVX2 b=4).
// q_hat sits in LDS and every lane reads the same coordinate at the same
// time (broadcast reads).
// ---------------------------------------------------------------------------

constexpr uint32_t kQInLds = 8192;  // largest dim whose q_hat is staged in LDS

template <int BITS>
__device__ __forceinline__ uint32_t code_of(const uint32_t* w, int j) {
    constexpr int CPB = 8 / BITS;
    const int byte_idx = j / CPB;
    const int shift = (byte_idx & 3) * 8 + (CPB - 1 - j % CPB) * BITS;
    return (w[byte_idx >> 2] >> shift) & ((1u << BITS) - 1);
}

// Scalar LUT tier (src/quant_kernels.rs:89-160): acc += q[p] * centre(code)
// in coordinate order; returns acc.
template <int BITS>
__device__ __forceinline__ float raw_scalar(const uint8_t* __restrict__ row, uint32_t d, const float* qs) {
    constexpr uint32_t CPB = 8 / BITS, MASK = (1u << BITS) - 1;
    const float half = ((float)(1u << BITS) - 1.0f) / 2.0f;
    const uint32_t bpv = d / CPB;
    float acc = 0.0f;
    uint32_t g = 0;
    if ((bpv & 3) == 0) {
        for (; g < bpv; g += 4) {
            uint32_t w = *reinterpret_cast<const uint32_t*>(row + g);
#pragma unroll
            for (int j = 0; j < (int)(4 * CPB); j++) {
                float c = __fadd_rn((float)code_of<BITS>(&w, j), -half);
                acc = __fadd_rn(acc, __fmul_rn(qs[g * CPB + j], c));
            }
        }
    } else {
        for (; g < bpv; g++) {
            uint32_t byte = row[g];
#pragma unroll
            for (int j = 0; j < (int)CPB; j++) {
                uint32_t code = (byte >> ((CPB - 1 - j) * BITS)) & MASK;
                acc = __fadd_rn(acc, __fmul_rn(qs[g * CPB + j], __fadd_rn((float)code, -half)));
            }
        }
    }
    return acc;
}

// AVX-512 tier (src/quant_kernels.rs:406-581), d % 64 == 0.
template <int BITS>
__device__ __forceinline__ float raw_avx512(const uint8_t* __restrict__ row, uint32_t d, const float* qs) {
    constexpr int NW = 2 * BITS;  // dwords per 64 codes
    float a[64];
#pragma unroll
    for (int i = 0; i < 64; i++) a[i] = 0.0f;
    const uint4* r4 = reinterpret_cast<const uint4*>(row);
    for (uint32_t c = 0; c < d / 64; c++) {
        uint32_t w[NW];
#pragma unroll
        for (int v = 0; v < NW / 4; v++) {
            uint4 x = r4[c * (NW / 4) + v];
            w[4 * v] = x.x;
            w[4 * v + 1] = x.y;
            w[4 * v + 2] = x.z;
            w[4 * v + 3] = x.w;
        }
        const float* qc = qs + c * 64;
#pragma unroll
        for (int j = 0; j < 64; j++) a[j] = __builtin_fmaf((float)code_of<BITS>(w, j), qc[j], a[j]);
    }
    float t[16];
#pragma unroll
    for (int l = 0; l < 16; l++) t[l] = __fadd_rn(__fadd_rn(a[l], a[16 + l]), __fadd_rn(a[32 + l], a[48 + l]));
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = __fadd_rn(t[j], t[j + 8]);
#pragma unroll
    for (int j = 0; j < 4; j++) t[j] = __fadd_rn(t[j], t[j + 4]);
#pragma unroll
    for (int j = 0; j < 2; j++) t[j] = __fadd_rn(t[j], t[j + 2]);
    return __fadd_rn(t[0], t[1]);
}

// One chunk of 16 codes (BITS = 2: one dword; BITS = 4: two; code l as
// code_of) multiplied into the 16 accumulators, a[l] = fma(c_l, q_l, a[l]),
// with each code read as an FP16 subnormal.  The code dword is masked so that
// each 16-bit half holds one code at bit t' <= 8 (BITS = 2) or <= 4 (BITS = 4);
// that half is the FP16 value c * 2^(t' - 24), exact, and the mixed-precision
// FMA (f32 x f16 + f32, one rounding) multiplies it by the staged
// q_l * 2^(24 - t') (a power-of-two scale, exact for |q_l| <= 1).  The
// product is c * q_l exactly, so every accumulator rounds as the plain f32
// FMA does.  qsc holds q pre-scaled by stage_q_mix.
template <int BITS>
__device__ __forceinline__ void chunk_fma_mix(const uint32_t* w, const float (&qv)[16], float (&a)[16]) {
    if (BITS == 2) {
        // code l sits at bit s = 8 (l / 4) + 2 (3 - l % 4) of the dword: half
        // s / 16, bit t = s % 16 of the half; t >= 10 is read shifted down by 10
        uint32_t m[8];
        const uint32_t x = w[0] >> 10;
#pragma unroll
        for (int t = 0; t < 16; t += 2) {
            m[t / 2] = t <= 8 ? (w[0] & (0x00030003u << t)) : (x & (0x00030003u << (t - 10)));
            asm volatile("" : "+v"(m[t / 2]));  // keeps the masks (no per-code bit-field extracts)
        }
#pragma unroll
        for (int l = 0; l < 16; l++) {
            const int sb = 8 * (l / 4) + 2 * (3 - l % 4), h = sb / 16, t = sb % 16;
            const uint32_t mm = m[t / 2];
            const uint16_t half = (uint16_t)(h ? (mm >> 16) : (mm & 0xffffu));
            a[l] = __builtin_fmaf(qv[l], (float)__builtin_bit_cast(_Float16, half), a[l]);
        }
    } else {
        // BITS = 4: code l is nibble s = 8 ((l % 8) / 2) + 4 (1 - l % 2) of
        // dword l / 8; t = s % 16 in {0, 4, 8, 12}, t >= 8 read shifted down by 8
#pragma unroll
        for (int dw = 0; dw < 2; dw++) {
            uint32_t m[4];
            const uint32_t x = w[dw] >> 8;
#pragma unroll
            for (int t = 0; t < 16; t += 4) {
                m[t / 4] = t <= 4 ? (w[dw] & (0x000f000fu << t)) : (x & (0x000f000fu << (t - 8)));
                asm volatile("" : "+v"(m[t / 4]));
            }
#pragma unroll
            for (int l8 = 0; l8 < 8; l8++) {
                const int sb = 8 * (l8 / 2) + 4 * (1 - l8 % 2), h = sb / 16, t = sb % 16;
                const uint32_t mm = m[t / 4];
                const uint16_t half = (uint16_t)(h ? (mm >> 16) : (mm & 0xffffu));
                const int l = 8 * dw + l8;
                a[l] = __builtin_fmaf(qv[l], (float)__builtin_bit_cast(_Float16, half), a[l]);
            }
        }
    }
}

// The scale chunk_fma_mix expects on coordinate i: 2^(24 - t'), t' the bit of
// its code inside the masked half (a function of i % 16 only).
template <int BITS>
__device__ __forceinline__ float mix_scale(uint32_t i) {
    const uint32_t l = i % 16;
    int tp;
    if (BITS == 2) {
        const int t = (8 * (l / 4) + 2 * (3 - l % 4)) % 16;
        tp = t <= 8 ? t : t - 10;
    } else {
        const int t = (8 * ((l % 8) / 2) + 4 * (1 - l % 2)) % 16;
        tp = t <= 4 ? t : t - 8;
    }
    return __int_as_float((127 + 24 - tp) << 23);
}

template <int BITS>
__device__ __forceinline__ void stage_q_mix(float* qs, const float* qg, uint32_t d) {
    for (uint32_t i = threadIdx.x; i < d; i += blockDim.x) qs[i] = __fmul_rn(qg[i], mix_scale<BITS>(i));
}

// x + (x of quad lane j ^ 1), then of lane j ^ 2, by DPP quad permutes (no LDS).
__device__ __forceinline__ float quad_sum(float x) {
    const float y = __fadd_rn(x, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, true)));
    return __fadd_rn(y, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(y), 0x4E, 0xF, 0xF, true)));
}

// AVX-512 tier on a quad of lanes, NC candidates per quad: lane j of the quad
// owns the reference's accumulator j (chunks c = j mod 4, 16 SIMD lanes each
// in registers) of every candidate, so (acc0+acc1)+(acc2+acc3) is two quad
// permutes and every lane of the quad ends with the same raw scores.  Each q
// value read from LDS serves all NC candidates.  BITS is 2 or 4 (the SIMD
// tiers), and qs is staged by stage_q_mix<BITS>.
template <int BITS, int NC>
__device__ __forceinline__ void raw_avx512_quad_n(const uint8_t* const (&rows)[NC], uint32_t d, const float* qs,
                                                  uint32_t j, float (&out)[NC]) {
    float a[NC][16];
#pragma unroll
    for (int r = 0; r < NC; r++)
#pragma unroll
        for (int l = 0; l < 16; l++) a[r][l] = 0.0f;
    // Lane j takes chunks j, j + 4, ... (d % 64 == 0, so every lane has d / 64
    // of them), U at a time: their code words are requested together and q is
    // read from LDS as float4, so one wait covers U chunks.
    constexpr int NW = BITS == 2 ? 1 : 2;
    constexpr uint32_t U = NC == 1 ? 8 : 4;
    const uint32_t per_lane = d / 64;
    auto load = [&](uint32_t (&w)[NW], const uint8_t* row, uint32_t c) {
        if (BITS == 2) {
            w[0] = *reinterpret_cast<const uint32_t*>(row + c * 4);
        } else {
            const uint2 x = *reinterpret_cast<const uint2*>(row + c * 8);
            w[0] = x.x;
            w[NW - 1] = x.y;
        }
    };
    auto chunk = [&](const uint32_t (&w)[NC][NW], uint32_t c) {
        const float4* q4 = reinterpret_cast<const float4*>(qs + c * 16);
        float qv[16];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float4 x = q4[v];
            qv[4 * v] = x.x;
            qv[4 * v + 1] = x.y;
            qv[4 * v + 2] = x.z;
            qv[4 * v + 3] = x.w;
        }
#pragma unroll
        for (int r = 0; r < NC; r++) chunk_fma_mix<BITS>(w[r], qv, a[r]);
    };
    uint32_t i = 0;
    if (per_lane == 16) {  // d = 1024: the whole row share in one batch of loads
        uint32_t w[16][NC][NW];
#pragma unroll
        for (uint32_t u = 0; u < 16; u++)
#pragma unroll
            for (int r = 0; r < NC; r++) load(w[u][r], rows[r], j + 4 * u);
#pragma unroll
        for (uint32_t u = 0; u < 16; u++) chunk(w[u], j + 4 * u);
        i = per_lane;
    }
    for (; i + U <= per_lane; i += U) {
        uint32_t w[U][NC][NW];
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
#pragma unroll
            for (int r = 0; r < NC; r++) load(w[u][r], rows[r], j + 4 * (i + u));
#pragma unroll
        for (uint32_t u = 0; u < U; u++) chunk(w[u], j + 4 * (i + u));
    }
    for (; i < per_lane; i++) {
        uint32_t w[NC][NW];
#pragma unroll
        for (int r = 0; r < NC; r++) load(w[r], rows[r], j + 4 * i);
        chunk(w, j + 4 * i);
    }
#pragma unroll
    for (int r = 0; r < NC; r++) {
#pragma unroll
        for (int l = 0; l < 16; l++) a[r][l] = quad_sum(a[r][l]);
#pragma unroll
        for (int l = 0; l < 8; l++) a[r][l] = __fadd_rn(a[r][l], a[r][l + 8]);
#pragma unroll
        for (int l = 0; l < 4; l++) a[r][l] = __fadd_rn(a[r][l], a[r][l + 4]);
#pragma unroll
        for (int l = 0; l < 2; l++) a[r][l] = __fadd_rn(a[r][l], a[r][l + 2]);
        out[r] = __fadd_rn(a[r][0], a[r][1]);
    }
}

template <int BITS>
__device__ __forceinline__ float raw_avx512_quad(const uint8_t* __restrict__ row, uint32_t d, const float* qs,
                                                 uint32_t j) {
    const uint8_t* rows[1] = {row};
    float out[1];
    raw_avx512_quad_n<BITS, 1>(rows, d, qs, j, out);
    return out[0];
}

// AVX2 tier (src/quant_kernels.rs:226-389): b=2 keeps 16 lanes (acc_hi codes
// 0-7, acc_lo codes 8-15 of each 4-byte chunk), b=4 keeps 8 lanes; then
// horizontal_sum_avx2's order.
template <int BITS>
__device__ __forceinline__ float raw_avx2(const uint8_t* __restrict__ row, uint32_t d, const float* qs) {
    constexpr int W = BITS == 2 ? 16 : 8;  // codes per 4-byte chunk
    float a[W];
#pragma unroll
    for (int i = 0; i < W; i++) a[i] = 0.0f;
    const uint32_t* r32 = reinterpret_cast<const uint32_t*>(row);
    for (uint32_t c = 0; c < d / W; c++) {
        uint32_t w = r32[c];
        const float* qc = qs + c * W;
#pragma unroll
        for (int j = 0; j < W; j++) a[j] = __builtin_fmaf((float)code_of<BITS>(&w, j), qc[j], a[j]);
    }
    float t[8];
#pragma unroll
    for (int l = 0; l < 8; l++) t[l] = W == 16 ? __fadd_rn(a[l], a[l + 8]) : a[l];
    float s0 = __fadd_rn(t[0], t[4]), s1 = __fadd_rn(t[1], t[5]), s2 = __fadd_rn(t[2], t[6]), s3 = __fadd_rn(t[3], t[7]);
    return __fadd_rn(__fadd_rn(s0, s1), __fadd_rn(s2, s3));
}

// Score exactly as the reference's TopK sees it: raw * (1/R) + offset.
template <int TIER, int BITS>
__device__ __forceinline__ float doc_score(const uint8_t* row, uint32_t d, const float* qs, float inv_norm, float off) {
    float raw;
    if (TIER == 2) raw = raw_avx512<BITS>(row, d, qs);
    else if (TIER == 1) raw = raw_avx2<BITS>(row, d, qs);
    else raw = raw_scalar<BITS>(row, d, qs);
    return __fadd_rn(__fmul_rn(raw, inv_norm), off);
}

__device__ __forceinline__ void stage_q(float* qs, const float* qg, uint32_t d) {
    for (uint32_t i = threadIdx.x; i < d; i += blockDim.x) qs[i] = qg[i];
}

// Stage 2 / subset rerank (src/quant.rs:1126-1281): grid (queries, splits);
// each workgroup scores one slice of a query's candidate row, one candidate
// per thread per round, tie key = global id.  With one split the workgroup
// writes the final row; otherwise it writes its slice's top-k keys to
// `partial` and k_merge_topk finishes the row.  `qmap` (optional) maps grid
// rows to query rows.  q_hat is staged in dynamic LDS (d floats).
template <int KW>
__device__ void wave_kmin(const uint64_t* a, uint32_t cnt, uint32_t k, uint64_t* out, uint32_t lane);
constexpr uint32_t kRerankSelK = 32;  // largest k the rerank selects from per-candidate key slots

template <int TIER, int BITS>
__global__ __launch_bounds__(256) void k_rerank(const uint8_t* __restrict__ codes, uint32_t d, float inv_norm,
                                                const float* __restrict__ qhat, const float* __restrict__ offset,
                                                const uint64_t* __restrict__ offsets,
                                                const uint32_t* __restrict__ cands, uint64_t row_m, uint64_t n_docs,
                                                const uint32_t* __restrict__ qmap, uint32_t out_k,
                                                float* __restrict__ out_s, int64_t* __restrict__ out_i,
                                                uint64_t* __restrict__ partial) {
    __shared__ TopKShared tk;
    extern __shared__ __align__(16) float qs[];
    const uint64_t qidx = blockIdx.x;
    const uint64_t q = qmap ? qmap[qidx] : qidx;
    const uint32_t split = blockIdx.y, nsplit = gridDim.y;
    const uint64_t b = offsets ? offsets[q] : q * row_m;
    const uint64_t e = offsets ? offsets[q + 1] : (q + 1) * row_m;
    const uint64_t per = (e - b + nsplit - 1) / nsplit;
    const uint64_t sb = min(e, b + split * per), se = min(e, sb + per);
    if (TIER == 2) stage_q_mix<BITS>(qs, qhat + q * d, d);
    else stage_q(qs, qhat + q * d, d);
    const float off = offset[q];
    const uint32_t bpv = d * BITS / 8;
    tk_init(tk, out_k);
    // Rows of at most TK_CAP candidates with k <= kRerankSelK (the two-stage
    // rerank: m = 1000, k = 10) keep every key at its candidate's slot of
    // tk.keys (no atomics, no barrier per round) and select the k smallest at
    // the end: each wave takes its quarter's k smallest (wave_kmin), one wave
    // the k smallest of those.  Same keys and order as the streaming top-k.
    const bool slots = TIER == 2 && se - sb <= TK_CAP && out_k <= kRerankSelK && blockDim.x == 256;
    if (TIER == 2) {
        // Two candidates per quad per round (quad and quad + 64), so every q
        // value read from LDS is used twice.
        constexpr uint32_t NC = 2;
        const uint32_t nquad = blockDim.x / 4, per_round = NC * nquad, quad = threadIdx.x >> 2, jl = threadIdx.x & 3;
        uint32_t id_next[NC];  // candidate ids run one round ahead
#pragma unroll
        for (uint32_t r = 0; r < NC; r++) id_next[r] = sb + quad + r * nquad < se ? cands[sb + quad + r * nquad] : 0u;
        for (uint64_t r0 = sb; r0 < se; r0 += per_round) {
            if (!slots) tk_reserve(tk, out_k, per_round);
            const uint8_t* rows[NC];
            uint32_t id[NC];
            bool ok[NC];
#pragma unroll
            for (uint32_t r = 0; r < NC; r++) {
                const uint64_t j = r0 + quad + r * nquad;
                id[r] = id_next[r];
                id_next[r] = j + per_round < se ? cands[j + per_round] : 0u;
                ok[r] = j < se && id[r] < n_docs;
                rows[r] = codes + (uint64_t)(ok[r] ? id[r] : 0u) * bpv;
            }
            float raw[NC];
            raw_avx512_quad_n<BITS, NC>(rows, d, qs, jl, raw);
            if (slots) {
#pragma unroll
                for (uint32_t r = 0; r < NC; r++) {
                    const uint64_t j = r0 + quad + r * nquad;
                    if (jl == 0 && j < se)
                        tk.keys[j - sb] = ok[r] ? score_key(__fadd_rn(__fmul_rn(raw[r], inv_norm), off), id[r]) : kKeyMax;
                }
            } else {
#pragma unroll
                for (uint32_t r = 0; r < NC; r++)
                    if (ok[r] && jl == 0)
                        tk_offer(tk, out_k, score_key(__fadd_rn(__fmul_rn(raw[r], inv_norm), off), id[r]));
            }
        }
    } else {
        for (uint64_t r0 = sb; r0 < se; r0 += blockDim.x) {
            tk_reserve(tk, out_k, blockDim.x);
            const uint64_t j = r0 + threadIdx.x;
            const uint32_t id = j < se ? cands[j] : 0u;
            if (j < se && id < n_docs) {
                float s = doc_score<TIER, BITS>(codes + (uint64_t)id * bpv, d, qs, inv_norm, off);
                tk_offer(tk, out_k, score_key(s, id));
            }
        }
    }
    __syncthreads();
    const uint64_t* best = tk.keys;  // the k smallest keys, ascending
    if (slots) {
        // four waves' k smallest (packed: wave w at w * out_k), then the result
        __shared__ uint64_t s_sel[5 * kRerankSelK];
        const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const uint32_t count = (uint32_t)(se - sb), quarter = (count + 3) / 4;  // <= 512 keys per wave
        const uint32_t w0 = min(count, wave * quarter), w1 = min(count, w0 + quarter);
        wave_kmin<8>(tk.keys + w0, w1 - w0, out_k, s_sel + wave * out_k, lane);
        __syncthreads();
        if (wave == 0) wave_kmin<2>(s_sel, 4 * out_k, out_k, s_sel + 4 * kRerankSelK, lane);
        __syncthreads();
        best = s_sel + 4 * kRerankSelK;
    } else {
        tk_flush(tk, out_k);
    }
    if (nsplit == 1) {
        float* os = out_s + q * out_k;
        int64_t* oi = out_i + q * out_k;
        for (uint32_t j = threadIdx.x; j < out_k; j += blockDim.x) {
            uint64_t key = best[j];
            os[j] = key == kKeyMax ? -INFINITY : key_score(key);
            oi[j] = key == kKeyMax ? -1 : (int64_t)(key & 0xffffffffu);
        }
    } else {
        uint64_t* pp = partial + (qidx * nsplit + split) * out_k;
        for (uint32_t j = threadIdx.x; j < out_k; j += blockDim.x) pp[j] = best[j];
    }
}

// Exclusive prefix sum over a 256-thread workgroup (wtot: 4 LDS words);
// *total (optional) gets the sum.  Synchronises.
__device__ uint32_t block_excl_scan256(uint32_t v, uint32_t* wtot, uint32_t* total = nullptr) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t incl = wave_excl_scan(v, lane) + v;
    __syncthreads();  // wtot may still be read by an earlier use
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    uint32_t base = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; w++) {
        const uint32_t t = wtot[w];
        base += w < wave ? t : 0u;
        all += t;
    }
    if (total) *total = all;
    return base + incl - v;
}

// Wave-level selection helpers (one wave, no barriers).
// The k smallest of cnt keys a[0..cnt) (cnt <= 64 * KW), ascending, into
// out[0..k) (kKeyMax past the last key): k rounds of a wave minimum, one
// occurrence removed per round.
template <int KW>
__device__ void wave_kmin(const uint64_t* a, uint32_t cnt, uint32_t k, uint64_t* out, uint32_t lane) {
    uint64_t r[KW];
#pragma unroll
    for (int j = 0; j < KW; j++) {
        const uint32_t i = lane + 64u * j;
        r[j] = i < cnt ? a[i] : kKeyMax;
    }
    for (uint32_t t = 0; t < k; t++) {
        uint64_t lm = r[0];
#pragma unroll
        for (int j = 1; j < KW; j++) lm = min(lm, r[j]);
        const uint64_t m = wave_min_u64(lm);
        const uint64_t own = __ballot(lm == m);
        if ((int)lane == __ffsll((long long)own) - 1) {
            bool done_one = false;
#pragma unroll
            for (int j = 0; j < KW; j++)
                if (!done_one && r[j] == m) {
                    r[j] = kKeyMax;
                    done_one = true;
                }
        }
        if (lane == 0) out[t] = m;
    }
}

// The top-m of one query's stage-1 survivors as a set (k_s1_select's rule,
// without its order), for a 256-thread workgroup whose threads hold the keys
// (hamming << 32 | id) in registers: thread t's slot j holds survivor
// t + 256 j, kKeyMax past the count (count >= m_eff >= 1); slots from
// `nslot` (wave-uniform) on hold no survivor and are not read; ids are below
// 2^idbits and distances at most hcap (the scan's survivor threshold).  The
// keys are histogrammed over kSelBins bins that follow the key order: bin =
// (hamming - hmin) << s | id >> (idbits - s), clamped to kSelBins - 1, with
// hmin the smallest survivor distance and s the most id bits for which the
// distance range hmin..hcap still fits (so that a bin
// holds one distance and a slice of the ids, and ties on one distance spread
// over 2^s bins).  The cut bin is where the cumulative count reaches m_eff;
// when it holds more keys than the top-m needs, its take-th smallest key is
// the one whose rank among the bin's keys (gathered in tie_buf, each thread
// counting the smaller ones) is take - 1.  A key is in the top-m iff its bin
// is below the cut, or it is in the cut bin and not above that key: the
// exact top-m by (hamming, id), as ids are unique.  Returns false (the query
// takes the exact path) when the cut bin holds more than kTailTieCap keys.
// Synchronises; every thread gets the same result.
constexpr uint32_t kTailTieCap = 2048, kTailWaveK = 32;

// Bits of the largest document id of an n-document index (1 to 32).
__host__ __device__ inline uint32_t id_bits(uint64_t n) {
    uint32_t b = 1;
    while (b < 32 && (n - 1) >> b) b++;
    return b;
}

struct S1CutShared {
    uint32_t bh[kSelBins];
    uint32_t wsum[4];
    uint32_t cut, before, hmin, ntie, nout;
    uint64_t cutkey;
};

struct S1Cut {
    uint32_t hbase, sbits, ishift, bin;
    uint64_t key;
    __device__ uint32_t bin_of(uint64_t k) const {
        const uint32_t dh = (uint32_t)(k >> 32) - hbase;
        return dh >= (kSelBins >> sbits) ? kSelBins - 1 : (dh << sbits | (uint32_t)((k & 0xffffffffull) >> ishift));
    }
    __device__ bool has(uint64_t k) const {
        const uint32_t b = bin_of(k);
        return k != kKeyMax && (b < bin || (b == bin && k <= key));
    }
};

// Appends this wave's lanes with `in` set to list[*count ...] (one LDS
// atomic per wave); returns the lane's slot (valid where `in`).
__device__ __forceinline__ uint32_t wave_append(bool in, uint32_t* count, uint32_t lane) {
    const uint64_t m = __ballot(in);
    uint32_t base = 0;
    if (lane == 0 && m) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, 0);
    return base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <uint32_t KPT>
__device__ bool s1_cut(const uint64_t (&kr)[KPT], uint32_t nslot, uint32_t m_eff, uint32_t idbits, uint32_t hcap,
                       S1CutShared& sc, uint64_t* tie_buf, S1Cut& out) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < kSelBins; i += blockDim.x) sc.bh[i] = 0;
    uint64_t kmin = kKeyMax;
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++)
        if (j < nslot) kmin = min(kmin, kr[j]);
    const uint32_t hmin = (uint32_t)(wave_min_u64(kmin) >> 32);
    if (threadIdx.x == 0) {
        sc.cutkey = kKeyMax;
        sc.hmin = 0xffffffffu;
        sc.ntie = 0;
        sc.nout = 0;
    }
    __syncthreads();
    if (lane == 0) atomicMin(&sc.hmin, hmin);
    __syncthreads();
    out.hbase = sc.hmin;  // count >= m_eff >= 1: some key is valid
    // id bits per bin: the most with (hcap - hmin + 1) << s <= kSelBins
    const uint32_t span = hcap >= out.hbase ? hcap - out.hbase + 1 : 1u;
    uint32_t sb = 0;
    while (sb < idbits && (span << (sb + 1)) <= kSelBins) sb++;
    out.sbits = sb;
    out.ishift = idbits - sb;
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++)
        if (j < nslot && kr[j] != kKeyMax) atomicAdd(&sc.bh[out.bin_of(kr[j])], 1u);
    __syncthreads();
    // exclusive bin starts: 8 bins per thread, wave scans, then 4 wave totals
    constexpr uint32_t BPT = kSelBins / 256;
    uint32_t c[BPT], local = 0;
#pragma unroll
    for (uint32_t b = 0; b < BPT; b++) {
        c[b] = sc.bh[threadIdx.x * BPT + b];
        local += c[b];
    }
    const uint32_t incl = wave_excl_scan(local, lane) + local;
    if (lane == 63) sc.wsum[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0;
    for (uint32_t w = 0; w < wave; w++) wbase += sc.wsum[w];
    uint32_t run = wbase + incl - local;
    if (run < m_eff && run + local >= m_eff) {
#pragma unroll
        for (uint32_t b = 0; b < BPT; b++) {
            if (run < m_eff && run + c[b] >= m_eff) {
                sc.cut = threadIdx.x * BPT + b;
                sc.before = run;
            }
            run += c[b];
        }
    }
    __syncthreads();
    out.bin = sc.cut;
    const uint32_t take = m_eff - sc.before, ncut = sc.bh[out.bin];
    if (take < ncut) {
        // the cut bin is split: its keys to LDS, and the take-th smallest
        if (ncut > kTailTieCap) return false;
#pragma unroll
        for (uint32_t j = 0; j < KPT; j++)
            if (j < nslot) {
                const bool in = kr[j] != kKeyMax && out.bin_of(kr[j]) == out.bin;
                const uint32_t pos = wave_append(in, &sc.ntie, lane);
                if (in) tie_buf[pos] = kr[j];
            }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < ncut; i += blockDim.x) {
            const uint64_t k = tie_buf[i];
            uint32_t rank = 0, t = 0;
            for (; t + 4 <= ncut; t += 4)  // every lane reads the same words
                rank += (tie_buf[t] < k) + (tie_buf[t + 1] < k) + (tie_buf[t + 2] < k) + (tie_buf[t + 3] < k);
            for (; t < ncut; t++) rank += tie_buf[t] < k;
            if (rank == take - 1) sc.cutkey = k;
        }
        __syncthreads();
    }
    out.key = sc.cutkey;
    return true;
}

// Stage-1 select when the caller needs the top-m as a set (the two-stage
// rerank, whose top-k breaks ties by document id, not by candidate
// position): one 256-thread workgroup per query, s1_cut's rule, candidates
// written in no particular order (each wave appends its keys as one run).
// Queries whose survivors cannot prove exactness (fewer than m_eff, more
// than cap, or a cut bin above kTailTieCap keys) get qstat bit 1, as from
// k_s1_select.  Document ids are below 2^idbits.
__global__ __launch_bounds__(256) void k_s1_pick(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ cnt,
                                                 const uint32_t* __restrict__ thr, uint32_t cap, uint32_t m_eff,
                                                 uint32_t idbits, uint32_t* __restrict__ cand,
                                                 uint32_t* __restrict__ qstat) {
    constexpr uint32_t KPT = kFastCap / 256;
    __shared__ S1CutShared sc;
    __shared__ uint64_t tie_buf[kTailTieCap];
    const uint64_t q = blockIdx.x;
    const uint32_t count = cnt[q];
    if (count < m_eff || count > cap) {
        if (threadIdx.x == 0) qstat[q] |= 2u;
        return;
    }
    const uint32_t nslot = (count + 255) / 256, lane = threadIdx.x & 63;
    const uint64_t* kq = keys + q * cap;
    uint64_t kr[KPT];
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++) {
        const uint32_t i = threadIdx.x + j * 256;
        kr[j] = i < count ? kq[i] : kKeyMax;
    }
    S1Cut cut;
    if (!s1_cut<KPT>(kr, nslot, m_eff, idbits, thr[q], sc, tie_buf, cut)) {
        if (threadIdx.x == 0) qstat[q] |= 2u;
        return;
    }
    // exactly m_eff keys qualify: each wave reserves its run, then writes it
    uint32_t inm = 0, n = 0;  // inm bit j: slot j is in the top-m
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++)
        if (j < nslot && cut.has(kr[j])) inm |= 1u << j;
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++)
        if (j < nslot) n += (uint32_t)__popcll(__ballot((inm >> j) & 1u));
    uint32_t base = 0;
    if (lane == 0 && n) base = atomicAdd(&sc.nout, n);
    base = __shfl(base, 0);
    uint32_t* cq = cand + q * m_eff;
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++)
        if (j < nslot) {
            const bool in = (inm >> j) & 1u;
            const uint64_t m = __ballot(in);
            const uint32_t pos =
                base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (in && pos < m_eff) cq[pos] = (uint32_t)(kr[j] & 0xffffffffu);
            base += (uint32_t)__popcll(m);
        }
}

// Stage-1 select, stage-2 rerank and the top-k merge of a small two-stage
// call in one launch: grid (nsplit, nq), 256 threads.
//  * select (s1_cut): every workgroup of query q reads the survivor keys
//    into registers and finds the top-m set by (hamming, id).
//  * rerank (k_rerank's scoring): workgroup `split` scores the candidates
//    held by threads t with t mod nsplit == split, and keeps their top-k.
//  * merge (k_merge_topk): the workgroup that arrives last at query q's
//    counter done[q] (zero at the start, reset by it) merges the nsplit
//    partial top-k rows and writes the output row.
// Small k (out_k <= kTailWaveK and at most 256 keys to choose from) is
// chosen by one wave (wave_kmin); larger ones by the block top-k.  Queries
// whose survivors cannot prove exactness (fewer than m_eff, more than cap,
// or a cut bin above kTailTieCap keys) get qstat bit 1 and no output, as
// from k_s1_select.  The last workgroup writes the query's final status word
// to qstat and, when hstat is not null, to that host-visible word.  Partial
// rows and counters pass between workgroups through device-scope atomics
// (no device-scope fence, which would write back the XCD's L2).  Dynamic
// LDS: s1_tail_lds(d) bytes (cut-bin keys, then q_hat).

template <int TIER, int BITS>
__global__ __launch_bounds__(256) void k_s1_tail(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ cnt,
                                                 const uint32_t* __restrict__ thr, uint32_t cap, uint32_t m_eff, const uint8_t* __restrict__ codes,
                                                 uint32_t d, float inv_norm, const float* __restrict__ qhat,
                                                 const float* __restrict__ offset, uint64_t n_docs, uint32_t out_k,
                                                 float* __restrict__ out_s, int64_t* __restrict__ out_i,
                                                 uint64_t* __restrict__ partial, uint32_t* __restrict__ done,
                                                 uint32_t* __restrict__ qstat, uint32_t* __restrict__ hstat) {
    constexpr uint32_t KPT = kFastCap / 256;
    __shared__ TopKShared tk;  // block top-k (large k); otherwise the scored keys
    __shared__ S1CutShared sc;
    __shared__ uint32_t s_list[kStage1MaxM];
    __shared__ uint64_t s_res[kTailWaveK];
    __shared__ uint32_t s_nlist;
    __shared__ int s_last;
    extern __shared__ __align__(16) uint64_t tail_dyn[];  // [kTailTieCap] cut-bin keys, then q_hat (d floats)
    float* qs = reinterpret_cast<float*>(tail_dyn + kTailTieCap);
    const uint32_t split = blockIdx.x, nsplit = gridDim.x;
    const uint64_t q = blockIdx.y;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // One round trip to memory: the count, every slot of the survivor
    // buffer (cap entries; those at or past the count are masked below) and
    // q_hat are requested together.
    const uint32_t count = cnt[q], hcap = thr[q];
    const uint64_t* kq = keys + q * cap;
    uint64_t kr[KPT];
#pragma unroll
    for (uint32_t j = 0; j < KPT; j++) kr[j] = kq[threadIdx.x + j * 256];
    if (TIER == 2) stage_q_mix<BITS>(qs, qhat + q * d, d);
    else stage_q(qs, qhat + q * d, d);
    bool fb = count < m_eff || count > cap;
    if (!fb) {
#pragma unroll
        for (uint32_t j = 0; j < KPT; j++)
            if (threadIdx.x + j * 256 >= count) kr[j] = kKeyMax;
        const uint32_t nslot = min(KPT, (count + 255) / 256);
        S1Cut cut;
        fb = !s1_cut<KPT>(kr, nslot, m_eff, id_bits(n_docs), hcap, sc, tail_dyn, cut);
        if (!fb) {
            const bool take_mine = threadIdx.x % nsplit == split;
            auto in = [&](uint64_t key) { return take_mine && cut.has(key); };
            uint32_t inm = 0;  // bit j: slot j is a candidate of this workgroup
#pragma unroll
            for (uint32_t j = 0; j < KPT; j++)
                if (j < nslot && in(kr[j])) inm |= 1u << j;
            uint32_t total = 0;
            uint32_t pos = block_excl_scan256((uint32_t)__popc(inm), sc.wsum, &total);
#pragma unroll
            for (uint32_t j = 0; j < KPT; j++)
                if ((inm >> j) & 1u) s_list[pos++] = (uint32_t)(kr[j] & 0xffffffffu);
            if (threadIdx.x == 0) s_nlist = total;
        }
        __syncthreads();  // s_list and s_nlist complete (also orders qs)
    }
    const bool wave_k = out_k <= kTailWaveK;
    uint64_t* sk = tk.keys;  // scored keys of this workgroup (wave-selected path)
    if (!fb) {
        const float off = offset[q];
        const uint32_t bpv = d * BITS / 8;
        const uint32_t nl = s_nlist;
        const bool small = wave_k && nl <= 256;
        if (!small) tk_init(tk, out_k);
        if (TIER == 2) {
            constexpr uint32_t NC = 2;
            const uint32_t nquad = blockDim.x / 4, per_round = NC * nquad, quad = threadIdx.x >> 2, jl = threadIdx.x & 3;
            for (uint32_t r0 = 0; r0 < nl; r0 += per_round) {
                if (!small) tk_reserve(tk, out_k, per_round);
                const uint8_t* rows[NC];
                uint32_t id[NC];
                bool ok[NC];
#pragma unroll
                for (uint32_t r = 0; r < NC; r++) {
                    const uint32_t j = r0 + quad + r * nquad;
                    id[r] = j < nl ? s_list[j] : 0u;
                    ok[r] = j < nl && id[r] < n_docs;
                    rows[r] = codes + (uint64_t)(ok[r] ? id[r] : 0u) * bpv;
                }
                float raw[NC];
                raw_avx512_quad_n<BITS, NC>(rows, d, qs, jl, raw);
#pragma unroll
                for (uint32_t r = 0; r < NC; r++) {
                    const uint32_t j = r0 + quad + r * nquad;
                    const uint64_t key =
                        ok[r] ? score_key(__fadd_rn(__fmul_rn(raw[r], inv_norm), off), id[r]) : kKeyMax;
                    if (jl == 0 && j < nl) {
                        if (small) sk[j] = key;
                        else if (ok[r]) tk_offer(tk, out_k, key);
                    }
                }
            }
        } else {
            for (uint32_t r0 = 0; r0 < nl; r0 += blockDim.x) {
                if (!small) tk_reserve(tk, out_k, blockDim.x);
                const uint32_t j = r0 + threadIdx.x;
                const uint32_t id = j < nl ? s_list[j] : 0u;
                if (j < nl) {
                    const uint64_t key =
                        id < n_docs
                            ? score_key(doc_score<TIER, BITS>(codes + (uint64_t)id * bpv, d, qs, inv_norm, off), id)
                            : kKeyMax;
                    if (small) sk[j] = key;
                    else if (id < n_docs) tk_offer(tk, out_k, key);
                }
            }
        }
        __syncthreads();
        if (small) {
            if (wave == 0) wave_kmin<4>(sk, nl, out_k, s_res, lane);
            __syncthreads();
        } else {
            tk_flush(tk, out_k);
        }
        const uint64_t* best = small ? s_res : tk.keys;
        if (nsplit > 1) {
            // device-scope stores: written through to where the last
            // workgroup, on any XCD, reads them
            uint64_t* pp = partial + (q * nsplit + split) * out_k;
            for (uint32_t j = threadIdx.x; j < out_k; j += blockDim.x)
                __hip_atomic_store(&pp[j], best[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            for (uint32_t j = threadIdx.x; j < out_k; j += blockDim.x) {
                const uint64_t key = best[j];
                out_s[q * out_k + j] = key == kKeyMax ? -INFINITY : key_score(key);
                out_i[q * out_k + j] = key == kKeyMax ? -1 : (int64_t)(key & 0xffffffffu);
            }
        }
    }
    // arrival (after this workgroup's stores completed); the last workgroup
    // of the query merges the partial rows
    if (nsplit > 1) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0)
            s_last = __hip_atomic_fetch_add(&done[q], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nsplit - 1;
        __syncthreads();
        if (!s_last) return;
        if (!fb) {
            const uint64_t* in = partial + q * nsplit * out_k;
            const uint32_t total = nsplit * out_k;
            if (wave_k && total <= 1024) {
                for (uint32_t j = threadIdx.x; j < total; j += blockDim.x)
                    sk[j] = __hip_atomic_load(&in[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __syncthreads();
                if (wave == 0) {
                    if (total <= 256) wave_kmin<4>(sk, total, out_k, s_res, lane);
                    else wave_kmin<16>(sk, total, out_k, s_res, lane);
                }
                __syncthreads();
            } else {
                tk_init(tk, out_k);
                for (uint32_t r0 = 0; r0 < total; r0 += blockDim.x) {
                    tk_reserve(tk, out_k, blockDim.x);
                    const uint32_t j = r0 + threadIdx.x;
                    if (j < total) {
                        const uint64_t key = __hip_atomic_load(&in[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (key != kKeyMax) tk_offer(tk, out_k, key);
                    }
                }
                __syncthreads();
                tk_flush(tk, out_k);
            }
            const uint64_t* best = (wave_k && total <= 1024) ? s_res : tk.keys;
            for (uint32_t j = threadIdx.x; j < out_k; j += blockDim.x) {
                const uint64_t key = best[j];
                out_s[q * out_k + j] = key == kKeyMax ? -INFINITY : key_score(key);
                out_i[q * out_k + j] = key == kKeyMax ? -1 : (int64_t)(key & 0xffffffffu);
            }
        }
        if (threadIdx.x == 0) __hip_atomic_store(&done[q], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) {
        const uint32_t st = qstat[q] | (fb ? 2u : 0u);
        qstat[q] = st;
        if (hstat) __hip_atomic_store(&hstat[q], st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Full asymmetric scan (src/quant.rs:713-847): grid (segments, queries), one
// document per thread per round; each workgroup keeps its segment's top-k and
// k_merge_topk combines the segments.
template <int TIER, int BITS>
__global__ __launch_bounds__(256) void k_fullscan_rq(const uint8_t* __restrict__ codes, uint64_t n, uint32_t d,
                                                     float inv_norm, const float* __restrict__ qhat,
                                                     const float* __restrict__ offset, uint64_t per_seg, uint32_t k,
                                                     uint64_t* __restrict__ partial) {
    __shared__ TopKShared tk;
    __shared__ float qs[kQInLds];
    const uint64_t q = blockIdx.y, seg = blockIdx.x, nseg = gridDim.x;
    const uint64_t begin = seg * per_seg, end = min(n, begin + per_seg);
    stage_q(qs, qhat + q * d, d);
    const float off = offset[q];
    const uint32_t bpv = d * BITS / 8;
    tk_init(tk, k);
    for (uint64_t r0 = begin; r0 < end; r0 += blockDim.x) {
        tk_reserve(tk, k, blockDim.x);
        const uint64_t di = r0 + threadIdx.x;
        if (di < end) tk_offer(tk, k, score_key(doc_score<TIER, BITS>(codes + di * bpv, d, qs, inv_norm, off), (uint32_t)di));
    }
    __syncthreads();
    tk_flush(tk, k);
    uint64_t* out = partial + (q * nseg + seg) * k;
    for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) out[j] = tk.keys[j];
}

// ---------------------------------------------------------------------------
// Symmetric scoring and the Rank index (src/quant.rs:629-695, src/rank.rs:291-471)
// ---------------------------------------------------------------------------

// centre(bucket(rank)) for the symmetric RankQuant search's query side.
__global__ void k_rank_centres(const uint16_t* __restrict__ ranks, uint64_t count, uint32_t d, uint32_t bits,
                               float* __restrict__ out) {
    const float half = ((float)(1u << bits) - 1.0f) / 2.0f;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = (uint32_t)(((uint64_t)ranks[i] << bits) / d);
        out[i] = __fadd_rn((float)b, -half);
    }
}

// Rank rows are u16 [n][d]; VEC loads 8 ranks at a time (d % 8 == 0).
template <bool VEC>
__device__ __forceinline__ void rank_row8(const uint16_t* __restrict__ row, uint32_t p, uint32_t* r) {
    if (VEC) {
        const uint4 v = *reinterpret_cast<const uint4*>(row + p);
        r[0] = v.x & 0xffffu; r[1] = v.x >> 16; r[2] = v.y & 0xffffu; r[3] = v.y >> 16;
        r[4] = v.z & 0xffffu; r[5] = v.z >> 16; r[6] = v.w & 0xffffu; r[7] = v.w >> 16;
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = row[p + j];
    }
}

// Rank::search: acc = sum (2 q_p - (d-1)) (2 r_p - (d-1)) exactly (i32 while
// d^3 fits, else i64), score = (f32)acc * 0.25 * inv_norm_sq.  Grid
// (segments, queries), one document per thread per round, segment top-k.
template <bool VEC, bool WIDE>
__global__ __launch_bounds__(256) void k_rank_sym(const uint16_t* __restrict__ ranks, uint64_t n, uint32_t d,
                                                  const uint16_t* __restrict__ qranks, float inv_norm_sq,
                                                  uint64_t per_seg, uint32_t k, uint64_t* __restrict__ partial) {
    __shared__ TopKShared tk;
    extern __shared__ __align__(16) int32_t qc[];
    const uint64_t q = blockIdx.y, seg = blockIdx.x, nseg = gridDim.x;
    const int32_t m2 = (int32_t)d - 1;
    for (uint32_t p = threadIdx.x; p < d; p += blockDim.x) qc[p] = 2 * (int32_t)qranks[q * d + p] - m2;
    const uint64_t begin = seg * per_seg, end = min(n, begin + per_seg);
    tk_init(tk, k);
    for (uint64_t r0 = begin; r0 < end; r0 += blockDim.x) {
        tk_reserve(tk, k, blockDim.x);
        const uint64_t di = r0 + threadIdx.x;
        if (di < end) {
            const uint16_t* row = ranks + di * d;
            int64_t acc64 = 0;
            int32_t acc32 = 0;
            uint32_t p = 0;
            for (; p + 8 <= d; p += 8) {
                uint32_t r[8];
                rank_row8<VEC>(row, p, r);
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int32_t dc = 2 * (int32_t)r[j] - m2;
                    if (WIDE) acc64 += (int64_t)qc[p + j] * dc;
                    else acc32 += qc[p + j] * dc;
                }
            }
            for (; p < d; p++) {
                const int32_t dc = 2 * (int32_t)row[p] - m2;
                if (WIDE) acc64 += (int64_t)qc[p] * dc;
                else acc32 += qc[p] * dc;
            }
            const float a = WIDE ? (float)acc64 : (float)acc32;
            const float sc = __fadd_rn(__fmul_rn(__fmul_rn(a, 0.25f), inv_norm_sq), 0.0f);
            tk_offer(tk, k, score_key(sc, (uint32_t)di));
        }
    }
    __syncthreads();
    tk_flush(tk, k);
    uint64_t* out = partial + (q * nseg + seg) * k;
    for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) out[j] = tk.keys[j];
}

// Rank::search_asymmetric: acc = sum q_unit[p] * (f32)r_p in coordinate order,
// score = (acc - mean * q_sum) * inv_norm (+ 0.0, the TopK offset).
template <bool VEC>
__global__ __launch_bounds__(256) void k_rank_asym(const uint16_t* __restrict__ ranks, uint64_t n, uint32_t d,
                                                   const float* __restrict__ qhat, const float* __restrict__ qsum,
                                                   float inv_norm, float mean, uint64_t per_seg, uint32_t k,
                                                   uint64_t* __restrict__ partial) {
    __shared__ TopKShared tk;
    extern __shared__ __align__(16) float qsh[];
    const uint64_t q = blockIdx.y, seg = blockIdx.x, nseg = gridDim.x;
    for (uint32_t p = threadIdx.x; p < d; p += blockDim.x) qsh[p] = qhat[q * d + p];
    const float shift = __fmul_rn(mean, qsum[q]);
    const uint64_t begin = seg * per_seg, end = min(n, begin + per_seg);
    tk_init(tk, k);
    for (uint64_t r0 = begin; r0 < end; r0 += blockDim.x) {
        tk_reserve(tk, k, blockDim.x);
        const uint64_t di = r0 + threadIdx.x;
        if (di < end) {
            const uint16_t* row = ranks + di * d;
            float acc = 0.0f;
            uint32_t p = 0;
            for (; p + 8 <= d; p += 8) {
                uint32_t r[8];
                rank_row8<VEC>(row, p, r);
#pragma unroll
                for (int j = 0; j < 8; j++) acc = __fadd_rn(acc, __fmul_rn(qsh[p + j], (float)r[j]));
            }
            for (; p < d; p++) acc = __fadd_rn(acc, __fmul_rn(qsh[p], (float)row[p]));
            const float sc = __fadd_rn(__fmul_rn(__fadd_rn(acc, -shift), inv_norm), 0.0f);
            tk_offer(tk, k, score_key(sc, (uint32_t)di));
        }
    }
    __syncthreads();
    tk_flush(tk, k);
    uint64_t* out = partial + (q * nseg + seg) * k;
    for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) out[j] = tk.keys[j];
}

// ---------------------------------------------------------------------------
// Device-side validation of loaded files (src/rank_io.rs:600-640, 764-816,
// 938-951): each kernel records the lowest invalid row with atomicMin; the
// host then rebuilds the reference's message from that one row.
// ---------------------------------------------------------------------------

// Zero nibbles of x, one bit (bit 3 of the nibble) each.
__device__ __forceinline__ uint32_t zero_nibbles(uint32_t x) {
    return ~(((x & 0x77777777u) + 0x77777777u) | x) & 0x88888888u;
}

// RankQuant rows: every bucket must hold exactly d / 2^bits codes.  One thread
// per row, the row read as dwords (bpv % 4 == 0 for every valid (d, bits)).
__global__ void k_check_rq_rows(const uint8_t* __restrict__ codes, uint64_t n, uint32_t d, uint32_t bits,
                                unsigned long long* __restrict__ first_bad) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t bpv = d * bits / 8, per = d >> bits;
    const uint32_t* row = reinterpret_cast<const uint32_t*>(codes + r * bpv);
    bool bad = false;
    if (bits == 1) {
        uint32_t ones = 0;
        for (uint32_t i = 0; i < bpv / 4; i++) ones += __popc(row[i]);
        bad = ones != per;
    } else if (bits == 2) {
        uint32_t c1 = 0, c2 = 0, c3 = 0;
        for (uint32_t i = 0; i < bpv / 4; i++) {
            const uint32_t w = row[i], hi = (w >> 1) & 0x55555555u, lo = w & 0x55555555u;
            c3 += __popc(hi & lo);
            c2 += __popc(hi & ~lo);
            c1 += __popc(~hi & lo & 0x55555555u);
        }
        bad = c1 != per || c2 != per || c3 != per;  // then bucket 0 holds per as well
    } else {
        for (uint32_t c = 0; c < 16 && !bad; c++) {
            const uint32_t rep = c * 0x11111111u;
            uint32_t cnt = 0;
            for (uint32_t i = 0; i < bpv / 4; i++) cnt += __popc(zero_nibbles(row[i] ^ rep));
            bad = cnt != per;
        }
    }
    if (bad) atomicMin(first_bad, (unsigned long long)r);
}

// Rank rows: a permutation of [0, d).  One workgroup per row (grid-stride),
// a d-bit LDS bitmap; a value >= d or a bit already set marks the row.
__global__ __launch_bounds__(256) void k_check_rank_rows(const uint16_t* __restrict__ ranks, uint64_t n, uint32_t d,
                                                         unsigned long long* __restrict__ first_bad) {
    extern __shared__ __align__(16) uint32_t seen[];
    __shared__ int s_bad;
    const uint32_t words = (d + 31) / 32;
    for (uint64_t r = blockIdx.x; r < n; r += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) seen[i] = 0;
        if (threadIdx.x == 0) s_bad = 0;
        __syncthreads();
        bool bad = false;
        for (uint32_t p = threadIdx.x; p < d; p += blockDim.x) {
            const uint32_t v = ranks[r * d + p];
            if (v >= d) {
                bad = true;
            } else {
                const uint32_t bit = 1u << (v & 31);
                bad |= (atomicOr(&seen[v >> 5], bit) & bit) != 0;
            }
        }
        if (bad) s_bad = 1;
        __syncthreads();
        if (threadIdx.x == 0 && s_bad) atomicMin(first_bad, (unsigned long long)r);
        __syncthreads();
    }
}

// Row-major bitmap rows [cnt][qpv] of documents first_doc.. -> interleaved
// words; with ntop > 0 also records the first row whose popcount != ntop.
__global__ void k_interleave_rows(const uint64_t* __restrict__ rows, uint64_t cnt, uint64_t first_doc, uint32_t qpv,
                                  uint64_t* __restrict__ words, uint32_t ntop,
                                  unsigned long long* __restrict__ first_bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const uint64_t doc = first_doc + i;
    uint32_t pop = 0;
    for (uint32_t w = 0; w < qpv; w++) {
        const uint64_t v = rows[i * qpv + w];
        pop += (uint32_t)__popcll(v);
        words[(((doc >> 6) * qpv + w) << 6) + (doc & 63)] = v;
    }
    if (ntop && pop != ntop) atomicMin(first_bad, (unsigned long long)doc);
}

// Moves document `from` onto `to` in the interleaved bitmap layout
// (swap_remove; one thread per word).
__global__ void k_move_doc_interleaved(uint64_t* __restrict__ words, uint32_t qpv, uint64_t from, uint64_t to) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < qpv) {
        const uint64_t src = (((from >> 6) * qpv + w) << 6) + (from & 63);
        const uint64_t dst = (((to >> 6) * qpv + w) << 6) + (to & 63);
        words[dst] = words[src];
        words[src] = 0;
    }
}

__global__ __launch_bounds__(256) void k_merge_topk(const uint64_t* __restrict__ partial, uint32_t nseg, uint32_t k,
                                                    const uint32_t* __restrict__ qmap, float* __restrict__ out_s,
                                                    int64_t* __restrict__ out_i) {
    __shared__ TopKShared tk;
    const uint64_t qidx = blockIdx.x;
    const uint64_t q = qmap ? qmap[qidx] : qidx;
    const uint64_t total = (uint64_t)nseg * k;
    const uint64_t* in = partial + qidx * total;
    if (total <= 256 && k <= 32) {
        // Few keys (one-query calls): wave 0 holds up to 4 per lane and takes
        // the k smallest one at a time (wave minimum, one occurrence removed
        // per round, so duplicate keys are kept like the LDS top-k keeps them).
        if (threadIdx.x >= 64) return;
        const uint32_t lane = threadIdx.x;
        uint64_t r[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t i = lane + 64u * j;
            r[j] = i < total ? in[i] : kKeyMax;
        }
        for (uint32_t j = 0; j < k; j++) {
            const uint64_t lm = min(min(r[0], r[1]), min(r[2], r[3]));
            uint64_t m = lm;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const uint64_t o = __shfl_xor(m, off);
                m = o < m ? o : m;
            }
            const uint64_t own = __ballot(lm == m);
            if ((int)lane == __ffsll((long long)own) - 1) {
                if (r[0] == m) r[0] = kKeyMax;
                else if (r[1] == m) r[1] = kKeyMax;
                else if (r[2] == m) r[2] = kKeyMax;
                else r[3] = kKeyMax;
            }
            if (lane == 0) {
                out_s[q * k + j] = m == kKeyMax ? -INFINITY : key_score(m);
                out_i[q * k + j] = m == kKeyMax ? -1 : (int64_t)(m & 0xffffffffu);
            }
        }
        return;
    }
    tk_init(tk, k);
    for (uint64_t r0 = 0; r0 < total; r0 += blockDim.x) {
        tk_reserve(tk, k, blockDim.x);
        uint64_t j = r0 + threadIdx.x;
        if (j < total) {
            uint64_t key = in[j];
            if (key != kKeyMax) tk_offer(tk, k, key);
        }
    }
    __syncthreads();
    tk_flush(tk, k);
    for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
        uint64_t key = tk.keys[j];
        out_s[q * k + j] = key == kKeyMax ? -INFINITY : key_score(key);
        out_i[q * k + j] = key == kKeyMax ? -1 : (int64_t)(key & 0xffffffffu);
    }
}

// ---------------------------------------------------------------------------
// Differentiable rerank: scores of given ids, and their gradient with respect
// to the raw query by re-decoding the selected rows' codes.
// ---------------------------------------------------------------------------

// Sum over the workgroup (256 threads); every thread gets the result.  Fixed
// reduction order, so the result is deterministic.
__device__ __forceinline__ float block_sum256(float v, float* red) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = __fadd_rn(v, __shfl_xor(v, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return __fadd_rn(__fadd_rn(red[0], red[1]), __fadd_rn(red[2], red[3]));
}

// out[q][i] = score of row ids[q][i] for query q, computed exactly as the
// search kernels compute it; id -1 gives -inf, any other id outside [0, n)
// also gives -inf and sets *bad.
template <int TIER, int BITS>
__global__ __launch_bounds__(256) void k_score_ids(const uint8_t* __restrict__ codes, uint64_t n_docs, uint32_t d,
                                                   float inv_norm, const float* __restrict__ qhat,
                                                   const float* __restrict__ offset, const int64_t* __restrict__ ids,
                                                   uint32_t k, float* __restrict__ out, uint32_t* __restrict__ bad) {
    extern __shared__ __align__(16) float qs[];
    const uint64_t q = blockIdx.x;
    if (TIER == 2) stage_q_mix<BITS>(qs, qhat + q * d, d);
    else stage_q(qs, qhat + q * d, d);
    __syncthreads();
    const float off = offset[q];
    const uint32_t bpv = d * BITS / 8;
    const int64_t* row = ids + q * k;
    float* o = out + q * k;
    if (TIER == 2) {
        const uint32_t quad = threadIdx.x >> 2, jl = threadIdx.x & 3;
        for (uint32_t r0 = 0; r0 < k; r0 += 64) {
            const uint32_t j = r0 + quad;
            const int64_t id = j < k ? row[j] : -1;
            const bool ok = id >= 0 && (uint64_t)id < n_docs;
            const float raw = raw_avx512_quad<BITS>(codes + (uint64_t)(ok ? id : 0) * bpv, d, qs, jl);
            if (j < k && jl == 0) {
                o[j] = ok ? __fadd_rn(__fmul_rn(raw, inv_norm), off) : -INFINITY;
                if (!ok && id != -1) bad[0] = 1u;
            }
        }
    } else {
        for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
            const int64_t id = row[j];
            const bool ok = id >= 0 && (uint64_t)id < n_docs;
            o[j] = ok ? doc_score<TIER, BITS>(codes + (uint64_t)id * bpv, d, qs, inv_norm, off) : -INFINITY;
            if (!ok && id != -1) bad[0] = 1u;
        }
    }
}

// ---------------------------------------------------------------------------
// Large k: score keys of whole candidate rows (or of every document), sorted
// per row by the caller, then gathered.  The keys order (score desc, id asc)
// exactly as TopK does, and the scores are the search kernels' own.
// ---------------------------------------------------------------------------
// keys[j] for j in query q's row [b, e): CSR (offsets), uniform rows of
// row_m candidates, or every document (cands == null: row_m = n_docs, id = j - b).
// Candidates outside [0, n_docs) get kKeyMax (dropped, as the rerank drops them).
template <int TIER, int BITS>
__global__ __launch_bounds__(256) void k_score_keys(const uint8_t* __restrict__ codes, uint64_t n_docs, uint32_t d,
                                                    float inv_norm, const float* __restrict__ qhat,
                                                    const float* __restrict__ offset,
                                                    const uint64_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ cands, uint64_t row_m,
                                                    uint64_t* __restrict__ keys) {
    extern __shared__ __align__(16) float qs[];
    const uint64_t q = blockIdx.y;
    if (TIER == 2) stage_q_mix<BITS>(qs, qhat + q * d, d);
    else stage_q(qs, qhat + q * d, d);
    __syncthreads();
    const float off = offset[q];
    const uint32_t bpv = d * BITS / 8;
    const uint64_t b = offsets ? offsets[q] : q * row_m;
    const uint64_t e = offsets ? offsets[q + 1] : (q + 1) * row_m;
    const uint64_t per = (e - b + gridDim.x - 1) / gridDim.x;
    const uint64_t sb = min(e, b + blockIdx.x * per), se = min(e, sb + per);
    if (TIER == 2) {
        const uint32_t quad = threadIdx.x >> 2, jl = threadIdx.x & 3;
        for (uint64_t r0 = sb; r0 < se; r0 += 64) {
            const uint64_t j = r0 + quad;
            const uint64_t id = j < se ? (cands ? (uint64_t)cands[j] : j - b) : 0;
            const bool ok = j < se && id < n_docs;
            const float raw = raw_avx512_quad<BITS>(codes + (ok ? id : 0) * bpv, d, qs, jl);
            if (j < se && jl == 0)
                keys[j] = ok ? score_key(__fadd_rn(__fmul_rn(raw, inv_norm), off), (uint32_t)id) : kKeyMax;
        }
    } else {
        for (uint64_t j = sb + threadIdx.x; j < se; j += blockDim.x) {
            const uint64_t id = cands ? (uint64_t)cands[j] : j - b;
            keys[j] = id < n_docs ? score_key(doc_score<TIER, BITS>(codes + id * bpv, d, qs, inv_norm, off), (uint32_t)id)
                                  : kKeyMax;
        }
    }
}

// Top-bucket Bitmap keys ((n_top - overlap) << 32 | row) of CSR rows; rows
// outside [0, n_docs) get kKeyMax.
__global__ __launch_bounds__(256) void k_bm_keys(const uint64_t* __restrict__ words, uint32_t qpv, uint64_t n_docs,
                                                 const uint64_t* __restrict__ qbits,
                                                 const uint64_t* __restrict__ offsets,
                                                 const uint32_t* __restrict__ rows, uint32_t ntop,
                                                 uint64_t* __restrict__ keys) {
    const uint64_t q = blockIdx.y;
    const uint64_t b = offsets[q], e = offsets[q + 1];
    const uint64_t* qb = qbits + q * qpv;
    for (uint64_t j = b + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < e; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = rows[j];
        if (r >= n_docs) {
            keys[j] = kKeyMax;
            continue;
        }
        const uint64_t* dw = words + (r / 64) * (uint64_t)qpv * 64 + (r % 64);
        uint32_t acc = 0;
        for (uint32_t w = 0; w < qpv; w++) acc += (uint32_t)__popcll(dw[(uint64_t)w * 64] & qb[w]);
        keys[j] = ((uint64_t)(ntop - acc) << 32) | r;
    }
}

// Row q of the output: the first min(out_k, row length) sorted keys of its
// row, decoded; the rest (and kKeyMax keys) padded -inf / -1.  bitmap_ntop > 0
// decodes Bitmap keys (score = overlap as f32), else score keys.
__global__ __launch_bounds__(256) void k_gather_keys(const uint64_t* __restrict__ sorted,
                                                     const uint64_t* __restrict__ offsets, uint64_t row_m,
                                                     uint32_t out_k, uint32_t bitmap_ntop, float* __restrict__ out_s,
                                                     int64_t* __restrict__ out_i) {
    const uint64_t q = blockIdx.y;
    const uint64_t b = offsets ? offsets[q] : q * row_m;
    const uint64_t e = offsets ? offsets[q + 1] : (q + 1) * row_m;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < out_k; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = b + j < e ? sorted[b + j] : kKeyMax;
        float sc;
        if (key == kKeyMax) sc = -INFINITY;
        else if (bitmap_ntop) sc = (float)(bitmap_ntop - (uint32_t)(key >> 32));
        else sc = key_score(key);
        out_s[q * out_k + j] = sc;
        out_i[q * out_k + j] = key == kKeyMax ? -1 : (int64_t)(key & 0xffffffffu);
    }
}

// Backward of k_score_ids through the query normalisation.  With
// s_i = sum_j qhat_j * c(code_ij) / R and qhat = x / ||x||:
//   g_j  = sum_i G_i * c(code_ij) / R          (re-decoded from the codes)
//   dx_j = (g_j - qhat_j * (qhat . g)) / ||x||  (zero when ||x|| <= 1e-12)
// One workgroup per query; rows (ids) and G staged in LDS, g in LDS (d floats).
template <int BITS>
__global__ __launch_bounds__(256) void k_score_ids_bwd(const uint8_t* __restrict__ codes, uint64_t n_docs,
                                                       uint32_t d, float inv_norm, const float* __restrict__ x,
                                                       const int64_t* __restrict__ ids, uint32_t k,
                                                       const float* __restrict__ grad, float* __restrict__ dx,
                                                       uint32_t* __restrict__ bad) {
    constexpr uint32_t CPB = 8 / BITS, MASK = (1u << BITS) - 1;
    const float half = ((float)(1u << BITS) - 1.0f) / 2.0f;
    extern __shared__ __align__(16) uint64_t smem[];
    uint32_t* sid = reinterpret_cast<uint32_t*>(smem);
    float* sg = reinterpret_cast<float*>(sid + k);
    float* gs = sg + k;  // d floats
    __shared__ float red[4];
    const uint64_t q = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
        const int64_t id = ids[q * k + i];
        const bool ok = id >= 0 && (uint64_t)id < n_docs;
        if (!ok && id != -1) bad[0] = 1u;
        sid[i] = ok ? (uint32_t)id : 0u;
        sg[i] = ok ? grad[q * k + i] : 0.0f;
    }
    const float* xq = x + q * d;
    float ss = 0.0f;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) ss = __builtin_fmaf(xq[j], xq[j], ss);
    const float norm = __fsqrt_rn(block_sum256(ss, red));  // also orders the LDS staging above
    const bool zero = !(norm > 1e-12f);
    const uint32_t bpv = d * BITS / 8;
    for (uint32_t byte = threadIdx.x; byte < bpv; byte += blockDim.x) {
        float acc[CPB];
#pragma unroll
        for (uint32_t c = 0; c < CPB; c++) acc[c] = 0.0f;
        const uint8_t* col = codes + byte;
#pragma unroll 8
        for (uint32_t i = 0; i < k; i++) {
            const uint32_t v = col[(uint64_t)sid[i] * bpv];
            const float g = sg[i];
#pragma unroll
            for (uint32_t c = 0; c < CPB; c++) {
                const float cen = __fadd_rn((float)((v >> ((CPB - 1 - c) * BITS)) & MASK), -half);
                acc[c] = __builtin_fmaf(g, cen, acc[c]);
            }
        }
#pragma unroll
        for (uint32_t c = 0; c < CPB; c++) gs[byte * CPB + c] = __fmul_rn(acc[c], inv_norm);
    }
    __syncthreads();
    float xg = 0.0f;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) xg = __builtin_fmaf(xq[j], gs[j], xg);
    const float inv = zero ? 0.0f : __fdiv_rn(1.0f, norm);
    const float dot = __fmul_rn(block_sum256(xg, red), inv);  // qhat . g
    float* o = dx + q * d;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        const float qh = __fmul_rn(xq[j], inv);
        o[j] = zero ? 0.0f : __fmul_rn(__fsub_rn(gs[j], __fmul_rn(qh, dot)), inv);
    }
}


// FastScan b=2 query LUT (src/fastscan.rs:130-173): one workgroup per query.
__global__ __launch_bounds__(256) void k_fastscan_prepare(const float* __restrict__ qhat, uint32_t d,
                                                          uint8_t* __restrict__ lut, float* __restrict__ bias_invq) {
    __shared__ float smin[256], smax[256];
    const float cen[4] = {-1.5f, -0.5f, 0.5f, 1.5f};
    const uint64_t q = blockIdx.x;
    const float* x = qhat + q * d;
    const uint32_t pairs = d / 2;
    float mn = INFINITY, mx = -INFINITY;
    for (uint32_t t = threadIdx.x; t < pairs * 16; t += blockDim.x) {
        uint32_t p = t >> 4, a = (t >> 2) & 3, c = t & 3;
        float v = __fadd_rn(__fmul_rn(x[2 * p], cen[a]), __fmul_rn(x[2 * p + 1], cen[c]));
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    const float gmin = smin[0], gmax = smax[0];
    float span = __fadd_rn(gmax, -gmin);
    if (!(span >= 1e-12f)) span = 1e-12f;
    const float scale = __fdiv_rn(255.0f, span);
    if (threadIdx.x == 0) {
        bias_invq[q * 2 + 0] = __fmul_rn((float)pairs, gmin);
        bias_invq[q * 2 + 1] = __fdiv_rn(1.0f, scale);
    }
    uint8_t* lq = lut + q * (uint64_t)pairs * 16;
    for (uint32_t t = threadIdx.x; t < pairs * 16; t += blockDim.x) {
        uint32_t p = t >> 4, a = (t >> 2) & 3, c = t & 3;
        float v = __fadd_rn(__fmul_rn(x[2 * p], cen[a]), __fmul_rn(x[2 * p + 1], cen[c]));
        float r = roundf(__fmul_rn(__fadd_rn(v, -gmin), scale));
        r = fminf(fmaxf(r, 0.0f), 255.0f);
        lq[t] = (uint8_t)r;
    }
}

// FastScan scan (src/fastscan.rs:184-347): exact u32 accumulation of u8 LUT
// entries, score = (bias_sum + acc*inv_q) * inv_norm.  One thread per doc.
__global__ __launch_bounds__(256) void k_fullscan_fs(const uint8_t* __restrict__ packed, uint64_t n, uint32_t d,
                                                     const uint8_t* __restrict__ lut, const float* __restrict__ bias_invq,
                                                     float inv_norm, uint64_t blocks_per_seg, uint32_t k,
                                                     uint64_t* __restrict__ partial) {
    __shared__ TopKShared tk;
    extern __shared__ __align__(16) uint8_t slut[];
    const uint64_t q = blockIdx.y, seg = blockIdx.x, nseg = gridDim.x;
    const uint32_t pairs = d / 2;
    const uint8_t* lq = lut + q * (uint64_t)pairs * 16;
    for (uint32_t i = threadIdx.x; i < pairs * 16; i += blockDim.x) slut[i] = lq[i];
    const float bias = bias_invq[q * 2 + 0], invq = bias_invq[q * 2 + 1];
    tk_init(tk, k);
    const uint64_t nblk = (n + 31) / 32;
    const uint64_t b0 = seg * blocks_per_seg, b1 = min(nblk, b0 + blocks_per_seg);
    for (uint64_t blk = b0; blk < b1; blk += blockDim.x / 32) {
        tk_reserve(tk, k, blockDim.x);
        uint64_t mb = blk + threadIdx.x / 32;
        uint32_t lane = threadIdx.x & 31;
        uint64_t doc = mb * 32 + lane;
        if (mb < b1 && doc < n) {
            const uint8_t* bp = packed + mb * (uint64_t)pairs * 32 + lane;
            uint32_t acc = 0;
            for (uint32_t p = 0; p < pairs; p++) acc += slut[p * 16 + bp[(uint64_t)p * 32]];
            float raw = __fadd_rn(bias, __fmul_rn((float)acc, invq));
            float s = __fadd_rn(__fmul_rn(raw, inv_norm), 0.0f);
            tk_offer(tk, k, score_key(s, (uint32_t)doc));
        }
    }
    __syncthreads();
    tk_flush(tk, k);
    uint64_t* out = partial + (q * nseg + seg) * k;
    for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) out[j] = tk.keys[j];
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

void check_launch() { OVX_HIP(hipGetLastError()); }

}  // namespace

int select_tier(uint32_t d, uint32_t bits) {
    if (bits != 2 && bits != 4) return 0;
    if (d % 64 == 0) return 2;
    if ((bits == 2 && d % 16 == 0) || (bits == 4 && d % 8 == 0)) return 1;
    return 0;
}

float rankquant_inv_norm(uint32_t d, uint32_t bits) {
    double nb = (double)(1u << bits);
    double var = (nb * nb - 1.0) / 12.0;
    float norm = (float)std::sqrt((double)d * var);
    return 1.0f / norm;
}

void launch_check_finite(const float* x, uint64_t count, uint32_t* flag, hipStream_t s) {
    if (count == 0) return;
    uint32_t grid = std::min<uint64_t>(4096, (count + 255) / 256);
    hipLaunchKernelGGL(k_check_finite, dim3(grid), dim3(256), 0, s, x, count, flag);
    check_launch();
}

// Row gather / scatter by an index list, for the exact recomputation of
// flagged queries (one launch each instead of one copy per row): gather
// writes dst row i = src row idx[i], scatter writes dst row idx[i] = src row
// i; rows of `row` elements.
template <typename T, bool GATHER>
__global__ __launch_bounds__(256) void k_rows_by_index(T* __restrict__ dst, const T* __restrict__ src,
                                                       const uint32_t* __restrict__ idx, uint64_t nrows, uint32_t row) {
    const uint64_t total = nrows * row;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / row, j = e % row, o = (uint64_t)idx[i] * row + j;
        if (GATHER) dst[e] = src[o];
        else dst[o] = src[e];
    }
}

template <typename T, bool GATHER>
static void launch_rows_by_index(T* dst, const T* src, const uint32_t* idx, uint64_t nrows, uint32_t row,
                                 hipStream_t s) {
    const uint64_t total = nrows * row;
    if (total == 0) return;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL((k_rows_by_index<T, GATHER>), dim3(grid), dim3(256), 0, s, dst, src, idx, nrows, row);
    check_launch();
}

void gather_rows_u64(uint64_t* dst, const uint64_t* src, const uint32_t* idx, uint64_t nrows, uint32_t row,
                     hipStream_t s) {
    launch_rows_by_index<uint64_t, true>(dst, src, idx, nrows, row, s);
}

void scatter_rows_u32(uint32_t* dst, const uint32_t* src, const uint32_t* idx, uint64_t nrows, uint32_t row,
                      hipStream_t s) {
    launch_rows_by_index<uint32_t, false>(dst, src, idx, nrows, row, s);
}

void launch_prepare_queries(const float* q, uint64_t nq, uint32_t d, uint32_t bits, int tier, float inv_norm,
                            float* qhat, float* offset, uint32_t* qstat, hipStream_t s) {
    if (nq == 0) return;
    float centre = ((float)(1u << bits) - 1.0f) / 2.0f;
    ktime_begin(kFamPrepare, s);
    hipLaunchKernelGGL(k_prepare, dim3((uint32_t)nq), dim3(256), 0, s, q, nq, d, tier, centre, inv_norm, qhat, offset,
                       qstat);
    check_launch();
    ktime_end(kFamPrepare, s);
}

void launch_sign_encode(const float* x, uint64_t n, uint32_t d, uint64_t* out, uint64_t first_doc, bool interleaved,
                        hipStream_t s) {
    if (n == 0) return;
    ktime_begin(kFamEncode, s);
    hipLaunchKernelGGL(k_sign_encode, dim3(cdiv(n, 4)), dim3(256), 0, s, x, n, d, out, first_doc, interleaved ? 1 : 0);
    check_launch();
    ktime_end(kFamEncode, s);
}

void launch_rank_encode(const float* x, uint64_t n, uint32_t d, uint32_t bits, uint8_t* codes, hipStream_t s) {
    if (n == 0) return;
    uint32_t P = 1;
    while (P < d) P <<= 1;
    size_t lds = (size_t)P * 8 + d;
    uint32_t grid = (uint32_t)std::min<uint64_t>(n, 1u << 20);
    ktime_begin(kFamEncode, s);
    hipLaunchKernelGGL(k_rank_encode<false>, dim3(grid), dim3(256), lds, s, x, n, d, bits, codes, nullptr);
    check_launch();
    ktime_end(kFamEncode, s);
}

static uint32_t next_pow2_host(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// One-pass encode of n rows (codes and/or interleaved sign words, non-finite
// flag), then the whole-row sort for the rows it listed.
template <int EPL, int BITS>
static void enc_launch(const float* x, uint64_t n, uint8_t* codes, uint64_t* words, uint64_t first_doc,
                       uint32_t* flag, uint32_t* fb_rows, uint32_t* fb_count, uint32_t grid, hipStream_t s) {
    const size_t lds = enc_lds_bytes<EPL>();
#define OVX_ENC(C, W)                                                                                               \
    hipLaunchKernelGGL((k_encode_rows<EPL, BITS, C, W>), dim3(grid), dim3(64), lds, s, x, n, codes, words, first_doc, \
                       flag, fb_rows, fb_count)
    if constexpr ((EPL * BITS) % 8 == 0 && EPL % 8 == 0) {
        if (codes && words) {
            OVX_ENC(true, true);
            return;
        }
    }
    if constexpr ((EPL * BITS) % 8 == 0) {
        if (codes) {
            OVX_ENC(true, false);
            return;
        }
    }
    if constexpr (EPL % 8 == 0 && BITS == 1) OVX_ENC(false, true);
#undef OVX_ENC
}

template <int EPL>
static void enc_epl(const float* x, uint64_t n, uint32_t bits, uint8_t* codes, uint64_t* words, uint64_t first_doc,
                    uint32_t* flag, uint32_t* fb_rows, uint32_t* fb_count, uint32_t grid, hipStream_t s) {
    if (!codes) bits = 1;
    if (bits == 1) enc_launch<EPL, 1>(x, n, codes, words, first_doc, flag, fb_rows, fb_count, grid, s);
    else if (bits == 2) enc_launch<EPL, 2>(x, n, codes, words, first_doc, flag, fb_rows, fb_count, grid, s);
    else enc_launch<EPL, 4>(x, n, codes, words, first_doc, flag, fb_rows, fb_count, grid, s);
}

bool encode_rows_supported(uint32_t d, uint32_t bits, bool codes, bool words) {
    if (!codes && !words) return false;
    if (d != 512 && d != 768 && d != 1024 && d != 2048) return false;
    const uint32_t epl = d / 64;
    if (codes && (bits != 1 && bits != 2 && bits != 4)) return false;
    if (codes && (epl * bits) % 8 != 0) return false;
    if (words && epl % 8 != 0) return false;
    return true;
}

size_t encode_rows_work_bytes(uint64_t n) { return align256((n + 64) * 4); }

bool launch_encode_rows(const float* x, uint64_t n, uint32_t d, uint32_t bits, uint8_t* codes, uint64_t* words,
                        uint64_t first_doc, uint32_t* flag, void* work, hipStream_t s) {
    if (!encode_rows_supported(d, bits, codes != nullptr, words != nullptr)) return false;
    if (n == 0) return true;
    if (n > 0xffffffffull) throw Error(kInternal, "one-pass encode: more than 2^32 rows in one launch");
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    uint32_t* fb_count = static_cast<uint32_t*>(work);
    uint32_t* fb_rows = fb_count + 64;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(n, (uint64_t)cus * 32);
    ktime_begin(kFamEncode, s);
    if (codes) OVX_HIP(hipMemsetAsync(fb_count, 0, 4, s));
    switch (d) {
        case 512: enc_epl<8>(x, n, bits, codes, words, first_doc, flag, fb_rows, fb_count, grid, s); break;
        case 768: enc_epl<12>(x, n, bits, codes, words, first_doc, flag, fb_rows, fb_count, grid, s); break;
        case 1024: enc_epl<16>(x, n, bits, codes, words, first_doc, flag, fb_rows, fb_count, grid, s); break;
        default: enc_epl<32>(x, n, bits, codes, words, first_doc, flag, fb_rows, fb_count, grid, s); break;
    }
    check_launch();
    if (codes) {  // rows the one-pass kernel listed: the whole-row sort
        const uint32_t P = next_pow2_host(d);
        const uint32_t g2 = (uint32_t)std::min<uint64_t>(n, (uint64_t)cus * 2);
        hipLaunchKernelGGL(k_rank_encode<false>, dim3(g2), dim3(256), (size_t)P * 8 + d, s, x, n, d, bits, codes,
                           nullptr, fb_rows, fb_count);
        check_launch();
    }
    ktime_end(kFamEncode, s);
    return true;
}

size_t rank_sort_work_bytes(uint64_t n, uint32_t d, hipStream_t s) {
    return 2 * align256(n * d * 8) + align256((n + 1) * 8) + segsort_temp_bytes(n * d, (uint32_t)n, nullptr, s);
}

void launch_rank_sort(const float* x, uint64_t n, uint32_t d, uint16_t* ranks, uint32_t bits, uint8_t* codes,
                      void* work, hipStream_t s) {
    if (n == 0) return;
    if (n > 65535) throw Error(kInternal, "rank sort: more than 65535 rows in one launch");
    char* c = static_cast<char*>(work);
    uint64_t* keys = reinterpret_cast<uint64_t*>(c);
    c += align256(n * d * 8);
    uint64_t* sorted = reinterpret_cast<uint64_t*>(c);
    c += align256(n * d * 8);
    uint64_t* offs = reinterpret_cast<uint64_t*>(c);
    c += align256((n + 1) * 8);
    const size_t tb = segsort_temp_bytes(n * d, (uint32_t)n, nullptr, s);
    ktime_begin(kFamEncode, s);
    hipLaunchKernelGGL(k_rank_keys, dim3((uint32_t)std::min<uint64_t>(65536, (n * d + 255) / 256)), dim3(256), 0, s,
                       x, n * d, d, keys);
    check_launch();
    hipLaunchKernelGGL(k_row_offsets, dim3((uint32_t)std::min<uint64_t>(1024, (n + 256) / 256)), dim3(256), 0, s,
                       offs, n, (uint64_t)d);
    check_launch();
    segsort_keys(c, tb, keys, sorted, n * d, (uint32_t)n, offs, s);
    hipLaunchKernelGGL(k_rank_scatter, dim3((uint32_t)std::min<uint64_t>(n, 65535)), dim3(256), ranks ? 0 : (size_t)d,
                       s, sorted, n, d, ranks, bits, codes);
    check_launch();
    ktime_end(kFamEncode, s);
}

size_t eval_search_work_bytes(uint64_t n, uint64_t nq, uint32_t d, hipStream_t s) {
    return align256(nq * (uint64_t)d * 2) + align256(nq * (uint64_t)d * 4) + 2 * align256(nq * n * 8) +
           align256((nq + 1) * 8) + segsort_temp_bytes(nq * n, (uint32_t)nq, nullptr, s);
}

void launch_eval_buckets(const uint16_t* ranks, uint64_t count, uint32_t d, uint32_t bits, uint8_t* out,
                         hipStream_t s) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_eval_buckets, dim3((uint32_t)std::min<uint64_t>(65536, (count + 255) / 256)), dim3(256), 0,
                       s, ranks, count, d, bits, out, nullptr);
    check_launch();
}

void launch_eval_search(const uint8_t* buckets, uint64_t n, uint32_t d, uint32_t bits, float inv_norm_sq,
                        const uint16_t* qranks, uint64_t nq, uint32_t k, float* out_s, int64_t* out_i, void* work,
                        hipStream_t s) {
    if (n == 0 || nq == 0 || k == 0) return;
    if (nq > 65535) throw Error(kInternal, "eval search: more than 65535 queries in one launch");
    char* c = static_cast<char*>(work);
    float* qcent = reinterpret_cast<float*>(c + align256(nq * (uint64_t)d * 2));
    c += align256(nq * (uint64_t)d * 2) + align256(nq * (uint64_t)d * 4);
    uint64_t* keys = reinterpret_cast<uint64_t*>(c);
    c += align256(nq * n * 8);
    uint64_t* sorted = reinterpret_cast<uint64_t*>(c);
    c += align256(nq * n * 8);
    uint64_t* offs = reinterpret_cast<uint64_t*>(c);
    c += align256((nq + 1) * 8);
    const size_t tb = segsort_temp_bytes(nq * n, (uint32_t)nq, nullptr, s);
    ktime_begin(kFamFullScan, s);
    hipLaunchKernelGGL(k_eval_buckets, dim3((uint32_t)std::min<uint64_t>(65536, (nq * d + 255) / 256)), dim3(256), 0,
                       s, qranks, nq * d, d, bits, nullptr, qcent);
    check_launch();
    const uint32_t gx = (uint32_t)std::min<uint64_t>(65535, std::max<uint64_t>(1, (n + 255) / 256 / 4));
    hipLaunchKernelGGL(k_eval_keys, dim3(gx, (uint32_t)nq), dim3(256), (size_t)d * 4, s, buckets, n, d, bits, qcent,
                       inv_norm_sq, keys);
    check_launch();
    hipLaunchKernelGGL(k_row_offsets, dim3((uint32_t)std::min<uint64_t>(1024, (nq + 256) / 256)), dim3(256), 0, s,
                       offs, nq, n);
    check_launch();
    segsort_keys(c, tb, keys, sorted, nq * n, (uint32_t)nq, offs, s);
    hipLaunchKernelGGL(k_gather_keys, dim3((k + 255) / 256, (uint32_t)nq), dim3(256), 0, s, sorted, offs, n, k, 0u,
                       out_s, out_i);
    check_launch();
    ktime_end(kFamFullScan, s);
}

void launch_sign_score_all(const uint64_t* words, uint64_t n, uint32_t d, const uint64_t* qbits, uint64_t nq,
                           uint32_t* out, hipStream_t s) {
    if (n == 0 || nq == 0) return;
    const uint64_t nblk = (n + 63) / 64;
    for (uint64_t q0 = 0; q0 < nq; q0 += 8 * 65535ull) {
        const uint64_t qc = std::min<uint64_t>(nq - q0, 8 * 65535ull);
        hipLaunchKernelGGL(k_sign_score_all, dim3((uint32_t)((nblk + 3) / 4), (uint32_t)((qc + 7) / 8)), dim3(256), 0,
                           s, words, n, d / 64, qbits + q0 * (d / 64), (uint32_t)qc, out + q0 * n);
        check_launch();
    }
}

void launch_rank_u16(const float* x, uint64_t n, uint32_t d, uint16_t* ranks, hipStream_t s) {
    if (n == 0) return;
    uint32_t P = 1;
    while (P < d) P <<= 1;
    size_t lds = (size_t)P * 8 + d;
    uint32_t grid = (uint32_t)std::min<uint64_t>(n, 1u << 20);
    ktime_begin(kFamEncode, s);
    hipLaunchKernelGGL(k_rank_encode<true>, dim3(grid), dim3(256), lds, s, x, n, d, 0u, nullptr, ranks);
    check_launch();
    ktime_end(kFamEncode, s);
}

void launch_fastscan_pack(const uint8_t* codes, uint64_t n, uint32_t d, uint8_t* packed, hipStream_t s) {
    if (n == 0) return;
    uint64_t total = n * (d / 2);
    uint32_t grid = (uint32_t)std::min<uint64_t>(65536, (total + 255) / 256);
    ktime_begin(kFamEncode, s);
    hipLaunchKernelGGL(k_fastscan_pack, dim3(grid), dim3(256), 0, s, codes, n, d, packed);
    check_launch();
    ktime_end(kFamEncode, s);
}


void launch_topbits_docs(const float* x, uint64_t n, uint32_t d, uint32_t ntop, uint64_t* out, uint64_t first_doc,
                         hipStream_t s) {
    if (n == 0) return;
    const uint32_t P = next_pow2_host(d);
    ktime_begin(kFamEncode, s);
    hipLaunchKernelGGL(k_topbits<false>, dim3((uint32_t)std::min<uint64_t>(n, 65535)), dim3(256),
                       (size_t)P * 8 + (size_t)(d / 64) * 8, s, x, n, d, ntop, out, first_doc, nullptr, 0, nullptr,
                       nullptr);
    check_launch();
    ktime_end(kFamEncode, s);
}

void launch_bitmap_query_bits(const float* queries, uint64_t nq, uint32_t d, uint32_t ntop, uint64_t* qbits,
                              uint32_t* qstat, hipStream_t s) {
    if (nq == 0) return;
    const uint32_t P = next_pow2_host(d);
    ktime_begin(kFamPrepare, s);
    hipLaunchKernelGGL(k_topbits<true>, dim3((uint32_t)std::min<uint64_t>(nq, 65535)), dim3(256),
                       (size_t)P * 8 + (size_t)(d / 64) * 8, s, queries, nq, d, ntop, qbits, 0, nullptr, 0, nullptr,
                       qstat);
    check_launch();
    ktime_end(kFamPrepare, s);
}

void launch_bm_subset(const uint64_t* words, uint32_t d, uint64_t n_docs, const uint64_t* qbits, uint64_t nq,
                      const uint64_t* offsets, const uint32_t* rows, uint32_t ntop, uint32_t out_k, float* out_s,
                      int64_t* out_i, hipStream_t s) {
    if (nq == 0 || out_k == 0) return;
    ktime_begin(kFamRerank, s);
    hipLaunchKernelGGL(k_bm_subset, dim3((uint32_t)nq), dim3(256), 0, s, words, d / 64, n_docs, qbits, offsets, rows,
                       ntop, out_k, out_s, out_i);
    check_launch();
    ktime_end(kFamRerank, s);
}

void launch_bm_overlap_rows(const uint64_t* words, uint32_t d, const uint64_t* q, const uint32_t* rows, uint64_t count,
                            uint32_t* out, hipStream_t s) {
    if (count == 0) return;
    uint32_t grid = (uint32_t)std::min<uint64_t>(4096, (count + 255) / 256);
    hipLaunchKernelGGL(k_bm_overlap_rows, dim3(grid), dim3(256), 0, s, words, d / 64, q, rows, count, out);
    check_launch();
}

void launch_bm_finish(const uint32_t* cand, const uint32_t* cand_h, uint64_t count, uint32_t ntop, float* out_s,
                      int64_t* out_i, hipStream_t s) {
    if (count == 0) return;
    uint32_t grid = (uint32_t)std::min<uint64_t>(4096, (count + 255) / 256);
    hipLaunchKernelGGL(k_bm_finish, dim3(grid), dim3(256), 0, s, cand, cand_h, count, ntop, out_s, out_i);
    check_launch();
}

Stage1Plan plan_stage1(uint64_t n, uint32_t d, uint64_t nq) {
    Stage1Plan p{};
    p.nbins = d + 1;
    p.hwords = (p.nbins + 1) / 2;
    uint32_t qt = nq <= 1 ? 1 : (nq <= 8 ? 4 : 16);
    while (qt > 1 && (size_t)qt * p.hwords * 4 > 65536) qt = qt == 16 ? 4 : 1;
    p.qt = qt;
    p.lds_bytes = qt * p.hwords * 4;
    uint64_t nblk = (n + 63) / 64;
    p.n_pad = nblk * 64;
    uint64_t tiles = (nq + qt - 1) / qt;
    uint64_t want = std::max<uint64_t>(1, (2048 + tiles - 1) / tiles);
    uint64_t nseg = std::min<uint64_t>(std::max<uint64_t>(1, nblk), want);
    uint64_t bps = (nblk + nseg - 1) / nseg;
    if (bps > 1023) bps = 1023;  // packed 16-bit LDS bins: <= 65472 docs per segment
    if (bps == 0) bps = 1;
    p.blocks_per_seg = (uint32_t)bps;
    p.nseg = (uint32_t)std::max<uint64_t>(1, (nblk + bps - 1) / bps);
    uint64_t want3 = std::max<uint64_t>(1, (2048 + nq - 1) / nq);
    uint64_t max3 = std::max<uint64_t>(1, (n + 4095) / 4096);
    p.nseg3 = (uint32_t)std::min(want3, max3);
    return p;
}

// m_eff above kStage1MaxM takes the ordered scatter (run offsets, no below /
// ties buffers); smaller m the compact + sort.
size_t stage1_work_bytes(const Stage1Plan& p, uint64_t nq, uint64_t m_eff) {
    const bool big = m_eff > kStage1MaxM;
    size_t b = 0;
    b += align256(nq * p.n_pad * 2);
    b += align256(nq * (uint64_t)p.nseg * p.nbins * 2);
    b += align256(nq * 16);
    b += align256(nq * 8);
    b += big ? align256(nq * (uint64_t)p.nseg * p.nbins * 4) : align256(nq * m_eff * 8);
    b += big ? 0 : align256(nq * (uint64_t)kTieCap * 4);
    return b;
}

void carve_stage1(const Stage1Plan& p, uint64_t nq, uint64_t m_eff, void* base, Stage1Work* w) {
    const bool big = m_eff > kStage1MaxM;
    char* c = static_cast<char*>(base);
    w->H = reinterpret_cast<uint16_t*>(c);
    c += align256(nq * p.n_pad * 2);
    w->hist_part = reinterpret_cast<uint16_t*>(c);
    c += align256(nq * (uint64_t)p.nseg * p.nbins * 2);
    w->thr = reinterpret_cast<uint32_t*>(c);
    c += align256(nq * 16);
    w->counters = reinterpret_cast<uint32_t*>(c);
    c += align256(nq * 8);
    w->below = nullptr;
    w->ties = nullptr;
    w->run_off = nullptr;
    if (big) {
        w->run_off = reinterpret_cast<uint32_t*>(c);
        return;
    }
    w->below = reinterpret_cast<uint64_t*>(c);
    c += align256(nq * m_eff * 8);
    w->ties = reinterpret_cast<uint32_t*>(c);
}

template <int QT, int OP>
static void launch_scan_qt(const uint64_t* words, uint64_t n, uint32_t qpv, const uint64_t* qbits, uint64_t nq,
                           const Stage1Plan& p, const Stage1Work& w, uint32_t ntop, hipStream_t s) {
    uint32_t nblk = (uint32_t)(p.n_pad / 64);
    dim3 grid((uint32_t)((nq + QT - 1) / QT), p.nseg);
#define OVX_SIGN_SCAN(W)                                                                                           \
    hipLaunchKernelGGL((k_sign_scan<QT, W, OP>), grid, dim3(256), p.lds_bytes, s, words, nblk, qpv, n, qbits,      \
                       (uint32_t)nq, p.blocks_per_seg, p.n_pad, w.H, w.hist_part, p.nseg, p.nbins, p.hwords, ntop)
    if (qpv % 4 == 0) OVX_SIGN_SCAN(4);
    else if (qpv % 2 == 0) OVX_SIGN_SCAN(2);
    else OVX_SIGN_SCAN(1);
#undef OVX_SIGN_SCAN
    check_launch();
}

template <int OP>
static void launch_scan_qts(const uint64_t* words, uint64_t n, uint32_t qpv, const uint64_t* qbits, uint64_t nq,
                            const Stage1Plan& p, const Stage1Work& w, uint32_t ntop, hipStream_t s) {
    if (p.qt == 16) launch_scan_qt<16, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
    else if (p.qt == 4) launch_scan_qt<4, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
    else launch_scan_qt<1, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
}

void launch_stage1(const uint64_t* words, uint64_t n, uint32_t d, const uint64_t* qbits, uint64_t nq, uint64_t m_eff,
                   const Stage1Plan& p, const Stage1Work& w, uint32_t* cand, uint32_t* cand_h, uint32_t ntop,
                   hipStream_t s) {
    if (nq == 0 || m_eff == 0) return;
    const uint32_t qpv = d / 64;
    ktime_begin(kFamSignScan, s);
    if (ntop) launch_scan_qts<1>(words, n, qpv, qbits, nq, p, w, ntop, s);
    else launch_scan_qts<0>(words, n, qpv, qbits, nq, p, w, 0, s);
    ktime_end(kFamSignScan, s);
    ktime_begin(kFamStage1Select, s);
    hipLaunchKernelGGL(k_stage1_threshold, dim3((uint32_t)nq), dim3(256), (size_t)p.nbins * 4, s, w.hist_part, p.nseg,
                       p.nbins, m_eff, w.thr, w.counters);
    check_launch();
    if (m_eff > kStage1MaxM) {
        hipLaunchKernelGGL(k_stage1_run_offsets, dim3((uint32_t)nq), dim3(256), (size_t)p.nbins * 4, s, w.hist_part,
                           p.nseg, p.nbins, w.thr, w.run_off);
        check_launch();
        hipLaunchKernelGGL(k_stage1_scatter, dim3(p.nseg, (uint32_t)nq), dim3(64), (size_t)p.nbins * 4, s, w.H, n,
                           p.n_pad, p.blocks_per_seg, p.nseg, p.nbins, w.thr, w.run_off, m_eff, cand, cand_h);
        check_launch();
        ktime_end(kFamStage1Select, s);
        return;
    }
    uint64_t per3 = (n + p.nseg3 - 1) / p.nseg3;
    per3 = (per3 + 7) & ~7ull;
    hipLaunchKernelGGL(k_stage1_compact, dim3(p.nseg3, (uint32_t)nq), dim3(256), 0, s, w.H, n, p.n_pad, per3, w.thr,
                       w.counters, m_eff, w.below, w.ties);
    check_launch();
    hipLaunchKernelGGL(k_stage1_finalize, dim3((uint32_t)nq), dim3(256), 0, s, w.H, n, p.n_pad, w.thr, w.below,
                       w.ties, m_eff, cand, cand_h);
    check_launch();
    ktime_end(kFamStage1Select, s);
}

FastPlan plan_stage1_fast(uint64_t n, uint32_t d, uint64_t nq, uint64_t m_eff) {
    FastPlan p{};
    p.nbins = d + 1;
    p.hwords = (p.nbins + 1) / 2;
    uint32_t qt = nq <= 1 ? 1 : (nq <= 8 ? 4 : 16);
    while (qt > 1 && (size_t)qt * p.hwords * 4 > 65536) qt = qt == 16 ? 4 : 1;
    p.qt = qt;
    p.lds_bytes = qt * p.hwords * 4;
    p.cap = kFastCap;
    const uint64_t nblk = (n + 63) / 64;
    const uint64_t tiles = (nq + qt - 1) / qt;
    if (n <= p.cap) {
        p.s_target = 0xffffffffu;
        p.stride = 1;
        p.nsample = 0;
        p.sparse = 0;
    } else {
        // Survivors to aim at: min(2 m, (m + cap) / 2); where that leaves
        // under 2x of room on either side (m above cap / 4) the geometric
        // mean of m and cap, the same factor below (count < m) and above
        // (count > cap), with a sample grown (a smaller stride) to about 200
        // sampled survivors so that its noise stays inside that room: at 1M x
        // 1024 and m = 4000, 36 of 1024 queries had gone to the exact path
        // (2.87 ms per call), none do now (0.93 ms).
        const bool tight = 4 * m_eff > p.cap;
        const double target_full = tight ? std::sqrt((double)m_eff * (double)p.cap)
                                         : std::min(2.0 * (double)m_eff, 0.5 * (double)(m_eff + p.cap));
        uint64_t stride = std::min<uint64_t>(64, std::max<uint64_t>(1, nblk / 256));
        if (tight) stride = std::min<uint64_t>(stride, std::max<uint64_t>(1, (uint64_t)(target_full / 200.0)));
        p.stride = (uint32_t)stride;
        p.nsample = (uint32_t)((nblk + stride - 1) / stride);
        double expect;  // survivors per query the threshold aims at
        if (stride == 1) {
            p.s_target = (uint32_t)m_eff;  // the sample is the corpus: exact threshold
            expect = (double)m_eff;
        } else {
            double sample_docs = std::min<double>((double)n, (double)p.nsample * 64.0);
            p.s_target = (uint32_t)std::ceil(target_full * sample_docs / (double)n) + 8;
            expect = target_full;
        }
        // Measured at 1024 and 200 queries on mixture data: an exact
        // threshold (stride 1) keeping one survivor in 12 documents stays on
        // the fast path; a sampled one, whose count overshoots its aim, needs
        // one in 14 (one in 12.5 has overflowed the lists).
        p.sparse = (double)n >= (stride == 1 ? 12.0 : 14.0) * expect ? 1u : 0u;
        uint64_t want = std::max<uint64_t>(1, (1024 + tiles - 1) / tiles);
        uint64_t nsseg = std::min<uint64_t>(std::max<uint64_t>(1, p.nsample), want);
        uint64_t sbps = (p.nsample + nsseg - 1) / nsseg;
        if (sbps > 1023) sbps = 1023;
        if (sbps == 0) sbps = 1;
        p.sample_bps = (uint32_t)sbps;
        p.nsseg = (uint32_t)std::max<uint64_t>(1, (p.nsample + sbps - 1) / sbps);
    }
    p.scan_qt = nq >= 256 ? 64 : (nq >= 32 ? 16 : (nq >= 8 ? 4 : 1));
    const uint64_t scan_tiles = (nq + p.scan_qt - 1) / p.scan_qt;
    uint64_t want = std::max<uint64_t>(1, (2048 + scan_tiles - 1) / scan_tiles);
    uint64_t nseg = std::min<uint64_t>(std::max<uint64_t>(1, nblk), want);
    uint64_t bps = std::max<uint64_t>(1, (nblk + nseg - 1) / nseg);
    p.scan_wgl = 0;
    if (p.scan_qt <= 4) {
        // Tiles of up to 4 queries collect survivors in a workgroup LDS list
        // of qt x 64 x bps entries (16 KiB: 32 blocks per segment at one
        // query), which no segment can overflow.
        bps = std::min<uint64_t>(std::max<uint64_t>(1, nblk), 32 / p.scan_qt);
        p.scan_wgl = (uint32_t)(p.scan_qt * 64 * bps);
    }
    p.scan_bps = (uint32_t)bps;
    p.nseg = (uint32_t)std::max<uint64_t>(1, (nblk + bps - 1) / bps);
    return p;
}

size_t stage1_fast_bytes(const FastPlan& p, uint64_t nq) {
    return align256(nq * (uint64_t)p.nbins * 4) + 2 * align256(nq * 4) + align256(nq * (uint64_t)p.cap * 8);
}

void carve_stage1_fast(const FastPlan& p, uint64_t nq, void* base, FastWork* w) {
    char* c = static_cast<char*>(base);
    w->hist = reinterpret_cast<uint32_t*>(c);
    c += align256(nq * (uint64_t)p.nbins * 4);
    w->thr = reinterpret_cast<uint32_t*>(c);
    c += align256(nq * 4);
    w->cnt = reinterpret_cast<uint32_t*>(c);
    c += align256(nq * 4);
    w->keys = reinterpret_cast<uint64_t*>(c);
}

template <int QT, int WCH, int OP>
static void launch_sample_qt(const uint64_t* words, uint64_t n, uint32_t qpv, const uint64_t* qbits, uint64_t nq,
                             const FastPlan& p, const FastWork& w, uint32_t ntop, hipStream_t s) {
    const uint32_t nblk = (uint32_t)((n + 63) / 64);
    const uint32_t tiles = (uint32_t)((nq + QT - 1) / QT);
    hipLaunchKernelGGL((k_s1_sample<QT, WCH, OP>), dim3(tiles, p.nsseg), dim3(256), p.lds_bytes, s, words, nblk, qpv,
                       n, qbits, (uint32_t)nq, p.stride, p.nsample, p.sample_bps, w.hist, p.nbins, p.hwords, ntop);
    check_launch();
}

template <int QT, int OP>
static void launch_sample_wch(const uint64_t* words, uint64_t n, uint32_t qpv, const uint64_t* qbits, uint64_t nq,
                              const FastPlan& p, const FastWork& w, uint32_t ntop, hipStream_t s) {
    if (qpv % 4 == 0) launch_sample_qt<QT, 4, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
    else if (qpv % 2 == 0) launch_sample_qt<QT, 2, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
    else launch_sample_qt<QT, 1, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
}

static bool use_mfma_scan(uint64_t nq, uint32_t qpv, const FastPlan& p);
// about one sample workgroup per CU: fewer leave CUs idle, more flush more
// (each workgroup's local threshold sits higher) -- 256 measured best at 1M x 1024
constexpr uint64_t kSampleWgs = 256;

template <int OP>
static void launch_sample(const uint64_t* words, uint64_t n, uint32_t qpv, const uint64_t* qbits, uint64_t nq,
                          const FastPlan& p, const FastWork& w, uint32_t ntop, hipStream_t s) {
    if (use_mfma_scan(nq, qpv, p) && qpv <= 16) {
        // matrix-core sample: 32 queries per workgroup, >= 8 sampled blocks
        // (one per wave) per workgroup, about kSampleWgs workgroups
        const uint32_t nblk = (uint32_t)((n + 63) / 64);
        const uint32_t tiles = (uint32_t)((nq + 31) / 32);
        const uint64_t wgs = p.tune.sample_wgs ? p.tune.sample_wgs : kSampleWgs;
        const uint64_t want = std::max<uint64_t>(1, (wgs + tiles - 1) / tiles);
        const uint64_t nseg = std::max<uint64_t>(1, std::min<uint64_t>(want, (p.nsample + 7) / 8));
        const uint32_t bps = (uint32_t)std::min<uint64_t>(1023, (p.nsample + nseg - 1) / nseg);
        const uint32_t segs = (p.nsample + bps - 1) / bps;
        const size_t lds = (size_t)qpv * 64 * 16 + (size_t)32 * p.hwords * 4;
#define OVX_SMPL(QPV)                                                                                            \
    hipLaunchKernelGGL((k_s1_sample_mfma<OP, QPV>), dim3(tiles, segs), dim3(kMfmaThreads), lds, s, words, nblk, n, \
                       qbits, (uint32_t)nq, p.stride, p.nsample, bps, w.hist, p.nbins, p.hwords, p.s_target, ntop)
        switch (qpv) {
            case 16: OVX_SMPL(16); break;
            case 12: OVX_SMPL(12); break;
            case 8: OVX_SMPL(8); break;
            case 4: OVX_SMPL(4); break;
            case 2: OVX_SMPL(2); break;
            default: OVX_SMPL(1); break;
        }
#undef OVX_SMPL
        check_launch();
        return;
    }
    if (p.qt == 16) launch_sample_wch<16, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
    else if (p.qt == 4) launch_sample_wch<4, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
    else launch_sample_wch<1, OP>(words, n, qpv, qbits, nq, p, w, ntop, s);
}

// The matrix-core scan pays off once a workgroup has a full 32-query tile;
// its query fragments live in LDS (nqt x d bytes).
static std::atomic<bool> g_mfma_scan{true};
bool set_mfma_scan(int value) {
    const bool prev = g_mfma_scan.load();
    if (value >= 0) g_mfma_scan.store(value != 0);
    return prev;
}
template <int NQT, int OP>
static void set_mfma_lds() {
#define OVX_SET(QPV)                                                                          \
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_s1_scan_mfma<NQT, OP, QPV>), \
                              hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024)
    OVX_SET(16);
    OVX_SET(12);
    OVX_SET(8);
    OVX_SET(4);
    OVX_SET(2);
    OVX_SET(1);
#undef OVX_SET
}
template <int OP>
static void set_sample_lds() {
#define OVX_SET(QPV)                                                                             \
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_s1_sample_mfma<OP, QPV>),        \
                              hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024)
    OVX_SET(16);
    OVX_SET(12);
    OVX_SET(8);
    OVX_SET(4);
    OVX_SET(2);
    OVX_SET(1);
#undef OVX_SET
}

// The matrix-core scan pads a batch to 32-query tiles; from 2 queries up that
// beats the VALU scan (whose work grows with every query; at 20M x 1024 the
// padded tile streams the codes at the same speed as a 1-query VALU pass).
// Query fragments live in LDS (16 * d bytes per 32-query tile, <= 64 KiB).
// Only when survivors are sparse (plan.sparse): with many survivors per block
// (every document, in an index of at most cap documents) the workgroup's LDS
// lists overflow and the tile goes to the exact path, while the vector-ALU
// scan appends to the query's buffer, which holds up to cap of them.
constexpr uint64_t kMfmaMinQueries = 2;
static bool use_mfma_scan(uint64_t nq, uint32_t qpv, const FastPlan& p) {
    // once per process, also when serving slots reach here concurrently
    static std::once_flag configured;
    std::call_once(configured, [] {
        set_mfma_lds<4, 0>();
        set_mfma_lds<4, 1>();
        set_sample_lds<0>();
        set_sample_lds<1>();
        set_mfma_lds<2, 0>();
        set_mfma_lds<2, 1>();
        (void)hipGetLastError();
    });
    return g_mfma_scan && nq >= kMfmaMinQueries && p.sparse &&
           (qpv == 16 || qpv == 12 || qpv == 8 || qpv == 4 || qpv == 2 || qpv == 1);
}

// Whether the vector-ALU scan derives the thresholds itself (tiles of at most
// 4 queries, one wave each), so that k_s1_threshold is not launched.
static bool scan_fuses_threshold(uint64_t nq, uint32_t qpv, const FastPlan& p) {
    return !use_mfma_scan(nq, qpv, p) && p.scan_qt <= 4;
}

// Whether launch_scan runs a PrepFuse on the scan's own workgroups.
static bool scan_fuses_prep(uint64_t nq, uint32_t qpv, const FastPlan& p) {
    const uint64_t tiles = (nq + p.scan_qt - 1) / p.scan_qt;
    return !use_mfma_scan(nq, qpv, p) && (uint64_t)p.nseg + (nq + tiles - 1) / tiles <= 65535;
}

template <int OP>
static void launch_scan(const uint64_t* words, uint64_t n, uint32_t qpv, const uint64_t* qbits, uint64_t nq,
                        const FastPlan& p, const FastWork& w, uint32_t ntop, hipStream_t s,
                        const PrepFuse* prep = nullptr) {
    const uint32_t nblk = (uint32_t)((n + 63) / 64);
    if (use_mfma_scan(nq, qpv, p)) {
        // Matrix-core scan: 32, 64 or 128 queries per workgroup of 8 waves,
        // the smallest tile that holds the batch: every extra query tile
        // re-reads the codes, and a large corpus does not stay in L2 between
        // the tiles of a segment.
        uint32_t nqt = nq > 64 ? 4 : (nq > 32 ? 2 : 1);
        if (p.tune.scan_nqt == 1 || p.tune.scan_nqt == 2 || p.tune.scan_nqt == 4) nqt = p.tune.scan_nqt;
        const uint32_t tiles_m = (uint32_t)((nq + 32 * nqt - 1) / (32 * nqt));
        const uint64_t pairs = p.tune.scan_pairs ? p.tune.scan_pairs : 1024;
        const uint64_t want = std::max<uint64_t>(1, (pairs + tiles_m - 1) / tiles_m);
        uint64_t nseg = std::min<uint64_t>(std::max<uint64_t>(1, (nblk + 7) / 8), want);
        if (nseg >= 8) nseg &= ~7ull;  // XCD-grouped segments (see the kernel)
        const uint32_t bps = (uint32_t)((nblk + nseg - 1) / nseg);
        const uint32_t segs = (nblk + bps - 1) / bps;
        const size_t lds = (size_t)nqt * qpv * 64 * 16;
        // Persistent pairs: with grouped segments, the workgroups per XCD are a
        // multiple of the tile count (each keeps one tile), at most the XCD's
        // resident capacity (32 CUs x workgroups per CU).
        uint32_t grid = tiles_m * segs;
        if (segs % 8 == 0 && !p.tune.scan_flat) {
            const uint32_t per_cu = nqt == 4 ? 1u : 2u;
            const uint32_t c = std::max<uint32_t>(1, 32 * per_cu / tiles_m);
            const uint32_t pairs = tiles_m * (segs / 8);
            grid = 8 * std::min<uint32_t>(pairs, tiles_m * c);
        }
#define OVX_MFMA(NQT, QPV)                                                                                         \
    hipLaunchKernelGGL((k_s1_scan_mfma<NQT, OP, QPV>), dim3(grid), dim3(kMfmaThreads), lds, s, words, nblk, n,     \
                       qbits, (uint32_t)nq, bps, w.thr, p.cap, w.keys, w.cnt, ntop, tiles_m, segs)
#define OVX_MFMA_Q(NQT)                        \
    switch (qpv) {                             \
        case 16: OVX_MFMA(NQT, 16); break;     \
        case 12: OVX_MFMA(NQT, 12); break;     \
        case 8: OVX_MFMA(NQT, 8); break;       \
        case 4: OVX_MFMA(NQT, 4); break;       \
        case 2: OVX_MFMA(NQT, 2); break;       \
        default: OVX_MFMA(NQT, 1); break;      \
    }
        if (nqt == 4) {
            OVX_MFMA_Q(4);
        } else if (nqt == 2) {
            OVX_MFMA_Q(2);
        } else {
            OVX_MFMA_Q(1);
        }
#undef OVX_MFMA_Q
#undef OVX_MFMA
        check_launch();
        return;
    }
    const uint32_t tiles = (uint32_t)((nq + p.scan_qt - 1) / p.scan_qt);
    const bool fuse = prep && scan_fuses_prep(nq, qpv, p);
    const uint32_t prep_rows = fuse ? (uint32_t)((nq + tiles - 1) / tiles) : 0u;
    const PrepFuse pf = fuse ? *prep : PrepFuse{};
    size_t lds = (size_t)p.scan_wgl * 8;
    if (fuse) lds = std::max<size_t>(lds, (size_t)kPrepChunk * 4);
    dim3 grid(tiles, p.nseg + prep_rows);
    const uint32_t* hist = scan_fuses_threshold(nq, qpv, p) ? w.hist : nullptr;
#define OVX_SCAN(W)                                                                                               \
    hipLaunchKernelGGL((k_s1_scan<W, OP>), grid, dim3(256), lds, s, words, nblk, qpv, n, qbits, (uint32_t)nq,     \
                       p.scan_qt, p.scan_bps, w.thr, p.cap, w.keys, w.cnt, ntop, p.scan_wgl, pf, prep_rows, hist, \
                       p.nbins, p.s_target)
    switch (qpv) {
        case 16: OVX_SCAN(16); break;
        case 12: OVX_SCAN(12); break;
        case 8: OVX_SCAN(8); break;
        case 4: OVX_SCAN(4); break;
        case 2: OVX_SCAN(2); break;
        case 1: OVX_SCAN(1); break;
        default: OVX_SCAN(0); break;
    }
#undef OVX_SCAN
    check_launch();
}

void launch_stage1_fast(const uint64_t* words, uint64_t n, uint32_t d, const float* queries, uint64_t nq,
                        uint64_t m_eff, const FastPlan& p, const FastWork& w, uint64_t* qbits, uint32_t* cand,
                        uint32_t* cand_h, uint32_t* qstat, uint32_t ntop, hipStream_t s, const PrepFuse* prep,
                        bool select, hipEvent_t before_scan, bool as_set) {
    if (nq == 0) return;
    const uint32_t qpv = d / 64;
    bool s1_prepped = false;
    if (prep && (m_eff == 0 || !scan_fuses_prep(nq, qpv, p))) {
        if (!ntop && !prep->qstat && prep->q == queries && prep->d == d) {
            const float centre = ((float)(1u << prep->bits) - 1.0f) / 2.0f;
            ktime_begin(kFamPrepare, s);
            hipLaunchKernelGGL(k_prepare_s1, dim3((uint32_t)nq), dim3(256), 0, s, queries, d, prep->tier, centre,
                               prep->inv_norm, prep->qhat, prep->offset, qbits, w.hist, p.nbins, w.cnt, qstat);
            check_launch();
            ktime_end(kFamPrepare, s);
            s1_prepped = true;
        } else {
            launch_prepare_queries(prep->q, nq, prep->d, prep->bits, prep->tier, prep->inv_norm, prep->qhat,
                                   prep->offset, prep->qstat, s);
        }
        prep = nullptr;
    }
    if (!s1_prepped) {  // otherwise k_prepare_s1 did k_s1_prep's work
        ktime_begin(kFamStage1Select, s);
        if (ntop) {
            const uint32_t P = next_pow2_host(d);
            hipLaunchKernelGGL(k_topbits<true>, dim3((uint32_t)std::min<uint64_t>(nq, 65535)), dim3(256),
                               (size_t)P * 8 + (size_t)qpv * 8, s, queries, nq, d, ntop, qbits, 0, w.hist, p.nbins,
                               w.cnt, qstat);
        } else {
            hipLaunchKernelGGL(k_s1_prep, dim3((uint32_t)nq), dim3(256), 0, s, queries, d, qbits, w.hist, p.nbins,
                               w.cnt, qstat);
        }
        check_launch();
        ktime_end(kFamStage1Select, s);
    }
    if (m_eff == 0) return;
    if (p.s_target != 0xffffffffu) {
        ktime_begin(kFamSample, s);
        if (ntop) launch_sample<1>(words, n, qpv, qbits, nq, p, w, ntop, s);
        else launch_sample<0>(words, n, qpv, qbits, nq, p, w, 0, s);
        ktime_end(kFamSample, s);
    }
    if (!scan_fuses_threshold(nq, qpv, p)) {
        hipLaunchKernelGGL(k_s1_threshold, dim3((uint32_t)nq), dim3(64), 0, s, w.hist, p.nbins, p.s_target, w.thr);
        check_launch();
    }
    if (before_scan) OVX_HIP(hipStreamWaitEvent(s, before_scan, 0));
    ktime_begin(kFamSignScan, s);
    if (ntop) launch_scan<1>(words, n, qpv, qbits, nq, p, w, ntop, s, prep);
    else launch_scan<0>(words, n, qpv, qbits, nq, p, w, 0, s, prep);
    ktime_end(kFamSignScan, s);
    if (!select) return;  // the caller's tail selects from the survivor keys
    ktime_begin(kFamStage1Select, s);
    if (as_set && !cand_h)
        hipLaunchKernelGGL(k_s1_pick, dim3((uint32_t)nq), dim3(256), 0, s, w.keys, w.cnt, w.thr, p.cap,
                           (uint32_t)m_eff, id_bits(n), cand, qstat);
    else
        hipLaunchKernelGGL(k_s1_select, dim3((uint32_t)nq), dim3(1024), 0, s, w.keys, w.cnt, w.thr, id_bits(n), p.cap,
                           m_eff, cand, cand_h, qstat);
    check_launch();
    ktime_end(kFamStage1Select, s);
}

#define OVX_TIER_BITS_DISPATCH(KERNEL, ...)                                      \
    do {                                                                        \
        if (tier == 2 && bits == 2) KERNEL(2, 2, __VA_ARGS__);                  \
        else if (tier == 2 && bits == 4) KERNEL(2, 4, __VA_ARGS__);             \
        else if (tier == 1 && bits == 2) KERNEL(1, 2, __VA_ARGS__);             \
        else if (tier == 1 && bits == 4) KERNEL(1, 4, __VA_ARGS__);             \
        else if (bits == 1) KERNEL(0, 1, __VA_ARGS__);                          \
        else if (bits == 2) KERNEL(0, 2, __VA_ARGS__);                          \
        else KERNEL(0, 4, __VA_ARGS__);                                         \
    } while (0)

// One-query sign-bitmap stage 1 (see k_s1_head): the head, then the
// vector-ALU scan reading the head's threshold, with the rerank's query
// preparation on extra workgroups.  hist and done are zero on entry and on
// return (in stream order).
void launch_stage1_plain(const uint64_t* words, uint64_t n, uint32_t d, const float* query, const FastPlan& p,
                         uint32_t* hist, uint32_t* done, uint32_t* thr, uint32_t* cnt, uint64_t* keys,
                         uint64_t* qbits, uint32_t* qstat, const PrepFuse& prep, hipStream_t s) {
    const uint32_t qpv = d / 64, nblk = (uint32_t)((n + 63) / 64);
    const bool sampled = p.s_target != 0xffffffffu;
    const uint32_t grid = sampled ? std::max<uint32_t>(1, p.nsseg) : 1u;
    const size_t lds = (size_t)std::max(p.hwords, p.nbins) * 4;
    const uint32_t nsample = sampled ? p.nsample : 0u, bps = sampled ? p.sample_bps : 1u;
    ktime_begin(kFamSample, s);
#define OVX_HEAD(W)                                                                                                \
    hipLaunchKernelGGL(k_s1_head<W>, dim3(grid), dim3(256), lds, s, words, nblk, qpv, n, query, p.stride, nsample, \
                       bps, hist, p.nbins, p.hwords, p.s_target, thr, cnt, qbits, qstat, done)
    if (qpv == 16) OVX_HEAD(16);
    else if (qpv == 8) OVX_HEAD(8);
    else OVX_HEAD(0);
#undef OVX_HEAD
    check_launch();
    ktime_end(kFamSample, s);
    const FastWork w{nullptr, thr, cnt, keys};
    ktime_begin(kFamSignScan, s);
    launch_scan<0>(words, n, qpv, qbits, 1, p, w, 0, s, &prep);
    ktime_end(kFamSignScan, s);
}

bool stage1_plain_supported(uint64_t n, uint32_t d, uint64_t m_eff, uint64_t out_k) {
    if (d % 64 || d > kSignMaxDim || m_eff < 1 || m_eff > kStage1MaxM || out_k < 1 || out_k > kTopKMax) return false;
    const FastPlan p = plan_stage1_fast(n, d, 1, m_eff);
    return p.cap == kFastCap && !use_mfma_scan(1, d / 64, p) && scan_fuses_prep(1, d / 64, p);
}

bool s1_tail_supported(uint64_t n, uint32_t d, uint64_t nq, uint64_t m_eff, uint64_t out_k) {
    if (d > kQInLds || m_eff < 1 || m_eff > kStage1MaxM || out_k < 1 || out_k > kTopKMax || nq > 65535) return false;
    return plan_stage1_fast(n, d, nq, m_eff).cap == kFastCap;
}

size_t s1_tail_lds(uint32_t d) { return (size_t)kTailTieCap * 8 + (size_t)d * 4; }

void launch_s1_tail(const uint64_t* keys, const uint32_t* cnt, const uint32_t* thr, uint32_t cap, uint64_t m_eff,
                    const uint8_t* codes,
                    uint32_t d, uint32_t bits, int tier, float inv_norm, const float* qhat, const float* offset,
                    uint64_t n_docs, uint64_t nq, uint32_t out_k, float* out_s, int64_t* out_i, uint32_t nsplit,
                    void* work, uint32_t* done, uint32_t* qstat, uint32_t* hstat, hipStream_t s) {
    if (nq == 0) return;
    if (nsplit < 1) nsplit = 1;
    const size_t lds = s1_tail_lds(d);
    static std::once_flag configured;
    std::call_once(configured, [] {
        const int bytes = (int)s1_tail_lds(kQInLds);
#define OVX_TAIL_ATTR(T, B, _) \
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_s1_tail<T, B>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes)
        OVX_TAIL_ATTR(2, 2, 0);
        OVX_TAIL_ATTR(2, 4, 0);
        OVX_TAIL_ATTR(1, 2, 0);
        OVX_TAIL_ATTR(1, 4, 0);
        OVX_TAIL_ATTR(0, 1, 0);
        OVX_TAIL_ATTR(0, 2, 0);
        OVX_TAIL_ATTR(0, 4, 0);
#undef OVX_TAIL_ATTR
        (void)hipGetLastError();
    });
    ktime_begin(kFamRerank, s);
#define OVX_TAIL(T, B, _)                                                                                          \
    hipLaunchKernelGGL((k_s1_tail<T, B>), dim3(nsplit, (uint32_t)nq), dim3(256), lds, s, keys, cnt, thr, cap,     \
                       (uint32_t)m_eff, codes, d, inv_norm, qhat, offset, n_docs, out_k, out_s, out_i,            \
                       static_cast<uint64_t*>(work), done, qstat, hstat)
    OVX_TIER_BITS_DISPATCH(OVX_TAIL, 0);
#undef OVX_TAIL
    check_launch();
    ktime_end(kFamRerank, s);
}

uint32_t rerank_nsplit(uint64_t nq, uint64_t avg_m) {
    if (nq >= 128) return 1;
    // one round (128 candidates: two per lane quad) per split workgroup
    uint64_t want = (avg_m + 127) / 128;
    uint64_t cap = std::max<uint64_t>(1, 512 / std::max<uint64_t>(1, nq));
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({want, cap, 32}));
}

size_t rerank_work_bytes(uint64_t nq, uint32_t out_k, uint32_t nsplit) {
    return nsplit > 1 ? align256(nq * (uint64_t)nsplit * out_k * 8) : 0;
}

void launch_rerank(const uint8_t* codes, uint32_t d, uint32_t bits, int tier, float inv_norm, const float* qhat,
                   const float* offset, uint64_t nq, const uint64_t* offsets, const uint32_t* cands, uint64_t row_m,
                   uint64_t n_docs, const uint32_t* qmap, uint32_t out_k, float* out_s, int64_t* out_i,
                   uint32_t nsplit, void* work, hipStream_t s) {
    if (nq == 0 || out_k == 0) return;
    if (nsplit < 1) nsplit = 1;
    dim3 grid((uint32_t)nq, nsplit);
    uint64_t* partial = static_cast<uint64_t*>(work);
    const size_t lds = (size_t)d * 4;
    ktime_begin(kFamRerank, s);
#define OVX_RERANK(T, B, _)                                                                                       \
    hipLaunchKernelGGL((k_rerank<T, B>), grid, dim3(256), lds, s, codes, d, inv_norm, qhat, offset, offsets, cands, \
                       row_m, n_docs, qmap, out_k, out_s, out_i, partial)
    OVX_TIER_BITS_DISPATCH(OVX_RERANK, 0);
#undef OVX_RERANK
    check_launch();
    if (nsplit > 1) {
        hipLaunchKernelGGL(k_merge_topk, dim3((uint32_t)nq), dim3(256), 0, s, partial, nsplit, out_k, qmap, out_s,
                           out_i);
        check_launch();
    }
    ktime_end(kFamRerank, s);
}

void launch_score_ids(const uint8_t* codes, uint64_t n_docs, uint32_t d, uint32_t bits, int tier, float inv_norm,
                      const float* qhat, const float* offset, uint64_t nq, const int64_t* ids, uint32_t k, float* out,
                      uint32_t* bad, hipStream_t s) {
    if (nq == 0 || k == 0) return;
    ktime_begin(kFamRerank, s);
#define OVX_SCORE_IDS(T, B, _)                                                                                    \
    hipLaunchKernelGGL((k_score_ids<T, B>), dim3((uint32_t)nq), dim3(256), (size_t)d * 4, s, codes, n_docs, d, \
                       inv_norm, qhat, offset, ids, k, out, bad)
    OVX_TIER_BITS_DISPATCH(OVX_SCORE_IDS, 0);
#undef OVX_SCORE_IDS
    check_launch();
    ktime_end(kFamRerank, s);
}

void launch_score_ids_backward(const uint8_t* codes, uint64_t n_docs, uint32_t d, uint32_t bits, float inv_norm,
                               const float* x, uint64_t nq, const int64_t* ids, uint32_t k, const float* grad,
                               float* dx, uint32_t* bad, hipStream_t s) {
    if (nq == 0) return;
    const size_t lds = (size_t)k * 8 + (size_t)d * 4;
    ktime_begin(kFamBackward, s);
    if (bits == 1)
        hipLaunchKernelGGL(k_score_ids_bwd<1>, dim3((uint32_t)nq), dim3(256), lds, s, codes, n_docs, d, inv_norm, x,
                           ids, k, grad, dx, bad);
    else if (bits == 2)
        hipLaunchKernelGGL(k_score_ids_bwd<2>, dim3((uint32_t)nq), dim3(256), lds, s, codes, n_docs, d, inv_norm, x,
                           ids, k, grad, dx, bad);
    else
        hipLaunchKernelGGL(k_score_ids_bwd<4>, dim3((uint32_t)nq), dim3(256), lds, s, codes, n_docs, d, inv_norm, x,
                           ids, k, grad, dx, bad);
    check_launch();
    ktime_end(kFamBackward, s);
}

static uint32_t fullscan_nseg(uint64_t n, uint64_t nq) {
    uint64_t want = std::max<uint64_t>(1, (1024 + nq - 1) / nq);
    uint64_t maxseg = std::max<uint64_t>(1, n / 1024);
    return (uint32_t)std::min(want, maxseg);
}

size_t fullscan_work_bytes(uint64_t n, uint64_t nq, uint32_t k) {
    return align256((size_t)nq * fullscan_nseg(n, nq) * k * 8);
}

void launch_fullscan_rq(const uint8_t* codes, uint64_t n, uint32_t d, uint32_t bits, int tier, float inv_norm,
                        const float* qhat, const float* offset, uint64_t nq, uint32_t k, void* work, float* out_s,
                        int64_t* out_i, hipStream_t s) {
    if (nq == 0 || k == 0) return;
    uint32_t nseg = fullscan_nseg(n, nq);
    uint64_t per = (n + nseg - 1) / nseg;
    uint64_t* partial = static_cast<uint64_t*>(work);
    dim3 grid(nseg, (uint32_t)nq);
    ktime_begin(kFamFullScan, s);
#define OVX_FULLSCAN(T, B, _)                                                                                   \
    hipLaunchKernelGGL((k_fullscan_rq<T, B>), grid, dim3(256), 0, s, codes, n, d, inv_norm, qhat, offset, per, k, \
                       partial)
    OVX_TIER_BITS_DISPATCH(OVX_FULLSCAN, 0);
#undef OVX_FULLSCAN
    check_launch();
    hipLaunchKernelGGL(k_merge_topk, dim3((uint32_t)nq), dim3(256), 0, s, partial, nseg, k, nullptr, out_s, out_i);
    check_launch();
    ktime_end(kFamFullScan, s);
}

void launch_rank_centres(const uint16_t* ranks, uint64_t count, uint32_t d, uint32_t bits, float* out,
                         hipStream_t s) {
    if (count == 0) return;
    uint32_t grid = (uint32_t)std::min<uint64_t>(4096, (count + 255) / 256);
    hipLaunchKernelGGL(k_rank_centres, dim3(grid), dim3(256), 0, s, ranks, count, d, bits, out);
    check_launch();
}

void launch_rank_search(const uint16_t* ranks, uint64_t n, uint32_t d, bool symmetric, const uint16_t* qranks,
                        const float* qhat, const float* qsum, uint64_t nq, uint32_t k, void* work, float* out_s,
                        int64_t* out_i, hipStream_t s) {
    if (nq == 0 || k == 0) return;
    const uint32_t nseg = fullscan_nseg(n, nq);
    const uint64_t per = (n + nseg - 1) / nseg;
    uint64_t* partial = static_cast<uint64_t*>(work);
    dim3 grid(nseg, (uint32_t)nq);
    const bool vec = d % 8 == 0;
    const double nd = (double)d;
    const float norm = (float)std::sqrt(nd * (nd * nd - 1.0) / 12.0);  // src/rank.rs:249-252
    ktime_begin(kFamFullScan, s);
    if (symmetric) {
        const float inv_norm_sq = 1.0f / (norm * norm);
        const bool wide = (double)d * (double)(d - 1) * (double)(d - 1) >= 2147483647.0;
        const size_t lds = (size_t)d * 4;
#define OVX_RSYM(V, W) \
    hipLaunchKernelGGL((k_rank_sym<V, W>), grid, dim3(256), lds, s, ranks, n, d, qranks, inv_norm_sq, per, k, partial)
        if (vec && !wide) OVX_RSYM(true, false);
        else if (vec) OVX_RSYM(true, true);
        else if (!wide) OVX_RSYM(false, false);
        else OVX_RSYM(false, true);
#undef OVX_RSYM
    } else {
        const float inv_norm = 1.0f / norm;
        const float mean = ((float)d - 1.0f) / 2.0f;
        const size_t lds = (size_t)d * 4;
        if (vec)
            hipLaunchKernelGGL(k_rank_asym<true>, grid, dim3(256), lds, s, ranks, n, d, qhat, qsum, inv_norm, mean, per,
                               k, partial);
        else
            hipLaunchKernelGGL(k_rank_asym<false>, grid, dim3(256), lds, s, ranks, n, d, qhat, qsum, inv_norm, mean,
                               per, k, partial);
    }
    check_launch();
    hipLaunchKernelGGL(k_merge_topk, dim3((uint32_t)nq), dim3(256), 0, s, partial, nseg, k, nullptr, out_s, out_i);
    check_launch();
    ktime_end(kFamFullScan, s);
}

void launch_prepare_sum(const float* q, uint64_t nq, uint32_t d, float* qhat, float* qsum, uint32_t* qstat,
                        hipStream_t s) {
    if (nq == 0) return;
    // k_prepare's SIMD-tier offset with centre -1 and scale 1 is the plain
    // sequential sum of q_hat (from -0.0).
    ktime_begin(kFamPrepare, s);
    hipLaunchKernelGGL(k_prepare, dim3((uint32_t)nq), dim3(256), 0, s, q, nq, d, 1, -1.0f, 1.0f, qhat, qsum, qstat);
    check_launch();
    ktime_end(kFamPrepare, s);
}

void launch_check_rq_rows(const uint8_t* codes, uint64_t n, uint32_t d, uint32_t bits, unsigned long long* first_bad,
                          hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_check_rq_rows, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, codes, n, d, bits,
                       first_bad);
    check_launch();
}

void launch_check_rank_rows(const uint16_t* ranks, uint64_t n, uint32_t d, unsigned long long* first_bad,
                            hipStream_t s) {
    if (n == 0) return;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(n, 1u << 16);
    hipLaunchKernelGGL(k_check_rank_rows, dim3(grid), dim3(256), (size_t)((d + 31) / 32) * 4, s, ranks, n, d,
                       first_bad);
    check_launch();
}

void launch_interleave_rows(const uint64_t* rows, uint64_t cnt, uint64_t first_doc, uint32_t qpv, uint64_t* words,
                            uint32_t ntop, unsigned long long* first_bad, hipStream_t s) {
    if (cnt == 0) return;
    hipLaunchKernelGGL(k_interleave_rows, dim3((uint32_t)((cnt + 255) / 256)), dim3(256), 0, s, rows, cnt, first_doc,
                       qpv, words, ntop, first_bad);
    check_launch();
}

void launch_move_doc_interleaved(uint64_t* words, uint32_t qpv, uint64_t from, uint64_t to, hipStream_t s) {
    hipLaunchKernelGGL(k_move_doc_interleaved, dim3((qpv + 63) / 64), dim3(64), 0, s, words, qpv, from, to);
    check_launch();
}

void launch_fastscan_prepare(const float* qhat, uint64_t nq, uint32_t d, uint8_t* lut, float* bias_invq,
                             hipStream_t s) {
    if (nq == 0) return;
    ktime_begin(kFamFastScan, s);
    hipLaunchKernelGGL(k_fastscan_prepare, dim3((uint32_t)nq), dim3(256), 0, s, qhat, d, lut, bias_invq);
    check_launch();
    ktime_end(kFamFastScan, s);
}


struct LargeKWork {
    uint64_t *keys, *sorted, *offs;
    void* temp;
    size_t temp_bytes;
};

static LargeKWork carve_large_k(void* work, uint64_t total, uint64_t nq, hipStream_t s) {
    LargeKWork w{};
    char* c = static_cast<char*>(work);
    w.keys = reinterpret_cast<uint64_t*>(c);
    c += align256(total * 8);
    w.sorted = reinterpret_cast<uint64_t*>(c);
    c += align256(total * 8);
    w.offs = reinterpret_cast<uint64_t*>(c);
    c += align256((nq + 1) * 8);
    w.temp = c;
    w.temp_bytes = segsort_temp_bytes(total, (uint32_t)nq, nullptr, s);
    return w;
}

size_t large_k_work_bytes(uint64_t total, uint64_t nq, hipStream_t s) {
    return 2 * align256(total * 8) + align256((nq + 1) * 8) + segsort_temp_bytes(total, (uint32_t)nq, nullptr, s);
}

// Rows split across workgroups of about 4096 entries each.
static uint32_t large_k_chunks(uint64_t total, uint64_t nq) {
    const uint64_t avg = (total + nq - 1) / std::max<uint64_t>(1, nq);
    return (uint32_t)std::min<uint64_t>(65535, std::max<uint64_t>(1, (avg + 4095) / 4096));
}

void launch_rerank_large(const uint8_t* codes, uint32_t d, uint32_t bits, int tier, float inv_norm, const float* qhat,
                         const float* offset, uint64_t nq, const uint64_t* offsets, const uint32_t* cands,
                         uint64_t row_m, uint64_t total, uint64_t n_docs, uint32_t out_k, float* out_s,
                         int64_t* out_i, void* work, hipStream_t s) {
    if (nq == 0 || out_k == 0) return;
    if (nq > 65535) throw Error(kInternal, "large-k rerank: more than 65535 queries in one launch");
    LargeKWork w = carve_large_k(work, total, nq, s);
    const uint64_t* segoff = offsets;
    if (!offsets) {
        hipLaunchKernelGGL(k_row_offsets, dim3((uint32_t)std::min<uint64_t>(1024, (nq + 256) / 256)), dim3(256), 0, s,
                           w.offs, nq, row_m);
        check_launch();
        segoff = w.offs;
    }
    ktime_begin(kFamRerank, s);
    const dim3 grid(large_k_chunks(total, nq), (uint32_t)nq);
#define OVX_KEYS(T, B, _)                                                                                           \
    hipLaunchKernelGGL((k_score_keys<T, B>), grid, dim3(256), (size_t)d * 4, s, codes, n_docs, d, inv_norm, qhat, \
                       offset, offsets, cands, row_m, w.keys)
    OVX_TIER_BITS_DISPATCH(OVX_KEYS, 0);
#undef OVX_KEYS
    check_launch();
    segsort_keys(w.temp, w.temp_bytes, w.keys, w.sorted, total, (uint32_t)nq, segoff, s);
    hipLaunchKernelGGL(k_gather_keys, dim3((out_k + 255) / 256, (uint32_t)nq), dim3(256), 0, s, w.sorted, segoff,
                       row_m, out_k, 0u, out_s, out_i);
    check_launch();
    ktime_end(kFamRerank, s);
}

void launch_bm_subset_large(const uint64_t* words, uint32_t d, uint64_t n_docs, const uint64_t* qbits, uint64_t nq,
                            const uint64_t* offsets, const uint32_t* rows, uint64_t total, uint32_t ntop,
                            uint32_t out_k, float* out_s, int64_t* out_i, void* work, hipStream_t s) {
    if (nq == 0 || out_k == 0) return;
    if (nq > 65535) throw Error(kInternal, "large-k Bitmap subset: more than 65535 queries in one launch");
    LargeKWork w = carve_large_k(work, total, nq, s);
    const dim3 grid(large_k_chunks(total, nq), (uint32_t)nq);
    hipLaunchKernelGGL(k_bm_keys, grid, dim3(256), 0, s, words, d / 64, n_docs, qbits, offsets, rows, ntop, w.keys);
    check_launch();
    segsort_keys(w.temp, w.temp_bytes, w.keys, w.sorted, total, (uint32_t)nq, offsets, s);
    hipLaunchKernelGGL(k_gather_keys, dim3((out_k + 255) / 256, (uint32_t)nq), dim3(256), 0, s, w.sorted, offsets,
                       (uint64_t)0, out_k, ntop, out_s, out_i);
    check_launch();
}

void launch_fullscan_fs(const uint8_t* packed, uint64_t n, uint32_t d, const uint8_t* lut, const float* bias_invq,
                        float inv_norm, uint64_t nq, uint32_t k, void* work, float* out_s, int64_t* out_i,
                        hipStream_t s) {
    if (nq == 0 || k == 0) return;
    uint32_t nseg = fullscan_nseg(n, nq);
    uint64_t nblk = (n + 31) / 32;
    uint64_t bps = (nblk + nseg - 1) / nseg;
    uint64_t* partial = static_cast<uint64_t*>(work);
    size_t lds = (size_t)(d / 2) * 16;
    ktime_begin(kFamFastScan, s);
    hipLaunchKernelGGL(k_fullscan_fs, dim3(nseg, (uint32_t)nq), dim3(256), lds, s, packed, n, d, lut, bias_invq,
                       inv_norm, bps, k, partial);
    check_launch();
    hipLaunchKernelGGL(k_merge_topk, dim3((uint32_t)nq), dim3(256), 0, s, partial, nseg, k, nullptr, out_s, out_i);
    check_launch();
    ktime_end(kFamFastScan, s);
}

}  // namespace ovx
