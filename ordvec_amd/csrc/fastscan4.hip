// FastScan b=4 on CDNA4 and its batched multi-index form.
//
// Scoring is the b=2 FastScan recipe (src/fastscan.rs:130-173, 316-476)
// carried to one 4-bit rank code per coordinate: a per-query u8 LUT of 16
// entries per coordinate under one global affine scale, exact u32
// accumulation, score = (bias_sum + acc * inv_q) * (1/R) + 0.0, top-k by
// (score desc, id asc).
//
// Layout choices for the GPU:
//  * codes: the RankQuant b=4 rows (first code in the high nibble), regrouped
//    into 64-document blocks of interleaved dwords, packed[(blk*wpd + w)*64 +
//    lane] = dword w of document blk*64 + lane (wpd = d/8).  A wave reads one
//    dword of 64 documents per load (256 contiguous bytes).
//  * LUTs: four queries interleaved per entry, lut[((g*d + j)*16 + c)*4 + s] =
//    entry c of coordinate j for query s of group g.  One 32-bit LDS read
//    fetches the entry for four queries; v_dot4_u32_u8 against a one-hot byte
//    selector adds each query's byte to its own exact u32 accumulator.
//
// The scan kernel walks a table of tiles (index, query groups, 64-document
// block range).  Launched with one workgroup per tile it is the plain
// per-index path; launched persistent (a fixed grid pulling tiles from an
// atomic counter) it runs every index of a batch in one launch.
#include "common.h"
#include "device_common.h"

#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>

namespace ovx {

namespace {

constexpr uint32_t kFs4TkCap = 512;  // k (<= 256) + one round of 256 offers
using Fs4TopK = TopKS<kFs4TkCap>;

// Per query group of up to four queries: min/max of the f32 LUT under the
// global affine quantisation, then the u8 entries in the interleaved layout.
__global__ __launch_bounds__(256) void k_fs4_lut(const float* __restrict__ qhat, uint32_t d,
                                                 const Fs4Group* __restrict__ groups, uint8_t* __restrict__ lut,
                                                 float* __restrict__ bias_invq) {
    __shared__ float smin[256], smax[256];
    const uint64_t g = blockIdx.x;
    const Fs4Group grp = groups[g];
    uint8_t* lg = lut + g * (uint64_t)d * 64;
    for (uint32_t s = 0; s < 4; s++) {
        if (s >= grp.count) {
            for (uint32_t t = threadIdx.x; t < d * 16; t += blockDim.x) lg[t * 4 + s] = 0;
            continue;
        }
        const uint64_t q = grp.q0 + s;
        const float* x = qhat + q * d;
        float mn = INFINITY, mx = -INFINITY;
        for (uint32_t t = threadIdx.x; t < d * 16; t += blockDim.x) {
            const float v = __fmul_rn(x[t >> 4], (float)(t & 15) - 7.5f);
            mn = fminf(mn, v);
            mx = fmaxf(mx, v);
        }
        smin[threadIdx.x] = mn;
        smax[threadIdx.x] = mx;
        __syncthreads();
        for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h) {
                smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + h]);
                smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + h]);
            }
            __syncthreads();
        }
        const float gmin = smin[0], gmax = smax[0];
        __syncthreads();
        float span = __fadd_rn(gmax, -gmin);
        if (!(span >= 1e-12f)) span = 1e-12f;
        const float scale = __fdiv_rn(255.0f, span);
        if (threadIdx.x == 0) {
            bias_invq[q * 2 + 0] = __fmul_rn((float)d, gmin);
            bias_invq[q * 2 + 1] = __fdiv_rn(1.0f, scale);
        }
        for (uint32_t t = threadIdx.x; t < d * 16; t += blockDim.x) {
            const float v = __fmul_rn(x[t >> 4], (float)(t & 15) - 7.5f);
            float r = roundf(__fmul_rn(__fadd_rn(v, -gmin), scale));
            r = fminf(fmaxf(r, 0.0f), 255.0f);
            lg[t * 4 + s] = (uint8_t)r;
        }
    }
}

typedef const __attribute__((address_space(1))) uint32_t gu32_t;

// The 8 coordinates of one code dword (w): coordinate 8w + c sits in byte c/2,
// high nibble first.  One 32-bit LDS read per coordinate and group fetches
// the entries of four queries; each dot4 adds one query's byte.
template <int QG>
__device__ __forceinline__ void fs4_dword(uint32_t v, uint32_t w, const uint32_t* slut, uint32_t d, uint32_t* acc) {
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t nib = __builtin_amdgcn_ubfe(v, (c >> 1) * 8 + ((c & 1) ? 0 : 4), 4);
        const uint32_t* row = slut + (8 * w + c) * 16 + nib;
#pragma unroll
        for (int g = 0; g < QG; g++) {
            const uint32_t e = row[g * d * 16];
            acc[4 * g + 0] = __builtin_amdgcn_udot4(e, 0x00000001u, acc[4 * g + 0], false);
            acc[4 * g + 1] = __builtin_amdgcn_udot4(e, 0x00000100u, acc[4 * g + 1], false);
            acc[4 * g + 2] = __builtin_amdgcn_udot4(e, 0x00010000u, acc[4 * g + 2], false);
            acc[4 * g + 3] = __builtin_amdgcn_udot4(e, 0x01000000u, acc[4 * g + 3], false);
        }
    }
}

// Tiles: QG groups (4*QG query slots) x a 64-document block range of one
// index.  counter == nullptr: workgroup b runs tile b once; otherwise
// workgroups pull tiles until the counter passes ntiles.  Each slot's top-kcap
// keys (score_key, index-local id) go to partial[(tile*4*QG + slot)*kcap ...].
template <int QG, int CH>
__global__ __launch_bounds__(256) void k_fs4_scan(const Fs4IndexDesc* __restrict__ indexes,
                                                  const Fs4Tile* __restrict__ tiles, uint32_t ntiles,
                                                  uint32_t* __restrict__ counter, const Fs4Group* __restrict__ groups,
                                                  const uint8_t* __restrict__ lut, const float* __restrict__ bias_invq,
                                                  uint32_t d, float inv_norm, uint32_t kcap,
                                                  uint64_t* __restrict__ partial) {
    constexpr int QT = 4 * QG;
    extern __shared__ __align__(16) uint64_t smem[];
    Fs4TopK* tk = reinterpret_cast<Fs4TopK*>(smem);
    uint32_t* slut = reinterpret_cast<uint32_t*>(tk + QT);
    __shared__ uint32_t s_tile;
    __shared__ float s_bias[QT], s_invq[QT];
    __shared__ uint32_t s_ok[QT];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t wpd = d / 8;
    for (uint32_t iter = 0;; iter++) {
        if (counter) {
            if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
            __syncthreads();
        } else if (threadIdx.x == 0) {
            s_tile = iter == 0 ? blockIdx.x : ntiles;
        }
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= ntiles) break;
        const Fs4Tile tile = tiles[t];
        // Stage the tile's LUTs (missing groups read as zeros) and slot constants.
        const uint4* src = reinterpret_cast<const uint4*>(lut + (uint64_t)tile.g0 * d * 64);
        uint4* dst = reinterpret_cast<uint4*>(slut);
        const uint32_t have = tile.ng * d * 4, total = QG * d * 4;
        for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) dst[i] = i < have ? src[i] : make_uint4(0, 0, 0, 0);
        if (threadIdx.x < QT) {
            const uint32_t g = threadIdx.x >> 2, s = threadIdx.x & 3;
            const bool ok = g < tile.ng && s < groups[tile.g0 + g].count;
            const uint64_t q = ok ? (uint64_t)groups[tile.g0 + g].q0 + s : 0;
            s_ok[threadIdx.x] = ok;
            s_bias[threadIdx.x] = ok ? bias_invq[q * 2 + 0] : 0.0f;
            s_invq[threadIdx.x] = ok ? bias_invq[q * 2 + 1] : 0.0f;
        }
        for (int s = 0; s < QT; s++) tk_init(tk[s], kcap);  // synchronises
        const Fs4IndexDesc ix = indexes[tile.index];
        for (uint32_t b = tile.blk0; b < tile.blk1; b += 4) {
            const uint32_t blk = b + wave;
            const uint64_t doc = (uint64_t)blk * 64 + lane;
            const bool valid = blk < tile.blk1 && doc < ix.n;
            uint32_t acc[QT];
#pragma unroll
            for (int s = 0; s < QT; s++) acc[s] = 0;
            if (blk < tile.blk1) {
                // Global (not flat) loads, CH dwords in flight ahead of the chunk
                // being scored (CH divides wpd).
                gu32_t* base = (gu32_t*)(ix.packed + (uint64_t)blk * wpd * 64 + lane);
                uint32_t cur[CH], nxt[CH];
#pragma unroll
                for (int i = 0; i < CH; i++) cur[i] = base[(uint64_t)i * 64];
                for (uint32_t w0 = 0; w0 < wpd; w0 += CH) {
                    // next chunk (the last one again at the end: loads stay unconditional)
                    const uint32_t w1 = min(w0 + CH, wpd - CH);
#pragma unroll
                    for (int i = 0; i < CH; i++) nxt[i] = base[(uint64_t)(w1 + i) * 64];
#pragma unroll
                    for (int i = 0; i < CH; i++) fs4_dword<QG>(cur[i], w0 + i, slut, d, acc);
#pragma unroll
                    for (int i = 0; i < CH; i++) cur[i] = nxt[i];
                }
            }
            for (int s = 0; s < QT; s++) tk_reserve(tk[s], kcap, blockDim.x);
#pragma unroll
            for (int s = 0; s < QT; s++) {
                if (valid && s_ok[s]) {
                    const float raw = __fadd_rn(s_bias[s], __fmul_rn((float)acc[s], s_invq[s]));
                    tk_offer(tk[s], kcap, score_key(__fadd_rn(__fmul_rn(raw, inv_norm), 0.0f), (uint32_t)doc));
                }
            }
        }
        __syncthreads();
        for (int s = 0; s < QT; s++) {
            tk_flush(tk[s], kcap);
            uint64_t* out = partial + ((uint64_t)t * QT + s) * kcap;
            for (uint32_t j = threadIdx.x; j < kcap; j += blockDim.x) out[j] = tk[s].keys[j];
        }
        __syncthreads();
    }
}

// ---- Matrix-core form (tiles of 16 queries) ----------------------------------
// The exact integer sum  acc = sum_j lut[j][code_j]  is a dot product over
// K = (coordinate, code) pairs: documents give one-hot rows (1.0 at the code
// of each coordinate), queries give their LUT rows.  A u8 LUT entry is split
// into base-16 digits lo + 16 hi (each 0..15, exact in FP8 e4m3), so one
// v_mfma_scale_f32_32x32x64_f8f6f4 with FP4 documents (A: 32 documents x 64 K)
// and FP8 digits (B: 64 K x 32 columns = 16 queries x {lo, hi}) covers four
// coordinates.  Every product is a digit and every partial sum an integer
// below 768 * 15, so the f32 accumulation is exact and lo + 16 hi equals the
// u32 sum of the VALU form bit for bit.
// K layout of one instruction (lane l, K half h = l / 32): FP4 document slot
// s (nibble s % 8 of dword s / 8) is coordinate 4 ks + 2 h + s / 16, code
// s % 16.  The two formats split K differently over the lane halves (FP8 byte
// b of half h is K = 32 h + b; FP4 slot s of half h is K = 16 (2 (s / 16) + h)
// + s % 16, measured), so the FP8 byte b of half h carries coordinate
// 4 ks + 2 (b / 16) + h, code b % 16.  B fragments ([query set][ks][half of
// the 32 bytes][lane][16 bytes]) are built once per batch from the u8 LUTs
// (k_fs4_frag) and staged through LDS 16 K-steps
// (32 KiB) at a time; each wave scores four 32-document groups (two 64-doc
// blocks) against every staged K-step.
constexpr uint32_t kFs4MfmaThreads = 512;
constexpr uint32_t kFs4MfmaQT = 16;
constexpr uint32_t kFs4ChunkKs = 16;                           // K-steps per staged chunk
constexpr uint32_t kFs4ChunkU4 = kFs4ChunkKs * 64 * 2;         // uint4 per chunk (32 KiB)
typedef int fs4_v8i __attribute__((ext_vector_type(8)));
typedef float fs4_v16f __attribute__((ext_vector_type(16)));

// e4m3 encodings of the integers 0..15 (all exact).
__constant__ uint8_t kE4m3Int[16] = {0x00, 0x38, 0x40, 0x44, 0x48, 0x4A, 0x4C, 0x4E,
                                     0x50, 0x51, 0x52, 0x53, 0x54, 0x55, 0x56, 0x57};

// Fragments of query set qs = groups [g0, g0 + gn) (up to 16 queries): block
// (qs, ks), thread t writes dwords t and t + 256 of the 512 (2 halves x 64
// lanes x 4).
__global__ __launch_bounds__(256) void k_fs4_frag(const uint8_t* __restrict__ lut, const Fs4Group* __restrict__ qsets,
                                                  uint32_t d, uint32_t* __restrict__ frag) {
    const uint32_t nks = d / 4;
    const uint32_t qs = blockIdx.x / nks, ks = blockIdx.x % nks;
    const Fs4Group set = qsets[qs];
    // K-step layout [half][lane][16 bytes]: a ds_read_b128 of one half reads
    // 16 consecutive bytes per lane (no LDS bank conflicts)
    uint32_t* out = frag + ((uint64_t)qs * nks + ks) * 512;
    for (uint32_t e = threadIdx.x; e < 512; e += blockDim.x) {
        const uint32_t half = e >> 8, lane = (e >> 2) & 63, dw = 4 * half + (e & 3);
        const uint32_t col = lane & 31, h = lane >> 5;
        const uint32_t slot = col & 15, hd = col >> 4, gi = slot >> 2, sq = slot & 3;
        uint32_t word = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            const uint32_t b = dw * 4 + i;
            // FP8 byte b of K half h is K = 32 h + b, the FP4 document slot
            // that pairs with it (16-wide K blocks alternate between the
            // halves for FP4): coordinate 4 ks + 2 (b / 16) + h, code b % 16
            const uint32_t j = 4 * ks + 2 * (b >> 4) + h, c = b & 15;
            uint32_t v = 0;
            if (gi < set.count) v = lut[(((uint64_t)(set.q0 + gi) * d + j) * 16 + c) * 4 + sq];
            word |= (uint32_t)kE4m3Int[hd ? (v >> 4) : (v & 15)] << (8 * i);
        }
        out[e] = word;
    }
}

// Make room for r offers in each of the n top-k lists (one barrier unless a
// list must flush).
template <class T>
__device__ void tk_reserve_all(T* tk, uint32_t n, uint32_t k, uint32_t r) {
    __syncthreads();
    for (uint32_t s = 0; s < n; s++)
        if (k + tk[s].nbuf + r > T::kCap) tk_flush(tk[s], k);
}

// Tiles (index, query set of up to 16 queries = tile.pad, 64-doc blocks
// [blk0, blk1)); counter as k_fs4_scan.  Outputs as k_fs4_scan with 16 slots.
__global__ __launch_bounds__(kFs4MfmaThreads, 1) void k_fs4_scan_mfma(
    const Fs4IndexDesc* __restrict__ indexes, const Fs4Tile* __restrict__ tiles, uint32_t ntiles,
    uint32_t* __restrict__ counter, const Fs4Group* __restrict__ groups, const uint4* __restrict__ frag,
    const float* __restrict__ bias_invq, uint32_t d, float inv_norm, uint32_t kcap, uint64_t* __restrict__ partial) {
    extern __shared__ __align__(16) uint64_t smem[];
    Fs4TopK* tk = reinterpret_cast<Fs4TopK*>(smem);
    uint4* sB = reinterpret_cast<uint4*>(tk + kFs4MfmaQT);
    __shared__ uint32_t s_tile;
    __shared__ float s_bias[kFs4MfmaQT], s_invq[kFs4MfmaQT];
    __shared__ uint32_t s_ok[kFs4MfmaQT];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
    const uint32_t wpd = d / 8, nks = d / 4, nchunk = nks / kFs4ChunkKs;
    for (uint32_t iter = 0;; iter++) {
        if (counter) {
            if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
            __syncthreads();
        } else if (threadIdx.x == 0) {
            s_tile = iter == 0 ? blockIdx.x : ntiles;
        }
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= ntiles) break;
        const Fs4Tile tile = tiles[t];
        if (threadIdx.x < kFs4MfmaQT) {
            const uint32_t g = threadIdx.x >> 2, s = threadIdx.x & 3;
            const bool ok = g < tile.ng && s < groups[tile.g0 + g].count;
            const uint64_t q = ok ? (uint64_t)groups[tile.g0 + g].q0 + s : 0;
            s_ok[threadIdx.x] = ok;
            s_bias[threadIdx.x] = ok ? bias_invq[q * 2 + 0] : 0.0f;
            s_invq[threadIdx.x] = ok ? bias_invq[q * 2 + 1] : 0.0f;
        }
        for (uint32_t s = 0; s < kFs4MfmaQT; s++) tk_init(tk[s], kcap);  // synchronises
        const Fs4IndexDesc ix = indexes[tile.index];
        const uint4* fq = frag + (uint64_t)tile.pad * nks * 128;  // 128 uint4 (2 KiB) per K-step
        for (uint32_t b0 = tile.blk0; b0 < tile.blk1; b0 += 16) {
            // this wave's blocks b0 + 2 wave + {0, 1}; an invalid block reads
            // the tile's first block (in bounds) and offers nothing
            uint32_t blk[2];
            bool bval[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const uint32_t b = b0 + 2 * wave + i;
                bval[i] = b < tile.blk1;
                blk[i] = bval[i] ? b : tile.blk0;
            }
            // dword w of group g's document: dp[g][w * 64]
            const uint32_t* dp[4];
#pragma unroll
            for (int g = 0; g < 4; g++) dp[g] = ix.packed + (uint64_t)blk[g >> 1] * wpd * 64 + 32 * (g & 1) + col;
            fs4_v16f acc[4];
#pragma unroll
            for (int g = 0; g < 4; g++)
#pragma unroll
                for (int v = 0; v < 16; v++) acc[g][v] = 0.0f;
            uint4 stage[4];
            uint32_t xn[4][8];
#pragma unroll
            for (int i = 0; i < 4; i++) stage[i] = fq[threadIdx.x + i * kFs4MfmaThreads];
#pragma unroll
            for (int g = 0; g < 4; g++)
#pragma unroll
                for (int i = 0; i < 8; i++) xn[g][i] = dp[g][(uint64_t)i * 64];
            for (uint32_t c = 0; c < nchunk; c++) {
                __syncthreads();  // every wave is done with the previous chunk
#pragma unroll
                for (int i = 0; i < 4; i++) sB[threadIdx.x + i * kFs4MfmaThreads] = stage[i];
                uint32_t xw[4][8];
#pragma unroll
                for (int g = 0; g < 4; g++)
#pragma unroll
                    for (int i = 0; i < 8; i++) xw[g][i] = xn[g][i] >> (8 * h);
                __syncthreads();
                // the next chunk (this one again after the last: in bounds, unused)
                const uint32_t cn = c + 1 < nchunk ? c + 1 : c;
#pragma unroll
                for (int i = 0; i < 4; i++) stage[i] = fq[(uint64_t)cn * kFs4ChunkU4 + threadIdx.x + i * kFs4MfmaThreads];
#pragma unroll
                for (int g = 0; g < 4; g++)
#pragma unroll
                    for (int i = 0; i < 8; i++) xn[g][i] = dp[g][(uint64_t)(cn * 8 + i) * 64];
#pragma unroll
                for (int ks = 0; ks < (int)kFs4ChunkKs; ks++) {
                    const uint4 q0 = sB[ks * 128 + lane], q1 = sB[ks * 128 + 64 + lane];
                    const fs4_v8i bq = {(int)q0.x, (int)q0.y, (int)q0.z, (int)q0.w,
                                        (int)q1.x, (int)q1.y, (int)q1.z, (int)q1.w};
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        // byte 2 (ks % 2) + h of dword ks / 2: codes of coordinates
                        // 4 ks + 2 h (high nibble) and 4 ks + 2 h + 1 (low nibble)
                        const uint32_t x = xw[g][ks >> 1] >> (16 * (ks & 1));
                        const uint32_t s1 = (x >> 2) & 0x3cu, s2 = (x << 2) & 0x3cu;
                        const uint64_t o1 = 2ull << s1, o2 = 2ull << s2;  // FP4 1.0 at nibble = code
                        const fs4_v8i a = {(int)(uint32_t)o1, (int)(uint32_t)(o1 >> 32), (int)(uint32_t)o2,
                                           (int)(uint32_t)(o2 >> 32), 0, 0, 0, 0};
                        acc[g] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, bq, acc[g], 4, 0, 0, 0x7f7f7f7f, 0,
                                                                                 0x7f7f7f7f);
                    }
                }
            }
            // lanes with column < 16 hold query col's low digit sums; column
            // col + 16 (lane ^ 16) the high digit sums
#pragma unroll
            for (int g = 0; g < 4; g++) {
                tk_reserve_all(tk, kFs4MfmaQT, kcap, 256);
                const uint32_t b = blk[g >> 1];
                const bool gv = bval[g >> 1];
                const bool mine = col < 16 && s_ok[col & 15];
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const float hi = __shfl_xor(acc[g][v], 16);
                    const uint64_t doc = (uint64_t)b * 64 + 32 * (g & 1) + 8 * (v / 4) + 4 * h + (v % 4);
                    if (mine && gv && doc < ix.n) {
                        const float total = __fadd_rn(acc[g][v], __fmul_rn(16.0f, hi));  // exact integer
                        const float raw = __fadd_rn(s_bias[col], __fmul_rn(total, s_invq[col]));
                        tk_offer(tk[col], kcap, score_key(__fadd_rn(__fmul_rn(raw, inv_norm), 0.0f), (uint32_t)doc));
                    }
                }
            }
        }
        __syncthreads();
        for (uint32_t s = 0; s < kFs4MfmaQT; s++) {
            tk_flush(tk[s], kcap);
            uint64_t* out = partial + ((uint64_t)t * kFs4MfmaQT + s) * kcap;
            for (uint32_t j = threadIdx.x; j < kcap; j += blockDim.x) out[j] = tk[s].keys[j];
        }
        __syncthreads();
    }
}

// Per query: merge its segments' top-kcap lists into the final top-k row
// (row stride kout; entries past k padded with -inf / -1).
__global__ __launch_bounds__(256) void k_fs4_merge(const uint64_t* __restrict__ partial, const Fs4Merge* __restrict__ m,
                                                   uint32_t kcap, uint32_t kout, float* __restrict__ out_s,
                                                   int64_t* __restrict__ out_i) {
    __shared__ TopKShared tk;
    const uint64_t q = blockIdx.x;
    const Fs4Merge mq = m[q];
    const uint32_t k = mq.k;
    if (k == 0 || mq.nseg == 0) {
        for (uint32_t j = threadIdx.x; j < kout; j += blockDim.x) {
            out_s[q * kout + j] = -INFINITY;
            out_i[q * kout + j] = -1;
        }
        return;
    }
    tk_init(tk, k);
    const uint64_t total = (uint64_t)mq.nseg * kcap;
    for (uint64_t r0 = 0; r0 < total; r0 += blockDim.x) {
        tk_reserve(tk, k, blockDim.x);
        const uint64_t i = r0 + threadIdx.x;
        if (i < total) {
            const uint64_t seg = i / kcap, j = i - seg * kcap;
            const uint64_t key = partial[((uint64_t)mq.prow + seg * mq.stride) * kcap + j];
            if (key != kKeyMax) tk_offer(tk, k, key);
        }
    }
    __syncthreads();
    tk_flush(tk, k);
    for (uint32_t j = threadIdx.x; j < kout; j += blockDim.x) {
        const uint64_t key = j < k ? tk.keys[j] : kKeyMax;
        out_s[q * kout + j] = key == kKeyMax ? -INFINITY : key_score(key);
        out_i[q * kout + j] = key == kKeyMax ? -1 : (int64_t)(key & 0xffffffffu);
    }
}

// RankQuant b=4 rows -> interleaved dword blocks (tail lanes of the last
// block are left as the caller cleared them).
__global__ void k_fs4_pack(const uint8_t* __restrict__ codes, uint64_t n, uint32_t d, uint32_t* __restrict__ out) {
    const uint32_t wpd = d / 8;
    const uint64_t total = n * wpd;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t doc = t / wpd;
        const uint32_t w = (uint32_t)(t - doc * wpd);
        out[((doc >> 6) * wpd + w) * 64 + (doc & 63)] = reinterpret_cast<const uint32_t*>(codes + doc * (d / 2))[w];
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t count, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = v;
}

void check() { OVX_HIP(hipGetLastError()); }

}  // namespace

size_t fs4_scan_lds_bytes(uint32_t qg, uint32_t d) {
    if (qg == 4) return (size_t)kFs4MfmaQT * sizeof(Fs4TopK) + (size_t)kFs4ChunkU4 * 16;
    return (size_t)4 * qg * sizeof(Fs4TopK) + (size_t)qg * d * 64;
}

bool fs4_mfma_ok(uint32_t d) { return d % (4 * kFs4ChunkKs) == 0 && d <= kFs4MaxDim; }

size_t fs4_frag_bytes(uint32_t d) { return (size_t)d * 512; }

void launch_fs4_frag(const uint8_t* lut, const Fs4Group* qsets, uint32_t nqsets, uint32_t d, uint32_t* frag,
                     hipStream_t s) {
    if (nqsets == 0) return;
    ktime_begin(kFamFastScan, s);
    hipLaunchKernelGGL(k_fs4_frag, dim3(nqsets * (d / 4)), dim3(256), 0, s, lut, qsets, d, frag);
    check();
    ktime_end(kFamFastScan, s);
}

void fs4_configure_lds() {
    static std::once_flag done;
    std::call_once(done, [] {
    // Larger dynamic LDS than the default 64 KiB (gfx950 has 160 KiB per CU).
#define OVX_FS4_LDS(QG, CH)                                                                       \
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fs4_scan<QG, CH>),                \
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)
    OVX_FS4_LDS(1, 8);
    OVX_FS4_LDS(1, 4);
    OVX_FS4_LDS(1, 2);
    OVX_FS4_LDS(1, 1);
    OVX_FS4_LDS(2, 8);
    OVX_FS4_LDS(2, 4);
    OVX_FS4_LDS(2, 2);
    OVX_FS4_LDS(2, 1);
#undef OVX_FS4_LDS
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fs4_scan_mfma),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
    });
}

void launch_fs4_pack(const uint8_t* codes, uint64_t n, uint32_t d, uint32_t* packed, hipStream_t s) {
    if (n == 0) return;
    const uint64_t total = n * (d / 8);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(65536, (total + 255) / 256);
    ktime_begin(kFamEncode, s);
    hipLaunchKernelGGL(k_fs4_pack, dim3(grid), dim3(256), 0, s, codes, n, d, packed);
    check();
    ktime_end(kFamEncode, s);
}

void launch_fs4_lut(const float* qhat, uint32_t d, const Fs4Group* groups, uint32_t ngroups, uint8_t* lut,
                    float* bias_invq, hipStream_t s) {
    if (ngroups == 0) return;
    ktime_begin(kFamFastScan, s);
    hipLaunchKernelGGL(k_fs4_lut, dim3(ngroups), dim3(256), 0, s, qhat, d, groups, lut, bias_invq);
    check();
    ktime_end(kFamFastScan, s);
}

void launch_fs4_scan(uint32_t qg, uint32_t grid, const Fs4IndexDesc* indexes, const Fs4Tile* tiles, uint32_t ntiles,
                     uint32_t* counter, const Fs4Group* groups, const uint8_t* lut, const float* bias_invq, uint32_t d,
                     float inv_norm, uint32_t kcap, uint64_t* partial, hipStream_t s) {
    if (ntiles == 0) return;
    fs4_configure_lds();
    const size_t lds = fs4_scan_lds_bytes(qg, d);
    ktime_begin(kFamFastScan, s);
    if (counter) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, counter, 1u, 0u);
        check();
    }
    if (qg == 4) {  // matrix-core tiles; `lut` holds the B fragments
        hipLaunchKernelGGL(k_fs4_scan_mfma, dim3(grid), dim3(kFs4MfmaThreads), lds, s, indexes, tiles, ntiles, counter,
                           groups, reinterpret_cast<const uint4*>(lut), bias_invq, d, inv_norm, kcap, partial);
        check();
        ktime_end(kFamFastScan, s);
        return;
    }
#define OVX_FS4_SCAN(QG, CH)                                                                                    \
    hipLaunchKernelGGL((k_fs4_scan<QG, CH>), dim3(grid), dim3(256), lds, s, indexes, tiles, ntiles, counter, groups, \
                       lut, bias_invq, d, inv_norm, kcap, partial)
    const uint32_t wpd = d / 8;
    if (qg == 2) {
        if (wpd % 8 == 0) OVX_FS4_SCAN(2, 8);
        else if (wpd % 4 == 0) OVX_FS4_SCAN(2, 4);
        else if (wpd % 2 == 0) OVX_FS4_SCAN(2, 2);
        else OVX_FS4_SCAN(2, 1);
    } else {
        if (wpd % 8 == 0) OVX_FS4_SCAN(1, 8);
        else if (wpd % 4 == 0) OVX_FS4_SCAN(1, 4);
        else if (wpd % 2 == 0) OVX_FS4_SCAN(1, 2);
        else OVX_FS4_SCAN(1, 1);
    }
#undef OVX_FS4_SCAN
    check();
    ktime_end(kFamFastScan, s);
}

void launch_fs4_merge(const uint64_t* partial, const Fs4Merge* merges, uint32_t nq, uint32_t kcap,
                      uint32_t kout, float* out_s, int64_t* out_i, hipStream_t s) {
    if (nq == 0) return;
    ktime_begin(kFamFastScan, s);
    hipLaunchKernelGGL(k_fs4_merge, dim3(nq), dim3(256), 0, s, partial, merges, kcap, kout, out_s, out_i);
    check();
    ktime_end(kFamFastScan, s);
}

}  // namespace ovx
