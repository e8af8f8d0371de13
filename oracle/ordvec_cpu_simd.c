/*
 * ordvec CPU oracle, AVX-512 tiers -- TEST INFRASTRUCTURE ONLY.
 *
 * The vector forms of two oracle loops, dispatched at run time the way the
 * reference dispatches its own kernels (src/quant.rs:300-329,
 * src/lib.rs:198-207): used when the host CPU has them, so the CPU baseline
 * timed beside the GPU runs the reference's AVX-512 path, not a scalar one.
 * Results are identical to the scalar restatements in ordvec_oracle.c (integer
 * Hamming sums; the same f32 operation order), which tests/test_oracle.py
 * checks on hosts that have the instructions.
 *
 *  - orc_simd_scan_block: sign_scan_collect_batched_avx512vpop
 *    (src/sign_bitmap.rs:845-955): eight queries per pass over a document,
 *    64-bit XOR + VPOPCNTQ per 512-bit word group, masked tail words, one
 *    transpose-and-add fold of the eight accumulators per document.
 *  - orc_simd_score_avx512: scan_b2/b4_asym_avx512 (src/quant_kernels.rs:
 *    406-581) for one document: 16 codes per chunk decoded by variable
 *    shifts of the big-endian chunk word, chunk c into accumulator c % 4,
 *    explicit FMA, (acc0 + acc1) + (acc2 + acc3), then the halving tree of
 *    _mm512_reduce_add_ps (8, 4, 2, 1 lanes).
 */
#include <immintrin.h>
#include <stdint.h>

#define SIMD_TARGET __attribute__((target("avx512f,avx512dq,avx512vpopcntdq")))

static int g_simd_state = -1;  /* -1 unknown, 0 off, 1 on */
static int g_simd_forced_off = 0;

int orc_simd_supported(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq") &&
           __builtin_cpu_supports("avx512vpopcntdq");
}

/* enable < 0: query; 0: force the scalar restatements; 1: use AVX-512 when
 * the CPU has it.  Returns whether the AVX-512 forms are in use. */
int orc_set_simd(int enable) {
    if (enable == 0) g_simd_forced_off = 1;
    if (enable == 1) g_simd_forced_off = 0;
    if (g_simd_state < 0) g_simd_state = orc_simd_supported();
    return g_simd_state && !g_simd_forced_off;
}

int orc_simd_active(void) { return orc_set_simd(-1); }

/* Fold eight u64x8 accumulators to their eight lane sums, in order. */
SIMD_TARGET static inline __m512i fold8(const __m512i a[8]) {
    /* pairs: [a_j(0+1), a_j+1(0+1), a_j(2+3), a_j+1(2+3), ...] */
    const __m512i s01 = _mm512_add_epi64(_mm512_unpacklo_epi64(a[0], a[1]), _mm512_unpackhi_epi64(a[0], a[1]));
    const __m512i s23 = _mm512_add_epi64(_mm512_unpacklo_epi64(a[2], a[3]), _mm512_unpackhi_epi64(a[2], a[3]));
    const __m512i s45 = _mm512_add_epi64(_mm512_unpacklo_epi64(a[4], a[5]), _mm512_unpackhi_epi64(a[4], a[5]));
    const __m512i s67 = _mm512_add_epi64(_mm512_unpacklo_epi64(a[6], a[7]), _mm512_unpackhi_epi64(a[6], a[7]));
    /* 128-bit lanes 0+1 and 2+3 of each pair vector: quarter sums */
    const __m512i u03 = _mm512_add_epi64(_mm512_shuffle_i64x2(s01, s23, 0x88), _mm512_shuffle_i64x2(s01, s23, 0xDD));
    const __m512i u47 = _mm512_add_epi64(_mm512_shuffle_i64x2(s45, s67, 0x88), _mm512_shuffle_i64x2(s45, s67, 0xDD));
    /* halves: [a0, a1, a2, a3, a4, a5, a6, a7] */
    return _mm512_add_epi64(_mm512_shuffle_i64x2(u03, u47, 0x88), _mm512_shuffle_i64x2(u03, u47, 0xDD));
}

/* One document against eight queries (q: 8 rows of qpv words): eight
 * accumulators kept in registers, the query words read from L1. */
#define ORC_ACC8(OP) OP(0) OP(1) OP(2) OP(3) OP(4) OP(5) OP(6) OP(7)

SIMD_TARGET static inline void scan_doc8(const uint64_t *row, uint32_t qpv, const uint64_t *q, uint32_t *out,
                                         uint64_t out_stride) {
    const uint32_t lanes = qpv / 8, rem = qpv % 8;
    __m512i a[8];
#define ORC_ZERO(t) a[t] = _mm512_setzero_si512();
    ORC_ACC8(ORC_ZERO)
#undef ORC_ZERO
    for (uint32_t l = 0; l < lanes; l++) {
        const __m512i dw = _mm512_loadu_si512((const void *)(row + (uint64_t)l * 8));
#define ORC_ACC(t)                                                                                   \
    a[t] = _mm512_add_epi64(a[t], _mm512_popcnt_epi64(_mm512_xor_si512(                             \
                                      dw, _mm512_loadu_si512((const void *)(q + (uint64_t)(t) * qpv + \
                                                                            (uint64_t)l * 8)))));
        ORC_ACC8(ORC_ACC)
#undef ORC_ACC
    }
    if (rem) {
        const __mmask8 m = (__mmask8)((1u << rem) - 1u);
        const __m512i dw = _mm512_maskz_loadu_epi64(m, row + (uint64_t)lanes * 8);
#define ORC_TAIL(t)                                                                               \
    a[t] = _mm512_add_epi64(a[t], _mm512_popcnt_epi64(_mm512_xor_si512(                          \
                                      dw, _mm512_maskz_loadu_epi64(m, q + (uint64_t)(t) * qpv + \
                                                                          (uint64_t)lanes * 8))));
        ORC_ACC8(ORC_TAIL)
#undef ORC_TAIL
    }
    uint64_t sums[8];
    _mm512_storeu_si512((void *)sums, fold8(a));
    for (int t = 0; t < 8; t++) out[(uint64_t)t * out_stride] = (uint32_t)sums[t];
}

/* scores[t * bn + d] = popcount(doc_d XOR query_t) for tq queries, eight
 * queries per pass over the block (a partial last group pads its query rows
 * with the group's first row and drops those sums). */
SIMD_TARGET void orc_simd_scan_block(const uint64_t *block, uint64_t bn, uint32_t qpv, const uint64_t *qb,
                                     uint64_t tq, uint32_t *scores) {
    uint64_t pad[8 * 64];  /* eight query rows of up to 64 words (dim <= 4096) */
    uint32_t junk[8];
    for (uint64_t t0 = 0; t0 < tq; t0 += 8) {
        const uint64_t tc = tq - t0 < 8 ? tq - t0 : 8;
        const uint64_t *q = qb + t0 * qpv;
        if (tc < 8 || qpv > 64) {
            if (qpv > 64) {  /* wide rows: one query at a time through the same kernel */
                for (uint64_t t = 0; t < tc; t++)
                    for (uint64_t d = 0; d < bn; d++) {
                        const uint64_t *row = block + d * qpv, *qq = qb + (t0 + t) * qpv;
                        uint32_t h = 0;
                        for (uint32_t w = 0; w < qpv; w++) h += (uint32_t)__builtin_popcountll(row[w] ^ qq[w]);
                        scores[(t0 + t) * bn + d] = h;
                    }
                continue;
            }
            for (uint64_t t = 0; t < 8; t++)
                for (uint32_t w = 0; w < qpv; w++) pad[t * qpv + w] = q[(t < tc ? t : 0) * qpv + w];
            for (uint64_t d = 0; d < bn; d++) {
                scan_doc8(block + d * qpv, qpv, pad, junk, 1);
                for (uint64_t t = 0; t < tc; t++) scores[(t0 + t) * bn + d] = junk[t];
            }
            continue;
        }
        for (uint64_t d = 0; d < bn; d++) scan_doc8(block + d * qpv, qpv, q, scores + t0 * bn + d, bn);
    }
}

/* Raw asymmetric dot product of one b=2 or b=4 row (dim % 64 == 0) times
 * scale, in the AVX-512 tier's order. */
SIMD_TARGET float orc_simd_score_avx512(const uint8_t *row, uint32_t d, uint32_t bits, const float *q,
                                        float scale) {
    __m512 acc[4] = {_mm512_setzero_ps(), _mm512_setzero_ps(), _mm512_setzero_ps(), _mm512_setzero_ps()};
    const uint32_t chunks = d / 16;
    if (bits == 2) {
        const __m512i sh = _mm512_setr_epi32(30, 28, 26, 24, 22, 20, 18, 16, 14, 12, 10, 8, 6, 4, 2, 0);
        const __m512i m3 = _mm512_set1_epi32(3);
        for (uint32_t c = 0; c < chunks; c++) {
            const uint8_t *b = row + c * 4;
            const uint32_t w = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
            const __m512i codes = _mm512_and_si512(_mm512_srlv_epi32(_mm512_set1_epi32((int)w), sh), m3);
            acc[c & 3] = _mm512_fmadd_ps(_mm512_cvtepi32_ps(codes), _mm512_loadu_ps(q + c * 16), acc[c & 3]);
        }
    } else {
        const __m512i sh = _mm512_setr_epi32(28, 24, 20, 16, 12, 8, 4, 0, 28, 24, 20, 16, 12, 8, 4, 0);
        const __m512i mf = _mm512_set1_epi32(0xF);
        for (uint32_t c = 0; c < chunks; c++) {
            const uint8_t *b = row + c * 8;
            const uint32_t lo = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
            const uint32_t hi = ((uint32_t)b[4] << 24) | ((uint32_t)b[5] << 16) | ((uint32_t)b[6] << 8) | b[7];
            const __m512i both = _mm512_mask_blend_epi32((__mmask16)0xFF00, _mm512_set1_epi32((int)lo),
                                                         _mm512_set1_epi32((int)hi));
            const __m512i codes = _mm512_and_si512(_mm512_srlv_epi32(both, sh), mf);
            acc[c & 3] = _mm512_fmadd_ps(_mm512_cvtepi32_ps(codes), _mm512_loadu_ps(q + c * 16), acc[c & 3]);
        }
    }
    const __m512 t = _mm512_add_ps(_mm512_add_ps(acc[0], acc[1]), _mm512_add_ps(acc[2], acc[3]));
    const __m256 u = _mm256_add_ps(_mm512_castps512_ps256(t), _mm512_extractf32x8_ps(t, 1));
    const __m128 w = _mm_add_ps(_mm256_castps256_ps128(u), _mm256_extractf128_ps(u, 1));
    const __m128 x = _mm_add_ps(w, _mm_movehl_ps(w, w));
    const __m128 r = _mm_add_ss(x, _mm_shuffle_ps(x, x, 0x55));
    return _mm_cvtss_f32(r) * scale;
}
