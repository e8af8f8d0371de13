/*
 * ordvec CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference (Project-Navi/ordvec 0.6.0) CPU path
 * for the ordinal retrieval hot path.  It is the checker for the HIP product
 * path and the CPU baseline timed beside it; it is never linked into, loaded
 * by, or called from the product library.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may use it.
 *
 * Every function cites the reference file:line it follows.  The f32 scoring
 * functions reproduce the reference's exact summation order per SIMD tier
 * (scalar LUT, AVX2, AVX-512) so results are bit-identical to what the
 * reference computes on an AVX-512F/DQ x86-64 host (the reference's benchmark
 * hosts).  Build with -ffp-contract=off: every fused multiply-add here is an
 * explicit fmaf() where the reference uses _mm*_fmadd_ps, and nothing else
 * may be contracted.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* AVX-512 forms of the scan and the AVX-512 tier (ordvec_cpu_simd.c), used
 * when the CPU has them; identical results. */
int orc_simd_active(void);
void orc_simd_scan_block(const uint64_t *block, uint64_t bn, uint32_t qpv, const uint64_t *qb, uint64_t tq,
                         uint32_t *scores);
float orc_simd_score_avx512(const uint8_t *row, uint32_t d, uint32_t bits, const float *q, float scale);

#define ORC_TIER_SCALAR 0
#define ORC_TIER_AVX2 1
#define ORC_TIER_AVX512 2

/* ------------------------------------------------------------------ */
/* f32 ordering helpers                                                 */
/* ------------------------------------------------------------------ */

/* IEEE totalOrder key: f32::total_cmp (Rust core) maps the bit pattern to a
 * signed integer by flipping the magnitude bits of negative values. */
static inline int32_t total_key(float x) {
    int32_t i;
    memcpy(&i, &x, 4);
    i ^= (int32_t)(((uint32_t)(i >> 31)) >> 1);
    return i;
}

static inline int total_cmp(float a, float b) {
    int32_t ka = total_key(a), kb = total_key(b);
    return (ka > kb) - (ka < kb);
}

/* src/util.rs:19-30 cmp_finite_f32_then_index: `==` first (so -0 == +0 and
 * ties go by index), otherwise total_cmp. */
static inline int cmp_value_then_index(float lv, uint32_t li, float rv, uint32_t ri) {
    if (lv == rv) return (li > ri) - (li < ri);
    return total_cmp(lv, rv);
}

/* ------------------------------------------------------------------ */
/* src/rank.rs primitives                                               */
/* ------------------------------------------------------------------ */

static const float *g_sort_vals; /* qsort has no context arg; guarded by call */

static int cmp_order(const void *a, const void *b) {
    uint16_t ia = *(const uint16_t *)a, ib = *(const uint16_t *)b;
    return cmp_value_then_index(g_sort_vals[ia], ia, g_sort_vals[ib], ib);
}

/* Merge sort on u16 orders with the (value, index) comparator; reentrant
 * (used by the OpenMP encode loop instead of qsort's global context). */
static void msort_order(uint16_t *a, uint16_t *tmp, uint32_t n, const float *v) {
    if (n <= 16) {
        for (uint32_t i = 1; i < n; i++) {
            uint16_t x = a[i];
            uint32_t j = i;
            while (j > 0 && cmp_value_then_index(v[a[j - 1]], a[j - 1], v[x], x) > 0) {
                a[j] = a[j - 1];
                j--;
            }
            a[j] = x;
        }
        return;
    }
    uint32_t h = n / 2;
    msort_order(a, tmp, h, v);
    msort_order(a + h, tmp, n - h, v);
    uint32_t i = 0, j = h, o = 0;
    while (i < h && j < n) {
        if (cmp_value_then_index(v[a[j]], a[j], v[a[i]], a[i]) < 0) tmp[o++] = a[j++];
        else tmp[o++] = a[i++];
    }
    while (i < h) tmp[o++] = a[i++];
    while (j < n) tmp[o++] = a[j++];
    memcpy(a, tmp, n * sizeof(uint16_t));
}

/* src/rank.rs:38-72 rank_transform[_into]: out[k] = position of v[k] in the
 * ascending (value, index) order = argsort(argsort(v)). The comparator is a
 * strict total order on (value, index) pairs, so any correct sort gives the
 * reference's result. */
void orc_rank_transform(const float *v, uint32_t d, uint16_t *out) {
    uint16_t *order = (uint16_t *)malloc(d * sizeof(uint16_t));
    uint16_t *tmp = (uint16_t *)malloc(d * sizeof(uint16_t));
    for (uint32_t i = 0; i < d; i++) order[i] = (uint16_t)i;
    msort_order(order, tmp, d, v);
    for (uint32_t r = 0; r < d; r++) out[order[r]] = (uint16_t)r;
    free(order);
    free(tmp);
}

/* qsort-based variant kept as an independent second formulation for tests. */
void orc_rank_transform_qsort(const float *v, uint32_t d, uint16_t *out) {
    uint16_t *order = (uint16_t *)malloc(d * sizeof(uint16_t));
    for (uint32_t i = 0; i < d; i++) order[i] = (uint16_t)i;
    g_sort_vals = v;
    qsort(order, d, sizeof(uint16_t), cmp_order);
    for (uint32_t r = 0; r < d; r++) out[order[r]] = (uint16_t)r;
    free(order);
}

/* src/rank.rs:89-95 rank_to_bucket: floor(rank * 2^bits / d), u64 product. */
uint8_t orc_rank_to_bucket(uint32_t rank, uint32_t d, uint32_t bits) {
    return (uint8_t)(((uint64_t)rank * (uint64_t)(1u << bits)) / (uint64_t)d);
}

/* src/rank.rs:128-156 pack_buckets: code i lands in byte i/cpb, first code in
 * the most significant bits. */
void orc_pack_buckets(const uint8_t *buckets, uint32_t d, uint32_t bits, uint8_t *out) {
    uint32_t cpb = 8 / bits;
    memset(out, 0, d / cpb);
    for (uint32_t i = 0; i < d; i++) {
        uint32_t shift = (cpb - 1 - i % cpb) * bits;
        out[i / cpb] |= (uint8_t)(buckets[i] << shift);
    }
}

/* src/rank.rs:163-181 unpack_buckets. */
void orc_unpack_buckets(const uint8_t *packed, uint32_t d, uint32_t bits, uint8_t *out) {
    uint32_t cpb = 8 / bits;
    uint8_t mask = (uint8_t)((1u << bits) - 1);
    for (uint32_t i = 0; i < d; i++) {
        uint32_t shift = (cpb - 1 - i % cpb) * bits;
        out[i] = (packed[i / cpb] >> shift) & mask;
    }
}

/* src/rank.rs:220-228 bucket_centre = b - (2^bits - 1)/2 in f32. */
float orc_bucket_centre(uint32_t b, uint32_t bits) {
    float n = (float)(1u << bits);
    return (float)b - (n - 1.0f) / 2.0f;
}

/* src/rank.rs:275-280 rankquant_norm = sqrt(d * (4^bits - 1)/12), f64 then f32. */
float orc_rankquant_norm(uint32_t d, uint32_t bits) {
    double n = (double)(1u << bits);
    double var = (n * n - 1.0) / 12.0;
    return (float)sqrt((double)d * var);
}

/* src/quant.rs:587-614 RankQuant::add: per row rank -> bucket -> pack. */
void orc_rq_encode(const float *x, uint64_t n, uint32_t d, uint32_t bits, uint8_t *out) {
    uint32_t bpv = d * bits / 8;
#pragma omp parallel
    {
        uint16_t *order = (uint16_t *)malloc(d * sizeof(uint16_t));
        uint16_t *tmp = (uint16_t *)malloc(d * sizeof(uint16_t));
        uint8_t *buck = (uint8_t *)malloc(d);
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)n; r++) {
            const float *v = x + (uint64_t)r * d;
            for (uint32_t i = 0; i < d; i++) order[i] = (uint16_t)i;
            msort_order(order, tmp, d, v);
            for (uint32_t rk = 0; rk < d; rk++) buck[order[rk]] = orc_rank_to_bucket(rk, d, bits);
            orc_pack_buckets(buck, d, bits, out + (uint64_t)r * bpv);
        }
        free(order);
        free(tmp);
        free(buck);
    }
}

/* src/sign_bitmap.rs:180-200 SignBitmap::add (and the query side 204-214):
 * bit j of u64 word j/64 (bit j%64) is set iff v[j] > 0.0. */
void orc_sign_encode(const float *x, uint64_t n, uint32_t d, uint64_t *out) {
    uint32_t qpv = d / 64;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)n; r++) {
        const float *v = x + (uint64_t)r * d;
        uint64_t *o = out + (uint64_t)r * qpv;
        for (uint32_t w = 0; w < qpv; w++) o[w] = 0;
        for (uint32_t j = 0; j < d; j++)
            if (v[j] > 0.0f) o[j / 64] |= (uint64_t)1 << (j % 64);
    }
}

/* src/util.rs:93-115 l2_normalise: norm = sqrt(sum x*x) summed in order
 * (Rust's f32 Sum folds from -0.0), zero vector when norm <= 1e-12, else
 * x * (1/norm). */
void orc_l2_normalise(const float *v, uint32_t d, float *out) {
    float s = -0.0f;
    for (uint32_t i = 0; i < d; i++) s = s + v[i] * v[i];
    float norm = sqrtf(s);
    if (norm <= 1e-12f) {
        for (uint32_t i = 0; i < d; i++) out[i] = 0.0f;
        return;
    }
    float inv = 1.0f / norm;
    for (uint32_t i = 0; i < d; i++) out[i] = v[i] * inv;
}

/* ------------------------------------------------------------------ */
/* Asymmetric scoring (src/quant_kernels.rs)                            */
/* ------------------------------------------------------------------ */

/* src/quant.rs:300-329 select_simd_tier, modelled for an x86-64 host with
 * AVX-512F+DQ and AVX2+FMA (the reference's benchmark hosts). */
int orc_select_tier(uint32_t d, uint32_t bits) {
    if (bits != 2 && bits != 4) return ORC_TIER_SCALAR;
    if (d % 64 == 0) return ORC_TIER_AVX512;
    if ((bits == 2 && d % 16 == 0) || (bits == 4 && d % 8 == 0)) return ORC_TIER_AVX2;
    return ORC_TIER_SCALAR;
}

static inline uint32_t code_at(const uint8_t *row, uint32_t p, uint32_t bits) {
    uint32_t cpb = 8 / bits;
    uint32_t shift = (cpb - 1 - p % cpb) * bits;
    return (row[p / cpb] >> shift) & ((1u << bits) - 1);
}

/* src/quant_kernels.rs:23-160 scalar LUT path: lut[c*nb + b] = q[c]*centre(b);
 * acc += lut[...] in coordinate order starting at 0.0; returns acc * scale. */
static float score_scalar(const uint8_t *row, uint32_t d, uint32_t bits, const float *lut, float scale) {
    uint32_t nb = 1u << bits;
    float acc = 0.0f;
    for (uint32_t p = 0; p < d; p++) acc = acc + lut[p * nb + code_at(row, p, bits)];
    return acc * scale;
}

/* src/quant_kernels.rs:406-581 AVX-512 b=2/b=4: 16 lanes, 4 accumulators
 * (chunk c of 16 codes goes to accumulator c%4, lane l holds coordinate
 * c*16+l), FMA(code, q, acc); then (acc0+acc1)+(acc2+acc3) per lane and the
 * _mm512_reduce_add_ps halving tree (8, 4, 2, 1). Returns raw * scale. */
static float score_avx512(const uint8_t *row, uint32_t d, uint32_t bits, const float *q, float scale) {
    float acc[4][16];
    memset(acc, 0, sizeof(acc));
    uint32_t chunks = d / 16;
    for (uint32_t c = 0; c < chunks; c++) {
        float *a = acc[c % 4];
        for (uint32_t l = 0; l < 16; l++) {
            uint32_t p = c * 16 + l;
            a[l] = fmaf((float)code_at(row, p, bits), q[p], a[l]);
        }
    }
    float t[16];
    for (uint32_t l = 0; l < 16; l++) t[l] = (acc[0][l] + acc[1][l]) + (acc[2][l] + acc[3][l]);
    float u[8], w[4], x[2];
    for (int j = 0; j < 8; j++) u[j] = t[j] + t[j + 8];
    for (int j = 0; j < 4; j++) w[j] = u[j] + u[j + 4];
    for (int j = 0; j < 2; j++) x[j] = w[j] + w[j + 2];
    float raw = x[0] + x[1];
    return raw * scale;
}

/* src/quant_kernels.rs:373-389 horizontal_sum_avx2. */
static float hsum_avx2(const float *t8) {
    float s[4];
    for (int j = 0; j < 4; j++) s[j] = t8[j] + t8[j + 4];
    float a = s[0] + s[1];
    float b = s[2] + s[3];
    return a + b;
}

/* src/quant_kernels.rs:226-369 AVX2 b=2 (two 8-lane accumulators: codes 0-7
 * and 8-15 of each 16-code chunk) and b=4 (one 8-lane accumulator over
 * 8-code chunks). Returns raw * scale. */
static float score_avx2(const uint8_t *row, uint32_t d, uint32_t bits, const float *q, float scale) {
    float t[8];
    if (bits == 2) {
        float hi[8] = {0}, lo[8] = {0};
        for (uint32_t c = 0; c < d / 16; c++) {
            for (uint32_t l = 0; l < 8; l++) {
                uint32_t p = c * 16 + l;
                hi[l] = fmaf((float)code_at(row, p, bits), q[p], hi[l]);
                lo[l] = fmaf((float)code_at(row, p + 8, bits), q[p + 8], lo[l]);
            }
        }
        for (int l = 0; l < 8; l++) t[l] = hi[l] + lo[l];
    } else {
        float a[8] = {0};
        for (uint32_t c = 0; c < d / 8; c++)
            for (uint32_t l = 0; l < 8; l++) {
                uint32_t p = c * 8 + l;
                a[l] = fmaf((float)code_at(row, p, bits), q[p], a[l]);
            }
        for (int l = 0; l < 8; l++) t[l] = a[l];
    }
    return hsum_avx2(t) * scale;
}

/* Per-query constants shared by every tier (src/quant.rs:734-798,1139-1145). */
typedef struct {
    float inv_norm;
    float offset; /* TopK::score_offset: centre offset for SIMD tiers, 0.0 otherwise */
    int tier;
} orc_qconst;

static orc_qconst query_consts(const float *q_unit, uint32_t d, uint32_t bits, float *lut) {
    orc_qconst c;
    c.inv_norm = 1.0f / orc_rankquant_norm(d, bits);
    c.tier = orc_select_tier(d, bits);
    c.offset = 0.0f;
    if (c.tier != ORC_TIER_SCALAR) {
        float centre = ((float)(1u << bits) - 1.0f) / 2.0f;
        float qs = -0.0f;
        for (uint32_t i = 0; i < d; i++) qs = qs + q_unit[i];
        c.offset = -centre * qs * c.inv_norm;
    } else if (lut) {
        uint32_t nb = 1u << bits;
        for (uint32_t p = 0; p < d; p++)
            for (uint32_t b = 0; b < nb; b++) lut[p * nb + b] = q_unit[p] * orc_bucket_centre(b, bits);
    }
    return c;
}

/* The exact score TopK sees for one document: kernel output + score_offset
 * (src/util.rs:451-452). */
static inline float doc_score(const uint8_t *row, uint32_t d, uint32_t bits, const float *q_unit,
                              const float *lut, const orc_qconst *c) {
    float s;
    if (c->tier == ORC_TIER_AVX512 && orc_simd_active()) s = orc_simd_score_avx512(row, d, bits, q_unit, c->inv_norm);
    else if (c->tier == ORC_TIER_AVX512) s = score_avx512(row, d, bits, q_unit, c->inv_norm);
    else if (c->tier == ORC_TIER_AVX2) s = score_avx2(row, d, bits, q_unit, c->inv_norm);
    else s = score_scalar(row, d, bits, lut, c->inv_norm);
    return s + c->offset;
}

/* Score rows (optionally gathered through ids) for one already-normalised
 * query, forcing a tier (-1 = reference dispatch). */
void orc_rq_scores(const uint8_t *codes, const uint32_t *ids, uint64_t n, uint32_t d, uint32_t bits,
                   const float *q_unit, int force_tier, float *out) {
    uint32_t nb = 1u << bits;
    float *lut = (float *)malloc((size_t)d * nb * sizeof(float));
    orc_qconst c = query_consts(q_unit, d, bits, lut);
    if (force_tier >= 0 && force_tier != c.tier) {
        c.tier = force_tier;
        c.offset = 0.0f;
        if (force_tier != ORC_TIER_SCALAR) {
            float centre = ((float)(1u << bits) - 1.0f) / 2.0f;
            float qs = -0.0f;
            for (uint32_t i = 0; i < d; i++) qs = qs + q_unit[i];
            c.offset = -centre * qs * c.inv_norm;
        } else {
            for (uint32_t p = 0; p < d; p++)
                for (uint32_t b = 0; b < nb; b++) lut[p * nb + b] = q_unit[p] * orc_bucket_centre(b, bits);
        }
    }
    uint32_t bpv = d * bits / 8;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t r = ids ? ids[i] : i;
        out[i] = doc_score(codes + r * bpv, d, bits, q_unit, lut, &c);
    }
    free(lut);
}

/* ------------------------------------------------------------------ */
/* TopK (src/util.rs:386-600)                                           */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t k, filled, worst_pos;
    float worst_val;
    int64_t worst_tie;
    float *scores;
    int64_t *idx;
    int64_t *tie;
} topk_t;

static void topk_init(topk_t *t, uint64_t k) {
    t->k = k;
    t->filled = 0;
    t->worst_pos = 0;
    t->worst_val = INFINITY;
    t->worst_tie = INT64_MAX;
    t->scores = (float *)malloc((k ? k : 1) * sizeof(float));
    t->idx = (int64_t *)malloc((k ? k : 1) * sizeof(int64_t));
    t->tie = (int64_t *)malloc((k ? k : 1) * sizeof(int64_t));
}

static void topk_free(topk_t *t) {
    free(t->scores);
    free(t->idx);
    free(t->tie);
}

/* src/util.rs:499-520 recompute_worst. */
static void topk_recompute_worst(topk_t *t) {
    float wv = INFINITY;
    int64_t wt = INT64_MIN;
    uint64_t wp = 0;
    for (uint64_t i = 0; i < t->filled; i++) {
        int c = total_cmp(t->scores[i], wv);
        int worse = c < 0 || (c == 0 && t->tie[i] > wt);
        if (worse) {
            wv = t->scores[i];
            wt = t->tie[i];
            wp = i;
        }
    }
    t->worst_val = wv;
    t->worst_tie = wt;
    t->worst_pos = wp;
}

/* src/util.rs:451-494 maybe_insert (score already includes the offset). */
static void topk_insert(topk_t *t, float score, int64_t id, int64_t tie) {
    if (t->filled < t->k) {
        t->scores[t->filled] = score;
        t->idx[t->filled] = id;
        t->tie[t->filled] = tie;
        t->filled++;
        if (t->filled == t->k) topk_recompute_worst(t);
    } else {
        int c = total_cmp(score, t->worst_val);
        int better = c > 0 || (c == 0 && tie < t->worst_tie);
        if (better) {
            t->scores[t->worst_pos] = score;
            t->idx[t->worst_pos] = id;
            t->tie[t->worst_pos] = tie;
            topk_recompute_worst(t);
        }
    }
}

typedef struct {
    float s;
    int64_t i, tie;
    uint64_t slot;
} fin_t;

static int fin_cmp(const void *a, const void *b) {
    const fin_t *x = (const fin_t *)a, *y = (const fin_t *)b;
    int c = total_cmp(y->s, x->s);
    if (c) return c;
    if (x->tie != y->tie) return x->tie < y->tie ? -1 : 1;
    return (x->slot > y->slot) - (x->slot < y->slot);
}

/* src/util.rs:554-600 finalize_into_with_scratch: (score desc, tie asc, slot
 * asc); positions past `filled` keep the -inf / -1 sentinels. */
static void topk_finalize(const topk_t *t, float *os, int64_t *oi, uint64_t out_len) {
    for (uint64_t j = 0; j < out_len; j++) {
        os[j] = -INFINITY;
        oi[j] = -1;
    }
    fin_t *f = (fin_t *)malloc((t->filled ? t->filled : 1) * sizeof(fin_t));
    for (uint64_t j = 0; j < t->filled; j++) {
        f[j].s = t->scores[j];
        f[j].i = t->idx[j];
        f[j].tie = t->tie[j];
        f[j].slot = j;
    }
    qsort(f, t->filled, sizeof(fin_t), fin_cmp);
    for (uint64_t j = 0; j < t->filled && j < out_len; j++) {
        os[j] = f[j].s;
        oi[j] = f[j].i;
    }
    free(f);
}

/* Exposed TopK over a dense score vector (tie key = tie_keys[i] or i). */
void orc_topk(const float *scores, const int64_t *tie_keys, uint64_t n, uint64_t k, float *os, int64_t *oi) {
    topk_t t;
    topk_init(&t, k);
    for (uint64_t i = 0; i < n; i++) topk_insert(&t, scores[i], (int64_t)i, tie_keys ? tie_keys[i] : (int64_t)i);
    topk_finalize(&t, os, oi, k);
    topk_free(&t);
}

/* ------------------------------------------------------------------ */
/* RankQuant searches                                                   */
/* ------------------------------------------------------------------ */

/* src/quant.rs:713-847 search_asymmetric (b in {1,2,4}); k is clamped to n by
 * the caller; outputs are nq*k. */
void orc_search_asymmetric(const uint8_t *codes, uint64_t n, uint32_t d, uint32_t bits, const float *queries,
                           uint64_t nq, uint64_t k, float *os, int64_t *oi) {
    if (k > n) k = n;
    if (k == 0) return;
    uint32_t bpv = d * bits / 8;
    uint32_t nb = 1u << bits;
#pragma omp parallel
    {
        float *qu = (float *)malloc(d * sizeof(float));
        float *lut = (float *)malloc((size_t)d * nb * sizeof(float));
        topk_t t;
        topk_init(&t, k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
            orc_l2_normalise(queries + (uint64_t)qi * d, d, qu);
            orc_qconst c = query_consts(qu, d, bits, lut);
            t.filled = 0;
            t.worst_pos = 0;
            t.worst_val = INFINITY;
            t.worst_tie = INT64_MAX;
            for (uint64_t di = 0; di < n; di++) {
                float s = doc_score(codes + di * bpv, d, bits, qu, lut, &c);
                topk_insert(&t, s, (int64_t)di, (int64_t)di);
            }
            topk_finalize(&t, os + (uint64_t)qi * k, oi + (uint64_t)qi * k, k);
        }
        topk_free(&t);
        free(qu);
        free(lut);
    }
}

/* src/quant.rs:1126-1281 subset_rerank_row_into for one normalised query:
 * tie key = global id, emitted index = local position mapped back to the
 * global id. */
static void subset_row(const uint8_t *codes, uint32_t d, uint32_t bits, const float *qu, float *lut,
                       const uint32_t *row, uint64_t m, uint64_t out_k, topk_t *t, float *os, int64_t *oi) {
    uint32_t bpv = d * bits / 8;
    orc_qconst c = query_consts(qu, d, bits, lut);
    t->k = out_k;
    t->filled = 0;
    t->worst_pos = 0;
    t->worst_val = INFINITY;
    t->worst_tie = INT64_MAX;
    for (uint64_t j = 0; j < m; j++) {
        float s = doc_score(codes + (uint64_t)row[j] * bpv, d, bits, qu, lut, &c);
        topk_insert(t, s, (int64_t)j, (int64_t)row[j]);
    }
    topk_finalize(t, os, oi, out_k);
    for (uint64_t j = 0; j < out_k; j++) oi[j] = oi[j] < 0 ? -1 : (int64_t)row[oi[j]];
}

/* src/quant.rs:1336-1394 search_asymmetric_subset_batched_serial_into.  The
 * caller validated the CSR; out_k = min(k, n); outputs nq*out_k; empty rows
 * padded with -inf / -1.  `threads` > 1 splits query ranges like a caller pool. */
void orc_subset_batched(const uint8_t *codes, uint64_t n, uint32_t d, uint32_t bits, const float *queries,
                        uint64_t nq, const uint64_t *offsets, const uint32_t *cands, uint64_t k, float *os,
                        int64_t *oi, int threads) {
    uint64_t out_k = k < n ? k : n;
    if (out_k == 0 || nq == 0) return;
    uint32_t nb = 1u << bits;
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel num_threads(threads)
#endif
    {
        float *qu = (float *)malloc(d * sizeof(float));
        float *lut = (float *)malloc((size_t)d * nb * sizeof(float));
        topk_t t;
        topk_init(&t, out_k);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
            uint64_t b = offsets[qi], e = offsets[qi + 1];
            float *s = os + (uint64_t)qi * out_k;
            int64_t *ix = oi + (uint64_t)qi * out_k;
            if (b == e) {
                for (uint64_t j = 0; j < out_k; j++) {
                    s[j] = -INFINITY;
                    ix[j] = -1;
                }
                continue;
            }
            orc_l2_normalise(queries + (uint64_t)qi * d, d, qu);
            uint64_t row_k = out_k;
            subset_row(codes, d, bits, qu, lut, cands + b, e - b, row_k, &t, s, ix);
        }
        topk_free(&t);
        free(qu);
        free(lut);
    }
}

/* src/quant.rs:1048-1083 search_asymmetric_subset (single query): returns
 * min(k, m) entries. */
void orc_search_subset(const uint8_t *codes, uint32_t d, uint32_t bits, const float *query, const uint32_t *rows,
                       uint64_t m, uint64_t k, float *os, int64_t *oi) {
    uint64_t out_k = k < m ? k : m;
    if (out_k == 0) return;
    float *qu = (float *)malloc(d * sizeof(float));
    float *lut = (float *)malloc((size_t)d * (1u << bits) * sizeof(float));
    topk_t t;
    topk_init(&t, out_k);
    orc_l2_normalise(query, d, qu);
    subset_row(codes, d, bits, qu, lut, rows, m, out_k, &t, os, oi);
    topk_free(&t);
    free(qu);
    free(lut);
}

/* ------------------------------------------------------------------ */
/* SignBitmap stage 1 (src/sign_bitmap.rs)                              */
/* ------------------------------------------------------------------ */

/* src/sign_bitmap.rs:674-753 / src/util.rs:192-238 xor_popcount scan. */
void orc_hamming_all(const uint64_t *bm, uint64_t n, uint32_t qpv, const uint64_t *qb, uint32_t *out) {
#pragma omp parallel for schedule(static)
    for (int64_t di = 0; di < (int64_t)n; di++) {
        const uint64_t *row = bm + (uint64_t)di * qpv;
        uint32_t h = 0;
        for (uint32_t w = 0; w < qpv; w++) h += (uint32_t)__builtin_popcountll(row[w] ^ qb[w]);
        out[di] = h;
    }
}

/* max-heap of (hamming, id) pairs: the retained set is the m smallest keys. */
static inline int key_gt(uint64_t a, uint64_t b) { return a > b; }

static void heap_push(uint64_t *h, uint64_t *len, uint64_t key) {
    uint64_t i = (*len)++;
    h[i] = key;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (!key_gt(h[i], h[p])) break;
        uint64_t t = h[i];
        h[i] = h[p];
        h[p] = t;
        i = p;
    }
}

static void heap_pop(uint64_t *h, uint64_t *len) {
    uint64_t n = --(*len);
    h[0] = h[n];
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < n && key_gt(h[l], h[m])) m = l;
        if (r < n && key_gt(h[r], h[m])) m = r;
        if (m == i) break;
        uint64_t t = h[i];
        h[i] = h[m];
        h[m] = t;
        i = m;
    }
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* CHUNK=8 batched scan of one doc block against a query tile
 * (src/sign_bitmap.rs:763-955): scores[ti*bn + d]. */
static void scan_block(const uint64_t *block, uint64_t bn, uint32_t qpv, const uint64_t *qb, uint64_t tq,
                       uint32_t *scores) {
    if (orc_simd_active()) {
        orc_simd_scan_block(block, bn, qpv, qb, tq, scores);
        return;
    }
    for (uint64_t t0 = 0; t0 < tq; t0 += 8) {
        uint64_t tc = tq - t0 < 8 ? tq - t0 : 8;
        for (uint64_t d = 0; d < bn; d++) {
            const uint64_t *row = block + d * qpv;
            for (uint64_t t = 0; t < tc; t++) {
                const uint64_t *q = qb + (t0 + t) * qpv;
                uint32_t h = 0;
                for (uint32_t w = 0; w < qpv; w++) h += (uint32_t)__builtin_popcountll(row[w] ^ q[w]);
                scores[(t0 + t) * bn + d] = h;
            }
        }
    }
}

/* src/sign_bitmap.rs:231-328 top_m_candidates_streamed: 256 KiB doc blocks,
 * 32-query tiles, per-query bounded max-heaps on (hamming, id) with the
 * worst-bound skip; rows sorted ascending.  out is nq*m_eff.  `threads` > 1
 * splits the query batch into ranges (a caller-owned pool), each range
 * running the serial streamed algorithm. */
void orc_sign_top_m(const uint64_t *bm, uint64_t n, uint32_t qpv, const uint64_t *qbits, uint64_t nq,
                    uint64_t m, uint32_t *out, int threads) {
    uint64_t m_eff = m < n ? m : n;
    if (m_eff == 0 || nq == 0) return;
    const uint64_t TILE = 32, BLOCK_BYTES = 256 * 1024;
    uint64_t block_docs = BLOCK_BYTES / ((uint64_t)qpv * 8);
    if (block_docs < 64) block_docs = 64;
    if (block_docs > n) block_docs = n;
    if (threads < 1) threads = 1;
    uint64_t per = (nq + (uint64_t)threads - 1) / (uint64_t)threads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static, 1)
#endif
    for (int64_t th = 0; th < threads; th++) {
        uint64_t q0 = (uint64_t)th * per;
        if (q0 >= nq) continue;
        uint64_t q1 = q0 + per < nq ? q0 + per : nq;
        uint64_t lnq = q1 - q0;
        uint64_t tile = TILE < lnq ? TILE : lnq;
        uint32_t *scores = (uint32_t *)malloc(tile * block_docs * sizeof(uint32_t));
        uint64_t *heaps = (uint64_t *)malloc(lnq * (m_eff + 1) * sizeof(uint64_t));
        uint64_t *lens = (uint64_t *)calloc(lnq, sizeof(uint64_t));
        uint32_t *worst = (uint32_t *)malloc(lnq * sizeof(uint32_t));
        for (uint64_t i = 0; i < lnq; i++) worst[i] = UINT32_MAX;
        for (uint64_t bs = 0; bs < n; bs += block_docs) {
            uint64_t bn = n - bs < block_docs ? n - bs : block_docs;
            const uint64_t *block = bm + bs * qpv;
            for (uint64_t ts = 0; ts < lnq; ts += tile) {
                uint64_t tq = lnq - ts < tile ? lnq - ts : tile;
                scan_block(block, bn, qpv, qbits + (q0 + ts) * qpv, tq, scores);
                for (uint64_t ti = 0; ti < tq; ti++) {
                    uint64_t qi = ts + ti;
                    uint64_t *h = heaps + qi * (m_eff + 1);
                    const uint32_t *row = scores + ti * bn;
                    for (uint64_t d = 0; d < bn; d++) {
                        if (row[d] >= worst[qi]) continue;
                        heap_push(h, &lens[qi], ((uint64_t)row[d] << 32) | (bs + d));
                        if (lens[qi] > m_eff) heap_pop(h, &lens[qi]);
                        if (lens[qi] == m_eff) worst[qi] = (uint32_t)(h[0] >> 32);
                    }
                }
            }
        }
        for (uint64_t qi = 0; qi < lnq; qi++) {
            uint64_t *h = heaps + qi * (m_eff + 1);
            qsort(h, lens[qi], sizeof(uint64_t), cmp_u64);
            for (uint64_t j = 0; j < m_eff; j++) out[(q0 + qi) * m_eff + j] = (uint32_t)(h[j] & 0xffffffffu);
        }
        free(scores);
        free(heaps);
        free(lens);
        free(worst);
    }
}

/* Brute force: dense hamming + full sort by (hamming, id) -- the reference's
 * own candgen oracle (tests/tiled_candgen.rs:41-49). */
void orc_sign_top_m_bruteforce(const uint64_t *bm, uint64_t n, uint32_t qpv, const uint64_t *qbits, uint64_t nq,
                               uint64_t m, uint32_t *out) {
    uint64_t m_eff = m < n ? m : n;
    if (m_eff == 0) return;
    uint32_t *h = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint64_t *keys = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (uint64_t qi = 0; qi < nq; qi++) {
        orc_hamming_all(bm, n, qpv, qbits + qi * qpv, h);
        for (uint64_t i = 0; i < n; i++) keys[i] = ((uint64_t)h[i] << 32) | i;
        qsort(keys, n, sizeof(uint64_t), cmp_u64);
        for (uint64_t j = 0; j < m_eff; j++) out[qi * m_eff + j] = (uint32_t)(keys[j] & 0xffffffffu);
    }
    free(h);
    free(keys);
}

/* ------------------------------------------------------------------ */
/* FastScan b=2 (src/fastscan.rs)                                       */
/* ------------------------------------------------------------------ */

/* src/fastscan.rs:87-110 pack_fastscan_b2 from unpacked buckets (n*d). */
void orc_fs_pack(const uint8_t *buckets, uint64_t n, uint32_t d, uint8_t *out) {
    uint32_t pairs = d / 2;
    uint64_t nblk = (n + 31) / 32;
    memset(out, 0, nblk * pairs * 32);
    for (uint64_t b = 0; b < nblk; b++) {
        uint64_t base = b * 32;
        uint64_t inb = n - base < 32 ? n - base : 32;
        for (uint64_t l = 0; l < inb; l++) {
            const uint8_t *row = buckets + (base + l) * d;
            for (uint32_t p = 0; p < pairs; p++)
                out[b * pairs * 32 + p * 32 + l] = (uint8_t)(((row[2 * p] & 3) << 2) | (row[2 * p + 1] & 3));
        }
    }
}

/* Rust f32::round: half away from zero. */
static inline float round_half_away(float x) { return roundf(x); }

/* src/fastscan.rs:130-173 build_fastscan_b2_query. */
void orc_fs_query(const float *q, uint32_t d, uint8_t *lut_u8, float *bias_sum, float *inv_q) {
    static const float centres[4] = {-1.5f, -0.5f, 0.5f, 1.5f};
    uint32_t pairs = d / 2;
    float *lut = (float *)malloc((size_t)pairs * 16 * sizeof(float));
    float gmin = INFINITY, gmax = -INFINITY;
    for (uint32_t p = 0; p < pairs; p++) {
        float qa = q[2 * p], qc = q[2 * p + 1];
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 4; c++) {
                float v = qa * centres[a] + qc * centres[c];
                lut[p * 16 + (a << 2 | c)] = v;
                if (v < gmin) gmin = v;
                if (v > gmax) gmax = v;
            }
    }
    float span = gmax - gmin;
    if (!(span >= 1e-12f)) span = 1e-12f; /* f32::max(span, 1e-12) */
    float scale = 255.0f / span;
    *inv_q = 1.0f / scale;
    *bias_sum = (float)pairs * gmin;
    for (uint32_t i = 0; i < pairs * 16; i++) {
        float v = round_half_away((lut[i] - gmin) * scale);
        if (v < 0.0f) v = 0.0f;
        if (v > 255.0f) v = 255.0f;
        lut_u8[i] = (uint8_t)v;
    }
    free(lut);
}

/* src/fastscan.rs:316-347 scalar scan + src/fastscan.rs:357-476 entry. */
void orc_fs_search(const uint8_t *packed, uint64_t n, uint32_t d, const float *queries, uint64_t nq, uint64_t k,
                   float *os, int64_t *oi) {
    if (k > n) k = n;
    if (k == 0) return;
    uint32_t pairs = d / 2;
    float inv_norm = 1.0f / orc_rankquant_norm(d, 2);
    uint64_t nblk = (n + 31) / 32;
#pragma omp parallel
    {
        float *qu = (float *)malloc(d * sizeof(float));
        uint8_t *lut = (uint8_t *)malloc((size_t)pairs * 16);
        topk_t t;
        topk_init(&t, k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
            float bias, invq;
            orc_l2_normalise(queries + (uint64_t)qi * d, d, qu);
            orc_fs_query(qu, d, lut, &bias, &invq);
            t.filled = 0;
            t.worst_pos = 0;
            t.worst_val = INFINITY;
            t.worst_tie = INT64_MAX;
            for (uint64_t b = 0; b < nblk; b++) {
                const uint8_t *blk = packed + b * pairs * 32;
                uint64_t base = b * 32, inb = n - base < 32 ? n - base : 32;
                uint32_t acc[32] = {0};
                for (uint32_t p = 0; p < pairs; p++)
                    for (uint64_t l = 0; l < inb; l++) acc[l] += lut[p * 16 + blk[p * 32 + l]];
                for (uint64_t l = 0; l < inb; l++) {
                    float raw = bias + (float)acc[l] * invq;
                    topk_insert(&t, raw * inv_norm + 0.0f, (int64_t)(base + l), (int64_t)(base + l));
                }
            }
            topk_finalize(&t, os + (uint64_t)qi * k, oi + (uint64_t)qi * k, k);
        }
        topk_free(&t);
        free(qu);
        free(lut);
    }
}

/* ------------------------------------------------------------------ */
/* Symmetric RankQuant search and the Rank index                         */
/* ------------------------------------------------------------------ */

static void topk_reset(topk_t *t) {
    t->filled = 0;
    t->worst_pos = 0;
    t->worst_val = INFINITY;
    t->worst_tie = INT64_MAX;
}

/* src/quant.rs:629-695 RankQuant::search (symmetric): the query is rank
 * transformed and bucketed; lut[p*nb + b] = centre(qbucket_p) * centre(b);
 * the scalar LUT scan (src/quant_kernels.rs:89-160) sums lut entries in
 * coordinate order and scales by 1 / (R * R) (f32). */
void orc_rq_search_symmetric(const uint8_t *codes, uint64_t n, uint32_t d, uint32_t bits, const float *queries,
                             uint64_t nq, uint64_t k, float *os, int64_t *oi) {
    if (k > n) k = n;
    if (k == 0) return;
    uint32_t bpv = d * bits / 8, nb = 1u << bits;
    float norm = orc_rankquant_norm(d, bits);
    float inv_norm_sq = 1.0f / (norm * norm);
#pragma omp parallel
    {
        uint16_t *rk = (uint16_t *)malloc(d * sizeof(uint16_t));
        float *lut = (float *)malloc((size_t)d * nb * sizeof(float));
        topk_t t;
        topk_init(&t, k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
            orc_rank_transform(queries + (uint64_t)qi * d, d, rk);
            for (uint32_t p = 0; p < d; p++) {
                float qc = orc_bucket_centre(orc_rank_to_bucket(rk[p], d, bits), bits);
                for (uint32_t b = 0; b < nb; b++) lut[p * nb + b] = qc * orc_bucket_centre(b, bits);
            }
            topk_reset(&t);
            for (uint64_t di = 0; di < n; di++) {
                const uint8_t *row = codes + di * bpv;
                float acc = 0.0f;
                for (uint32_t p = 0; p < d; p++) acc += lut[p * nb + code_at(row, p, bits)];
                topk_insert(&t, acc * inv_norm_sq + 0.0f, (int64_t)di, (int64_t)di);
            }
            topk_finalize(&t, os + (uint64_t)qi * k, oi + (uint64_t)qi * k, k);
        }
        topk_free(&t);
        free(rk);
        free(lut);
    }
}

/* src/quant.rs:1541-1610 rankquant_eval_search: corpus and queries
 * rank-bucketed at `bits` (1..8); score = (sum_j centre(bq_j) * centre(bd_j),
 * sequential) * (1 / norm^2) with norm = sqrt(sum_rank centre(bucket)^2) in
 * f64 (src/quant.rs:106-117); TopK by (score desc, id asc). */
void orc_eval_search(const float *corpus, uint64_t n, const float *queries, uint64_t nq, uint32_t d, uint32_t bits,
                     uint64_t k, float *os, int64_t *oi) {
    if (k > n) k = n;
    if (k == 0) return;
    double acc2 = 0.0;
    for (uint32_t r = 0; r < d; r++) {
        double c = (double)orc_bucket_centre(orc_rank_to_bucket(r, d, bits), bits);
        acc2 += c * c;
    }
    float norm = (float)sqrt(acc2);
    float inv_norm_sq = 1.0f / (norm * norm);
    uint8_t *db = (uint8_t *)malloc(n * d);
    uint16_t *rk = (uint16_t *)malloc(d * sizeof(uint16_t));
    for (uint64_t i = 0; i < n; i++) {
        orc_rank_transform(corpus + i * d, d, rk);
        for (uint32_t j = 0; j < d; j++) db[i * d + j] = orc_rank_to_bucket(rk[j], d, bits);
    }
    float *qc = (float *)malloc(d * sizeof(float));
    topk_t t;
    topk_init(&t, k);
    for (uint64_t qi = 0; qi < nq; qi++) {
        orc_rank_transform(queries + qi * d, d, rk);
        for (uint32_t j = 0; j < d; j++) qc[j] = orc_bucket_centre(orc_rank_to_bucket(rk[j], d, bits), bits);
        topk_reset(&t);
        for (uint64_t i = 0; i < n; i++) {
            float acc = -0.0f;
            for (uint32_t j = 0; j < d; j++) acc = acc + qc[j] * orc_bucket_centre(db[i * d + j], bits);
            topk_insert(&t, acc * inv_norm_sq, (int64_t)i, (int64_t)i);
        }
        topk_finalize(&t, os + qi * k, oi + qi * k, k);
    }
    topk_free(&t);
    free(db);
    free(rk);
    free(qc);
}

/* src/rank.rs:249-252 rank_norm = sqrt(d (d^2 - 1) / 12), f64 then f32. */
float orc_rank_norm(uint32_t d) {
    double x = (double)d;
    return (float)sqrt(x * (x * x - 1.0) / 12.0);
}

/* src/rank.rs:340-405 Rank::search (symmetric): exact i64 sum of centred
 * doubled ranks, score = (f32)acc * 0.25 * (1 / (norm * norm)). */
void orc_rank_search(const uint16_t *ranks, uint64_t n, uint32_t d, const float *queries, uint64_t nq, uint64_t k,
                     float *os, int64_t *oi) {
    if (k > n) k = n;
    if (k == 0) return;
    int32_t mean_2x = (int32_t)d - 1;
    float norm = orc_rank_norm(d);
    float inv_norm_sq = 1.0f / (norm * norm);
#pragma omp parallel
    {
        uint16_t *rk = (uint16_t *)malloc(d * sizeof(uint16_t));
        topk_t t;
        topk_init(&t, k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
            orc_rank_transform(queries + (uint64_t)qi * d, d, rk);
            topk_reset(&t);
            for (uint64_t di = 0; di < n; di++) {
                const uint16_t *doc = ranks + di * d;
                int64_t acc = 0;
                for (uint32_t p = 0; p < d; p++)
                    acc += (int64_t)(2 * (int32_t)rk[p] - mean_2x) * (int64_t)(2 * (int32_t)doc[p] - mean_2x);
                float s = (float)acc * 0.25f * inv_norm_sq;
                topk_insert(&t, s + 0.0f, (int64_t)di, (int64_t)di);
            }
            topk_finalize(&t, os + (uint64_t)qi * k, oi + (uint64_t)qi * k, k);
        }
        topk_free(&t);
        free(rk);
    }
}

/* src/rank.rs:420-471 Rank::search_asymmetric: q_unit = l2_normalise(q),
 * q_sum = sum q_unit (from -0.0), acc = sum q_unit[p] * rank[p] in order,
 * score = (acc - mean * q_sum) * (1 / norm), mean = (d - 1) / 2. */
void orc_rank_search_asymmetric(const uint16_t *ranks, uint64_t n, uint32_t d, const float *queries, uint64_t nq,
                                uint64_t k, float *os, int64_t *oi) {
    if (k > n) k = n;
    if (k == 0) return;
    float inv_norm = 1.0f / orc_rank_norm(d);
    float mean = ((float)d - 1.0f) / 2.0f;
#pragma omp parallel
    {
        float *qu = (float *)malloc(d * sizeof(float));
        topk_t t;
        topk_init(&t, k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
            orc_l2_normalise(queries + (uint64_t)qi * d, d, qu);
            float q_sum = -0.0f;
            for (uint32_t p = 0; p < d; p++) q_sum += qu[p];
            topk_reset(&t);
            for (uint64_t di = 0; di < n; di++) {
                const uint16_t *doc = ranks + di * d;
                float acc = 0.0f;
                for (uint32_t p = 0; p < d; p++) acc += qu[p] * (float)doc[p];
                float s = (acc - mean * q_sum) * inv_norm;
                topk_insert(&t, s + 0.0f, (int64_t)di, (int64_t)di);
            }
            topk_finalize(&t, os + (uint64_t)qi * k, oi + (uint64_t)qi * k, k);
        }
        topk_free(&t);
        free(qu);
    }
}

/* ------------------------------------------------------------------ */
/* FastScan b=4: the b=2 recipe (src/fastscan.rs:130-173, 316-476) carried
 * to one 4-bit rank code per coordinate.  The reference ships b=2 only; this
 * is the extension the b=4 GPU path implements, stated once here:
 *   lut_f[j*16 + c] = q_unit[j] * centre(c, 4)      (centre = c - 7.5)
 *   g_min/g_max over all entries, span = max(g_max - g_min, 1e-12),
 *   scale = 255/span, inv_q = 1/scale, bias_sum = d * g_min,
 *   lut_u8 = clamp(round_half_away((v - g_min) * scale), 0, 255),
 *   acc = sum_j lut_u8[j*16 + code_j] (exact u32),
 *   score = (bias_sum + acc * inv_q) * (1 / rankquant_norm(d, 4)) + 0.0,
 * top-k by TopK (score desc, id asc).  Codes are RankQuant b=4 rows (two
 * codes per byte, first code in the high nibble). */
/* ------------------------------------------------------------------ */

void orc_fs4_query(const float *q, uint32_t d, uint8_t *lut_u8, float *bias_sum, float *inv_q) {
    float *lut = (float *)malloc((size_t)d * 16 * sizeof(float));
    float gmin = INFINITY, gmax = -INFINITY;
    for (uint32_t j = 0; j < d; j++)
        for (uint32_t c = 0; c < 16; c++) {
            float v = q[j] * ((float)c - 7.5f);
            lut[j * 16 + c] = v;
            if (v < gmin) gmin = v;
            if (v > gmax) gmax = v;
        }
    float span = gmax - gmin;
    if (!(span >= 1e-12f)) span = 1e-12f;
    float scale = 255.0f / span;
    *inv_q = 1.0f / scale;
    *bias_sum = (float)d * gmin;
    for (uint32_t i = 0; i < d * 16; i++) {
        float v = round_half_away((lut[i] - gmin) * scale);
        if (v < 0.0f) v = 0.0f;
        if (v > 255.0f) v = 255.0f;
        lut_u8[i] = (uint8_t)v;
    }
    free(lut);
}

void orc_fs4_search(const uint8_t *codes, uint64_t n, uint32_t d, const float *queries, uint64_t nq, uint64_t k,
                    float *os, int64_t *oi) {
    if (k > n) k = n;
    if (k == 0) return;
    uint32_t bpv = d / 2;
    float inv_norm = 1.0f / orc_rankquant_norm(d, 4);
#pragma omp parallel
    {
        float *qu = (float *)malloc(d * sizeof(float));
        uint8_t *lut = (uint8_t *)malloc((size_t)d * 16);
        topk_t t;
        topk_init(&t, k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
            float bias, invq;
            orc_l2_normalise(queries + (uint64_t)qi * d, d, qu);
            orc_fs4_query(qu, d, lut, &bias, &invq);
            t.filled = 0;
            t.worst_pos = 0;
            t.worst_val = INFINITY;
            t.worst_tie = INT64_MAX;
            for (uint64_t i = 0; i < n; i++) {
                const uint8_t *row = codes + i * bpv;
                uint32_t acc = 0;
                for (uint32_t b = 0; b < bpv; b++)
                    acc += lut[(2 * b) * 16 + (row[b] >> 4)] + lut[(2 * b + 1) * 16 + (row[b] & 15)];
                float raw = bias + (float)acc * invq;
                topk_insert(&t, raw * inv_norm + 0.0f, (int64_t)i, (int64_t)i);
            }
            topk_finalize(&t, os + (uint64_t)qi * k, oi + (uint64_t)qi * k, k);
        }
        topk_free(&t);
        free(qu);
        free(lut);
    }
}

/* ------------------------------------------------------------------ */
/* Bitmap (top-bucket, AND-popcount): src/bitmap.rs                     */
/* ------------------------------------------------------------------ */

/* src/bitmap.rs:140-162 Bitmap::add: bit j set iff rank_transform(v)[j] >=
 * d - n_top.  Row-major u64 words, bit j of word j/64. */
void orc_bitmap_encode(const float *x, uint64_t n, uint32_t d, uint32_t ntop, uint64_t *out) {
    uint32_t qpv = d / 64;
#pragma omp parallel
    {
        uint16_t *rk = (uint16_t *)malloc(d * sizeof(uint16_t));
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)n; r++) {
            uint64_t *o = out + (uint64_t)r * qpv;
            memset(o, 0, qpv * 8);
            orc_rank_transform(x + (uint64_t)r * d, d, rk);
            for (uint32_t j = 0; j < d; j++)
                if (rk[j] >= d - ntop) o[j / 64] |= 1ull << (j % 64);
        }
        free(rk);
    }
}

/* src/bitmap.rs:164-187 build_query_bitmap_fp32: a stable sort of the
 * coordinates by value descending (partial_cmp, so -0 == +0 and ties keep
 * index order); bits for the first n_top.  Insertion into a stable merge. */
static void msort_desc(uint16_t *a, uint16_t *tmp, uint32_t n, const float *v) {
    if (n <= 1) return;
    uint32_t h = n / 2;
    msort_desc(a, tmp, h, v);
    msort_desc(a + h, tmp, n - h, v);
    uint32_t i = 0, j = h, o = 0;
    while (i < h && j < n) {
        /* take the right element only if strictly greater (stability) */
        if (v[a[j]] > v[a[i]]) tmp[o++] = a[j++];
        else tmp[o++] = a[i++];
    }
    while (i < h) tmp[o++] = a[i++];
    while (j < n) tmp[o++] = a[j++];
    memcpy(a, tmp, n * sizeof(uint16_t));
}

void orc_bitmap_query(const float *q, uint64_t nq, uint32_t d, uint32_t ntop, uint64_t *out) {
    uint32_t qpv = d / 64;
    uint16_t *idx = (uint16_t *)malloc(d * sizeof(uint16_t));
    uint16_t *tmp = (uint16_t *)malloc(d * sizeof(uint16_t));
    for (uint64_t r = 0; r < nq; r++) {
        const float *v = q + r * d;
        uint64_t *o = out + r * qpv;
        for (uint32_t i = 0; i < d; i++) idx[i] = (uint16_t)i;
        msort_desc(idx, tmp, d, v);
        memset(o, 0, qpv * 8);
        for (uint32_t t = 0; t < ntop; t++) o[idx[t] / 64] |= 1ull << (idx[t] % 64);
    }
    free(idx);
    free(tmp);
}

/* src/util.rs and_popcount over rows: out[q*n + i] = popcount(doc_i AND q). */
void orc_and_popcount_all(const uint64_t *bm, uint64_t n, uint32_t qpv, const uint64_t *qb, uint64_t nq,
                          uint32_t *out) {
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (int64_t)nq; q++)
        for (uint64_t i = 0; i < n; i++) {
            uint32_t c = 0;
            for (uint32_t w = 0; w < qpv; w++) c += (uint32_t)__builtin_popcountll(bm[i * qpv + w] & qb[q * qpv + w]);
            out[q * n + i] = c;
        }
}

/* ------------------------------------------------------------------ */
/* Score gradient (the reference has no backward pass)                  */
/* ------------------------------------------------------------------ */

/* Analytic gradient, in f64, of the asymmetric score of fixed rows with
 * respect to the raw query x.  The score is the closed form of the reference's
 * kernels, s_i = sum_j qhat_j * centre(code_ij) / R (src/quant_kernels.rs:23-160;
 * the SIMD tiers' centre-drop plus offset, src/quant.rs:788-798, is the same
 * function), with qhat = x / ||x|| (src/util.rs:84-115) and R from
 * src/rank.rs:275-280.  Given G_i = dL/ds_i:
 *   g = sum_i G_i centre(code_i) / R,  dx = (g - qhat (qhat . g)) / ||x||,
 * and dx = 0 when ||x|| <= 1e-12 (the normalised query is then constant).
 * ids[q*k + i] == -1 contributes nothing. */
void orc_score_ids_grad(const uint8_t *codes, uint32_t d, uint32_t bits, const float *x, uint64_t nq,
                        const int64_t *ids, uint64_t k, const double *G, double *dx) {
    uint32_t bpv = d * bits / 8;
    double R = sqrt((double)d * ((double)(1u << bits) * (double)(1u << bits) - 1.0) / 12.0);
    double half = ((double)(1u << bits) - 1.0) / 2.0;
    double *g = (double *)malloc(sizeof(double) * d);
    for (uint64_t q = 0; q < nq; q++) {
        const float *xq = x + q * d;
        double *o = dx + q * d;
        double ss = 0.0;
        for (uint32_t j = 0; j < d; j++) ss += (double)xq[j] * (double)xq[j];
        double norm = sqrt(ss);
        if (norm <= 1e-12) {
            for (uint32_t j = 0; j < d; j++) o[j] = 0.0;
            continue;
        }
        for (uint32_t j = 0; j < d; j++) g[j] = 0.0;
        for (uint64_t i = 0; i < k; i++) {
            int64_t id = ids[q * k + i];
            if (id < 0) continue;
            const uint8_t *row = codes + (uint64_t)id * bpv;
            double gi = G[q * k + i] / R;
            for (uint32_t j = 0; j < d; j++) g[j] += gi * ((double)code_at(row, j, bits) - half);
        }
        double dot = 0.0;
        for (uint32_t j = 0; j < d; j++) dot += (double)xq[j] / norm * g[j];
        for (uint32_t j = 0; j < d; j++) o[j] = (g[j] - (double)xq[j] / norm * dot) / norm;
    }
    free(g);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
