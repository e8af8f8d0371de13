/*
 * ordvec C ABI for the MI355X engine.
 *
 * Part 1 (ABI v1) is the drop-in boundary: the same symbols, struct layouts
 * (128 / 128 / 24 / 184 bytes), status codes, capability bits and validation
 * order as the reference's ordvec-ffi (ordvec-ffi/include/ordvec.h:1-260,
 * implementation ordvec-ffi/src/lib.rs:600-998).  A C or cgo caller of the
 * reference links against libordvec_amd.so instead and recompiles nothing.
 *
 * Part 2 (additive, ORDVEC_ABI_EXT_VERSION) exposes what ABI v1 cannot
 * express: building indexes, batched and two-stage search, the .ovsb/.ovfs
 * formats and device-resident (HBM) buffers.  Every entry point names the
 * reference function whose semantics it carries.
 *
 * Threading: every function may be called from any thread.  Calls are
 * serialised on one HIP stream per process; index handles may be shared for
 * concurrent searches but must not be freed while a search runs.  Indexes a
 * running server reads are pinned: add and swap_remove on them fail with
 * BAD_ARGUMENT and free is deferred until the last such server stops.
 */
#ifndef ORDVEC_H
#define ORDVEC_H

#include <stdint.h>

#define ORDVEC_ABI_VERSION 1
#define ORDVEC_ABI_EXT_VERSION 1

#define ORDVEC_CAP_FULL_SEARCH (1 << 0)
#define ORDVEC_CAP_SUBSET_SEARCH (1 << 1)
#define ORDVEC_CAP_STATS (1 << 2)
#define ORDVEC_CAP_ID_EQUALS_ROW_ID (1 << 3)

#define ORDVEC_SEARCH_FLAG_NONE 0

typedef uint32_t ordvec_status_t;
typedef uint32_t ordvec_index_kind_t;

#define ORDVEC_STATUS_OK 0
#define ORDVEC_STATUS_NULL_POINTER 1
#define ORDVEC_STATUS_BAD_ARGUMENT 2
#define ORDVEC_STATUS_BAD_STRUCT_SIZE 3
#define ORDVEC_STATUS_UNSUPPORTED_FORMAT 4
#define ORDVEC_STATUS_CORRUPT_INDEX 5
#define ORDVEC_STATUS_IO 6
#define ORDVEC_STATUS_DIM_MISMATCH 7
#define ORDVEC_STATUS_NONFINITE_QUERY 8
#define ORDVEC_STATUS_ROW_ID_OUT_OF_RANGE 9
#define ORDVEC_STATUS_BUFFER_TOO_SMALL 10
#define ORDVEC_STATUS_UNSUPPORTED_OPERATION 11
#define ORDVEC_STATUS_PANIC 12
#define ORDVEC_STATUS_INTERNAL 13

#define ORDVEC_INDEX_KIND_UNKNOWN 0
#define ORDVEC_INDEX_KIND_RANK_QUANT 1
#define ORDVEC_INDEX_KIND_BITMAP 2

typedef struct ordvec_index_t ordvec_index_t;

/* ordvec-ffi/src/lib.rs:56-70 (128 bytes). */
typedef struct {
    uint64_t struct_size;
    uint32_t kind;
    uint32_t format_version;
    uint64_t dim;
    uint32_t bit_width;
    uint32_t n_top;
    uint64_t vector_count;
    uint64_t bytes_per_vec;
    uint64_t source_file_size_bytes;
    uint64_t capabilities;
    uint64_t reserved[8];
} ordvec_index_info_t;

/* ordvec-ffi/src/lib.rs:72-86 (128 bytes).  candidate_rows are global row
 * ids, may be unsorted and may repeat; duplicates are scored separately. */
typedef struct {
    uint64_t struct_size;
    const float *query;
    uint64_t dim;
    uint64_t k;
    const uint32_t *candidate_rows;
    uint64_t candidate_count;
    uint64_t flags;
    uint64_t user_tag;
    uint64_t reserved[8];
} ordvec_search_params_t;

/* ordvec-ffi/src/lib.rs:97-124 (184 bytes).  This engine also fills the
 * fields ABI v1 reserved: prepare_ns (host validation and upload), score_ns
 * and select_ns (GPU time of the scoring and selection kernels) and
 * bytes_read (algorithmic HBM bytes the kernels read). */
typedef struct {
    uint64_t struct_size;
    uint32_t abi_version;
    uint32_t kind;
    uint64_t dim;
    uint32_t bit_width;
    uint32_t n_top;
    uint64_t k;
    uint64_t user_tag;
    uint64_t vector_count;
    uint64_t candidate_count;
    uint64_t returned_count;
    uint64_t total_ns;
    uint64_t prepare_ns;
    uint64_t score_ns;
    uint64_t select_ns;
    uint64_t vectors_scored;
    uint64_t bytes_read;
    uint64_t reserved[8];
} ordvec_search_stats_t;

/* ordvec-ffi/src/lib.rs:88-95 (24 bytes). */
typedef struct {
    uint64_t row_id;
    uint64_t id;
    float score;
    uint32_t reserved;
} ordvec_hit_t;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Part 1: ABI v1 (ordvec-ffi/include/ordvec.h:142-254) ---------------- */

/* ordvec-ffi/src/lib.rs:600-603 */
uint32_t ordvec_abi_version(void);
/* ordvec-ffi/src/lib.rs:605-608 */
const char *ordvec_version_string(void);
/* ordvec-ffi/src/lib.rs:610-628 */
const char *ordvec_status_name(ordvec_status_t status);
/* ordvec-ffi/src/lib.rs:630-633: thread-local, cleared on success. */
const char *ordvec_last_error(void);
/* ordvec-ffi/src/lib.rs:635-638: also selects and initialises the GPU. */
ordvec_status_t ordvec_init(void);
/* ordvec-ffi/src/lib.rs:640-653 */
void ordvec_index_info_init(ordvec_index_info_t *info);
/* ordvec-ffi/src/lib.rs:655-668 */
void ordvec_search_params_init(ordvec_search_params_t *params);
/* ordvec-ffi/src/lib.rs:670-683 */
void ordvec_search_stats_init(ordvec_search_stats_t *stats);
/* ordvec-ffi/src/lib.rs:685-776: .ovrq (RankQuant) and .ovbm (Bitmap) load into
 * HBM; .ovr, .ovsb and .ovfs return UNSUPPORTED_FORMAT as in the reference (use
 * the Part 2 loaders for .ovsb/.ovfs). */
ordvec_status_t ordvec_index_load(const char *path, uint64_t flags, ordvec_index_t **out);
/* ordvec-ffi/src/lib.rs:778-840 */
ordvec_status_t ordvec_index_probe(const char *path, uint64_t flags, ordvec_index_info_t *info_out);
/* ordvec-ffi/src/lib.rs:842-875 */
ordvec_status_t ordvec_index_info(const ordvec_index_t *index, ordvec_index_info_t *info_out);
/* ordvec-ffi/src/lib.rs:877-892 */
void ordvec_index_free(ordvec_index_t *index);
/* ordvec-ffi/src/lib.rs:894-998: synchronous single-query search. */
ordvec_status_t ordvec_index_search(const ordvec_index_t *index, const ordvec_search_params_t *params,
                                    ordvec_hit_t *hits_out, uint64_t hits_capacity, uint64_t *returned_out,
                                    ordvec_search_stats_t *stats_out);

/* ---- Part 2: additive extension ------------------------------------------ */

/* Flags for Part 2 calls (bitwise OR). */
#define ORDVEC_FLAG_NONE 0
/* Array arguments are device (HBM) pointers; work is enqueued on the library
 * stream and the call returns without waiting for the GPU (a non-finite check
 * on device inputs still synchronises unless ORDVEC_FLAG_ASSUME_FINITE). */
#define ORDVEC_FLAG_DEVICE_PTRS 1
/* Caller guarantees inputs are finite; skips the NaN/Inf scan. */
#define ORDVEC_FLAG_ASSUME_FINITE 2
/* ordvec_fastscan_batch_run and ordvec_two_stage_batch_run only: enqueue the
 * launch sequence eagerly instead of replaying its captured HIP graphs (same
 * results). */
#define ORDVEC_FLAG_NO_GRAPH 4

/* Engine limits of this library.  Searches with k above ORDVEC_MAX_TOPK and
 * stage-1 candidate counts above ORDVEC_MAX_STAGE1_M run too, on slower
 * paths (whole candidate rows sorted; every query on the exact stage-1
 * pipeline); the limits still bind the serving loop, the allocation-free
 * entry point, Rank searches and FastScan. */
#define ORDVEC_MAX_TOPK 1024
#define ORDVEC_MAX_STAGE1_M 4096
#define ORDVEC_MAX_ENCODE_DIM 4096
#define ORDVEC_MAX_SIGN_DIM 8192

typedef struct ordvec_rankquant_t ordvec_rankquant_t;
typedef struct ordvec_sign_bitmap_t ordvec_sign_bitmap_t;
typedef struct ordvec_fastscan_t ordvec_fastscan_t;

/* Stream used by every call (a hipStream_t; NULL = the default stream).  The
 * previous stream is synchronised first. */
ordvec_status_t ordvec_set_stream(void *hip_stream);
/* Wait for all work the library enqueued. */
ordvec_status_t ordvec_synchronize(void);
/* Queries whose stage-1 candidates were recomputed by the exact fallback
 * (survivor buffer underfull or overflowed) since the library was loaded, on
 * the engine's default context: a monitoring counter (no reference
 * counterpart). */
ordvec_status_t ordvec_stage1_exact_queries(uint64_t *count);

/* Event timing of the engine's kernel families (no reference counterpart;
 * the reference's stats carry wall-clock only).  enable != 0 resets the
 * counters and records a hipEvent pair around every launch of a family;
 * read synchronises and returns launches and summed GPU milliseconds. */
#define ORDVEC_KERNEL_SIGN_SCAN 0
#define ORDVEC_KERNEL_STAGE1_SELECT 1
#define ORDVEC_KERNEL_RERANK 2
#define ORDVEC_KERNEL_FULL_SCAN 3
#define ORDVEC_KERNEL_FASTSCAN 4
#define ORDVEC_KERNEL_ENCODE 5
#define ORDVEC_KERNEL_PREPARE 6
#define ORDVEC_KERNEL_STAGE1_SAMPLE 7
#define ORDVEC_KERNEL_BACKWARD 8
ordvec_status_t ordvec_kernel_timing_enable(int enable);
/* As ordvec_kernel_timing_enable, for the families whose bit (1 << family)
 * is set in mask only, so that a timed run carries the events of the kernels
 * it reports and no others; mask 0 turns timing off. */
ordvec_status_t ordvec_kernel_timing_enable_families(uint32_t mask);
ordvec_status_t ordvec_kernel_timing_read(uint32_t family, uint64_t *launches, double *total_ms);
/* Execution-path switches (no counterpart in the reference): each selects
 * between two paths of this engine that return identical results, for tests
 * that compare them and for measurement.  Sets option to value (0 or 1) and,
 * when previous is not NULL, stores the value it replaces; value -1 only
 * reads.  Process-wide; calls in flight keep the path they started on.
 *  ORDVEC_OPTION_ONE_QUERY_PATH: 1 (default) runs a one-query two-stage call
 *    in three launches (stage-1 head, scan, select+rerank+merge tail); 0 runs
 *    it through the general batched pipeline.
 *  ORDVEC_OPTION_MFMA_SCAN: 1 (default) runs batched stage-1 scans of two or
 *    more queries on the matrix cores; 0 uses the vector-ALU scan.
 *  ORDVEC_OPTION_BATCH_TAIL: 1 (default) runs the stage-1 select and the
 *    rerank of a two-stage batch of 2 to 127 queries in one launch; 0 in
 *    separate launches (larger batches always use separate launches).
 *  ORDVEC_OPTION_SIDE_PREP: 1 (default) prepares the queries of a batch of
 *    256 or more for the rerank on a second stream beside the stage-1 sample;
 *    0 in line. */
#define ORDVEC_OPTION_ONE_QUERY_PATH 1
#define ORDVEC_OPTION_MFMA_SCAN 2
#define ORDVEC_OPTION_BATCH_TAIL 3
#define ORDVEC_OPTION_SIDE_PREP 4
ordvec_status_t ordvec_engine_option(uint32_t option, int64_t value, int64_t *previous);

/* RankQuant: src/quant.rs:438-491 new, 587-614 add, 713-847 search_asymmetric. */
ordvec_status_t ordvec_rankquant_new(uint64_t dim, uint32_t bits, ordvec_rankquant_t **out);
ordvec_status_t ordvec_rankquant_add(ordvec_rankquant_t *index, const float *vectors, uint64_t n, uint64_t flags);
ordvec_status_t ordvec_rankquant_info(const ordvec_rankquant_t *index, uint64_t *dim, uint32_t *bits,
                                      uint64_t *n_vectors);
/* Outputs are nq x min(k, n) (row-major), like SearchResults (src/lib.rs:314). */
ordvec_status_t ordvec_rankquant_search_asymmetric(const ordvec_rankquant_t *index, const float *queries,
                                                   uint64_t nq, uint64_t k, float *out_scores, int64_t *out_ids,
                                                   uint64_t flags);
/* src/quant.rs:1048-1083: one query, returns min(k, n_rows) hits in *returned. */
ordvec_status_t ordvec_rankquant_search_subset(const ordvec_rankquant_t *index, const float *query,
                                               const uint32_t *rows, uint64_t n_rows, uint64_t k,
                                               float *out_scores, int64_t *out_ids, uint64_t *returned,
                                               uint64_t flags);
/* src/quant.rs:1336-1394: CSR rows (offsets has nq+1 entries), outputs
 * nq x min(k, n), underfull rows padded with -inf / -1. */
ordvec_status_t ordvec_rankquant_search_subset_batched(const ordvec_rankquant_t *index, const float *queries,
                                                       uint64_t nq, const uint64_t *offsets,
                                                       const uint32_t *candidates, uint64_t n_candidates,
                                                       uint64_t k, float *out_scores, int64_t *out_ids,
                                                       uint64_t flags);
/* src/rank.rs:38-72 rank_transform of n rows of `dim` values (dim <= 65535):
 * out_ranks[r][k] = position of v[k] in the row's stable ascending order
 * (ties, including -0.0 / +0.0, by index). */
ordvec_status_t ordvec_rank_transform(const float *vectors, uint64_t n, uint64_t dim, uint16_t *out_ranks,
                                      uint64_t flags);
/* src/quant.rs:1541-1610 rankquant_eval_search: the eval-only symmetric
 * search at any bit width 1..8 (non-persistable widths included), corpus and
 * queries rank-bucketed on the fly; score = sum_j centre(bq_j) centre(bd_j) /
 * norm^2 with the empirical norm; outputs nq x min(k, n) by (score desc, id
 * asc).  dim <= 8192 on the GPU. */
ordvec_status_t ordvec_rankquant_eval_search(const float *corpus, uint64_t n, const float *queries, uint64_t nq,
                                             uint64_t dim, uint32_t bits, uint64_t k, float *out_scores,
                                             int64_t *out_ids, uint64_t flags);
/* Copy the packed codes (n x dim*bits/8, .ovrq payload order). */
ordvec_status_t ordvec_rankquant_codes(const ordvec_rankquant_t *index, uint8_t *out, uint64_t capacity,
                                       uint64_t flags);
/* src/quant.rs:892-1017: .ovrq persistence. */
ordvec_status_t ordvec_rankquant_write(const ordvec_rankquant_t *index, const char *path);
ordvec_status_t ordvec_rankquant_load(const char *path, ordvec_rankquant_t **out);
void ordvec_rankquant_free(ordvec_rankquant_t *index);
/* src/quant.rs:629-695 search (symmetric): the query is rank-bucketed and
 * scored through the scalar LUT scan, scale 1 / (R * R); outputs nq x
 * min(k, n); dim <= 4096. */
ordvec_status_t ordvec_rankquant_search(const ordvec_rankquant_t *index, const float *queries, uint64_t nq, uint64_t k,
                                        float *out_scores, int64_t *out_ids, uint64_t flags);
/* src/quant.rs:869-881 swap_remove: the last row moves into idx; *moved_from
 * (optional) receives its old index (n - 1).  idx >= n is BAD_ARGUMENT. */
ordvec_status_t ordvec_rankquant_swap_remove(ordvec_rankquant_t *index, uint64_t idx, uint64_t *moved_from);

/* Rank (full-precision u16 rank vectors): src/rank.rs:291-584.  GPU add and
 * symmetric search: dim <= 4096; asymmetric search: dim <= 8192. */
typedef struct ordvec_rank_t ordvec_rank_t;
/* src/rank.rs:317-324 new: 2 <= dim <= 65535. */
ordvec_status_t ordvec_rank_new(uint64_t dim, ordvec_rank_t **out);
/* src/rank.rs:337-356 add (rank transform per row). */
ordvec_status_t ordvec_rank_add(ordvec_rank_t *index, const float *vectors, uint64_t n, uint64_t flags);
ordvec_status_t ordvec_rank_info(const ordvec_rank_t *index, uint64_t *dim, uint64_t *n_vectors);
/* src/rank.rs:364-412 search (Spearman): nq x min(k, n). */
ordvec_status_t ordvec_rank_search(const ordvec_rank_t *index, const float *queries, uint64_t nq, uint64_t k,
                                   float *out_scores, int64_t *out_ids, uint64_t flags);
/* src/rank.rs:429-480 search_asymmetric: nq x min(k, n). */
ordvec_status_t ordvec_rank_search_asymmetric(const ordvec_rank_t *index, const float *queries, uint64_t nq,
                                              uint64_t k, float *out_scores, int64_t *out_ids, uint64_t flags);
/* Copy the rank vectors (n x dim u16, .ovr payload order). */
ordvec_status_t ordvec_rank_ranks(const ordvec_rank_t *index, uint16_t *out, uint64_t capacity, uint64_t flags);
/* src/rank.rs:506-517 swap_remove (as ordvec_rankquant_swap_remove). */
ordvec_status_t ordvec_rank_swap_remove(ordvec_rank_t *index, uint64_t idx, uint64_t *moved_from);
/* src/rank_io.rs:514-640: .ovr persistence (load checks every row is a
 * permutation of [0, dim)). */
ordvec_status_t ordvec_rank_write(const ordvec_rank_t *index, const char *path);
ordvec_status_t ordvec_rank_load(const char *path, ordvec_rank_t **out);
void ordvec_rank_free(ordvec_rank_t *index);

/* SignBitmap: src/sign_bitmap.rs:143-200 new/add, 330-483 top-m candidates. */
ordvec_status_t ordvec_sign_bitmap_new(uint64_t dim, ordvec_sign_bitmap_t **out);
ordvec_status_t ordvec_sign_bitmap_add(ordvec_sign_bitmap_t *index, const float *vectors, uint64_t n,
                                       uint64_t flags);
ordvec_status_t ordvec_sign_bitmap_info(const ordvec_sign_bitmap_t *index, uint64_t *dim, uint64_t *n_vectors);
/* Exact top-m by (hamming asc, id asc), rows sorted: out is nq x min(m, n);
 * out_hamming (optional) receives the matching distances. */
ordvec_status_t ordvec_sign_bitmap_top_m(const ordvec_sign_bitmap_t *index, const float *queries, uint64_t nq,
                                         uint64_t m, uint32_t *out_candidates, uint32_t *out_hamming,
                                         uint64_t flags);
/* src/sign_bitmap.rs:204-214 build_query_bitmap for nq queries: out_words
 * nq x dim/64 u64 (bit j of word j/64 set iff q[j] > 0).  Non-finite queries
 * are BAD_ARGUMENT (the reference asserts finiteness). */
ordvec_status_t ordvec_sign_bitmap_query_bits(const ordvec_sign_bitmap_t *index, const float *queries, uint64_t nq,
                                              uint64_t *out_words, uint64_t flags);
/* src/sign_bitmap.rs:492-548 score_all / score_all_batched_flat: out_scores
 * nq x n u32 sign agreements dim - popcount(doc XOR q), by document id. */
ordvec_status_t ordvec_sign_bitmap_score_all(const ordvec_sign_bitmap_t *index, const float *queries, uint64_t nq,
                                             uint32_t *out_scores, uint64_t flags);
/* Copy the bitmaps row-major (n x dim/64 u64, .ovsb payload order). */
ordvec_status_t ordvec_sign_bitmap_words(const ordvec_sign_bitmap_t *index, uint64_t *out, uint64_t capacity,
                                         uint64_t flags);
/* src/sign_bitmap.rs:593-660: .ovsb persistence. */
ordvec_status_t ordvec_sign_bitmap_write(const ordvec_sign_bitmap_t *index, const char *path);
ordvec_status_t ordvec_sign_bitmap_load(const char *path, ordvec_sign_bitmap_t **out);
void ordvec_sign_bitmap_free(ordvec_sign_bitmap_t *index);
/* src/sign_bitmap.rs:578-590 swap_remove (as ordvec_rankquant_swap_remove). */
ordvec_status_t ordvec_sign_bitmap_swap_remove(ordvec_sign_bitmap_t *index, uint64_t idx, uint64_t *moved_from);

/* Bitmap (top-bucket, constant-weight): src/bitmap.rs.  Bit j of a document
 * is set iff coordinate j is among its n_top highest ranks; a query's bits are
 * its n_top largest values; score = popcount(Q AND D).  The same indexes load
 * through ABI v1 (ordvec_index_load on .ovbm) where ordvec_index_search runs
 * Bitmap::search / Bitmap::search_subset.  GPU encode and search: dim <= 4096. */
typedef struct ordvec_bitmap_t ordvec_bitmap_t;
/* src/bitmap.rs:71-136 validate_params / new. */
ordvec_status_t ordvec_bitmap_new(uint64_t dim, uint64_t n_top, ordvec_bitmap_t **out);
/* src/bitmap.rs:140-162. */
ordvec_status_t ordvec_bitmap_add(ordvec_bitmap_t *index, const float *vectors, uint64_t n, uint64_t flags);
ordvec_status_t ordvec_bitmap_info(const ordvec_bitmap_t *index, uint64_t *dim, uint64_t *n_top, uint64_t *n_vectors);
/* src/bitmap.rs:202-250 search: nq x min(k, n) overlaps (as f32) and ids by
 * (overlap desc, id asc); k <= 4096. */
ordvec_status_t ordvec_bitmap_search(const ordvec_bitmap_t *index, const float *queries, uint64_t nq, uint64_t k,
                                     float *out_scores, int64_t *out_ids, uint64_t flags);
/* src/bitmap.rs:264-362 top_m_candidates(_batched): nq x min(m, n) ids. */
ordvec_status_t ordvec_bitmap_top_m(const ordvec_bitmap_t *index, const float *queries, uint64_t nq, uint64_t m,
                                    uint32_t *out_candidates, uint64_t flags);
/* src/bitmap.rs:477-505 search_subset per query over CSR rows (host arrays):
 * rows nq x k, the first min(k, row count) filled (count in out_counts,
 * optional), the rest -inf / -1; k <= 1024. */
ordvec_status_t ordvec_bitmap_search_subset_batched(const ordvec_bitmap_t *index, const float *queries, uint64_t nq,
                                                    const uint64_t *offsets, const uint32_t *rows, uint64_t n_rows,
                                                    uint64_t k, float *out_scores, int64_t *out_ids,
                                                    uint64_t *out_counts, uint64_t flags);
/* src/bitmap.rs:164-187 build_query_bitmap_fp32: nq x dim/64 u64. */
ordvec_status_t ordvec_bitmap_query_bits(const ordvec_bitmap_t *index, const float *queries, uint64_t nq,
                                         uint64_t *out_words, uint64_t flags);
/* src/bitmap.rs:425-460 body_overlap_scores_subset: out[i] = popcount(doc
 * rows[i] AND q_words) (host arrays; q_words = dim/64 u64). */
ordvec_status_t ordvec_bitmap_overlap_subset(const ordvec_bitmap_t *index, const uint64_t *q_words,
                                             const uint32_t *rows, uint64_t n_rows, uint32_t *out);
/* Row-major copy of the document bitmaps (.ovbm payload order). */
ordvec_status_t ordvec_bitmap_words(const ordvec_bitmap_t *index, uint64_t *out, uint64_t capacity);
/* src/rank_io.rs:826-957: .ovbm persistence (load checks every row has n_top bits). */
ordvec_status_t ordvec_bitmap_write(const ordvec_bitmap_t *index, const char *path);
ordvec_status_t ordvec_bitmap_load(const char *path, ordvec_bitmap_t **out);
void ordvec_bitmap_free(ordvec_bitmap_t *index);
/* src/bitmap.rs:515-527 swap_remove (as ordvec_rankquant_swap_remove). */
ordvec_status_t ordvec_bitmap_swap_remove(ordvec_bitmap_t *index, uint64_t idx, uint64_t *moved_from);

/* SignBitmap::add then RankQuant::add (src/sign_bitmap.rs:180-200,
 * src/quant.rs:587-614) of the same vectors in one pass over them: each row is
 * read once for its sign bits, its rank codes and the finite check (no
 * reference counterpart as a single call; the results equal the two adds).
 * Both indexes must have the same dim; a non-finite value leaves both
 * unchanged (BAD_ARGUMENT). */
ordvec_status_t ordvec_two_stage_add(ordvec_sign_bitmap_t *sign, ordvec_rankquant_t *rq, const float *vectors,
                                     uint64_t n, uint64_t flags);

/* Two-stage search in one call: SignBitmap::top_m_candidates_batched_serial_csr
 * followed by RankQuant::search_asymmetric_subset_batched_serial_into
 * (src/sign_bitmap.rs:453-483, src/quant.rs:1336-1394).  Both indexes must
 * hold the same documents.  Outputs nq x min(k, n). */
ordvec_status_t ordvec_two_stage_search(const ordvec_sign_bitmap_t *sign, const ordvec_rankquant_t *rq,
                                        const float *queries, uint64_t nq, uint64_t m, uint64_t k,
                                        float *out_scores, int64_t *out_ids, uint64_t flags);

/* TwoStageCandidatePolicy (src/quant.rs:235-262): how many stage-1
 * candidates a two-stage search takes for a given k.  max_candidates applies
 * only when has_max_candidates != 0 (the reference's Option).  The default is
 * {256, 32, none}. */
typedef struct {
    uint64_t min_candidates;
    uint64_t k_multiplier;
    uint64_t max_candidates;
    uint32_t has_max_candidates;
    uint32_t reserved;
} ordvec_two_stage_policy_t;
/* src/quant.rs:252-262 Default. */
void ordvec_two_stage_policy_default(ordvec_two_stage_policy_t *policy);
/* src/quant.rs:241-250 candidate_count: 0 when k or search_space is 0, else
 * min(max(min_candidates, k * k_multiplier (saturating)), max_candidates if
 * set, search_space).  A NULL policy is the default. */
uint64_t ordvec_two_stage_policy_candidate_count(const ordvec_two_stage_policy_t *policy, uint64_t k,
                                                 uint64_t search_space);
/* RankQuant::try_search_with_sign_probe[_with_policy] (src/quant.rs:1453-1497):
 * one query of `dim` values; m = candidate_count(k, n) sign-bitmap candidates,
 * then search_asymmetric_subset over them.  Writes min(k, m) hits (capacity
 * k) and their count to *returned.  policy NULL = the default policy.
 * Errors in the reference's order: sign dim or length mismatch (BAD_ARGUMENT),
 * dim != index dim (DIM_MISMATCH), non-finite query (NONFINITE_QUERY). */
ordvec_status_t ordvec_rankquant_search_with_sign_probe(const ordvec_rankquant_t *rq, const ordvec_sign_bitmap_t *sign,
                                                        const float *query, uint64_t dim, uint64_t k,
                                                        const ordvec_two_stage_policy_t *policy, float *out_scores,
                                                        int64_t *out_ids, uint64_t *returned, uint64_t flags);
/* ordvec_two_stage_search with m = candidate_count(k, n) of the policy (NULL
 * = default) for a batch of queries: outputs nq x min(k, n), rows padded with
 * -inf / -1 past min(k, m). */
ordvec_status_t ordvec_two_stage_search_policy(const ordvec_sign_bitmap_t *sign, const ordvec_rankquant_t *rq,
                                               const float *queries, uint64_t nq, uint64_t k,
                                               const ordvec_two_stage_policy_t *policy, float *out_scores,
                                               int64_t *out_ids, uint64_t flags);

/* Allocation-free two-stage search.  No reference counterpart: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same semantics and
 * results as ordvec_two_stage_search with ORDVEC_FLAG_DEVICE_PTRS (implied:
 * queries and outputs are device pointers), but the call allocates nothing:
 * its scratch is a fixed array of ORDVEC_NOALLOC_SCRATCH_BYTES per GPU,
 * reserved by the library until the process ends.  A call whose scratch
 * (growing with nq, m and the index size) would not fit, or with more than
 * ORDVEC_NOALLOC_MAX_QUERIES queries, fails with UNSUPPORTED_OPERATION before
 * any work is enqueued.  Returns once the results are written.  Calls share
 * the GPU's scratch: calls that overlap on one GPU (another thread, another
 * stream) race; the library's own calls are serialised by its lock.  Call
 * ordvec_init once first so that the scratch address is looked up at set-up
 * (otherwise the first call on each device does it). */
#define ORDVEC_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define ORDVEC_NOALLOC_MAX_QUERIES 1024
ordvec_status_t ordvec_two_stage_search_noalloc(const ordvec_sign_bitmap_t *sign, const ordvec_rankquant_t *rq,
                                                const float *queries, uint64_t nq, uint64_t m, uint64_t k,
                                                float *out_scores, int64_t *out_ids, uint64_t flags);

/* RankQuantFastscan b=2: src/fastscan.rs:571-705. */
ordvec_status_t ordvec_fastscan_new(uint64_t dim, ordvec_fastscan_t **out);
ordvec_status_t ordvec_fastscan_add(ordvec_fastscan_t *index, const float *vectors, uint64_t n, uint64_t flags);
ordvec_status_t ordvec_fastscan_search(const ordvec_fastscan_t *index, const float *queries, uint64_t nq,
                                       uint64_t k, float *out_scores, int64_t *out_ids, uint64_t flags);
ordvec_status_t ordvec_fastscan_info(const ordvec_fastscan_t *index, uint64_t *dim, uint64_t *n_vectors);
ordvec_status_t ordvec_fastscan_write(const ordvec_fastscan_t *index, const char *path);
ordvec_status_t ordvec_fastscan_load(const char *path, ordvec_fastscan_t **out);
void ordvec_fastscan_free(ordvec_fastscan_t *index);

/* FastScan b=4 (no reference counterpart: the reference's FastScan is b=2).
 * The b=2 recipe carried to one 4-bit rank code per coordinate: per query a
 * u8 LUT of 16 entries per coordinate, entry c = q_unit[j] * (c - 7.5) under
 * one global affine scale (round half away from zero), exact u32 sums, score
 * = (d * g_min + acc / scale) / rankquant_norm(d, 4), top-k by (score desc,
 * id asc).  bits = 2 is RankQuantFastscan; bits = 4 needs dim % 8 == 0 and
 * dim <= 2048, k <= 256, one add() as for b=2, and has no .ovfs form. */
ordvec_status_t ordvec_fastscan_new_bits(uint64_t dim, uint32_t bits, ordvec_fastscan_t **out);
ordvec_status_t ordvec_fastscan_bits(const ordvec_fastscan_t *index, uint32_t *bits);

/* Batched multi-index FastScan b=4: many indexes of one dim, each with its
 * own query group (queries_per_index[i] rows, in index order, of the query
 * array) and k.  The whole batch is one launch sequence (normalise, LUTs, one
 * persistent scan kernel pulling (index, query group, document range) tiles,
 * per-query merge) captured once in a HIP graph and replayed by run().
 * Results equal ordvec_fastscan_search on each index.  Outputs are
 * total_queries x kout (kout = max over indexes of min(k_i, n_i)); row q holds
 * min(k_i, n_i) entries, the rest -inf / -1.  autotune() times tile shapes
 * (queries per tile, 64-document blocks per tile, persistent grid) on the
 * given queries, keeps the fastest, re-captures the graph and remembers the
 * choice for batches of the same shape class. */
typedef struct ordvec_fastscan_batch_t ordvec_fastscan_batch_t;
ordvec_status_t ordvec_fastscan_batch_new(const ordvec_fastscan_t *const *indexes, uint64_t n_indexes,
                                          const uint32_t *queries_per_index, const uint32_t *k_per_index,
                                          ordvec_fastscan_batch_t **out);
ordvec_status_t ordvec_fastscan_batch_autotune(ordvec_fastscan_batch_t *batch, const float *queries, uint32_t reps,
                                               uint64_t flags);
ordvec_status_t ordvec_fastscan_batch_run(ordvec_fastscan_batch_t *batch, const float *queries, float *out_scores,
                                          int64_t *out_ids, uint64_t flags);
ordvec_status_t ordvec_fastscan_batch_config(ordvec_fastscan_batch_t *batch, uint64_t *total_queries, uint32_t *kout,
                                             uint32_t *queries_per_tile, uint32_t *blocks_per_tile, uint32_t *grid,
                                             uint64_t *n_tiles, uint64_t *n_tuned);
/* Pins the tile shape (no reference counterpart): queries_per_tile 4 or 8
 * (LDS LUT lookups), 16 (matrix-core tiles, dim % 64 == 0; each index's last
 * query set padded to 16) or 20 / 24 (whole 16-query sets on the matrix
 * cores, each index's remaining queries on LUT tiles of 4 / 8), 64-document
 * blocks per tile >= 1, persistent grid >= 1; re-captures the graph.
 * BAD_ARGUMENT for a shape this dim cannot run.  config() and tuning report
 * queries_per_tile in the same encoding. */
ordvec_status_t ordvec_fastscan_batch_set_config(ordvec_fastscan_batch_t *batch, uint32_t queries_per_tile,
                                                 uint32_t blocks_per_tile, uint32_t grid);
/* Row i of the last autotune: the configuration and its best time (ms). */
ordvec_status_t ordvec_fastscan_batch_tuning(ordvec_fastscan_batch_t *batch, uint64_t i, uint32_t *queries_per_tile,
                                             uint32_t *blocks_per_tile, uint32_t *grid, double *ms);
void ordvec_fastscan_batch_free(ordvec_fastscan_batch_t *batch);

/* ---- Two-stage batch plans (HIP graphs, autotuned launch shapes) -----------
 * No reference counterpart: the batched form of ordvec_two_stage_search
 * (src/sign_bitmap.rs:453-483 then src/quant.rs:1336-1394) for a fixed
 * (sign bitmap, RankQuant, m, k).  A plan owns a workspace sized once for
 * max_queries and, per (query count, query buffer, output buffers), the call's
 * launch sequence captured in two HIP graphs: stage 1 (query preparation,
 * sign words, sampled threshold, persistent matrix-core scan, select) up to
 * the copy of the status words, then the rerank.  run() replays them on the
 * library stream, waits for the status copy only (device pointers: the rerank
 * completes in stream order), and recomputes queries whose stage-1 survivors
 * could not prove exactness with the exact pipeline, as the one-call search
 * does; results equal ordvec_two_stage_search's.  The last four graph pairs
 * are kept.  autotune() times the stage-1 launch shapes on the given queries
 * (nq <= max_queries; host or device per flags), keeps the fastest and drops
 * the graphs.  While a plan lives, add and swap_remove on its two indexes
 * fail and their free is deferred to the plan's free.  One thread at a time
 * per plan.  m <= ORDVEC_MAX_STAGE1_M, k <= ORDVEC_MAX_TOPK,
 * max_queries <= 65536; outputs nq x min(k, n). */
typedef struct ordvec_two_stage_batch_t ordvec_two_stage_batch_t;
/* Launch shapes of a plan (0 = the built-in rule; none changes results). */
typedef struct ordvec_two_stage_shape_t {
    uint32_t scan_tiles;  /* 32-query tiles per matrix-core scan workgroup: 1, 2 or 4 */
    uint32_t scan_pairs;  /* target (document segment, query tile) pairs of the scan (rule: 1024) */
    uint32_t scan_flat;   /* 1: one scan workgroup per pair instead of the persistent XCD-grouped grid */
    uint32_t sample_wgs;  /* target workgroups of the matrix-core sample (rule: 256) */
    uint32_t side_prep;   /* 1: batches of 256+ prepare the rerank's queries on a second stream */
    uint32_t reserved[3]; /* zero */
} ordvec_two_stage_shape_t;
ordvec_status_t ordvec_two_stage_batch_new(const ordvec_sign_bitmap_t *sign, const ordvec_rankquant_t *rq,
                                           uint64_t max_queries, uint64_t m, uint64_t k,
                                           ordvec_two_stage_batch_t **out);
ordvec_status_t ordvec_two_stage_batch_run(ordvec_two_stage_batch_t *batch, const float *queries, uint64_t nq,
                                           float *out_scores, int64_t *out_ids, uint64_t flags);
ordvec_status_t ordvec_two_stage_batch_autotune(ordvec_two_stage_batch_t *batch, const float *queries, uint64_t nq,
                                                uint32_t reps, uint64_t flags);
ordvec_status_t ordvec_two_stage_batch_shape(const ordvec_two_stage_batch_t *batch, ordvec_two_stage_shape_t *shape);
/* Pins the shapes (BAD_ARGUMENT for scan_tiles outside {0, 1, 2, 4}, flags
 * other than 0 / 1 or nonzero reserved words) and drops the graphs. */
ordvec_status_t ordvec_two_stage_batch_set_shape(ordvec_two_stage_batch_t *batch,
                                                 const ordvec_two_stage_shape_t *shape);
/* Row i of the last autotune: the shapes and the time per call (ms). */
ordvec_status_t ordvec_two_stage_batch_tuning(const ordvec_two_stage_batch_t *batch, uint64_t i,
                                              ordvec_two_stage_shape_t *shape, double *ms);
/* Counters: graph pairs captured, graph calls, eager calls (ORDVEC_FLAG_NO_GRAPH),
 * queries recomputed by the exact stage-1 pipeline, rows of the last autotune. */
ordvec_status_t ordvec_two_stage_batch_stats(const ordvec_two_stage_batch_t *batch, uint64_t *graph_captures,
                                             uint64_t *graph_calls, uint64_t *eager_calls, uint64_t *exact_queries,
                                             uint64_t *tuned_shapes);
void ordvec_two_stage_batch_free(ordvec_two_stage_batch_t *batch);

/* ---- Differentiable rerank -------------------------------------------------
 * No reference counterpart: the forward is the reference's asymmetric score
 * (src/quant.rs:1126-1281 scoring, src/quant_kernels.rs) evaluated for given
 * ids instead of a top-k, and the backward gives its gradient, so a caller
 * can train query-side parameters through a fixed candidate selection.
 * ids: [nq][k] int64 in [-1, n) (-1 = no row: score -inf, no gradient).
 * score_ids: out_scores[nq][k], bit-identical to the search kernels' scores.
 * score_ids_backward: grad_queries[nq][dim] = dL/dqueries (raw, unnormalised
 * queries) given grad_scores[nq][k] = dL/dscores, re-decoding the rows' codes
 * (no score matrix is kept).  k <= 4096.  Out-of-range ids fail with
 * ROW_ID_OUT_OF_RANGE (device pointers: after the kernels ran). */
ordvec_status_t ordvec_rankquant_score_ids(const ordvec_rankquant_t *index, const float *queries, uint64_t nq,
                                           const int64_t *ids, uint64_t k, float *out_scores, uint64_t flags);
ordvec_status_t ordvec_rankquant_score_ids_backward(const ordvec_rankquant_t *index, const float *queries,
                                                    uint64_t nq, const int64_t *ids, uint64_t k,
                                                    const float *grad_scores, float *grad_queries, uint64_t flags);

/* ---- Serving: continuous batching over resident indexes -------------------
 * No reference counterpart (ordvec serves in-process, one call at a time);
 * built on ordvec_two_stage_search's semantics.  `slots` worker contexts (each
 * a HIP stream with its own workspace) take pending requests in arrival order
 * and coalesce them up to the batch capacity that fits each slot's share of
 * `memory_budget_bytes` (and `max_batch_queries`); batches of different slots
 * overlap on the GPU.  A request's result is independent of the mix: rows are
 * the first min(k, n) entries of the batch's (score desc, id asc) order. */
typedef struct ordvec_server_t ordvec_server_t;

typedef struct {
    uint64_t struct_size;
    uint64_t requests;
    uint64_t queries;
    uint64_t errors;
    double duration_s;
    double offered_qps;   /* queries per second offered (open loop) */
    double achieved_qps;  /* queries completed per second */
    double achieved_rps;  /* requests completed per second */
    double p50_ms, p90_ms, p99_ms, p999_ms, max_ms, mean_ms;  /* request latency, arrival -> result */
    double batches;
    double mean_batch_queries;
    double p99_queue_ms;           /* arrival -> its batch starts */
    double p99_service_ms;         /* batch start -> results copied back */
    uint64_t exact_path_queries;   /* queries whose stage 1 was recomputed exactly */
    uint64_t pool_bytes;           /* device memory the server lends to batches (its budget) */
    uint64_t peak_pool_bytes;      /* most of it held by batches at one time during the run */
    uint64_t overlapped_batches;   /* batches started while another was still on the GPU */
    uint64_t max_batches_in_flight;
    uint64_t workspace_overflows;  /* batch calls that needed more than their share (allocated) */
} ordvec_serve_report_t;

ordvec_status_t ordvec_server_start(const ordvec_sign_bitmap_t *sign, const ordvec_rankquant_t *rq, uint64_t m,
                                    uint64_t max_k, uint32_t slots, uint64_t max_batch_queries,
                                    uint64_t memory_budget_bytes, ordvec_server_t **out);
/* Batch capacity chosen at start and the number of batches/queries run so far. */
ordvec_status_t ordvec_server_info(ordvec_server_t *server, uint64_t *max_batch_queries, uint64_t *batches,
                                   uint64_t *batched_queries);
/* Host buffers: queries [nq][dim]; outputs [nq][min(k, n)] must stay valid
 * until ordvec_server_wait returns for the ticket. */
ordvec_status_t ordvec_server_submit(ordvec_server_t *server, const float *queries, uint64_t nq, uint64_t k,
                                     float *out_scores, int64_t *out_ids, uint64_t *ticket);
/* Blocks until the request completes; reports submit -> completion latency
 * and the size of the batch it ran in. */
ordvec_status_t ordvec_server_wait(ordvec_server_t *server, uint64_t ticket, uint64_t *latency_ns,
                                   uint64_t *batch_queries);
/* Open-loop load: Poisson request arrivals at offered_qps / mean request size
 * per second for duration_s, request sizes uniform in [min_nq, max_nq]
 * (queries taken from query_pool [pool_n][dim]), k uniform in [min_k, max_k];
 * latency counts from the scheduled arrival. */
ordvec_status_t ordvec_server_open_loop(ordvec_server_t *server, const float *query_pool, uint64_t pool_n,
                                        double offered_qps, double duration_s, uint32_t min_nq, uint32_t max_nq,
                                        uint32_t min_k, uint32_t max_k, uint64_t seed,
                                        ordvec_serve_report_t *report);
/* Drains pending requests, joins the slots and frees their resources. */
void ordvec_server_stop(ordvec_server_t *server);

#ifdef __cplusplus
}
#endif

#endif /* ORDVEC_H */
