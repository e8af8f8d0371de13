// Shared declarations for the MI355X ordinal-retrieval engine.
//
// Layouts in HBM (all device-resident, uploaded once):
//   RankQuant codes   row-major, n x bpv bytes, MSB-first packing
//                     (bytes identical to the .ovrq payload, src/rank.rs:128-156)
//   SignBitmap words  doc-interleaved: 64-doc blocks, [block][word][lane] u64,
//                     so one wave's load of word w is 512 contiguous bytes
//   FastScan payload  block-32 [block][pair][lane] nibble bytes
//                     (identical to the .ovfs payload, src/fastscan.rs:87-110)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace ovx {

// Status codes mirror ordvec-ffi (ordvec.h ORDVEC_STATUS_*).
enum Status : uint32_t {
    kOk = 0,
    kNullPointer = 1,
    kBadArgument = 2,
    kBadStructSize = 3,
    kUnsupportedFormat = 4,
    kCorruptIndex = 5,
    kIo = 6,
    kDimMismatch = 7,
    kNonfiniteQuery = 8,
    kRowIdOutOfRange = 9,
    kBufferTooSmall = 10,
    kUnsupportedOperation = 11,
    kPanic = 12,
    kInternal = 13,
};

struct Error : std::runtime_error {
    uint32_t status;
    Error(uint32_t s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define OVX_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess)                                                              \
            throw ::ovx::Error(::ovx::kInternal, std::string("HIP error: ") +              \
                                                     hipGetErrorString(_e) + " at " #call); \
    } while (0)

// Compile-time engine limits (documented in the header and the design notes).
constexpr uint32_t kWave = 64;
constexpr uint32_t kTopKMax = 1024;      // largest k a single block-level top-k keeps
constexpr uint32_t kStage1MaxM = 4096;   // largest stage-1 candidate count per query
constexpr uint32_t kTieCap = 4096;       // boundary-tier ids buffered per query before
                                         // the ordered fallback scan is used
constexpr uint32_t kEncodeMaxDim = 4096; // LDS rank sort capacity per row
constexpr uint32_t kSignMaxDim = 8192;   // LDS Hamming histogram capacity

// ---- optional per-kernel event timing (capi.cpp) ----------------------------
// Families of launches whose GPU time can be read back through the C ABI.
enum KernelFamily : int {
    kFamSignScan = 0,    // stage-1 XOR-popcount scan (k_sign_scan)
    kFamStage1Select,    // stage-1 threshold + compaction + finalize
    kFamRerank,          // stage-2 / subset rerank
    kFamFullScan,        // full asymmetric scans and their merge
    kFamFastScan,        // FastScan LUT build, scan and merge
    kFamEncode,          // rank / sign / FastScan encode
    kFamPrepare,         // query normalisation
    kFamSample,          // stage-1 sampled threshold pass
    kFamBackward,        // score-by-id backward (differentiable rerank)
    kFamCount
};
void ktime_begin(int family, hipStream_t s);
void ktime_end(int family, hipStream_t s);

// ---- kernel launchers (kernels.hip) ----------------------------------------

struct QueryPrep {
    float* qhat;     // [nq][d] L2-normalised queries
    float* offset;   // [nq] TopK score offset (centre term for SIMD tiers, else 0)
    uint64_t* qbits; // [nq][qpv] sign bits (may be null)
};

int select_tier(uint32_t d, uint32_t bits);
float rankquant_inv_norm(uint32_t d, uint32_t bits);

void launch_check_finite(const float* x, uint64_t count, uint32_t* flag, hipStream_t s);
// Scores of given ids (n_docs bounds the ids; -1 gives -inf; others outside
// [0, n_docs) give -inf and set *bad).
void launch_score_ids(const uint8_t* codes, uint64_t n_docs, uint32_t d, uint32_t bits, int tier, float inv_norm,
                      const float* qhat, const float* offset, uint64_t nq, const int64_t* ids, uint32_t k, float* out,
                      uint32_t* bad, hipStream_t s);
// dL/dx of the scores above given dL/dscores, through x -> x / ||x||.
void launch_score_ids_backward(const uint8_t* codes, uint64_t n_docs, uint32_t d, uint32_t bits, float inv_norm,
                               const float* x, uint64_t nq, const int64_t* ids, uint32_t k, const float* grad,
                               float* dx, uint32_t* bad, hipStream_t s);
// dst row i = src row idx[i] / dst row idx[i] = src row i, rows of `row`
// elements, in one launch.
void gather_rows_u64(uint64_t* dst, const uint64_t* src, const uint32_t* idx, uint64_t nrows, uint32_t row,
                     hipStream_t s);
void scatter_rows_u32(uint32_t* dst, const uint32_t* src, const uint32_t* idx, uint64_t nrows, uint32_t row,
                      hipStream_t s);
void launch_prepare_queries(const float* q, uint64_t nq, uint32_t d, uint32_t bits, int tier, float inv_norm,
                            float* qhat, float* offset, uint32_t* qstat, hipStream_t s);
void launch_sign_encode(const float* x, uint64_t n, uint32_t d, uint64_t* out, uint64_t first_doc, bool interleaved,
                        hipStream_t s);
// Rank transform of n rows of any d <= 65535 by a per-row sort: u16 ranks
// (ranks != null) or RankQuant codes of `bits` (codes != null); work =
// rank_sort_work_bytes(n, d), n <= 65535 rows per launch.
size_t rank_sort_work_bytes(uint64_t n, uint32_t d, hipStream_t s);
void launch_rank_sort(const float* x, uint64_t n, uint32_t d, uint16_t* ranks, uint32_t bits, uint8_t* codes,
                      void* work, hipStream_t s);
// rankquant_eval_search: u8 buckets of width `bits` from u16 ranks; the
// search itself (query ranks qranks [nq][d], placed by the caller at the start
// of work = eval_search_work_bytes; top-k by (score desc, id asc)).
void launch_eval_buckets(const uint16_t* ranks, uint64_t count, uint32_t d, uint32_t bits, uint8_t* out,
                         hipStream_t s);
size_t eval_search_work_bytes(uint64_t n, uint64_t nq, uint32_t d, hipStream_t s);
void launch_eval_search(const uint8_t* buckets, uint64_t n, uint32_t d, uint32_t bits, float inv_norm_sq,
                        const uint16_t* qranks, uint64_t nq, uint32_t k, float* out_s, int64_t* out_i, void* work,
                        hipStream_t s);
// out[q][doc] = d - popcount(doc XOR q) for every document (interleaved words).
void launch_sign_score_all(const uint64_t* words, uint64_t n, uint32_t d, const uint64_t* qbits, uint64_t nq,
                           uint32_t* out, hipStream_t s);
void launch_rank_encode(const float* x, uint64_t n, uint32_t d, uint32_t bits, uint8_t* codes, hipStream_t s);
// One pass over n f32 rows: RankQuant codes (codes != null, row-major; rows
// whose values do not bin finely enough are finished by the whole-row sort
// in a second launch), sign words (words != null, interleaved layout from
// document first_doc), and flag |= 1 when any value is non-finite.  Returns false (nothing launched)
// for a (d, bits) without a one-pass form; the callers then use the kernels
// above.
bool launch_encode_rows(const float* x, uint64_t n, uint32_t d, uint32_t bits, uint8_t* codes, uint64_t* words,
                        uint64_t first_doc, uint32_t* flag, void* work, hipStream_t s);
// Device scratch of launch_encode_rows for n rows (rows handed to the
// whole-row sort).
size_t encode_rows_work_bytes(uint64_t n);
bool encode_rows_supported(uint32_t d, uint32_t bits, bool codes, bool words);
void launch_fastscan_pack(const uint8_t* codes, uint64_t n, uint32_t d, uint8_t* packed, hipStream_t s);
// Rank vectors (u16 [n][d]) of rows; centre(bucket(rank)) of rank vectors.
void launch_rank_u16(const float* x, uint64_t n, uint32_t d, uint16_t* ranks, hipStream_t s);
void launch_rank_centres(const uint16_t* ranks, uint64_t count, uint32_t d, uint32_t bits, float* out,
                         hipStream_t s);
// Rank index full search: symmetric (query ranks) or asymmetric (q_hat and
// its sum); work = fullscan_work_bytes(n, nq, k).
void launch_rank_search(const uint16_t* ranks, uint64_t n, uint32_t d, bool symmetric, const uint16_t* qranks,
                        const float* qhat, const float* qsum, uint64_t nq, uint32_t k, void* work, float* out_s,
                        int64_t* out_i, hipStream_t s);
// q_hat = l2_normalise(q) and q_sum = sum q_hat (sequential, from -0.0).
void launch_prepare_sum(const float* q, uint64_t nq, uint32_t d, float* qhat, float* qsum, uint32_t* qstat,
                        hipStream_t s);
// Loaded-file validation on the device: lowest invalid row -> *first_bad
// (atomicMin; the caller initialises it to ~0).
void launch_check_rq_rows(const uint8_t* codes, uint64_t n, uint32_t d, uint32_t bits, unsigned long long* first_bad,
                          hipStream_t s);
void launch_check_rank_rows(const uint16_t* ranks, uint64_t n, uint32_t d, unsigned long long* first_bad,
                            hipStream_t s);
// Row-major bitmaps -> interleaved layout (documents first_doc..); ntop > 0
// also checks every row's popcount.
void launch_interleave_rows(const uint64_t* rows, uint64_t cnt, uint64_t first_doc, uint32_t qpv, uint64_t* words,
                            uint32_t ntop, unsigned long long* first_bad, hipStream_t s);
// swap_remove on the interleaved bitmap layout: document `from` -> `to`.
void launch_move_doc_interleaved(uint64_t* words, uint32_t qpv, uint64_t from, uint64_t to, hipStream_t s);
// Top-bucket bitmaps of documents (interleaved layout), d <= kEncodeMaxDim.
void launch_topbits_docs(const float* x, uint64_t n, uint32_t d, uint32_t ntop, uint64_t* out, uint64_t first_doc,
                         hipStream_t s);
// build_query_bitmap_fp32 for each query (qstat bit 0: non-finite input).
void launch_bitmap_query_bits(const float* queries, uint64_t nq, uint32_t d, uint32_t ntop, uint64_t* qbits,
                              uint32_t* qstat, hipStream_t s);
// Bitmap subset search: CSR rows per query -> top-k by (overlap desc, row asc).
void launch_bm_subset(const uint64_t* words, uint32_t d, uint64_t n_docs, const uint64_t* qbits, uint64_t nq,
                      const uint64_t* offsets, const uint32_t* rows, uint32_t ntop, uint32_t out_k, float* out_s,
                      int64_t* out_i, hipStream_t s);
// Overlap of listed rows (rows < n) with one query bitmap.
void launch_bm_overlap_rows(const uint64_t* words, uint32_t d, const uint64_t* q, const uint32_t* rows, uint64_t count,
                            uint32_t* out, hipStream_t s);
// (id, distance) rows -> (overlap as f32, i64 id).
void launch_bm_finish(const uint32_t* cand, const uint32_t* cand_h, uint64_t count, uint32_t ntop, float* out_s,
                      int64_t* out_i, hipStream_t s);

// Stage 1: exact top-m by (hamming asc, id asc) for nq queries over the
// interleaved sign corpus.  Writes cand[nq][m_eff] (sorted) and, when
// cand_h != null, the matching Hamming distances.
struct Stage1Work {
    uint16_t* H;          // [nq][n_pad]
    uint16_t* hist_part;  // [nq][nseg][nbins]
    uint32_t* thr;        // [nq][4]
    uint32_t* counters;   // [nq][2]
    uint64_t* below;      // [nq][m_eff]       (m_eff <= kStage1MaxM)
    uint32_t* ties;       // [nq][kTieCap]     (m_eff <= kStage1MaxM)
    uint32_t* run_off;    // [nq][nseg][nbins] (m_eff > kStage1MaxM: ordered scatter)
};
struct Stage1Plan {
    uint32_t qt, nseg, blocks_per_seg, nbins, hwords, lds_bytes, nseg3;
    uint64_t n_pad;
};
Stage1Plan plan_stage1(uint64_t n, uint32_t d, uint64_t nq);
size_t stage1_work_bytes(const Stage1Plan& p, uint64_t nq, uint64_t m_eff);
void carve_stage1(const Stage1Plan& p, uint64_t nq, uint64_t m_eff, void* base, Stage1Work* w);
void launch_stage1(const uint64_t* words, uint64_t n, uint32_t d, const uint64_t* qbits, uint64_t nq, uint64_t m_eff,
                   const Stage1Plan& p, const Stage1Work& w, uint32_t* cand, uint32_t* cand_h, uint32_t ntop,
                   hipStream_t s);
// ntop == 0: sign bitmaps, distance = popcount(d XOR q); ntop > 0: top-bucket
// bitmaps, distance = ntop - popcount(d AND q) (queries then get their top-ntop
// bits in the fast path's prep).

// Stage 1 fast path (sampled threshold + survivor compaction); queries whose
// survivor buffer cannot prove exactness get qstat bit 1 and are recomputed
// by launch_stage1 (the exact histogram pipeline).
constexpr uint32_t kFastCap = 8192;  // survivor keys per query (64 KB LDS sort)
// Launch shapes of the batched stage 1 that do not change its results
// (0 = the built-in rule); a two-stage batch plan times them and keeps the
// fastest.
struct FastTune {
    uint32_t scan_nqt;    // matrix-core scan: 32-query tiles per workgroup (1, 2 or 4)
    uint32_t scan_pairs;  // matrix-core scan: target (document segment, query tile) pairs (rule: 1024)
    uint32_t scan_flat;   // 1: one workgroup per pair instead of the persistent XCD-grouped grid
    uint32_t sample_wgs;  // matrix-core sample: target workgroups (rule: 256)
};
struct FastPlan {
    uint32_t qt, nbins, hwords, lds_bytes;
    uint32_t stride, nsample, sample_bps, nsseg;
    uint32_t scan_qt, scan_bps, nseg;
    uint32_t scan_wgl;  // VALU scan: survivor list entries per workgroup (0: global appends)
    uint32_t s_target;  // ~0u: admit every document
    uint32_t cap;
    uint32_t sparse;    // expected survivors at most 1 in 12-14 documents (the matrix-core scan's lists hold them)
    FastTune tune;      // zero: built-in rules
};
struct FastWork {
    uint32_t* hist;  // [nq][nbins]
    uint32_t* thr;   // [nq]
    uint32_t* cnt;   // [nq]
    uint64_t* keys;  // [nq][cap]
};
FastPlan plan_stage1_fast(uint64_t n, uint32_t d, uint64_t nq, uint64_t m_eff);
size_t stage1_fast_bytes(const FastPlan& p, uint64_t nq);
void carve_stage1_fast(const FastPlan& p, uint64_t nq, void* base, FastWork* w);
// The rerank's query preparation (launch_prepare_queries' arguments for the
// same nq queries).  Handed to launch_stage1_fast, it runs on extra
// workgroups of the vector-ALU scan, beside the scan instead of before it;
// where the matrix-core scan is used it is a launch of its own.
struct PrepFuse {
    const float* q;
    float* qhat;
    float* offset;
    uint32_t* qstat;  // may be null
    uint32_t d, bits;
    int tier;
    float inv_norm;
};
// as_set: the caller needs each query's top-m candidates as a set (cand in
// no particular order, cand_h null), as the two-stage rerank does.
void launch_stage1_fast(const uint64_t* words, uint64_t n, uint32_t d, const float* queries, uint64_t nq,
                        uint64_t m_eff, const FastPlan& p, const FastWork& w, uint64_t* qbits, uint32_t* cand,
                        uint32_t* cand_h, uint32_t* qstat, uint32_t ntop, hipStream_t s,
                        const PrepFuse* prep = nullptr, bool select = true, hipEvent_t before_scan = nullptr,
                        bool as_set = false);

// One-query two-stage path: stage 1 in two launches (head: sign words,
// sampled histogram and threshold; then the vector-ALU scan with the rerank's
// query preparation beside it) and the tail (select + rerank + merge) in one.
// hist ([nbins]) and done (one word for the head, nq for the tail) are device
// words that are zero before and after each launch.  The tail writes each
// query's final status word to qstat and, when hstat is not null, to that
// host-visible array.
bool stage1_plain_supported(uint64_t n, uint32_t d, uint64_t m_eff, uint64_t out_k);
// Whether a batch can take stage 1 without its select and the tail after it
// (survivor buffers of kFastCap keys, m_eff and out_k within the tail's limits).
bool s1_tail_supported(uint64_t n, uint32_t d, uint64_t nq, uint64_t m_eff, uint64_t out_k);
// Whether batched stage-1 scans use the matrix cores; value 0 / 1 sets it,
// -1 only reads.  Returns the value before the call.
bool set_mfma_scan(int value);
void launch_stage1_plain(const uint64_t* words, uint64_t n, uint32_t d, const float* query, const FastPlan& p,
                         uint32_t* hist, uint32_t* done, uint32_t* thr, uint32_t* cnt, uint64_t* keys,
                         uint64_t* qbits, uint32_t* qstat, const PrepFuse& prep, hipStream_t s);
void launch_s1_tail(const uint64_t* keys, const uint32_t* cnt, const uint32_t* thr, uint32_t cap, uint64_t m_eff,
                    const uint8_t* codes,
                    uint32_t d, uint32_t bits, int tier, float inv_norm, const float* qhat, const float* offset,
                    uint64_t n_docs, uint64_t nq, uint32_t out_k, float* out_s, int64_t* out_i, uint32_t nsplit,
                    void* work, uint32_t* done, uint32_t* qstat, uint32_t* hstat, hipStream_t s);

// Stage 2 / subset rerank over candidate rows.  Rows are CSR (offsets != null)
// or uniform (row q = cands[q*row_m .. (q+1)*row_m)).  Output nq x out_k.
// Small batches split each row across `nsplit` workgroups (partial top-k in
// `work`, rerank_work_bytes) and merge; rerank_nsplit picks the split.
uint32_t rerank_nsplit(uint64_t nq, uint64_t avg_m);
size_t rerank_work_bytes(uint64_t nq, uint32_t out_k, uint32_t nsplit);
void launch_rerank(const uint8_t* codes, uint32_t d, uint32_t bits, int tier, float inv_norm, const float* qhat,
                   const float* offset, uint64_t nq, const uint64_t* offsets, const uint32_t* cands, uint64_t row_m,
                   uint64_t n_docs, const uint32_t* qmap, uint32_t out_k, float* out_s, int64_t* out_i,
                   uint32_t nsplit, void* work, hipStream_t s);

// Per-row ascending sort of 64-bit keys (segsort.hip); rows are
// [offsets[i], offsets[i+1]) of a device array of nseg + 1 offsets.
size_t segsort_temp_bytes(uint64_t total, uint32_t nseg, const uint64_t* offsets, hipStream_t s);
void segsort_keys(void* temp, size_t temp_bytes, const uint64_t* in, uint64_t* out, uint64_t total, uint32_t nseg,
                  const uint64_t* offsets, hipStream_t s);

// Large k (out_k above kTopKMax): score keys of whole rows, per-row sort,
// gather of the first out_k.  Rows are CSR (offsets, device, nq + 1 entries),
// uniform rows of row_m candidates, or every document (cands == null, row_m =
// n_docs); total = number of entries over all rows; work =
// large_k_work_bytes(total, nq).
size_t large_k_work_bytes(uint64_t total, uint64_t nq, hipStream_t s);
void launch_rerank_large(const uint8_t* codes, uint32_t d, uint32_t bits, int tier, float inv_norm, const float* qhat,
                         const float* offset, uint64_t nq, const uint64_t* offsets, const uint32_t* cands,
                         uint64_t row_m, uint64_t total, uint64_t n_docs, uint32_t out_k, float* out_s,
                         int64_t* out_i, void* work, hipStream_t s);
void launch_bm_subset_large(const uint64_t* words, uint32_t d, uint64_t n_docs, const uint64_t* qbits, uint64_t nq,
                            const uint64_t* offsets, const uint32_t* rows, uint64_t total, uint32_t ntop,
                            uint32_t out_k, float* out_s, int64_t* out_i, void* work, hipStream_t s);

// Full scans: per (query, segment) block top-k into partial keys, then merge.
size_t fullscan_work_bytes(uint64_t n, uint64_t nq, uint32_t k);
void launch_fullscan_rq(const uint8_t* codes, uint64_t n, uint32_t d, uint32_t bits, int tier, float inv_norm,
                        const float* qhat, const float* offset, uint64_t nq, uint32_t k, void* work, float* out_s,
                        int64_t* out_i, hipStream_t s);
void launch_fastscan_prepare(const float* qhat, uint64_t nq, uint32_t d, uint8_t* lut, float* bias_invq,
                             hipStream_t s);
void launch_fullscan_fs(const uint8_t* packed, uint64_t n, uint32_t d, const uint8_t* lut, const float* bias_invq,
                        float inv_norm, uint64_t nq, uint32_t k, void* work, float* out_s, int64_t* out_i,
                        hipStream_t s);

// ---- FastScan b=4 and the batched multi-index form ----------------------------
// One index of a batch: codes in 64-document blocks of interleaved dwords.
struct Fs4IndexDesc {
    const uint32_t* packed;
    uint64_t n;
};
// Up to four queries (consecutive rows q0 .. q0+count-1) sharing one LUT group.
struct Fs4Group {
    uint32_t q0, count;
};
// Work unit: groups g0 .. g0+ng-1 (ng <= QG) of one index x 64-doc blocks [blk0, blk1).
struct Fs4Tile {
    uint32_t index, g0, ng, blk0, blk1, pad;
};
// Per query: its partial top-k rows are prow + seg * stride for seg < nseg.
struct Fs4Merge {
    uint32_t prow, nseg, stride, k;
};
constexpr uint32_t kFs4MaxK = 256;
constexpr uint32_t kFs4MaxDim = 2048;

size_t fs4_scan_lds_bytes(uint32_t qg, uint32_t d);
// Matrix-core tiles (qg == 4: 16 queries): dims they support, the B-fragment
// bytes per 16-query set, and the kernel building them from the u8 LUTs
// (query sets: Fs4Group{first LUT group, number of groups}).
bool fs4_mfma_ok(uint32_t d);
size_t fs4_frag_bytes(uint32_t d);
void launch_fs4_frag(const uint8_t* lut, const Fs4Group* qsets, uint32_t nqsets, uint32_t d, uint32_t* frag,
                     hipStream_t s);
void launch_fs4_pack(const uint8_t* codes, uint64_t n, uint32_t d, uint32_t* packed, hipStream_t s);
void launch_fs4_lut(const float* qhat, uint32_t d, const Fs4Group* groups, uint32_t ngroups, uint8_t* lut,
                    float* bias_invq, hipStream_t s);
// counter != nullptr: persistent launch of `grid` workgroups pulling tiles
// (the counter is reset inside the launch sequence, graph-safe).
void launch_fs4_scan(uint32_t qg, uint32_t grid, const Fs4IndexDesc* indexes, const Fs4Tile* tiles, uint32_t ntiles,
                     uint32_t* counter, const Fs4Group* groups, const uint8_t* lut, const float* bias_invq, uint32_t d,
                     float inv_norm, uint32_t kcap, uint64_t* partial, hipStream_t s);
void launch_fs4_merge(const uint64_t* partial, const Fs4Merge* merges, uint32_t nq, uint32_t kcap,
                      uint32_t kout, float* out_s, int64_t* out_i, hipStream_t s);

}  // namespace ovx
