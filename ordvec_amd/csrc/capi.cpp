// Host side of the engine: the ordvec C ABI (v1 drop-in plus the additive
// extension), index handles resident in HBM, persisted formats, validation
// and the per-call device pipelines.  Errors cross the ABI as status codes
// with a thread-local message, mirroring ordvec-ffi/src/lib.rs:195-239.
#include "ordvec.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "common.h"

namespace ovx {
namespace {

// ---------------------------------------------------------------------------
// Errors (ordvec-ffi/src/lib.rs:195-239)
// ---------------------------------------------------------------------------
thread_local std::string g_last_error;

template <typename F>
ordvec_status_t boundary(F&& f) {
    try {
        f();
        g_last_error.clear();
        return ORDVEC_STATUS_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return ORDVEC_STATUS_PANIC;
    } catch (...) {
        g_last_error = "unknown C++ exception crossing ordvec C ABI";
        return ORDVEC_STATUS_PANIC;
    }
}

[[noreturn]] void fail(uint32_t status, const std::string& msg) { throw Error(status, msg); }

#define NONNULL(p, name) \
    if (!(p)) fail(kNullPointer, name " pointer is NULL")

// Limits shared with the reference loaders (src/rank_io.rs:124-138,268-286).
constexpr uint64_t kMaxDim = 65535;
constexpr uint64_t kMaxSignDim = 1ull << 24;
constexpr uint64_t kMaxVectors = 64ull * 1024 * 1024;
constexpr uint64_t kMaxPayload = 128ull * 1024 * 1024 * 1024;

// ---------------------------------------------------------------------------
// Engine: one stream, a grow-only workspace and staging area, one mutex.
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t size = 0;
    void* get(size_t bytes, hipStream_t stream) {
        if (bytes <= size && p) return p;
        if (p) {
            OVX_HIP(hipStreamSynchronize(stream));
            OVX_HIP(hipFree(p));
            p = nullptr;
        }
        size_t want = std::max(bytes, size + size / 2);
        if (hipMalloc(&p, std::max<size_t>(want, 256)) != hipSuccess) {
            (void)hipGetLastError();
            size = 0;
            if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
                (void)hipGetLastError();
                p = nullptr;
                fail(kInternal, "device workspace allocation of " + std::to_string(bytes) + " bytes failed");
            }
            want = bytes;
        }
        size = std::max<size_t>(want, 256);
        return p;
    }
};

// An execution context: one HIP stream with its own workspaces and
// host-visible status buffer.  API calls use the engine's default context
// (serialised by the engine mutex); serving slots each own one.
struct Ctx {
    hipStream_t stream = nullptr;
    uint64_t exact_queries = 0;  // queries whose stage 1 went through the exact recomputation
    bool timed = false;  // kernel-family timing applies (API calls and two-stage batch plans; not serving slots)
    DevBuf ws, ws2;  // per-call workspace; ws2 serves the exact stage-1 recomputation
    uint32_t* pinned = nullptr;
    size_t pinned_count = 0;
    hipEvent_t status_ev = nullptr;  // marks the copy of a call's status words to `pinned`
    // Device words that stay zero between calls: the one-query path's sample
    // histogram and arrival counters (the kernels that use them clear them
    // before they exit).  Zeroed once, when allocated.
    uint32_t* zw = nullptr;
    size_t zw_count = 0;

    // A second stream for work that runs beside the main one within a call
    // (forked from and joined back into `stream`; default priority).
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    // A region lent to this context for one call (a serving batch's share of
    // the server's pool); calls that need more fall back to `ws`.
    char* lent = nullptr;
    size_t lent_size = 0;
    uint64_t lent_overflows = 0;

    void* workspace(size_t bytes) {
        if (lent) {
            if (bytes <= lent_size) return lent;
            lent_overflows++;
        }
        return ws.get(bytes, stream);
    }
    hipStream_t fork() {
        if (!side) {
            OVX_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
            OVX_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
            OVX_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        }
        OVX_HIP(hipEventRecord(ev_fork, stream));
        OVX_HIP(hipStreamWaitEvent(side, ev_fork, 0));
        return side;
    }
    // Marks the end of the side stream's work; the main stream waits for it
    // where the returned event is waited on.
    hipEvent_t join_event() {
        OVX_HIP(hipEventRecord(ev_join, side));
        return ev_join;
    }
    uint32_t* zero_words(size_t count) {
        if (count > zw_count) {
            if (zw) {
                OVX_HIP(hipStreamSynchronize(stream));
                OVX_HIP(hipFree(zw));
                zw = nullptr;
                zw_count = 0;
            }
            const size_t want = std::max<size_t>(count, 16384);
            if (hipMalloc(&zw, want * 4) != hipSuccess) {
                (void)hipGetLastError();
                zw = nullptr;
                fail(kInternal, "device allocation of the zeroed counter words failed");
            }
            OVX_HIP(hipMemsetAsync(zw, 0, want * 4, stream));
            zw_count = want;
        }
        return zw;
    }
    void* workspace2(size_t bytes) { return ws2.get(bytes, stream); }
    uint32_t* pinned_u32(size_t count) {
        if (count > pinned_count) {
            if (pinned) {
                OVX_HIP(hipStreamSynchronize(stream));
                OVX_HIP(hipHostFree(pinned));
                pinned = nullptr;
            }
            size_t want = std::max<size_t>(count, 1024);
            OVX_HIP(hipHostMalloc((void**)&pinned, want * 4, hipHostMallocDefault));
            pinned_count = want;
        }
        return pinned;
    }
    void release() {
        if (ws.p) (void)hipFree(ws.p);
        if (ws2.p) (void)hipFree(ws2.p);
        if (pinned) (void)hipHostFree(pinned);
        if (zw) (void)hipFree(zw);
        zw = nullptr;
        zw_count = 0;
        if (status_ev) (void)hipEventDestroy(status_ev);
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
            (void)hipEventDestroy(ev_fork);
            (void)hipEventDestroy(ev_join);
        }
        side = nullptr;
        ev_fork = ev_join = nullptr;
        status_ev = nullptr;
        ws.p = ws2.p = nullptr;
        ws.size = ws2.size = 0;
        pinned = nullptr;
        pinned_count = 0;
    }
};

struct Engine {
    std::recursive_mutex mu;
    bool inited = false;
    Ctx def;
    uint32_t* d_flag = nullptr;

    void init() {
        if (inited) return;
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n <= 0)
            fail(kInternal, std::string("no HIP device is visible (hipGetDeviceCount: ") + hipGetErrorString(e) +
                                ", count " + std::to_string(n) + ")");
        OVX_HIP(hipMalloc(&d_flag, 256));
        inited = true;
    }
};

Engine& eng() {
    static Engine* e = new Engine();
    return *e;
}

thread_local Ctx* t_ctx = nullptr;

// The context the calling thread runs on (a serving slot, or the default).
Ctx& cur() { return t_ctx ? *t_ctx : eng().def; }

struct CtxScope {
    Ctx* prev;
    explicit CtxScope(Ctx* c) : prev(t_ctx) { t_ctx = c; }
    ~CtxScope() { t_ctx = prev; }
};

struct Carver {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(uint64_t count) {
        T* p = reinterpret_cast<T*>(base + off);
        off += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

bool host_all_finite(const float* x, uint64_t n) {
    for (uint64_t i = 0; i < n; i++)
        if (!std::isfinite(x[i])) return false;
    return true;
}

// Device finite check (synchronises); returns true when every value is finite.
bool device_all_finite(const float* x, uint64_t n) {
    Engine& E = eng();
    hipStream_t st = cur().stream;
    OVX_HIP(hipMemsetAsync(E.d_flag, 0, 4, st));
    launch_check_finite(x, n, E.d_flag, st);
    uint32_t flag = 0;
    OVX_HIP(hipMemcpyAsync(&flag, E.d_flag, 4, hipMemcpyDeviceToHost, st));
    OVX_HIP(hipStreamSynchronize(st));
    return flag == 0;
}

void check_finite(const float* x, uint64_t n, uint64_t flags, uint32_t status, const char* what) {
    if (flags & ORDVEC_FLAG_ASSUME_FINITE) return;
    bool ok = (flags & ORDVEC_FLAG_DEVICE_PTRS) ? device_all_finite(x, n) : host_all_finite(x, n);
    if (!ok) fail(status, std::string(what) + " contains NaN or infinity");
}

void h2d(void* dst, const void* src, size_t bytes) {
    if (bytes) OVX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, cur().stream));
}
void d2h(void* dst, const void* src, size_t bytes) {
    if (bytes) OVX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, cur().stream));
}
void sync() { OVX_HIP(hipStreamSynchronize(cur().stream)); }

// ---------------------------------------------------------------------------
// Index handles
// ---------------------------------------------------------------------------
// Grows a row-major device buffer of `n` rows to hold `want` rows.
template <typename T>
void grow_rows(T*& buf, uint64_t& cap, uint64_t n, uint64_t row_bytes, uint64_t want, const char* what) {
    if (want <= cap) return;
    uint64_t nc = std::max<uint64_t>(want, cap + cap / 2);
    T* p = nullptr;
    if (hipMalloc(&p, std::max<uint64_t>(16, nc * row_bytes)) != hipSuccess) {
        (void)hipGetLastError();
        fail(kInternal, std::string("HBM allocation for ") + what + " failed");
    }
    if (buf) {
        OVX_HIP(hipMemcpyAsync(p, buf, n * row_bytes, hipMemcpyDeviceToDevice, cur().stream));
        sync();
        OVX_HIP(hipFree(buf));
    }
    buf = p;
    cap = nc;
}

// Every index counts the running servers and live two-stage batch plans that
// read it (engine mutex held): while the count is nonzero, add and
// swap_remove fail and free is deferred to the last server's stop or plan's
// free, so no in-flight batch sees its rows move.
struct Served {
    uint32_t serving = 0;
    bool free_pending = false;
};

void check_not_serving(const Served& ix, const char* what) {
    if (ix.serving)
        fail(kBadArgument, std::string(what) + " is not allowed while a running server or a two-stage batch plan "
                                               "reads this index (stop the server or free the plan first)");
}

struct RqIndex : Served {
    uint32_t dim = 0, bits = 0, bpv = 0;
    uint64_t n = 0, cap = 0;
    uint8_t* codes = nullptr;
    ~RqIndex() {
        if (codes) (void)hipFree(codes);
    }
    void reserve(uint64_t want) { grow_rows(codes, cap, n, bpv, want, "RankQuant codes"); }
};

// Rank (src/rank.rs:291-305): u16 rank vectors, row-major [n][dim].
struct RkIndex {
    uint32_t dim = 0;
    uint64_t n = 0, cap = 0;
    uint16_t* ranks = nullptr;
    ~RkIndex() {
        if (ranks) (void)hipFree(ranks);
    }
    uint64_t row_bytes() const { return (uint64_t)dim * 2; }
    void reserve(uint64_t want) { grow_rows(ranks, cap, n, row_bytes(), want, "Rank vectors"); }
};

// Sign bitmap (ntop == 0) or top-bucket Bitmap (ntop > 0): same interleaved
// storage, different per-word statistic.
struct SbIndex : Served {
    uint32_t dim = 0, qpv = 0, ntop = 0;
    uint64_t n = 0, cap_blocks = 0;
    uint64_t* words = nullptr;  // interleaved [block][word][64]
    ~SbIndex() {
        if (words) (void)hipFree(words);
    }
    uint64_t block_words() const { return (uint64_t)qpv * 64; }
    void reserve(uint64_t want_docs) {
        uint64_t want = (want_docs + 63) / 64;
        if (want <= cap_blocks) return;
        uint64_t nc = std::max<uint64_t>(want, cap_blocks + cap_blocks / 2);
        uint64_t* p = nullptr;
        if (hipMalloc(&p, std::max<uint64_t>(8, nc * block_words() * 8)) != hipSuccess) {
            (void)hipGetLastError();
            fail(kInternal, "HBM allocation for SignBitmap words failed");
        }
        OVX_HIP(hipMemsetAsync(p, 0, nc * block_words() * 8, cur().stream));
        if (words) {
            OVX_HIP(hipMemcpyAsync(p, words, cap_blocks * block_words() * 8, hipMemcpyDeviceToDevice, cur().stream));
            sync();
            OVX_HIP(hipFree(words));
        }
        words = p;
        cap_blocks = nc;
    }
};

// RankQuantFastscan: b=2 in the reference's block-32 nibble layout, or b=4
// (the extension) as RankQuant b=4 rows regrouped into 64-document blocks of
// interleaved dwords.
struct FsIndex {
    uint32_t dim = 0, bits = 2;
    uint64_t n = 0;
    uint8_t* packed = nullptr;    // b = 2
    uint32_t* packed4 = nullptr;  // b = 4
    ~FsIndex() {
        if (packed) (void)hipFree(packed);
        if (packed4) (void)hipFree(packed4);
    }
    uint64_t payload() const { return ((n + 31) / 32) * (uint64_t)(dim / 2) * 32; }
    uint64_t payload4() const { return ((n + 63) / 64) * (uint64_t)(dim / 8) * 64 * 4; }
    uint32_t blocks64() const { return (uint32_t)((n + 63) / 64); }
};

// ABI v1 handle: a RankQuant (.ovrq) or a top-bucket Bitmap (.ovbm).
struct IndexHandle {
    RqIndex* rq = nullptr;
    SbIndex* bm = nullptr;
    uint64_t source_file_size_bytes = 0;
    ~IndexHandle() {
        delete rq;
        delete bm;
    }
};

// ---------------------------------------------------------------------------
// Parameter validation (src/quant.rs:375-420, src/sign_bitmap.rs:115-139,
// src/fastscan.rs:571-593)
// ---------------------------------------------------------------------------
void validate_rq_params(uint64_t dim, uint32_t bits) {
    if (bits != 1 && bits != 2 && bits != 4)
        fail(kBadArgument, "bits must be 1, 2, or 4 (b=8 is not persistable and not offered on the GPU engine)");
    if (dim < 2) fail(kBadArgument, "dim must be >= 2");
    if (dim > kMaxDim) fail(kBadArgument, "dim must fit in u16");
    uint32_t cpb = 8 / bits;
    if (dim % cpb) fail(kBadArgument, "dim must be a multiple of " + std::to_string(cpb) + " for bits = " + std::to_string(bits));
    uint32_t nb = 1u << bits;
    if (dim % nb)
        fail(kBadArgument, "dim must be divisible by 2^bits = " + std::to_string(nb) +
                               " so every bucket receives exactly dim / 2^bits rank entries");
}

void validate_sign_dim(uint64_t dim) {
    if (dim == 0) fail(kBadArgument, "dim must be > 0");
    if (dim % 64) fail(kBadArgument, "dim must be a multiple of 64");
    if (dim > kMaxSignDim) fail(kBadArgument, "dim must be <= MAX_SIGN_BITMAP_DIM (= 16777216)");
}

void check_new_count(uint64_t cur, uint64_t add) {
    if (cur + add > kMaxVectors)
        fail(kBadArgument, "index would exceed MAX_VECTORS (" + std::to_string(kMaxVectors) + "); had " +
                               std::to_string(cur) + ", adding " + std::to_string(add));
}

// ---------------------------------------------------------------------------
// Encode paths (upload in chunks, encode on the GPU into the resident index)
// ---------------------------------------------------------------------------
constexpr uint64_t kChunkBytes = 256ull << 20;

template <typename Fn>
void for_each_chunk(const float* vectors, uint64_t n, uint32_t dim, uint64_t flags, Fn&& fn) {
    if (n == 0) return;
    if (flags & ORDVEC_FLAG_DEVICE_PTRS) {
        fn(vectors, (uint64_t)0, n);
        return;
    }
    uint64_t rows = std::max<uint64_t>(1, kChunkBytes / ((uint64_t)dim * 4));
    float* stage = static_cast<float*>(cur().workspace(rows * dim * 4));
    for (uint64_t r0 = 0; r0 < n; r0 += rows) {
        uint64_t cnt = std::min(rows, n - r0);
        h2d(stage, vectors + r0 * dim, cnt * dim * 4);
        fn(stage, r0, cnt);
        sync();
    }
}

// One pass over the vectors into the rows [n, n + count) of a RankQuant
// and/or a sign bitmap (k_encode_rows: codes, sign words and the non-finite
// check together).  The indexes' row counts move only once every value was
// found finite, so a failed add leaves them unchanged.
void add_one_pass(RqIndex* rq, SbIndex* sb, const float* vectors, uint64_t n, uint64_t flags) {
    const uint32_t dim = rq ? rq->dim : sb->dim;
    if (!(flags & (ORDVEC_FLAG_DEVICE_PTRS | ORDVEC_FLAG_ASSUME_FINITE)) && !host_all_finite(vectors, n * dim))
        fail(kBadArgument, "vectors contains NaN or infinity");
    if (rq) check_new_count(rq->n, n);
    if (sb) check_new_count(sb->n, n);
    if (rq) rq->reserve(rq->n + n);
    if (sb) sb->reserve(sb->n + n);
    Engine& E = eng();
    hipStream_t s = cur().stream;
    const uint64_t max_rows = (flags & ORDVEC_FLAG_DEVICE_PTRS) ? n : std::max<uint64_t>(1, kChunkBytes / ((uint64_t)dim * 4));
    void* work = cur().workspace2(encode_rows_work_bytes(std::min(n, max_rows)));
    OVX_HIP(hipMemsetAsync(E.d_flag, 0, 4, s));
    for_each_chunk(vectors, n, dim, flags, [&](const float* dv, uint64_t r0, uint64_t cnt) {
        launch_encode_rows(dv, cnt, dim, rq ? rq->bits : 1, rq ? rq->codes + (rq->n + r0) * rq->bpv : nullptr,
                           sb ? sb->words : nullptr, sb ? sb->n + r0 : 0, E.d_flag, work, s);
    });
    uint32_t flag = 0;
    OVX_HIP(hipMemcpyAsync(&flag, E.d_flag, 4, hipMemcpyDeviceToHost, s));
    OVX_HIP(hipStreamSynchronize(s));
    if (flag && !(flags & ORDVEC_FLAG_ASSUME_FINITE)) fail(kBadArgument, "vectors contains NaN or infinity");
    if (rq) rq->n += n;
    if (sb) sb->n += n;
}

// The one-pass kernel reads rows with 16-byte loads: device rows must start
// 16-byte aligned (host rows are staged into the aligned workspace).
bool one_pass_ok(uint32_t dim, uint32_t bits, bool codes, bool words, const float* vectors, uint64_t flags) {
    return encode_rows_supported(dim, bits, codes, words) &&
           (!(flags & ORDVEC_FLAG_DEVICE_PTRS) || reinterpret_cast<uintptr_t>(vectors) % 16 == 0);
}

// Rank transform of cnt device rows into u16 ranks (ranks != null) or
// RankQuant codes: the per-row LDS sort up to kEncodeMaxDim, a per-row
// segmented sort (second workspace) above, up to the format limit of 65535.
void rank_rows(const float* dv, uint64_t cnt, uint32_t d, uint16_t* ranks, uint32_t bits, uint8_t* codes) {
    hipStream_t s = cur().stream;
    if (d <= kEncodeMaxDim) {
        if (ranks) launch_rank_u16(dv, cnt, d, ranks, s);
        else launch_rank_encode(dv, cnt, d, bits, codes, s);
        return;
    }
    const uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(65535, (1ull << 30) / (16ull * d)));
    for (uint64_t r0 = 0; r0 < cnt; r0 += per) {
        const uint64_t c = std::min(per, cnt - r0);
        void* work = cur().workspace2(rank_sort_work_bytes(c, d, s));
        launch_rank_sort(dv + r0 * d, c, d, ranks ? ranks + r0 * d : nullptr, bits,
                         codes ? codes + r0 * (d * bits / 8) : nullptr, work, s);
    }
}

void rq_add(RqIndex* ix, const float* vectors, uint64_t n, uint64_t flags) {
    if (one_pass_ok(ix->dim, ix->bits, true, false, vectors, flags)) {
        eng().init();
        add_one_pass(ix, nullptr, vectors, n, flags);
        return;
    }
    check_finite(vectors, n * ix->dim, flags, kBadArgument, "vectors");
    check_new_count(ix->n, n);
    ix->reserve(ix->n + n);
    uint64_t base = ix->n;
    for_each_chunk(vectors, n, ix->dim, flags, [&](const float* dv, uint64_t r0, uint64_t cnt) {
        rank_rows(dv, cnt, ix->dim, nullptr, ix->bits, ix->codes + (base + r0) * ix->bpv);
    });
    ix->n += n;
    if (!(flags & ORDVEC_FLAG_DEVICE_PTRS)) sync();
}

void sb_add(SbIndex* ix, const float* vectors, uint64_t n, uint64_t flags) {
    if (ix->dim > kSignMaxDim) fail(kUnsupportedOperation, "GPU sign scan supports dim <= " + std::to_string(kSignMaxDim));
    if (one_pass_ok(ix->dim, 1, false, true, vectors, flags)) {
        eng().init();
        add_one_pass(nullptr, ix, vectors, n, flags);
        return;
    }
    check_finite(vectors, n * ix->dim, flags, kBadArgument, "vectors");
    check_new_count(ix->n, n);
    ix->reserve(ix->n + n);
    uint64_t base = ix->n;
    for_each_chunk(vectors, n, ix->dim, flags, [&](const float* dv, uint64_t r0, uint64_t cnt) {
        launch_sign_encode(dv, cnt, ix->dim, ix->words, base + r0, true, cur().stream);
    });
    ix->n += n;
    if (!(flags & ORDVEC_FLAG_DEVICE_PTRS)) sync();
}

// ---------------------------------------------------------------------------
// Search pipelines.  *_dev functions enqueue work only (device pointers, no
// synchronisation).  Each query carries a status word: bit 0 = non-finite
// input, bit 1 = stage-1 survivor buffer could not prove exactness (the query
// is then recomputed by the exact histogram pipeline).
// ---------------------------------------------------------------------------
constexpr uint32_t kStatNonfinite = 1, kStatFallback = 2;

const float* stage_queries(Carver& cv, const float* q, uint64_t count, uint64_t flags) {
    if (flags & ORDVEC_FLAG_DEVICE_PTRS) return q;
    float* d = cv.take<float>(count);
    h2d(d, q, count * 4);
    return d;
}

// Waits for the stream and returns the per-query status words.
std::vector<uint32_t> read_qstat(const uint32_t* dstat, uint64_t nq) {
    uint32_t* h = cur().pinned_u32(nq);
    d2h(h, dstat, nq * 4);
    sync();
    return std::vector<uint32_t>(h, h + nq);
}

constexpr uint64_t kEarlyStatusQueries = 64;

// The same words, copied back from this point of the stream on the side
// stream, so that the work enqueued after it does not wait for the copy;
// qstat_wait blocks until that copy has finished.
uint32_t* qstat_copy_async(const uint32_t* dstat, uint64_t nq) {
    Ctx& c = cur();
    uint32_t* h = c.pinned_u32(nq);
    hipStream_t side = c.fork();
    OVX_HIP(hipMemcpyAsync(h, dstat, nq * 4, hipMemcpyDeviceToHost, side));
    if (!c.status_ev) OVX_HIP(hipEventCreateWithFlags(&c.status_ev, hipEventDisableTiming));
    OVX_HIP(hipEventRecord(c.status_ev, side));
    return h;
}

std::vector<uint32_t> qstat_wait(const uint32_t* h, uint64_t nq) {
    OVX_HIP(hipEventSynchronize(cur().status_ev));
    return std::vector<uint32_t>(h, h + nq);
}

bool any_bit(const std::vector<uint32_t>& st, uint32_t bit) {
    for (uint32_t v : st)
        if (v & bit) return true;
    return false;
}

void check_k(uint64_t k) {
    if (k > kTopKMax) fail(kUnsupportedOperation, "k above " + std::to_string(kTopKMax) + " is not supported by the GPU engine");
}

void check_search_dim(uint32_t d) {
    if (d > 8192) fail(kUnsupportedOperation, "GPU search stages q_hat in LDS and supports dim <= 8192");
}

void check_m(uint64_t m_eff) {
    if (m_eff > kStage1MaxM)
        fail(kUnsupportedOperation, "stage-1 candidate count above " + std::to_string(kStage1MaxM) +
                                        " is not supported by the GPU engine");
}

size_t full_search_bytes(const RqIndex* ix, uint64_t nq, uint64_t kk) {
    return al(nq * ix->dim * 4) + al(nq * 4) * 2 + fullscan_work_bytes(ix->n, nq, (uint32_t)kk);
}

void full_search_dev(const RqIndex* ix, const float* dq, uint64_t nq, uint64_t kk, float* ds, int64_t* di,
                     uint32_t* qstat, Carver& cv) {
    const uint32_t d = ix->dim;
    float* qhat = cv.take<float>(nq * d);
    float* off = cv.take<float>(nq);
    void* work = cv.take<char>(fullscan_work_bytes(ix->n, nq, (uint32_t)kk));
    const int tier = select_tier(d, ix->bits);
    const float inv_norm = rankquant_inv_norm(d, ix->bits);
    hipStream_t s = cur().stream;
    launch_prepare_queries(dq, nq, d, ix->bits, tier, inv_norm, qhat, off, qstat, s);
    launch_fullscan_rq(ix->codes, ix->n, d, ix->bits, tier, inv_norm, qhat, off, nq, (uint32_t)kk, work, ds, di, s);
}

size_t subset_search_bytes(const RqIndex* ix, uint64_t nq, uint64_t out_k, uint64_t avg_m) {
    return al(nq * ix->dim * 4) + al(nq * 4) * 2 +
           rerank_work_bytes(nq, (uint32_t)out_k, rerank_nsplit(nq, avg_m));
}

void subset_search_dev(const RqIndex* ix, const float* dq, uint64_t nq, const uint64_t* doff, const uint32_t* dc,
                       uint64_t row_m, uint64_t avg_m, uint64_t out_k, float* ds, int64_t* di, uint32_t* qstat,
                       Carver& cv) {
    const uint32_t d = ix->dim;
    float* qhat = cv.take<float>(nq * d);
    float* off = cv.take<float>(nq);
    const uint32_t nsplit = rerank_nsplit(nq, avg_m);
    void* work = cv.take<char>(rerank_work_bytes(nq, (uint32_t)out_k, nsplit));
    const int tier = select_tier(d, ix->bits);
    const float inv_norm = rankquant_inv_norm(d, ix->bits);
    hipStream_t s = cur().stream;
    launch_prepare_queries(dq, nq, d, ix->bits, tier, inv_norm, qhat, off, qstat, s);
    launch_rerank(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, nq, doff, dc, row_m, ix->n, nullptr,
                  (uint32_t)out_k, ds, di, nsplit, work, s);
}

struct Stage1Bufs {
    uint64_t* qbits;
    FastPlan plan;
    FastWork work;
};

size_t stage1_bytes(const SbIndex* ix, uint64_t nq, uint64_t m_eff) {
    FastPlan p = plan_stage1_fast(ix->n, ix->dim, nq, m_eff);
    return al(nq * ix->qpv * 8) + al(stage1_fast_bytes(p, nq));
}

Stage1Bufs stage1_dev(const SbIndex* ix, const float* dq, uint64_t nq, uint64_t m_eff, uint32_t* cand,
                      uint32_t* cand_h, uint32_t* qstat, Carver& cv, const PrepFuse* prep = nullptr,
                      bool select = true, hipEvent_t before_scan = nullptr, const FastTune* tune = nullptr,
                      bool as_set = false) {
    Stage1Bufs b{};
    b.qbits = cv.take<uint64_t>(nq * ix->qpv);
    b.plan = plan_stage1_fast(ix->n, ix->dim, nq, m_eff);
    if (tune) b.plan.tune = *tune;
    void* base = cv.take<char>(stage1_fast_bytes(b.plan, nq));
    carve_stage1_fast(b.plan, nq, base, &b.work);
    launch_stage1_fast(ix->words, ix->n, ix->dim, dq, nq, m_eff, b.plan, b.work, b.qbits, cand, cand_h, qstat,
                       ix->ntop, cur().stream, prep, select, before_scan, as_set);
    return b;
}

std::vector<uint32_t> flagged(const std::vector<uint32_t>& st, uint32_t bit) {
    std::vector<uint32_t> f;
    for (uint32_t i = 0; i < st.size(); i++)
        if (st[i] & bit) f.push_back(i);
    return f;
}

void d2d(void* dst, const void* src, size_t bytes) {
    if (bytes) OVX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, cur().stream));
}

// Host-facing wrappers: host or device pointers per `flags`; host-pointer
// calls return after the results are copied back.
void host_finite_or_fail(const float* x, uint64_t n, uint64_t flags, uint32_t status, const char* what) {
    if ((flags & (ORDVEC_FLAG_DEVICE_PTRS | ORDVEC_FLAG_ASSUME_FINITE)) == 0 && !host_all_finite(x, n))
        fail(status, std::string(what) + " contains NaN or infinity");
}

// Host pointers: always wait (results).  Device pointers: wait only to read
// the non-finite bit, unless the caller vouched for the inputs.
void finish_status(const uint32_t* qstat, uint64_t nq, uint64_t flags, uint32_t status, const char* what) {
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    if (dev && (flags & ORDVEC_FLAG_ASSUME_FINITE)) return;
    std::vector<uint32_t> st = read_qstat(qstat, nq);
    if (dev && any_bit(st, kStatNonfinite)) fail(status, std::string(what) + " contains NaN or infinity");
}

// Full search with k above kTopKMax: every document's score key, a per-query
// sort, the first k gathered; queries in chunks whose keys fit kLargeKBytes.
constexpr size_t kLargeKBytes = 1ull << 30;

void rq_full_search_large(const RqIndex* ix, const float* queries, uint64_t nq, uint64_t kk, float* out_s,
                          int64_t* out_i, uint64_t flags) {
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    const uint32_t d = ix->dim;
    hipStream_t s = cur().stream;
    const uint64_t qc = std::max<uint64_t>(1, std::min<uint64_t>({nq, 65535, kLargeKBytes / (16 * ix->n + 1)}));
    const int tier = select_tier(d, ix->bits);
    const float inv_norm = rankquant_inv_norm(d, ix->bits);
    for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
        const uint64_t cn = std::min(qc, nq - q0);
        const size_t lk = large_k_work_bytes(cn * ix->n, cn, s);
        size_t bytes = al(cn * d * 4) + 2 * al(cn * 4) + lk +
                       (dev ? 0 : al(cn * d * 4) + al(cn * kk * 4) + al(cn * kk * 8));
        Carver cv{static_cast<char*>(cur().workspace(bytes))};
        const float* dq = stage_queries(cv, queries + q0 * d, cn * d, flags);
        float* qhat = cv.take<float>(cn * d);
        float* off = cv.take<float>(cn);
        uint32_t* qstat = cv.take<uint32_t>(cn);
        void* work = cv.take<char>(lk);
        float* ds = dev ? out_s + q0 * kk : cv.take<float>(cn * kk);
        int64_t* di = dev ? out_i + q0 * kk : cv.take<int64_t>(cn * kk);
        launch_prepare_queries(dq, cn, d, ix->bits, tier, inv_norm, qhat, off, qstat, s);
        launch_rerank_large(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, cn, nullptr, nullptr, ix->n,
                            cn * ix->n, ix->n, (uint32_t)kk, ds, di, work, s);
        if (!dev) {
            d2h(out_s + q0 * kk, ds, cn * kk * 4);
            d2h(out_i + q0 * kk, di, cn * kk * 8);
        }
        finish_status(qstat, cn, flags, kNonfiniteQuery, "queries");
    }
}

void rq_full_search(const RqIndex* ix, const float* queries, uint64_t nq, uint64_t k, float* out_s, int64_t* out_i,
                    uint64_t flags) {
    const uint64_t kk = std::min<uint64_t>(k, ix->n);
    if (nq == 0 || kk == 0) return;
    check_search_dim(ix->dim);
    host_finite_or_fail(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
    if (kk > kTopKMax) {
        rq_full_search_large(ix, queries, nq, kk, out_s, out_i, flags);
        return;
    }
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    size_t bytes = full_search_bytes(ix, nq, kk) + al(nq * 4) +
                   (dev ? 0 : al(nq * ix->dim * 4) + al(nq * kk * 4) + al(nq * kk * 8));
    Carver cv{static_cast<char*>(cur().workspace(bytes))};
    const float* dq = stage_queries(cv, queries, nq * ix->dim, flags);
    uint32_t* qstat = cv.take<uint32_t>(nq);
    float* ds = dev ? out_s : cv.take<float>(nq * kk);
    int64_t* di = dev ? out_i : cv.take<int64_t>(nq * kk);
    full_search_dev(ix, dq, nq, kk, ds, di, qstat, cv);
    if (!dev) {
        d2h(out_s, ds, nq * kk * 4);
        d2h(out_i, di, nq * kk * 8);
    }
    finish_status(qstat, nq, flags, kNonfiniteQuery, "queries");
}

// CSR (offsets != null) or uniform rows; outputs nq x out_k.
// The large-k paths size their key buffers by the entry count and hand the
// offsets to a per-row sort, so device CSR offsets are read back and checked
// (monotone, from 0 to n_entries) before any kernel uses them.
void check_device_csr(const uint64_t* offsets, uint64_t nq, uint64_t n_entries) {
    std::vector<uint64_t> h(nq + 1);
    d2h(h.data(), offsets, (nq + 1) * 8);
    sync();
    if (h[0] != 0) fail(kBadArgument, "candidate_offsets[0] must be 0");
    if (h[nq] != n_entries) fail(kBadArgument, "candidate_offsets[nq] must equal candidates.len()");
    for (uint64_t q = 0; q < nq; q++)
        if (h[q + 1] < h[q]) fail(kBadArgument, "candidate_offsets must be monotonic non-decreasing");
}

void rq_subset_search(const RqIndex* ix, const float* queries, uint64_t nq, const uint64_t* offsets,
                      const uint32_t* cands, uint64_t n_cands, uint64_t row_m, uint64_t out_k, float* out_s,
                      int64_t* out_i, uint64_t flags) {
    if (nq == 0 || out_k == 0) return;
    check_search_dim(ix->dim);
    if (out_k > kTopKMax && nq > 65535)
        fail(kUnsupportedOperation, "subset search with k above " + std::to_string(kTopKMax) +
                                        " takes at most 65535 queries per call");
    host_finite_or_fail(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    const bool big_k = out_k > kTopKMax;
    if (big_k && dev && offsets) check_device_csr(offsets, nq, n_cands);
    const uint64_t avg_m = offsets ? (n_cands + nq - 1) / nq : row_m;
    const size_t lk = big_k ? large_k_work_bytes(n_cands, nq, cur().stream) : 0;
    size_t bytes = subset_search_bytes(ix, nq, out_k, avg_m) + al(nq * 4) + lk +
                   (dev ? 0 : al(nq * ix->dim * 4) + al((nq + 1) * 8) + al(n_cands * 4) + al(nq * out_k * 4) +
                                  al(nq * out_k * 8));
    Carver cv{static_cast<char*>(cur().workspace(bytes))};
    const float* dq = stage_queries(cv, queries, nq * ix->dim, flags);
    uint32_t* qstat = cv.take<uint32_t>(nq);
    const uint64_t* doff = offsets;
    const uint32_t* dc = cands;
    float* ds = out_s;
    int64_t* di = out_i;
    if (!dev) {
        if (offsets) {
            uint64_t* t = cv.take<uint64_t>(nq + 1);
            h2d(t, offsets, (nq + 1) * 8);
            doff = t;
        }
        uint32_t* tc = cv.take<uint32_t>(n_cands);
        h2d(tc, cands, n_cands * 4);
        dc = tc;
        ds = cv.take<float>(nq * out_k);
        di = cv.take<int64_t>(nq * out_k);
    }
    if (big_k) {
        const uint32_t d = ix->dim;
        float* qhat = cv.take<float>(nq * d);
        float* off = cv.take<float>(nq);
        void* work = cv.take<char>(lk);
        const int tier = select_tier(d, ix->bits);
        const float inv_norm = rankquant_inv_norm(d, ix->bits);
        launch_prepare_queries(dq, nq, d, ix->bits, tier, inv_norm, qhat, off, qstat, cur().stream);
        launch_rerank_large(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, nq, doff, dc, row_m, n_cands, ix->n,
                            (uint32_t)out_k, ds, di, work, cur().stream);
    } else {
        subset_search_dev(ix, dq, nq, doff, dc, row_m, avg_m, out_k, ds, di, qstat, cv);
    }
    if (!dev) {
        d2h(out_s, ds, nq * out_k * 4);
        d2h(out_i, di, nq * out_k * 8);
    }
    finish_status(qstat, nq, flags, kNonfiniteQuery, "queries");
}

// Top-bucket Bitmap add (src/bitmap.rs:140-162): rank each row, set the bits
// of its n_top highest-ranked coordinates.
// ---------------------------------------------------------------------------
// Rank index and symmetric RankQuant search (src/rank.rs:329-471,
// src/quant.rs:629-695).  Queries are processed in chunks so the grid's
// query dimension and the workspace stay bounded.
// ---------------------------------------------------------------------------
constexpr uint64_t kQueryChunk = 16384;

void rank_add(RkIndex* ix, const float* vectors, uint64_t n, uint64_t flags) {
    check_finite(vectors, n * ix->dim, flags, kBadArgument, "vectors");
    check_new_count(ix->n, n);
    ix->reserve(ix->n + n);
    const uint64_t base = ix->n;
    for_each_chunk(vectors, n, ix->dim, flags, [&](const float* dv, uint64_t r0, uint64_t cnt) {
        rank_rows(dv, cnt, ix->dim, ix->ranks + (base + r0) * ix->dim, 0, nullptr);
    });
    ix->n += n;
    if (!(flags & ORDVEC_FLAG_DEVICE_PTRS)) sync();
}

// Shared driver: `enqueue(dq, cn, kk, ds, di, qstat, cv)` scores one chunk of
// queries; `per_query_status` says whether it fills qstat (asymmetric paths)
// or the queries were checked up front (rank-transformed paths).
template <typename Enqueue>
void chunked_full_search(uint32_t d, uint64_t n, const float* queries, uint64_t nq, uint64_t k, float* out_s,
                         int64_t* out_i, uint64_t flags, bool per_query_status, size_t extra_bytes_per_query,
                         Enqueue&& enqueue) {
    const uint64_t kk = std::min<uint64_t>(k, n);
    if (nq == 0 || kk == 0) return;
    check_k(kk);
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    if (per_query_status) host_finite_or_fail(queries, nq * d, flags, kNonfiniteQuery, "queries");
    else check_finite(queries, nq * d, flags, kNonfiniteQuery, "queries");
    for (uint64_t q0 = 0; q0 < nq; q0 += kQueryChunk) {
        const uint64_t cn = std::min<uint64_t>(kQueryChunk, nq - q0);
        // + 1024: alignment slack of the enqueue's up to four carved buffers
        size_t bytes = al(cn * 4) + 1024 + al(cn * extra_bytes_per_query) + fullscan_work_bytes(n, cn, (uint32_t)kk) +
                       (dev ? 0 : al(cn * d * 4) + al(cn * kk * 4) + al(cn * kk * 8));
        Carver cv{static_cast<char*>(cur().workspace(bytes))};
        const float* dq = stage_queries(cv, queries + q0 * d, cn * d, flags);
        uint32_t* qstat = cv.take<uint32_t>(cn);
        float* ds = dev ? out_s + q0 * kk : cv.take<float>(cn * kk);
        int64_t* di = dev ? out_i + q0 * kk : cv.take<int64_t>(cn * kk);
        enqueue(dq, cn, (uint32_t)kk, ds, di, qstat, cv);
        if (!dev) {
            d2h(out_s + q0 * kk, ds, cn * kk * 4);
            d2h(out_i + q0 * kk, di, cn * kk * 8);
        }
        if (per_query_status) finish_status(qstat, cn, flags, kNonfiniteQuery, "queries");
        else if (!dev) sync();
    }
}

void rank_search(const RkIndex* ix, const float* queries, uint64_t nq, uint64_t k, bool symmetric, float* out_s,
                 int64_t* out_i, uint64_t flags) {
    const uint32_t d = ix->dim;
    check_search_dim(d);
    const size_t extra = symmetric ? (size_t)d * 2 : (size_t)d * 4 + 4;
    chunked_full_search(d, ix->n, queries, nq, k, out_s, out_i, flags, !symmetric, extra,
                        [&](const float* dq, uint64_t cn, uint32_t kk, float* ds, int64_t* di, uint32_t* qstat,
                            Carver& cv) {
                            hipStream_t s = cur().stream;
                            void* work = cv.take<char>(fullscan_work_bytes(ix->n, cn, kk));
                            if (symmetric) {
                                uint16_t* qr = cv.take<uint16_t>(cn * d);
                                rank_rows(dq, cn, d, qr, 0, nullptr);
                                launch_rank_search(ix->ranks, ix->n, d, true, qr, nullptr, nullptr, cn, kk, work, ds,
                                                   di, s);
                            } else {
                                float* qhat = cv.take<float>(cn * d);
                                float* qsum = cv.take<float>(cn);
                                launch_prepare_sum(dq, cn, d, qhat, qsum, qstat, s);
                                launch_rank_search(ix->ranks, ix->n, d, false, nullptr, qhat, qsum, cn, kk, work, ds,
                                                   di, s);
                            }
                        });
}

// RankQuant::search: rank-bucket the query, q[p] = centre(bucket_p), then the
// scalar LUT scan with scale 1 / (R * R) and a zero offset.
void rq_symmetric_search(const RqIndex* ix, const float* queries, uint64_t nq, uint64_t k, float* out_s,
                         int64_t* out_i, uint64_t flags) {
    const uint32_t d = ix->dim;
    const double nb = (double)(1u << ix->bits);  // src/rank.rs:275-280 rankquant_norm
    const float R = (float)std::sqrt((double)d * ((nb * nb - 1.0) / 12.0));
    const float inv_norm_sq = 1.0f / (R * R);
    chunked_full_search(d, ix->n, queries, nq, k, out_s, out_i, flags, false, (size_t)d * 6 + 4,
                        [&](const float* dq, uint64_t cn, uint32_t kk, float* ds, int64_t* di, uint32_t*, Carver& cv) {
                            hipStream_t s = cur().stream;
                            void* work = cv.take<char>(fullscan_work_bytes(ix->n, cn, kk));
                            uint16_t* qr = cv.take<uint16_t>(cn * d);
                            float* qc = cv.take<float>(cn * d);
                            float* zero = cv.take<float>(cn);
                            rank_rows(dq, cn, d, qr, 0, nullptr);
                            launch_rank_centres(qr, cn * d, d, ix->bits, qc, s);
                            OVX_HIP(hipMemsetAsync(zero, 0, cn * 4, s));
                            launch_fullscan_rq(ix->codes, ix->n, d, ix->bits, 0, inv_norm_sq, qc, zero, cn, kk, work,
                                               ds, di, s);
                        });
}

void bm_add(SbIndex* ix, const float* vectors, uint64_t n, uint64_t flags) {
    if (ix->dim > kEncodeMaxDim) fail(kUnsupportedOperation, "GPU Bitmap encode supports dim <= " + std::to_string(kEncodeMaxDim));
    check_finite(vectors, n * ix->dim, flags, kBadArgument, "vectors");
    check_new_count(ix->n, n);
    ix->reserve(ix->n + n);
    uint64_t base = ix->n;
    for_each_chunk(vectors, n, ix->dim, flags, [&](const float* dv, uint64_t r0, uint64_t cnt) {
        launch_topbits_docs(dv, cnt, ix->dim, ix->ntop, ix->words, base + r0, cur().stream);
    });
    ix->n += n;
    if (!(flags & ORDVEC_FLAG_DEVICE_PTRS)) sync();
}

void check_bitmap_query_dim(const SbIndex* ix) {
    if (ix->dim > kEncodeMaxDim) fail(kUnsupportedOperation, "GPU Bitmap search supports dim <= " + std::to_string(kEncodeMaxDim));
}

void stage1_fix(const SbIndex* ix, const uint64_t* qbits, const std::vector<uint32_t>& fl, uint64_t m_eff,
                uint32_t* cand, uint32_t* cand_h);

// The fast stage-1 path handles m_eff up to kStage1MaxM; above it every query
// takes the exact pipeline (ordered scatter, no cap on m).
uint64_t stage1_fast_m(uint64_t m_eff) { return m_eff > kStage1MaxM ? 0 : m_eff; }

// Stage 1 to completion on the stream: the fast path, then the exact
// recomputation of the queries it flagged (or of every query when m_eff >
// kStage1MaxM).  Synchronises once to read the status words; fails on
// non-finite device queries unless the caller vouched for them.
void stage1_complete(const SbIndex* ix, const float* dq, uint64_t nq, uint64_t m_eff, uint32_t* cand,
                     uint32_t* cand_h, uint32_t* qstat, Carver& cv, uint64_t flags, const PrepFuse* prep = nullptr) {
    Stage1Bufs b = stage1_dev(ix, dq, nq, stage1_fast_m(m_eff), cand, cand_h, qstat, cv, prep);
    std::vector<uint32_t> st = read_qstat(qstat, nq);
    if ((flags & ORDVEC_FLAG_DEVICE_PTRS) && !(flags & ORDVEC_FLAG_ASSUME_FINITE) && any_bit(st, kStatNonfinite))
        fail(kNonfiniteQuery, "queries contain NaN or infinity");
    std::vector<uint32_t> fl;
    if (stage1_fast_m(m_eff) == 0) {
        fl.resize(nq);
        for (uint64_t q = 0; q < nq; q++) fl[q] = (uint32_t)q;
    } else {
        fl = flagged(st, kStatFallback);
    }
    if (!fl.empty()) {
        cur().exact_queries += fl.size();
        stage1_fix(ix, b.qbits, fl, m_eff, cand, cand_h);
    }
}

// Exact top-m by (distance, id) for a sign bitmap (Hamming) or a Bitmap
// (n_top - overlap): src/sign_bitmap.rs:231-483, src/bitmap.rs:264-362.
// Outputs nq x m_eff ids and, when requested, distances.
void sb_top_m(const SbIndex* ix, const float* queries, uint64_t nq, uint64_t m_eff, uint32_t* out_c, uint32_t* out_h,
              uint64_t flags) {
    if (ix->ntop) check_bitmap_query_dim(ix);
    else if (ix->dim > kSignMaxDim) fail(kUnsupportedOperation, "GPU sign scan supports dim <= 8192");
    host_finite_or_fail(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    size_t bytes = stage1_bytes(ix, nq, stage1_fast_m(m_eff)) + al(nq * 4) +
                   (dev ? 0 : al(nq * ix->dim * 4) + 2 * al(nq * m_eff * 4));
    Carver cv{static_cast<char*>(cur().workspace(bytes))};
    const float* dq = stage_queries(cv, queries, nq * ix->dim, flags);
    uint32_t* qstat = cv.take<uint32_t>(nq);
    uint32_t* dc = dev ? out_c : cv.take<uint32_t>(nq * m_eff);
    uint32_t* dh = dev ? out_h : (out_h ? cv.take<uint32_t>(nq * m_eff) : nullptr);
    stage1_complete(ix, dq, nq, m_eff, dc, dh, qstat, cv, flags);
    if (!dev) {
        d2h(out_c, dc, nq * m_eff * 4);
        if (out_h) d2h(out_h, dh, nq * m_eff * 4);
    }
    sync();
}

// Bitmap::search (src/bitmap.rs:202-250): top-k_eff by (overlap desc, id asc),
// scores = overlap as f32.  k_eff = min(k, n) > 0.
void bm_search(const SbIndex* ix, const float* queries, uint64_t nq, uint64_t k_eff, float* out_s, int64_t* out_i,
               uint64_t flags) {
    check_bitmap_query_dim(ix);
    host_finite_or_fail(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    size_t bytes = stage1_bytes(ix, nq, stage1_fast_m(k_eff)) + al(nq * 4) + 2 * al(nq * k_eff * 4) +
                   (dev ? 0 : al(nq * ix->dim * 4) + al(nq * k_eff * 4) + al(nq * k_eff * 8));
    Carver cv{static_cast<char*>(cur().workspace(bytes))};
    const float* dq = stage_queries(cv, queries, nq * ix->dim, flags);
    uint32_t* qstat = cv.take<uint32_t>(nq);
    uint32_t* c = cv.take<uint32_t>(nq * k_eff);
    uint32_t* h = cv.take<uint32_t>(nq * k_eff);
    float* ds = dev ? out_s : cv.take<float>(nq * k_eff);
    int64_t* di = dev ? out_i : cv.take<int64_t>(nq * k_eff);
    stage1_complete(ix, dq, nq, k_eff, c, h, qstat, cv, flags);
    launch_bm_finish(c, h, nq * k_eff, ix->ntop, ds, di, cur().stream);
    if (!dev) {
        d2h(out_s, ds, nq * k_eff * 4);
        d2h(out_i, di, nq * k_eff * 8);
    }
    sync();
}

// Bitmap::search_subset per query (src/bitmap.rs:477-505) over CSR rows.
void bm_subset(const SbIndex* ix, const float* queries, uint64_t nq, const uint64_t* offsets, const uint32_t* rows,
               uint64_t n_rows, uint64_t out_k, float* out_s, int64_t* out_i, uint64_t flags) {
    if (nq == 0 || out_k == 0) return;
    check_bitmap_query_dim(ix);
    const bool big_k = out_k > kTopKMax;
    if (big_k && nq > 65535)
        fail(kUnsupportedOperation, "Bitmap subset search with k above " + std::to_string(kTopKMax) +
                                        " takes at most 65535 queries per call");
    host_finite_or_fail(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    if (big_k && dev) check_device_csr(offsets, nq, n_rows);
    const size_t lk = big_k ? large_k_work_bytes(n_rows, nq, cur().stream) : 0;
    size_t bytes = al(nq * ix->qpv * 8) + al(nq * 4) + lk +
                   (dev ? 0 : al(nq * ix->dim * 4) + al((nq + 1) * 8) + al(n_rows * 4) + al(nq * out_k * 12));
    Carver cv{static_cast<char*>(cur().workspace(bytes))};
    const float* dq = stage_queries(cv, queries, nq * ix->dim, flags);
    uint64_t* qb = cv.take<uint64_t>(nq * ix->qpv);
    uint32_t* qstat = cv.take<uint32_t>(nq);
    const uint64_t* doff = offsets;
    const uint32_t* dr = rows;
    float* ds = out_s;
    int64_t* di = out_i;
    if (!dev) {
        uint64_t* t = cv.take<uint64_t>(nq + 1);
        h2d(t, offsets, (nq + 1) * 8);
        doff = t;
        uint32_t* tr = cv.take<uint32_t>(n_rows);
        h2d(tr, rows, n_rows * 4);
        dr = tr;
        ds = cv.take<float>(nq * out_k);
        di = cv.take<int64_t>(nq * out_k);
    }
    launch_bitmap_query_bits(dq, nq, ix->dim, ix->ntop, qb, qstat, cur().stream);
    if (big_k)
        launch_bm_subset_large(ix->words, ix->dim, ix->n, qb, nq, doff, dr, n_rows, ix->ntop, (uint32_t)out_k, ds, di,
                               cv.take<char>(lk), cur().stream);
    else
        launch_bm_subset(ix->words, ix->dim, ix->n, qb, nq, doff, dr, ix->ntop, (uint32_t)out_k, ds, di, cur().stream);
    if (!dev) {
        d2h(out_s, ds, nq * out_k * 4);
        d2h(out_i, di, nq * out_k * 8);
    }
    finish_status(qstat, nq, flags, kNonfiniteQuery, "queries");
}

// Device bytes the exact recomputation of nf queries needs.
size_t stage1_fix_bytes(const SbIndex* ix, uint64_t nf, uint64_t m_eff) {
    Stage1Plan plan = plan_stage1(ix->n, ix->dim, nf);
    return al(nf * 4) + al(nf * ix->qpv * 8) + al(stage1_work_bytes(plan, nf, m_eff)) + 2 * al(nf * m_eff * 4);
}

// Exact recomputation of the nf queries listed in fl (histogram pipeline) in
// the scratch at `base` (stage1_fix_bytes); rewrites their rows of cand /
// cand_h in place.
void stage1_fix_into(const SbIndex* ix, const uint64_t* qbits, const uint32_t* fl, uint64_t nf, uint64_t m_eff,
                     uint32_t* cand, uint32_t* cand_h, char* base) {
    const uint64_t qpv = ix->qpv;
    Stage1Plan plan = plan_stage1(ix->n, ix->dim, nf);
    Carver cv{base};
    uint32_t* didx = cv.take<uint32_t>(nf);
    uint64_t* qb = cv.take<uint64_t>(nf * qpv);
    h2d(didx, fl, nf * 4);
    gather_rows_u64(qb, qbits, didx, nf, (uint32_t)qpv, cur().stream);
    Stage1Work w{};
    carve_stage1(plan, nf, m_eff, cv.take<char>(stage1_work_bytes(plan, nf, m_eff)), &w);
    uint32_t* cc = cv.take<uint32_t>(nf * m_eff);
    uint32_t* ch = cv.take<uint32_t>(nf * m_eff);
    launch_stage1(ix->words, ix->n, ix->dim, qb, nf, m_eff, plan, w, cc, cand_h ? ch : nullptr, ix->ntop,
                  cur().stream);
    scatter_rows_u32(cand, cc, didx, nf, (uint32_t)m_eff, cur().stream);
    if (cand_h) scatter_rows_u32(cand_h, ch, didx, nf, (uint32_t)m_eff, cur().stream);
}

// The same in the context's second workspace, in chunks of queries whose
// scratch (2 bytes per document per query and the segment histograms) stays
// within kFixChunkBytes (at least one query per chunk).
constexpr size_t kFixChunkBytes = 1ull << 30;

void stage1_fix(const SbIndex* ix, const uint64_t* qbits, const std::vector<uint32_t>& fl, uint64_t m_eff,
                uint32_t* cand, uint32_t* cand_h) {
    const uint64_t nf = fl.size();
    uint64_t c = nf;
    while (c > 1 && stage1_fix_bytes(ix, c, m_eff) > kFixChunkBytes) c = (c + 1) / 2;
    char* base = static_cast<char*>(cur().workspace2(stage1_fix_bytes(ix, c, m_eff)));
    for (uint64_t f0 = 0; f0 < nf; f0 += c)
        stage1_fix_into(ix, qbits, fl.data() + f0, std::min(c, nf - f0), m_eff, cand, cand_h, base);
}

// Two-stage search past the fast path's limits: m_eff > kStage1MaxM (every
// query takes the exact stage-1 pipeline) or out_k > kTopKMax (the rerank
// sorts whole candidate rows).  Stage 1 completes (one status wait) before
// the rerank is enqueued.
void two_stage_big(const SbIndex* sb, const RqIndex* ix, const float* queries, uint64_t nq, uint64_t m_eff,
                   uint64_t out_k, float* out_scores, int64_t* out_ids, uint64_t flags) {
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    const bool big_k = out_k > kTopKMax;
    const uint32_t d = ix->dim;
    hipStream_t s = cur().stream;
    const uint32_t nsplit = big_k ? 1 : rerank_nsplit(nq, m_eff);
    const size_t rr = big_k ? large_k_work_bytes(nq * m_eff, nq, s) : rerank_work_bytes(nq, (uint32_t)out_k, nsplit);
    size_t bytes = al(nq * d * 4) + 2 * al(nq * 4) + al(nq * m_eff * 4) + stage1_bytes(sb, nq, stage1_fast_m(m_eff)) +
                   (dev ? 0 : al(nq * d * 4) + al(nq * out_k * 4) + al(nq * out_k * 8)) + rr;
    Carver cv{static_cast<char*>(cur().workspace(bytes))};
    const float* dq = stage_queries(cv, queries, nq * d, flags);
    float* qhat = cv.take<float>(nq * d);
    float* off = cv.take<float>(nq);
    uint32_t* qstat = cv.take<uint32_t>(nq);
    uint32_t* cand = cv.take<uint32_t>(nq * m_eff);
    void* rwork = cv.take<char>(rr);
    float* ds = dev ? out_scores : cv.take<float>(nq * out_k);
    int64_t* di = dev ? out_ids : cv.take<int64_t>(nq * out_k);
    const int tier = select_tier(d, ix->bits);
    const float inv_norm = rankquant_inv_norm(d, ix->bits);
    const PrepFuse prep{dq, qhat, off, nullptr, d, ix->bits, tier, inv_norm};
    stage1_complete(sb, dq, nq, m_eff, cand, nullptr, qstat, cv, flags, &prep);
    if (big_k)
        launch_rerank_large(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, nq, nullptr, cand, m_eff, nq * m_eff,
                            ix->n, (uint32_t)out_k, ds, di, rwork, s);
    else
        launch_rerank(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, nq, nullptr, cand, m_eff, ix->n, nullptr,
                      (uint32_t)out_k, ds, di, nsplit, rwork, s);
    if (!dev) {
        d2h(out_scores, ds, nq * out_k * 4);
        d2h(out_ids, di, nq * out_k * 8);
        sync();
    }
}

// Two-stage search on the calling thread's context (src/sign_bitmap.rs:453-483
// then src/quant.rs:1336-1394).  Outputs nq x min(k, n).
// After a two-stage batch's status words are back: the non-finite error for
// device queries, then the exact recomputation of the queries whose stage-1
// survivors could not prove exactness (their rows rewritten in ds/di and,
// for host outputs, copied back).
void finish_two_stage(const SbIndex* sb, const RqIndex* ix, const uint64_t* qbits, const std::vector<uint32_t>& st,
                      uint64_t nq, uint64_t m_eff, uint64_t out_k, int tier, float inv_norm, const float* qhat,
                      const float* off, uint32_t* cand, uint32_t* qmap, float* ds, int64_t* di, float* out_scores,
                      int64_t* out_ids, uint64_t flags) {
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    if (dev && !(flags & ORDVEC_FLAG_ASSUME_FINITE) && any_bit(st, kStatNonfinite))
        fail(kNonfiniteQuery, "queries contain NaN or infinity");
    std::vector<uint32_t> fl = flagged(st, kStatFallback);
    if (fl.empty()) return;
    cur().exact_queries += fl.size();
    stage1_fix(sb, qbits, fl, m_eff, cand, nullptr);
    h2d(qmap, fl.data(), fl.size() * 4);
    launch_rerank(ix->codes, ix->dim, ix->bits, tier, inv_norm, qhat, off, fl.size(), nullptr, cand, m_eff, ix->n,
                  qmap, (uint32_t)out_k, ds, di, 1, nullptr, cur().stream);
    if (!dev)
        for (uint32_t q : fl) {
            d2h(out_scores + (uint64_t)q * out_k, ds + (uint64_t)q * out_k, out_k * 4);
            d2h(out_ids + (uint64_t)q * out_k, di + (uint64_t)q * out_k, out_k * 8);
        }
    sync();
}

// ORDVEC_OPTION_ONE_QUERY_PATH: one-query calls take two_stage_plain.
std::atomic<bool> g_plain_path{true};
// ORDVEC_OPTION_BATCH_TAIL: batches select and rerank in one launch.
std::atomic<bool> g_batch_tail{true};
// ORDVEC_OPTION_SIDE_PREP: batches prepare their queries on the side stream.
std::atomic<bool> g_side_prep{true};

// Batches of this many queries or more keep separate select and rerank
// launches: the select then runs with 1024 threads per query, and the rerank
// with every query's workgroup resident at once.
constexpr uint64_t kTailMaxBatch = 128;

// From this batch size the rerank's query preparation runs on the side
// stream (smaller batches lose more to the fork and join than they overlap).
constexpr uint64_t kSidePrepMin = 256;

// Layout of a context's zero words: the head's histogram (up to 8193 bins),
// the head's arrival counter, then one tail counter per query.
constexpr size_t kZwHeadDone = 8200, kZwTailDone = 8208;

// One query (the plain form): three launches (stage-1 head, scan, tail) and
// one wait; the tail writes the status word straight to pinned host memory.
// A query whose survivors cannot prove exactness is recomputed as in
// two_stage_run.
void two_stage_plain(const SbIndex* sb, const RqIndex* ix, const float* queries, uint64_t m_eff, uint64_t out_k,
                     float* out_scores, int64_t* out_ids, uint64_t flags) {
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    const uint32_t d = ix->dim, qpv = d / 64;
    const FastPlan p = plan_stage1_fast(sb->n, d, 1, m_eff);
    const uint32_t nsplit = rerank_nsplit(1, m_eff);
    Ctx& c = cur();
    uint32_t* zw = c.zero_words(kZwTailDone + 1);
    uint32_t* hist = zw;
    uint32_t* done_head = zw + kZwHeadDone;
    uint32_t* done_tail = zw + kZwTailDone;
    uint32_t* hstat = c.pinned_u32(1);
    hstat[0] = 0xffffffffu;  // overwritten by the tail
    const size_t bytes = al(d * 4) + 4 * al(4) + al((uint64_t)p.cap * 8) + al(qpv * 8) + al(m_eff * 4) +
                         rerank_work_bytes(1, (uint32_t)out_k, nsplit) +
                         (dev ? 0 : al(d * 4) + al(out_k * 4) + al(out_k * 8));
    Carver cv{static_cast<char*>(c.workspace(bytes))};
    const float* dq = stage_queries(cv, queries, d, flags);
    float* qhat = cv.take<float>(d);
    float* off = cv.take<float>(1);
    uint32_t* qstat = cv.take<uint32_t>(1);
    uint32_t* thr = cv.take<uint32_t>(1);
    uint32_t* cnt = cv.take<uint32_t>(1);
    uint64_t* skeys = cv.take<uint64_t>(p.cap);
    uint64_t* qbits = cv.take<uint64_t>(qpv);
    uint32_t* cand = cv.take<uint32_t>(m_eff);
    void* rwork = cv.take<char>(rerank_work_bytes(1, (uint32_t)out_k, nsplit));
    float* ds = dev ? out_scores : cv.take<float>(out_k);
    int64_t* di = dev ? out_ids : cv.take<int64_t>(out_k);
    const int tier = select_tier(d, ix->bits);
    const float inv_norm = rankquant_inv_norm(d, ix->bits);
    hipStream_t s = c.stream;
    const PrepFuse prep{dq, qhat, off, nullptr, d, ix->bits, tier, inv_norm};
    launch_stage1_plain(sb->words, sb->n, d, dq, p, hist, done_head, thr, cnt, skeys, qbits, qstat, prep, s);
    launch_s1_tail(skeys, cnt, thr, p.cap, m_eff, ix->codes, d, ix->bits, tier, inv_norm, qhat, off, ix->n, 1,
                   (uint32_t)out_k, ds, di, nsplit, rwork, done_tail, qstat, hstat, s);
    if (!dev) {
        d2h(out_scores, ds, out_k * 4);
        d2h(out_ids, di, out_k * 8);
    }
    sync();
    const uint32_t st = __atomic_load_n(&hstat[0], __ATOMIC_ACQUIRE);
    if (st == 0xffffffffu) fail(kInternal, "one-query two-stage path: the tail did not report a status");
    if (dev && !(flags & ORDVEC_FLAG_ASSUME_FINITE) && (st & kStatNonfinite))
        fail(kNonfiniteQuery, "queries contain NaN or infinity");
    if (st & kStatFallback) {
        cur().exact_queries += 1;
        const std::vector<uint32_t> fl{0u};
        stage1_fix(sb, qbits, fl, m_eff, cand, nullptr);
        launch_rerank(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, 1, nullptr, cand, m_eff, ix->n, nullptr,
                      (uint32_t)out_k, ds, di, 1, nullptr, s);
        if (!dev) {
            d2h(out_scores, ds, out_k * 4);
            d2h(out_ids, di, out_k * 8);
        }
        sync();
    }
}

void two_stage_run(const SbIndex* sb, const RqIndex* ix, const float* queries, uint64_t nq, uint64_t m, uint64_t k,
                   float* out_scores, int64_t* out_ids, uint64_t flags) {
    if (sb->dim != ix->dim) fail(kBadArgument, "sign_probe.dim must match RankQuant dim " + std::to_string(ix->dim));
    if (sb->n != ix->n) fail(kBadArgument, "sign_probe.len must match RankQuant len " + std::to_string(ix->n));
    const uint64_t out_k = std::min<uint64_t>(k, ix->n);
    const uint64_t m_eff = std::min<uint64_t>(m, ix->n);
    if (nq == 0 || out_k == 0) return;
    NONNULL(queries, "queries");
    NONNULL(out_scores, "out_scores");
    NONNULL(out_ids, "out_ids");
    check_search_dim(ix->dim);
    if (ix->dim > kSignMaxDim) fail(kUnsupportedOperation, "GPU sign scan supports dim <= 8192");
    host_finite_or_fail(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
    eng().init();
    if (m_eff > kStage1MaxM || out_k > kTopKMax) {
        two_stage_big(sb, ix, queries, nq, m_eff, out_k, out_scores, out_ids, flags);
        return;
    }
    if (nq == 1 && sb->ntop == 0 && g_plain_path && stage1_plain_supported(sb->n, ix->dim, m_eff, out_k)) {
        two_stage_plain(sb, ix, queries, m_eff, out_k, out_scores, out_ids, flags);
        return;
    }
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    const uint32_t d = ix->dim;
    const uint32_t nsplit = rerank_nsplit(nq, m_eff);
    size_t bytes = al(nq * d * 4) + 2 * al(nq * 4) + al(nq * m_eff * 4) + stage1_bytes(sb, nq, m_eff) +
                   (dev ? 0 : al(nq * d * 4) + al(nq * out_k * 4) + al(nq * out_k * 8)) + al(nq * 4) +
                   rerank_work_bytes(nq, (uint32_t)out_k, nsplit);
    Carver cv{static_cast<char*>(cur().workspace(bytes))};
    const float* dq = stage_queries(cv, queries, nq * d, flags);
    float* qhat = cv.take<float>(nq * d);
    float* off = cv.take<float>(nq);
    uint32_t* qstat = cv.take<uint32_t>(nq);
    uint32_t* qmap = cv.take<uint32_t>(nq);
    uint32_t* cand = cv.take<uint32_t>(nq * m_eff);
    void* rwork = cv.take<char>(rerank_work_bytes(nq, (uint32_t)out_k, nsplit));
    float* ds = dev ? out_scores : cv.take<float>(nq * out_k);
    int64_t* di = dev ? out_ids : cv.take<int64_t>(nq * out_k);
    const int tier = select_tier(d, ix->bits);
    const float inv_norm = rankquant_inv_norm(d, ix->bits);
    hipStream_t s = cur().stream;
    const PrepFuse prep{dq, qhat, off, nullptr, d, ix->bits, tier, inv_norm};
    // Large batches: the rerank's query preparation (sequential sums, one
    // workgroup per query) runs on a second stream beside the stage-1 sample,
    // and the scan waits for it (beside the scan itself it slows the scan more
    // than it saves).
    const bool side_prep = nq >= kSidePrepMin && sb->ntop == 0 && g_side_prep;
    if (sb->ntop == 0 && g_batch_tail && nq < kTailMaxBatch && s1_tail_supported(sb->n, d, nq, m_eff, out_k)) {
        // stage 1 without its select, then select + rerank (+ merge) in one
        // launch that writes the status words straight to pinned host memory
        hipEvent_t prepped = nullptr;
        if (side_prep) {
            launch_prepare_queries(dq, nq, d, ix->bits, tier, inv_norm, qhat, off, nullptr, cur().fork());
            prepped = cur().join_event();
        }
        Stage1Bufs b = stage1_dev(sb, dq, nq, m_eff, cand, nullptr, qstat, cv, side_prep ? nullptr : &prep, false,
                                  prepped);
        Ctx& c = cur();
        uint32_t* zw = c.zero_words(kZwTailDone + nq);
        uint32_t* hst = c.pinned_u32(nq);
        std::fill(hst, hst + nq, 0xffffffffu);
        launch_s1_tail(b.work.keys, b.work.cnt, b.work.thr, b.plan.cap, m_eff, ix->codes, d, ix->bits, tier, inv_norm, qhat, off,
                       ix->n, nq, (uint32_t)out_k, ds, di, nsplit, rwork, zw + kZwTailDone, qstat, hst, s);
        if (!dev) {
            d2h(out_scores, ds, nq * out_k * 4);
            d2h(out_ids, di, nq * out_k * 8);
        }
        sync();
        std::vector<uint32_t> st(nq);
        for (uint64_t q = 0; q < nq; q++) st[q] = __atomic_load_n(&hst[q], __ATOMIC_ACQUIRE);
        if (std::find(st.begin(), st.end(), 0xffffffffu) != st.end())
            fail(kInternal, "two-stage tail: a query has no status");
        finish_two_stage(sb, ix, b.qbits, st, nq, m_eff, out_k, tier, inv_norm, qhat, off, cand, qmap, ds, di,
                         out_scores, out_ids, flags);
        return;
    }
    hipEvent_t prepped = nullptr;
    if (side_prep) {
        launch_prepare_queries(dq, nq, d, ix->bits, tier, inv_norm, qhat, off, nullptr, cur().fork());
        prepped = cur().join_event();
    }
    Stage1Bufs b = stage1_dev(sb, dq, nq, m_eff, cand, nullptr, qstat, cv, side_prep ? nullptr : &prep, true, prepped,
                              nullptr, true);
    // Device pointers, batches of kEarlyStatusQueries or more: the status
    // words are final after stage 1, so the call waits for their copy only
    // and returns while the rerank runs (results complete in stream order,
    // like every device-pointer call); the next call's launches then overlap
    // the rerank.  Smaller calls wait once, at the end (a second blocking wait
    // costs a one-query call more than the overlap gains).
    const bool early = dev && nq >= kEarlyStatusQueries;
    const uint32_t* hst = early ? qstat_copy_async(qstat, nq) : nullptr;
    launch_rerank(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, nq, nullptr, cand, m_eff, ix->n, nullptr,
                  (uint32_t)out_k, ds, di, nsplit, rwork, s);
    if (!dev) {
        d2h(out_scores, ds, nq * out_k * 4);
        d2h(out_ids, di, nq * out_k * 8);
    }
    std::vector<uint32_t> st = early ? qstat_wait(hst, nq) : read_qstat(qstat, nq);
    finish_two_stage(sb, ix, b.qbits, st, nq, m_eff, out_k, tier, inv_norm, qhat, off, cand, qmap, ds, di, out_scores,
                     out_ids, flags);
}

// ---------------------------------------------------------------------------
// Allocation-free two-stage search, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops).  Every scratch
// buffer of two_stage_run (stage-1 histograms, thresholds, counters and
// survivor keys, candidates, q_hat, rerank partials, and the exact
// recomputation of flagged queries) is carved from `scratch`, a fixed
// per-device array; the status words come back through a fixed host buffer.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
constexpr uint64_t kNoAllocMaxQueries = 1024;
static_assert(kNoAllocScratchBytes == (size_t)ORDVEC_NOALLOC_SCRATCH_BYTES, "header and library disagree");
static_assert(kNoAllocMaxQueries == (uint64_t)ORDVEC_NOALLOC_MAX_QUERIES, "header and library disagree");

// One copy per GPU, held until the process ends.  In host code `scratch` only
// names the symbol: its device address comes from hipGetSymbolAddress, once
// per device (noalloc_scratch), and only that address is passed to kernels.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (engine mutex held).
char* noalloc_scratch() {
    int dev = 0;
    OVX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(kInternal, "device ordinal out of range for the scratch table");
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(kInternal, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    return g_scratch_ptr[dev];
}

// Kernel-family timing records events (which it creates on first use); an
// allocation-free call records none.
thread_local bool t_no_events = false;
struct NoEventsScope {
    bool prev;
    NoEventsScope() : prev(t_no_events) { t_no_events = true; }
    ~NoEventsScope() { t_no_events = prev; }
};

thread_local uint32_t t_na_status[kNoAllocMaxQueries];
thread_local uint32_t t_na_flagged[kNoAllocMaxQueries];

// Main-path scratch of one allocation-free call (everything but the exact
// recomputation).
size_t noalloc_main_bytes(const SbIndex* sb, const RqIndex* ix, uint64_t nq, uint64_t m_eff, uint64_t out_k) {
    const uint32_t nsplit = rerank_nsplit(nq, m_eff);
    return al(nq * ix->dim * 4) + 3 * al(nq * 4) + al(nq * m_eff * 4) + stage1_bytes(sb, nq, m_eff) +
           rerank_work_bytes(nq, (uint32_t)out_k, nsplit);
}

// two_stage_run's semantics on device pointers, with no allocation: returns
// once the results are written.
void two_stage_noalloc(const SbIndex* sb, const RqIndex* ix, const float* queries, uint64_t nq, uint64_t m,
                       uint64_t k, float* out_scores, int64_t* out_ids, uint64_t flags) {
    if (sb->dim != ix->dim) fail(kBadArgument, "sign_probe.dim must match RankQuant dim " + std::to_string(ix->dim));
    if (sb->n != ix->n) fail(kBadArgument, "sign_probe.len must match RankQuant len " + std::to_string(ix->n));
    const uint64_t out_k = std::min<uint64_t>(k, ix->n);
    const uint64_t m_eff = std::min<uint64_t>(m, ix->n);
    if (nq == 0 || out_k == 0) return;
    NONNULL(queries, "queries");
    NONNULL(out_scores, "out_scores");
    NONNULL(out_ids, "out_ids");
    check_m(m_eff);
    check_k(out_k);
    check_search_dim(ix->dim);
    if (ix->dim > kSignMaxDim) fail(kUnsupportedOperation, "GPU sign scan supports dim <= 8192");
    if (nq > kNoAllocMaxQueries)
        fail(kUnsupportedOperation, "the allocation-free two-stage search takes at most " +
                                        std::to_string(kNoAllocMaxQueries) + " queries per call, got " +
                                        std::to_string(nq));
    const size_t main_bytes = noalloc_main_bytes(sb, ix, nq, m_eff, out_k);
    const size_t fix1 = stage1_fix_bytes(sb, 1, m_eff);
    if (main_bytes + fix1 > kNoAllocScratchBytes)
        fail(kUnsupportedOperation, "the allocation-free two-stage search needs " +
                                        std::to_string(main_bytes + fix1) + " bytes of scratch for " +
                                        std::to_string(nq) + " queries over " + std::to_string(ix->n) +
                                        " documents (m = " + std::to_string(m_eff) + "); its scratch holds " +
                                        std::to_string(kNoAllocScratchBytes) + " bytes: use fewer queries per call");
    char* base = noalloc_scratch();
    NoEventsScope no_events;
    const uint32_t d = ix->dim;
    const uint32_t nsplit = rerank_nsplit(nq, m_eff);
    Carver cv{base};
    float* qhat = cv.take<float>(nq * d);
    float* off = cv.take<float>(nq);
    uint32_t* qstat = cv.take<uint32_t>(nq);
    uint32_t* qmap = cv.take<uint32_t>(nq);
    uint32_t* cand = cv.take<uint32_t>(nq * m_eff);
    void* rwork = cv.take<char>(rerank_work_bytes(nq, (uint32_t)out_k, nsplit));
    const int tier = select_tier(d, ix->bits);
    const float inv_norm = rankquant_inv_norm(d, ix->bits);
    hipStream_t s = cur().stream;
    if (cv.off + stage1_bytes(sb, nq, m_eff) != main_bytes) fail(kInternal, "allocation-free scratch plan out of step");
    const PrepFuse prep{queries, qhat, off, nullptr, d, ix->bits, tier, inv_norm};
    Stage1Bufs b = stage1_dev(sb, queries, nq, m_eff, cand, nullptr, qstat, cv, &prep, true, nullptr, nullptr, true);
    launch_rerank(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, nq, nullptr, cand, m_eff, ix->n, nullptr,
                  (uint32_t)out_k, out_scores, out_ids, nsplit, rwork, s);
    OVX_HIP(hipMemcpyAsync(t_na_status, qstat, nq * 4, hipMemcpyDeviceToHost, s));
    OVX_HIP(hipStreamSynchronize(s));
    uint64_t nf = 0;
    bool nonfinite = false;
    for (uint64_t q = 0; q < nq; q++) {
        nonfinite |= (t_na_status[q] & kStatNonfinite) != 0;
        if (t_na_status[q] & kStatFallback) t_na_flagged[nf++] = (uint32_t)q;
    }
    if (!(flags & ORDVEC_FLAG_ASSUME_FINITE) && nonfinite) fail(kNonfiniteQuery, "queries contain NaN or infinity");
    if (nf == 0) return;
    cur().exact_queries += nf;
    // Recompute the flagged queries in as few chunks as the scratch left over
    // holds (at least one query: checked above).
    char* fix_base = base + al(main_bytes);
    const size_t room = kNoAllocScratchBytes - al(main_bytes);
    for (uint64_t f0 = 0; f0 < nf;) {
        uint64_t c = nf - f0;
        while (c > 1 && stage1_fix_bytes(sb, c, m_eff) > room) c = (c + 1) / 2;
        stage1_fix_into(sb, b.qbits, t_na_flagged + f0, c, m_eff, cand, nullptr, fix_base);
        h2d(qmap, t_na_flagged + f0, c * 4);
        launch_rerank(ix->codes, d, ix->bits, tier, inv_norm, qhat, off, c, nullptr, cand, m_eff, ix->n, qmap,
                      (uint32_t)out_k, out_scores, out_ids, 1, nullptr, s);
        OVX_HIP(hipStreamSynchronize(s));
        f0 += c;
    }
}

// ---------------------------------------------------------------------------
// Two-stage batch plans: one (sign bitmap, RankQuant, m, k) pairing with its
// own context and a workspace sized once for its largest batch.  The launch
// sequence of a batched two-stage call is captured in two HIP graphs per
// (query count, query buffer, output buffers): stage 1 up to the copy of the
// status words, then the rerank.  A call replays both, waits for the status
// copy only and recomputes the queries stage 1 flagged behind the rerank, as
// two_stage_run does; results are the same as two_stage_run's.  The stage-1
// launch shapes (FastTune) and the side-stream query preparation are picked
// by timing them on real queries.
// ---------------------------------------------------------------------------
struct TsShape {
    FastTune t{};
    bool side_prep = true;
};

struct TsTuneRow {
    TsShape c;
    double ms;
};

struct TsGraph {
    uint64_t nq = 0;
    const float* q = nullptr;
    float* s = nullptr;
    int64_t* i = nullptr;
    hipGraphExec_t stage1 = nullptr, rerank = nullptr;
    uint64_t used = 0;
    void drop() {
        if (stage1) (void)hipGraphExecDestroy(stage1);
        if (rerank) (void)hipGraphExecDestroy(rerank);
        stage1 = rerank = nullptr;
    }
};

constexpr size_t kTsGraphs = 4;  // graph pairs kept per plan (least recently used one replaced)

// Buffers of one call of nq queries, carved from the plan's workspace in a
// fixed order (so the graphs and the eager fallback see the same addresses).
struct TsBufs {
    float* qhat;
    float* off;
    uint32_t* qstat;
    uint32_t* qmap;
    uint32_t* cand;
    void* rwork;
    uint64_t* qbits;
    uint32_t nsplit;
    Carver stage1;  // positioned at stage 1's buffers (the query sign words first)
};

struct TsBatch {
    Ctx ctx;                   // its stream is the engine stream of each call
    hipStream_t cap = nullptr; // capture stream
    SbIndex* sb = nullptr;
    RqIndex* rq = nullptr;
    uint64_t max_nq = 0, m_eff = 0, out_k = 0;
    TsShape shape;
    std::vector<TsGraph> graphs;
    uint64_t clock = 0;
    std::vector<TsTuneRow> tuning;
    char* ws = nullptr;        // staging for host-pointer calls, then the per-call buffers
    size_t ws_bytes = 0;
    float* hq = nullptr;
    float* hs = nullptr;
    int64_t* hi = nullptr;
    uint32_t* hstat = nullptr; // pinned status words
    hipEvent_t st_ev = nullptr;
    uint64_t captures = 0, replays = 0, eager = 0;

    int tier() const { return select_tier(rq->dim, rq->bits); }
    float inv_norm() const { return rankquant_inv_norm(rq->dim, rq->bits); }
    size_t stage_bytes() const { return al(max_nq * rq->dim * 4) + al(max_nq * out_k * 4) + al(max_nq * out_k * 8); }
    size_t call_bytes(uint64_t nq) const {
        const uint32_t nsplit = rerank_nsplit(nq, m_eff);
        return al(nq * rq->dim * 4) + 4 * al(nq * 4) + al(nq * m_eff * 4) +
               rerank_work_bytes(nq, (uint32_t)out_k, nsplit) + stage1_bytes(sb, nq, m_eff);
    }
    TsBufs carve(uint64_t nq) {
        TsBufs u{};
        Carver cv{ws + stage_bytes()};
        u.qhat = cv.take<float>(nq * rq->dim);
        u.off = cv.take<float>(nq);
        u.qstat = cv.take<uint32_t>(nq);
        u.qmap = cv.take<uint32_t>(nq);
        u.cand = cv.take<uint32_t>(nq * m_eff);
        u.nsplit = rerank_nsplit(nq, m_eff);
        u.rwork = cv.take<char>(rerank_work_bytes(nq, (uint32_t)out_k, u.nsplit));
        u.stage1 = cv;
        Carver peek = cv;
        u.qbits = peek.take<uint64_t>(nq * sb->qpv);  // stage1_dev takes these first
        return u;
    }
    void drop_graphs() {
        for (TsGraph& g : graphs) g.drop();
        graphs.clear();
    }
    ~TsBatch() {
        {
            CtxScope scope(&ctx);
            if (ctx.stream) (void)hipStreamSynchronize(ctx.stream);
            if (cap) (void)hipStreamSynchronize(cap);
            drop_graphs();
            if (ws) (void)hipFree(ws);
            if (st_ev) (void)hipEventDestroy(st_ev);
            ctx.stream = nullptr;  // the engine's stream, not ours
            ctx.release();
        }
        if (cap) (void)hipStreamDestroy(cap);
    }
};

// Stage 1 of a call on cur().stream (dq: device queries): the rerank's query
// preparation (on the side stream for batches of kSidePrepMin+ when the shape
// says so), sign words, sampled threshold, scan, select, then the status
// words copied into the plan's pinned words.
void ts_enqueue_stage1(TsBatch* b, const float* dq, uint64_t nq, TsBufs& u, const TsShape& sh) {
    const uint32_t d = b->rq->dim;
    const PrepFuse prep{dq, u.qhat, u.off, nullptr, d, b->rq->bits, b->tier(), b->inv_norm()};
    const bool side = sh.side_prep && nq >= kSidePrepMin;
    hipEvent_t prepped = nullptr;
    if (side) {
        launch_prepare_queries(dq, nq, d, b->rq->bits, b->tier(), b->inv_norm(), u.qhat, u.off, nullptr, cur().fork());
        prepped = cur().join_event();
    }
    Carver cv = u.stage1;
    Stage1Bufs s1 = stage1_dev(b->sb, dq, nq, b->m_eff, u.cand, nullptr, u.qstat, cv, side ? nullptr : &prep, true,
                               prepped, &sh.t, true);
    if (s1.qbits != u.qbits) fail(kInternal, "two-stage batch plan: buffer layout out of step");
    OVX_HIP(hipMemcpyAsync(b->hstat, u.qstat, nq * 4, hipMemcpyDeviceToHost, cur().stream));
}

void ts_enqueue_rerank(TsBatch* b, const TsBufs& u, uint64_t nq, float* ds, int64_t* di) {
    launch_rerank(b->rq->codes, b->rq->dim, b->rq->bits, b->tier(), b->inv_norm(), u.qhat, u.off, nq, nullptr, u.cand,
                  b->m_eff, b->rq->n, nullptr, (uint32_t)b->out_k, ds, di, u.nsplit, u.rwork, cur().stream);
}

// The graphs for (nq, dq, ds, di), captured on first use on the plan's
// capture stream (kernel-family events are not recorded inside graphs).
TsGraph& ts_graph(TsBatch* b, uint64_t nq, const float* dq, float* ds, int64_t* di) {
    for (TsGraph& g : b->graphs)
        if (g.nq == nq && g.q == dq && g.s == ds && g.i == di) {
            g.used = ++b->clock;
            return g;
        }
    if (b->graphs.size() >= kTsGraphs) {
        size_t lru = 0;
        for (size_t j = 1; j < b->graphs.size(); j++)
            if (b->graphs[j].used < b->graphs[lru].used) lru = j;
        b->graphs[lru].drop();
        b->graphs.erase(b->graphs.begin() + (ptrdiff_t)lru);
    }
    TsGraph g;
    g.nq = nq;
    g.q = dq;
    g.s = ds;
    g.i = di;
    hipStream_t run = b->ctx.stream;
    b->ctx.stream = b->cap;
    NoEventsScope no_events;
    // Two graphs: stage 1 with the status copy, then the rerank; the call
    // records the status event between their launches and waits for it only
    // (an event record captured inside one graph is refused by this HIP).
    auto capture = [&](auto&& enqueue, hipGraphExec_t* out) {
        OVX_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        hipGraph_t graph = nullptr;
        try {
            enqueue();
        } catch (...) {
            (void)hipStreamEndCapture(b->cap, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        OVX_HIP(hipStreamEndCapture(b->cap, &graph));
        const hipError_t e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        OVX_HIP(e);
    };
    try {
        TsBufs u = b->carve(nq);
        capture([&] { ts_enqueue_stage1(b, dq, nq, u, b->shape); }, &g.stage1);
        capture([&] { ts_enqueue_rerank(b, u, nq, ds, di); }, &g.rerank);
    } catch (...) {
        b->ctx.stream = run;
        g.drop();
        throw;
    }
    b->ctx.stream = run;
    b->captures++;
    g.used = ++b->clock;
    b->graphs.push_back(g);
    return b->graphs.back();
}

// One call: nq queries (host or device per flags) -> nq x out_k results.
void ts_run(TsBatch* b, const float* queries, uint64_t nq, float* out_scores, int64_t* out_ids, uint64_t flags) {
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    const uint32_t d = b->rq->dim;
    host_finite_or_fail(queries, nq * d, flags, kNonfiniteQuery, "queries");
    b->ctx.stream = eng().def.stream;
    CtxScope scope(&b->ctx);
    hipStream_t s = b->ctx.stream;
    const float* dq = queries;
    float* ds = out_scores;
    int64_t* di = out_ids;
    if (!dev) {
        h2d(b->hq, queries, nq * d * 4);
        dq = b->hq;
        ds = b->hs;
        di = b->hi;
    }
    TsBufs u = b->carve(nq);
    if (flags & ORDVEC_FLAG_NO_GRAPH) {
        ts_enqueue_stage1(b, dq, nq, u, b->shape);
        OVX_HIP(hipEventRecord(b->st_ev, s));
        ts_enqueue_rerank(b, u, nq, ds, di);
        b->eager++;
    } else {
        TsGraph& g = ts_graph(b, nq, dq, ds, di);
        OVX_HIP(hipGraphLaunch(g.stage1, s));
        OVX_HIP(hipEventRecord(b->st_ev, s));
        OVX_HIP(hipGraphLaunch(g.rerank, s));
        b->replays++;
    }
    if (!dev) {
        d2h(out_scores, ds, nq * b->out_k * 4);
        d2h(out_ids, di, nq * b->out_k * 8);
    }
    OVX_HIP(hipEventSynchronize(b->st_ev));
    const std::vector<uint32_t> st(b->hstat, b->hstat + nq);
    finish_two_stage(b->sb, b->rq, u.qbits, st, nq, b->m_eff, b->out_k, b->tier(), b->inv_norm(), u.qhat, u.off,
                     u.cand, u.qmap, ds, di, out_scores, out_ids, flags);
    if (!dev) sync();
}

// Times candidate shapes on nq device queries (reps graph calls each after a
// warm one, mean per call) and keeps the fastest: first the scan's
// (tiles per workgroup, pairs, persistent or flat grid), then the sample's
// workgroups and the side-stream preparation with that scan.
void ts_tune(TsBatch* b, const float* dq, uint64_t nq, uint32_t reps) {
    b->ctx.stream = eng().def.stream;
    CtxScope scope(&b->ctx);
    NoEventsScope no_events;
    hipStream_t s = b->ctx.stream;
    float* ds = b->hs;
    int64_t* di = b->hi;
    const TsShape before = b->shape;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    OVX_HIP(hipEventCreate(&e0));
    OVX_HIP(hipEventCreate(&e1));
    // A shape's time per call: its graphs replayed as run() replays them
    // (status event between the two, waited for), after one warm call.
    auto time_shape = [&](const TsShape& sh) {
        b->drop_graphs();
        b->shape = sh;
        TsGraph& g = ts_graph(b, nq, dq, ds, di);
        for (uint32_t r = 0; r <= reps; r++) {
            if (r == 1) OVX_HIP(hipEventRecord(e0, s));
            OVX_HIP(hipGraphLaunch(g.stage1, s));
            OVX_HIP(hipEventRecord(b->st_ev, s));
            OVX_HIP(hipGraphLaunch(g.rerank, s));
            OVX_HIP(hipEventSynchronize(b->st_ev));
        }
        OVX_HIP(hipEventRecord(e1, s));
        OVX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        OVX_HIP(hipEventElapsedTime(&ms, e0, e1));
        b->tuning.push_back(TsTuneRow{sh, (double)ms / reps});
        return (double)ms / reps;
    };
    b->tuning.clear();
    try {
        // single queries run the vector-ALU scan, which these shapes do not touch
        std::vector<uint32_t> nqts;
        if (nq > 64) nqts = {2u, 4u};
        else if (nq > 32) nqts = {1u, 2u};
        else if (nq >= 2) nqts = {1u};
        TsShape best;
        double best_ms = time_shape(best);  // the built-in rules
        for (uint32_t nqt : nqts)
            for (uint32_t pairs : {512u, 1024u, 2048u})
                for (uint32_t flat : {0u, 1u}) {
                    TsShape sh;
                    sh.t.scan_nqt = nqt;
                    sh.t.scan_pairs = pairs;
                    sh.t.scan_flat = flat;
                    const double ms = time_shape(sh);
                    if (ms < best_ms) best_ms = ms, best = sh;
                }
        TsShape base = best;
        for (uint32_t wgs : {128u, 512u})
            for (bool side : {true, false}) {
                TsShape sh = base;
                sh.t.sample_wgs = wgs;
                sh.side_prep = side;
                const double ms = time_shape(sh);
                if (ms < best_ms) best_ms = ms, best = sh;
            }
        {
            TsShape sh = base;
            sh.side_prep = !base.side_prep;
            const double ms = time_shape(sh);
            if (ms < best_ms) best_ms = ms, best = sh;
        }
        b->shape = best;
    } catch (...) {
        b->shape = before;
        b->drop_graphs();
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    OVX_HIP(hipStreamSynchronize(s));
    b->drop_graphs();
}

// ---------------------------------------------------------------------------
// Serving: continuous batching over resident indexes.
//
// Each slot is a worker thread with its own execution context (stream,
// workspaces, pinned staging).  A free slot takes pending requests in arrival
// order until the next one would exceed the slot's batch capacity (derived
// from its share of the memory budget), runs one two-stage search for the
// coalesced batch with the largest requested k, and hands each request the
// first min(k, n) entries of its rows -- the (score desc, id asc) order makes a
// smaller k a prefix of a larger one, so results do not depend on the mix.
// Slots run concurrently, so batches overlap on the GPU.
// ---------------------------------------------------------------------------
uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct ServeRequest {
    const float* queries = nullptr;  // host, [nq][d]
    uint64_t nq = 0, k = 0;
    float* out_s = nullptr;  // host, [nq][min(k, n)]
    int64_t* out_i = nullptr;
    std::vector<float> own_s;  // buffers owned by the open-loop driver
    std::vector<int64_t> own_i;
    uint64_t t_arrival_ns = 0, t_start_ns = 0, t_done_ns = 0;
    uint32_t status = 0;
    std::string error;
    bool done = false;
    uint64_t batch_queries = 0;
};

struct ServeSlot {
    Ctx ctx;
    std::thread th;
    float* q_stage = nullptr;  // pinned host staging
    float* s_stage = nullptr;
    int64_t* i_stage = nullptr;
};

// The server's device memory: one allocation of the memory budget, lent
// out in 256-byte-aligned regions (first fit over a sorted free list) to
// batches in flight and taken back when they finish.
struct ServePool {
    char* base = nullptr;
    size_t size = 0, in_use = 0, peak = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void init(size_t bytes) {
        if (hipMalloc(&base, bytes) != hipSuccess) {
            (void)hipGetLastError();
            base = nullptr;
            fail(kInternal, "device allocation of the serving pool (" + std::to_string(bytes) + " bytes) failed");
        }
        size = bytes;
        free_[0] = bytes;
    }
    size_t largest() const {
        size_t b = 0;
        for (auto& f : free_) b = std::max(b, f.second);
        return b;
    }
    char* take(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            in_use += bytes;
            peak = std::max(peak, in_use);
            return base + off;
        }
        return nullptr;
    }
    void give(char* p, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        size_t off = (size_t)(p - base);
        in_use -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    void release() {
        if (base) (void)hipFree(base);
        base = nullptr;
        free_.clear();
    }
};

struct Server {
    const SbIndex* sb = nullptr;
    const RqIndex* rq = nullptr;
    uint64_t m = 0, max_batch = 0, kcap = 0;
    std::vector<std::unique_ptr<ServeSlot>> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeRequest*> pending;
    std::unordered_map<uint64_t, ServeRequest*> tickets;
    uint64_t next_ticket = 1;
    bool stopping = false;
    uint64_t batches = 0, batched_queries = 0;
    // Admission by device memory: a batch of q queries holds ws_need[q] bytes
    // of the pool while it runs; a free slot takes pending requests in order
    // while the batch's need fits the largest free region.
    ServePool pool;
    std::vector<size_t> ws_need;  // [0, max_batch]
    uint64_t in_flight = 0, max_in_flight = 0, overlapped = 0;

    void worker(ServeSlot* slot) {
        CtxScope scope(&slot->ctx);
        const uint64_t d = rq->dim;
        for (;;) {
            std::vector<ServeRequest*> batch;
            uint64_t q = 0, kmax = 0;
            char* region = nullptr;
            size_t region_bytes = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    cv_work.wait(lk, [&] { return stopping || !pending.empty(); });
                    if (pending.empty()) return;  // stopping and drained
                    const size_t room = pool.largest();
                    while (!pending.empty()) {
                        ServeRequest* r = pending.front();
                        if (q + r->nq > max_batch || ws_need[q + r->nq] > room) break;
                        batch.push_back(r);
                        pending.pop_front();
                        q += r->nq;
                        kmax = std::max(kmax, r->k);
                    }
                    if (!batch.empty()) break;
                    // the pool is held by batches in flight: wait for one to end
                    cv_work.wait(lk);
                }
                region_bytes = ws_need[q];
                region = pool.take(region_bytes);
                batches++;
                batched_queries += q;
                if (in_flight > 0) overlapped++;
                in_flight++;
                max_in_flight = std::max(max_in_flight, in_flight);
            }
            slot->ctx.lent = region;
            slot->ctx.lent_size = region_bytes;
            uint32_t status = 0;
            std::string err;
            const uint64_t t_start = now_ns();
            const uint64_t out_k = std::min<uint64_t>(kmax, rq->n);
            try {
                uint64_t o = 0;
                for (ServeRequest* r : batch) {
                    std::memcpy(slot->q_stage + o * d, r->queries, r->nq * d * 4);
                    o += r->nq;
                }
                // Queries were checked for NaN/inf when they were submitted.
                two_stage_run(sb, rq, slot->q_stage, q, m, kmax, slot->s_stage, slot->i_stage,
                              ORDVEC_FLAG_ASSUME_FINITE);
                o = 0;
                for (ServeRequest* r : batch) {
                    const uint64_t rk = std::min<uint64_t>(r->k, rq->n);
                    for (uint64_t i = 0; i < r->nq; i++) {
                        std::memcpy(r->out_s + i * rk, slot->s_stage + (o + i) * out_k, rk * 4);
                        std::memcpy(r->out_i + i * rk, slot->i_stage + (o + i) * out_k, rk * 8);
                    }
                    o += r->nq;
                }
            } catch (const Error& e) {
                status = e.status;
                err = e.what();
            } catch (const std::exception& e) {
                status = kPanic;
                err = e.what();
            }
            const uint64_t t = now_ns();
            slot->ctx.lent = nullptr;
            slot->ctx.lent_size = 0;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (ServeRequest* r : batch) {
                    r->status = status;
                    r->error = err;
                    r->t_start_ns = t_start;
                    r->t_done_ns = t;
                    r->batch_queries = q;
                    r->done = true;
                }
                if (region) pool.give(region, region_bytes);  // the batch's work has completed
                in_flight--;
            }
            cv_done.notify_all();
            cv_work.notify_all();  // memory came back: waiting slots may admit now
        }
    }

    void enqueue(ServeRequest* r) {
        {
            std::lock_guard<std::mutex> lk(mu);
            pending.push_back(r);
        }
        cv_work.notify_one();
    }

    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stopping = true;
        }
        cv_work.notify_all();
        for (auto& sl : slots)
            if (sl->th.joinable()) sl->th.join();
        for (auto& sl : slots) {
            {
                CtxScope scope(&sl->ctx);
                (void)hipStreamSynchronize(sl->ctx.stream);
            }
            if (sl->q_stage) (void)hipHostFree(sl->q_stage);
            if (sl->s_stage) (void)hipHostFree(sl->s_stage);
            if (sl->i_stage) (void)hipHostFree(sl->i_stage);
            sl->ctx.release();
            if (sl->ctx.stream) (void)hipStreamDestroy(sl->ctx.stream);
        }
        slots.clear();
        pool.release();
    }
};

// Device bytes one two-stage call of nq queries needs (both workspaces).
size_t two_stage_bytes(const SbIndex* sb, const RqIndex* rq, uint64_t nq, uint64_t m, uint64_t k) {
    const uint64_t d = rq->dim, m_eff = std::min<uint64_t>(m, rq->n), out_k = std::min<uint64_t>(k, rq->n);
    const uint32_t nsplit = rerank_nsplit(nq, m_eff);
    return al(nq * d * 4) * 2 + 3 * al(nq * 4) + al(nq * m_eff * 4) + stage1_bytes(sb, nq, m_eff) +
           al(nq * out_k * 4) + al(nq * out_k * 8) + rerank_work_bytes(nq, (uint32_t)out_k, nsplit);
}

// Workspace that covers every batch size up to cap (the split rerank of small
// batches can need more than a larger batch).
size_t serve_ws_bytes(const SbIndex* sb, const RqIndex* rq, uint64_t cap, uint64_t m, uint64_t k) {
    size_t b = two_stage_bytes(sb, rq, cap, m, k);
    for (uint64_t nq = 1; nq <= std::min<uint64_t>(cap, 512); nq++) b = std::max(b, two_stage_bytes(sb, rq, nq, m, k));
    return b;
}

// ---------------------------------------------------------------------------
// FastScan b=4: tile tables, the plain per-index search and the batched
// multi-index form (persistent scan captured in a HIP graph, autotuned tiles).
// ---------------------------------------------------------------------------
static uint32_t device_cus() {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus <= 0) {
        (void)hipGetLastError();
        cus = 256;
    }
    return (uint32_t)cus;
}

constexpr uint64_t kFs4MfmaMinDocs = 16384;  // plain search: matrix-core tiles from this index size

struct Fs4Config {
    uint32_t qg = 1;          // LUT groups of four queries per tile (1 or 2; 4: matrix-core tiles)
    uint32_t seg_blocks = 64; // 64-document blocks per tile
    uint32_t grid = 512;      // persistent workgroups
    uint32_t rem_qg = 0;      // qg == 4: 0 pads each index's last query set to 16; 1 or 2 runs
                              // only whole 16-query sets on the matrix cores and the rest on
                              // LUT tiles of 4 * rem_qg queries
    // queries_per_tile as the C ABI reports it: 4, 8, 16, or 20 / 24 for the mixed form
    uint32_t qpt() const { return 4 * qg + 4 * rem_qg; }
};
Fs4Config fs4_config_of(uint32_t qpt, uint32_t seg_blocks, uint32_t grid) {
    Fs4Config c;
    c.qg = qpt >= 16 ? 4 : qpt / 4;
    c.rem_qg = qpt > 16 ? (qpt - 16) / 4 : 0;
    c.seg_blocks = seg_blocks;
    c.grid = grid;
    return c;
}

struct Fs4Tables {
    std::vector<Fs4IndexDesc> idx;
    std::vector<Fs4Group> groups;
    std::vector<Fs4Tile> tiles;   // tiles of qg (matrix-core tiles when qg == 4)
    std::vector<Fs4Tile> tiles2;  // mixed form: LUT tiles of rem_qg for each index's last queries
    std::vector<Fs4Merge> merges; // one per query, in query order
    std::vector<Fs4Group> qsets;  // matrix-core tiles: {first group, groups} per query set
    uint32_t nq = 0, kcap = 1;
    uint64_t rows1 = 0, rows2 = 0;  // partial rows of tiles / tiles2 (tiles2's follow tiles')
};

// Queries of index i are rows [q_base_i, q_base_i + nq_i) of the batch.
Fs4Tables fs4_tables(const std::vector<const FsIndex*>& ixs, const std::vector<uint32_t>& nq_per,
                     const std::vector<uint32_t>& k_per, const Fs4Config& c) {
    Fs4Tables t;
    uint32_t q_base = 0;
    std::vector<bool> second;  // merges whose rows index tiles2
    for (size_t i = 0; i < ixs.size(); i++) {
        const FsIndex* ix = ixs[i];
        t.idx.push_back(Fs4IndexDesc{ix->packed4, ix->n});
        const uint32_t nq = nq_per[i];
        const uint32_t k = (uint32_t)std::min<uint64_t>(k_per[i], ix->n);
        t.kcap = std::max(t.kcap, k);
        const uint32_t g_first = (uint32_t)t.groups.size();
        for (uint32_t q = 0; q < nq; q += 4) t.groups.push_back(Fs4Group{q_base + q, std::min(4u, nq - q)});
        const uint32_t ng = (uint32_t)t.groups.size() - g_first;
        const uint32_t nblk = ix->blocks64();
        const uint32_t nseg = nblk ? (nblk + c.seg_blocks - 1) / c.seg_blocks : 0;
        // groups [0, ng1) go in tiles of qg groups, the rest in tiles2 of rem_qg
        const uint32_t ng1 = c.qg == 4 && c.rem_qg ? nq / 16 * 4 : ng;
        auto emit = [&](uint32_t gs, uint32_t gn, uint32_t qg, bool two) {
            std::vector<Fs4Tile>& tl = two ? t.tiles2 : t.tiles;
            const uint32_t t0 = (uint32_t)tl.size();
            const uint32_t qset = (uint32_t)t.qsets.size();
            if (qg == 4) t.qsets.push_back(Fs4Group{g_first + gs, gn});
            for (uint32_t seg = 0; seg < nseg; seg++)
                tl.push_back(Fs4Tile{(uint32_t)i, g_first + gs, gn, seg * c.seg_blocks,
                                     std::min(nblk, (seg + 1) * c.seg_blocks), qg == 4 ? qset : 0u});
            for (uint32_t g = 0; g < gn; g++)
                for (uint32_t s = 0; s < t.groups[g_first + gs + g].count; s++) {
                    t.merges.push_back(Fs4Merge{t0 * 4 * qg + g * 4 + s, nseg, 4 * qg, nseg ? k : 0});
                    second.push_back(two);
                }
        };
        for (uint32_t gs = 0; gs < ng1; gs += c.qg) emit(gs, std::min(c.qg, ng1 - gs), c.qg, false);
        for (uint32_t gs = ng1; gs < ng; gs += c.rem_qg) emit(gs, std::min(c.rem_qg, ng - gs), c.rem_qg, true);
        q_base += nq;
    }
    t.nq = q_base;
    t.rows1 = (uint64_t)t.tiles.size() * 4 * c.qg;
    t.rows2 = (uint64_t)t.tiles2.size() * 4 * (c.rem_qg ? c.rem_qg : 1);
    for (size_t q = 0; q < t.merges.size(); q++)
        if (second[q]) t.merges[q].prow += (uint32_t)t.rows1;
    return t;
}

// Device copy of the tables plus every per-batch buffer, one allocation.
struct Fs4Dev {
    DevBuf buf;
    Fs4IndexDesc* idx = nullptr;
    Fs4Group* groups = nullptr;
    Fs4Tile* tiles = nullptr;
    Fs4Tile* tiles2 = nullptr;
    Fs4Merge* merges = nullptr;
    float *dq = nullptr, *qhat = nullptr, *off = nullptr, *bq = nullptr, *out_s = nullptr;
    uint32_t *qstat = nullptr, *counter = nullptr;
    uint8_t* lut = nullptr;
    Fs4Group* qsets = nullptr;
    uint32_t* frag = nullptr;
    uint64_t* partial = nullptr;
    int64_t* out_i = nullptr;
    size_t partial_bytes = 0;
    uint64_t rows1 = 0;
    uint32_t ntiles = 0, ntiles2 = 0, ngroups = 0, nqsets = 0, nq = 0, kcap = 1;
};

size_t fs4_dev_bytes(const Fs4Tables& t, uint32_t d, uint32_t qg, size_t partial_floor) {
    (void)qg;
    const uint64_t nq = t.nq;
    const size_t part = std::max<size_t>(partial_floor, (size_t)(t.rows1 + t.rows2) * t.kcap * 8);
    return al(t.idx.size() * sizeof(Fs4IndexDesc)) + al(t.groups.size() * sizeof(Fs4Group)) +
           al(t.tiles.size() * sizeof(Fs4Tile)) + al(t.tiles2.size() * sizeof(Fs4Tile)) +
           al(t.merges.size() * sizeof(Fs4Merge)) + 3 * al(nq * d * 4) +
           al(nq * 4) * 2 + al(nq * 8) + al(t.groups.size() * (uint64_t)d * 64) + al(part) + al(256) +
           al(nq * t.kcap * 4) + al(nq * t.kcap * 8) + al(t.qsets.size() * sizeof(Fs4Group)) +
           al(t.qsets.size() * fs4_frag_bytes(d));
}

// Carves the batch buffers out of `base` (fs4_dev_bytes long) and uploads the
// tables.
void fs4_upload(Fs4Dev& dv, const Fs4Tables& t, uint32_t d, uint32_t qg, size_t partial_floor, void* base) {
    Carver cv{static_cast<char*>(base)};
    const uint64_t nq = t.nq;
    dv.idx = cv.take<Fs4IndexDesc>(t.idx.size());
    dv.groups = cv.take<Fs4Group>(t.groups.size());
    dv.tiles = cv.take<Fs4Tile>(t.tiles.size());
    dv.tiles2 = cv.take<Fs4Tile>(t.tiles2.size());
    dv.merges = cv.take<Fs4Merge>(t.merges.size());
    dv.dq = cv.take<float>(nq * d);
    dv.qhat = cv.take<float>(nq * d);
    cv.take<float>(nq * d);
    dv.off = cv.take<float>(nq);
    dv.qstat = cv.take<uint32_t>(nq);
    dv.bq = cv.take<float>(nq * 2);
    dv.lut = cv.take<uint8_t>(t.groups.size() * (uint64_t)d * 64);
    (void)qg;
    dv.partial_bytes = std::max<size_t>(partial_floor, (size_t)(t.rows1 + t.rows2) * t.kcap * 8);
    dv.rows1 = t.rows1;
    dv.partial = cv.take<uint64_t>(dv.partial_bytes / 8);
    dv.counter = cv.take<uint32_t>(64);
    dv.out_s = cv.take<float>(nq * t.kcap);
    dv.out_i = cv.take<int64_t>(nq * t.kcap);
    dv.qsets = cv.take<Fs4Group>(t.qsets.size());
    dv.frag = reinterpret_cast<uint32_t*>(cv.take<char>(t.qsets.size() * fs4_frag_bytes(d)));
    dv.nqsets = (uint32_t)t.qsets.size();
    if (dv.nqsets) h2d(dv.qsets, t.qsets.data(), t.qsets.size() * sizeof(Fs4Group));
    h2d(dv.idx, t.idx.data(), t.idx.size() * sizeof(Fs4IndexDesc));
    h2d(dv.groups, t.groups.data(), t.groups.size() * sizeof(Fs4Group));
    h2d(dv.tiles, t.tiles.data(), t.tiles.size() * sizeof(Fs4Tile));
    if (!t.tiles2.empty()) h2d(dv.tiles2, t.tiles2.data(), t.tiles2.size() * sizeof(Fs4Tile));
    dv.ntiles2 = (uint32_t)t.tiles2.size();
    h2d(dv.merges, t.merges.data(), t.merges.size() * sizeof(Fs4Merge));
    dv.ntiles = (uint32_t)t.tiles.size();
    dv.ngroups = (uint32_t)t.groups.size();
    dv.nq = t.nq;
    dv.kcap = t.kcap;
}

// The launch sequence of one batch: normalise, LUTs, scan, merge.
// persistent: one fixed-grid launch pulling tiles; else one workgroup per tile.
void fs4_enqueue(const Fs4Dev& dv, uint32_t d, const Fs4Config& c, bool persistent, bool prepare = true) {
    hipStream_t s = cur().stream;
    const float inv_norm = rankquant_inv_norm(d, 4);
    if (prepare) {
        launch_prepare_queries(dv.dq, dv.nq, d, 4, 0, inv_norm, dv.qhat, dv.off, dv.qstat, s);
        launch_fs4_lut(dv.qhat, d, dv.groups, dv.ngroups, dv.lut, dv.bq, s);
        if (c.qg == 4) launch_fs4_frag(dv.lut, dv.qsets, dv.nqsets, d, dv.frag, s);
    }
    const uint8_t* tables = c.qg == 4 ? reinterpret_cast<const uint8_t*>(dv.frag) : dv.lut;
    launch_fs4_scan(c.qg, persistent ? c.grid : dv.ntiles, dv.idx, dv.tiles, dv.ntiles,
                    persistent ? dv.counter : nullptr, dv.groups, tables, dv.bq, d, inv_norm, dv.kcap, dv.partial, s);
    if (dv.ntiles2)  // mixed form: each index's last queries on LUT tiles
        launch_fs4_scan(c.rem_qg, persistent ? c.grid : dv.ntiles2, dv.idx, dv.tiles2, dv.ntiles2,
                        persistent ? dv.counter + 32 : nullptr, dv.groups, dv.lut, dv.bq, d, inv_norm, dv.kcap,
                        dv.partial + dv.rows1 * dv.kcap, s);
    launch_fs4_merge(dv.partial, dv.merges, dv.nq, dv.kcap, dv.kcap, dv.out_s, dv.out_i, s);
}

bool fs4_config_fits(uint32_t qg, uint32_t d) {
    if (qg == 4 && !fs4_mfma_ok(d)) return false;
    return fs4_scan_lds_bytes(qg, d) + 1024 <= 160 * 1024;
}
bool fs4_config_fits(const Fs4Config& c, uint32_t d) {
    return fs4_config_fits(c.qg, d) && (c.rem_qg == 0 || fs4_config_fits(c.rem_qg, d));
}

// Plain search of one b=4 index (nq queries, top-k).
// Plain search of one b=4 index (nq queries, top-k).  Whole sets of 16
// queries run on matrix-core tiles (the autotuner's faster shape for full
// sets), the rest on LDS-LUT tiles of 4 or 8 queries: a partly filled 16-query
// tile costs as much as a full one.
void fs4_plain_search(const FsIndex* ix, const float* queries, uint64_t nq, uint64_t kk, float* out_s, int64_t* out_i,
                      uint64_t flags) {
    const uint32_t d = ix->dim;
    // (small indexes stay on LUT tiles: a 16-query tile's fixed cost -- 16
    // top-k lists to clear and sort -- outweighs its scan there)
    const uint64_t n16 = fs4_config_fits(4, d) && ix->n >= kFs4MfmaMinDocs ? nq / 16 * 16 : 0;
    struct Part {
        Fs4Config c;
        Fs4Tables t;
        Fs4Dev dv;
        uint64_t q0, nq;
        size_t bytes;
    };
    std::vector<Part> parts;
    for (int pi = 0; pi < 2; pi++) {
        const uint64_t q0 = pi == 0 ? 0 : n16, cnt = pi == 0 ? n16 : nq - n16;
        if (cnt == 0) continue;
        Part p;
        p.q0 = q0;
        p.nq = cnt;
        p.c.qg = pi == 0 ? 4 : (cnt > 4 && fs4_config_fits(2, d) ? 2 : 1);
        const uint64_t gsets = (cnt + 4 * p.c.qg - 1) / (4 * p.c.qg);
        if (p.c.qg == 4) {  // about one tile per CU, whole 16-block (1024-doc) steps
            const uint64_t want = std::max<uint64_t>(1, (device_cus() + gsets - 1) / gsets);
            const uint64_t sb = std::max<uint64_t>(16, (ix->blocks64() + want - 1) / want);
            p.c.seg_blocks = (uint32_t)((sb + 15) / 16 * 16);
        } else {
            const uint64_t want = std::max<uint64_t>(1, (2048 + gsets - 1) / gsets);
            p.c.seg_blocks = (uint32_t)std::max<uint64_t>(1, (ix->blocks64() + want - 1) / want);
        }
        p.t = fs4_tables({ix}, {(uint32_t)cnt}, {(uint32_t)kk}, p.c);
        p.bytes = fs4_dev_bytes(p.t, d, p.c.qg, 0);
        parts.push_back(std::move(p));
    }
    size_t total = 0;
    for (const Part& p : parts) total += p.bytes;
    char* base = static_cast<char*>(cur().workspace(total));
    const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
    for (Part& p : parts) {
        fs4_upload(p.dv, p.t, d, p.c.qg, 0, base);
        base += p.bytes;
        const float* q = queries + p.q0 * d;
        if (dev) d2d(p.dv.dq, q, p.nq * d * 4);
        else h2d(p.dv.dq, q, p.nq * d * 4);
        fs4_enqueue(p.dv, d, p.c, false);
        const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        OVX_HIP(hipMemcpy2DAsync(out_s + p.q0 * kk, kk * 4, p.dv.out_s, p.dv.kcap * 4, kk * 4, p.nq, kind, cur().stream));
        OVX_HIP(hipMemcpy2DAsync(out_i + p.q0 * kk, kk * 8, p.dv.out_i, p.dv.kcap * 8, kk * 8, p.nq, kind, cur().stream));
    }
    bool bad = false;
    for (const Part& p : parts) bad |= any_bit(read_qstat(p.dv.qstat, p.nq), kStatNonfinite);
    if (dev && !(flags & ORDVEC_FLAG_ASSUME_FINITE) && bad) fail(kNonfiniteQuery, "queries contain NaN or infinity");
}

// Autotuned configurations, keyed by the batch's shape class.
struct Fs4TuneKey {
    uint32_t d, kcap, log2_docs, log2_queries;
    bool operator<(const Fs4TuneKey& o) const {
        return std::tie(d, kcap, log2_docs, log2_queries) < std::tie(o.d, o.kcap, o.log2_docs, o.log2_queries);
    }
};
std::mutex g_tune_mu;
std::map<Fs4TuneKey, Fs4Config>& tune_cache() {
    static std::map<Fs4TuneKey, Fs4Config>* m = new std::map<Fs4TuneKey, Fs4Config>();
    return *m;
}
uint32_t ilog2(uint64_t x) {
    uint32_t r = 0;
    while (x > 1) {
        x >>= 1;
        r++;
    }
    return r;
}

struct Fs4TuneRow {
    Fs4Config c;
    double ms;
};

struct Fs4Batch {
    Ctx ctx;
    std::vector<const FsIndex*> ixs;
    std::vector<uint32_t> nq_per, k_per;
    uint32_t d = 0;
    Fs4Config cfg;
    Fs4Tables tables;
    Fs4Dev dv;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<Fs4TuneRow> tuning;
    Fs4TuneKey key{};

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    // Tables for `c`, and the graph of the whole launch sequence.
    void build(const Fs4Config& c) {
        drop_graph();
        cfg = c;
        tables = fs4_tables(ixs, nq_per, k_per, c);
        const size_t bytes = fs4_dev_bytes(tables, d, c.qg, 0);
        if (bytes > dv.buf.size && dv.buf.p) {
            sync();
            (void)hipFree(dv.buf.p);
            dv.buf.p = nullptr;
            dv.buf.size = 0;
        }
        fs4_upload(dv, tables, d, c.qg, 0, dv.buf.get(bytes, cur().stream));
        sync();
        OVX_HIP(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
        try {
            fs4_enqueue(dv, d, cfg, true);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(ctx.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        OVX_HIP(hipStreamEndCapture(ctx.stream, &graph));
        OVX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    ~Fs4Batch() {
        {
            CtxScope scope(&ctx);
            (void)hipStreamSynchronize(ctx.stream);
            drop_graph();
            if (dv.buf.p) (void)hipFree(dv.buf.p);
            ctx.release();
        }
        if (ctx.stream) (void)hipStreamDestroy(ctx.stream);
    }
};

// ---------------------------------------------------------------------------
// Persisted formats (src/rank_io.rs; docs/PERSISTED_FORMAT.md)
// ---------------------------------------------------------------------------
struct Reader {
    std::vector<uint8_t> buf;
    size_t pos = 0;
    std::string label;
    void need(size_t k, const char* field) {
        if (pos + k > buf.size()) fail(kCorruptIndex, label + " header truncated while reading " + field);
    }
    uint8_t u8(const char* f) {
        need(1, f);
        return buf[pos++];
    }
    uint32_t u32(const char* f) {
        need(4, f);
        uint32_t v = (uint32_t)buf[pos] | ((uint32_t)buf[pos + 1] << 8) | ((uint32_t)buf[pos + 2] << 16) |
                     ((uint32_t)buf[pos + 3] << 24);
        pos += 4;
        return v;
    }
    void version() {
        uint8_t v = u8("version");
        if (v != 1) fail(kCorruptIndex, "unsupported " + label + " version: " + std::to_string(v));
    }
    void payload_exact(uint64_t bytes) {
        if (bytes > kMaxPayload) fail(kCorruptIndex, "payload " + std::to_string(bytes) + " B exceeds MAX_PAYLOAD");
        uint64_t remaining = buf.size() - pos;
        if (bytes > remaining)
            fail(kCorruptIndex, label + " payload truncated: header declares " + std::to_string(bytes) +
                                    " B but file has " + std::to_string(remaining) + " B remaining");
        if (bytes < remaining)
            fail(kCorruptIndex, label + " payload has trailing bytes: header declares " + std::to_string(bytes) +
                                    " B but file has " + std::to_string(remaining) + " B remaining");
    }
};

// Largest index file: the longest header (17 B, .ovbm) plus MAX_PAYLOAD.
constexpr uint64_t kMaxFileBytes = kMaxPayload + 17;

// Opens a regular file for reading and returns its size.  Directories, FIFOs
// and devices are IO errors (a read of them cannot fill an index).
FILE* open_index_file(const char* path, uint64_t& size) {
    FILE* f = std::fopen(path, "rb");
    if (!f) fail(kIo, std::string("open index: ") + std::strerror(errno));
    struct stat st;
    if (fstat(fileno(f), &st) != 0 || !S_ISREG(st.st_mode)) {
        std::fclose(f);
        fail(kIo, "open index: not a regular file");
    }
    size = (uint64_t)st.st_size;
    return f;
}

// The first min(size, limit) bytes of a file; a whole-file read (no limit)
// refuses files larger than any valid index before allocating.
std::vector<uint8_t> read_file(const char* path, size_t limit = SIZE_MAX) {
    uint64_t sz = 0;
    std::unique_ptr<FILE, int (*)(FILE*)> f(open_index_file(path, sz), &std::fclose);
    if (limit == SIZE_MAX && sz > kMaxFileBytes)
        fail(kCorruptIndex, "index file of " + std::to_string(sz) + " B exceeds header + MAX_PAYLOAD");
    const size_t want = (size_t)std::min<uint64_t>(sz, limit);
    std::vector<uint8_t> b(want);
    const size_t got = want ? std::fread(b.data(), 1, want, f.get()) : 0;
    if (got != want) fail(kIo, "short read");
    return b;
}

uint64_t file_size(const char* path) {
    uint64_t sz = 0;
    std::fclose(open_index_file(path, sz));
    return sz;
}

void write_file(const char* path, const std::vector<uint8_t>& hdr, const void* payload, uint64_t bytes) {
    FILE* f = std::fopen(path, "wb");
    if (!f) fail(kIo, std::string("create: ") + std::strerror(errno));
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    if (ok && bytes) ok = std::fwrite(payload, 1, bytes, f) == bytes;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) fail(kIo, "write failed");
}

void put_u32(std::vector<uint8_t>& h, uint32_t v) {
    for (int i = 0; i < 4; i++) h.push_back((uint8_t)(v >> (8 * i)));
}

bool magic_is(const std::vector<uint8_t>& b, const char* a, const char* legacy) {
    if (b.size() < 4) return false;
    return std::memcmp(b.data(), a, 4) == 0 || (legacy && std::memcmp(b.data(), legacy, 4) == 0);
}

// src/rank_io.rs:713-818: header, exact payload, per-row constant composition.
struct OvrqParsed {
    uint32_t bits, dim;
    uint64_t n;
    size_t payload_off;
};

OvrqParsed parse_ovrq_header(Reader& r, bool check_payload_len) {
    r.label = "OVRQ";
    r.pos = 4;
    r.version();
    uint8_t bits = r.u8("bits");
    if (bits != 1 && bits != 2 && bits != 4)
        fail(kCorruptIndex, "unsupported OVRQ bits: " + std::to_string(bits) + " (expected 1, 2, or 4)");
    uint32_t dim = r.u32("dim");
    if (dim < 2 || dim > kMaxDim) fail(kCorruptIndex, "dim " + std::to_string(dim) + " out of range [2, 65535]");
    uint32_t nb = 1u << bits;
    if (dim % nb)
        fail(kCorruptIndex, "OVRQ dim " + std::to_string(dim) + " is not a multiple of 2^bits = " + std::to_string(nb) +
                                "; constant-composition invariant violated");
    if (dim % (8 / bits)) fail(kCorruptIndex, "OVRQ dim is not a multiple of codes_per_byte");
    uint32_t n = r.u32("n_vectors");
    if (n > kMaxVectors) fail(kCorruptIndex, "n_vectors " + std::to_string(n) + " exceeds MAX_VECTORS");
    uint64_t payload = (uint64_t)n * (dim * bits / 8);
    if (check_payload_len) r.payload_exact(payload);
    return {bits, dim, n, r.pos};
}

// Files whose payload reaches this size are validated on the GPU after the
// upload (the reference's per-row checks, the same first-row message).
constexpr uint64_t kDeviceCheckBytes = 64ull << 20;

// Device slot for a first-invalid-row search, set to "none".
unsigned long long* first_bad_slot() {
    auto* p = static_cast<unsigned long long*>(cur().workspace2(256));
    OVX_HIP(hipMemsetAsync(p, 0xff, 8, cur().stream));
    return p;
}

uint64_t read_first_bad(const unsigned long long* p) {
    uint64_t v = 0;
    d2h(&v, p, 8);
    sync();
    return v;
}

// src/rank_io.rs:764-816: the reference's message for a row whose bucket
// counts are not all dim / 2^bits (empty when the row is valid).
std::string ovrq_row_error(const uint8_t* rp, uint64_t row, uint32_t dim, uint32_t bits) {
    const uint32_t bpv = dim * bits / 8, cpb = 8 / bits, nb = 1u << bits, per = dim / nb;
    uint32_t hist[16] = {0};
    for (uint32_t g = 0; g < bpv; g++)
        for (uint32_t slot = 0; slot < cpb; slot++) hist[(rp[g] >> ((cpb - 1 - slot) * bits)) & (nb - 1)]++;
    for (uint32_t b = 0; b < nb; b++)
        if (hist[b] != per)
            return "OVRQ row " + std::to_string(row) + " violates constant composition: bucket " + std::to_string(b) +
                   " has " + std::to_string(hist[b]) + " codes, expected " + std::to_string(per) + " (= dim / 2^bits)";
    return "";
}

RqIndex* load_ovrq(const char* path) {
    Reader r;
    r.buf = read_file(path);
    if (!magic_is(r.buf, "OVRQ", "TVRQ")) fail(kCorruptIndex, "not an OVRQ/TVRQ (RankQuant) file: wrong magic");
    OvrqParsed h = parse_ovrq_header(r, true);
    const uint8_t* pay = r.buf.data() + h.payload_off;
    const uint32_t bpv = h.dim * h.bits / 8;
    const bool on_device = (uint64_t)h.n * bpv >= kDeviceCheckBytes;
    // Per-row constant composition, first bad row reported.
    if (!on_device)
        for (uint64_t row = 0; row < h.n; row++) {
            std::string e = ovrq_row_error(pay + row * bpv, row, h.dim, h.bits);
            if (!e.empty()) fail(kCorruptIndex, e);
        }
    eng().init();
    RqIndex* ix = new RqIndex();
    ix->dim = h.dim;
    ix->bits = h.bits;
    ix->bpv = bpv;
    try {
        ix->reserve(h.n);
        h2d(ix->codes, pay, h.n * bpv);
        if (on_device) {
            unsigned long long* fb = first_bad_slot();
            launch_check_rq_rows(ix->codes, h.n, h.dim, h.bits, fb, cur().stream);
            const uint64_t bad = read_first_bad(fb);
            if (bad < h.n) fail(kCorruptIndex, ovrq_row_error(pay + bad * bpv, bad, h.dim, h.bits));
        }
        sync();
    } catch (...) {
        delete ix;
        throw;
    }
    ix->n = h.n;
    return ix;
}

void write_ovrq(const RqIndex* ix, const char* path) {
    std::vector<uint8_t> host(ix->n * ix->bpv);
    d2h(host.data(), ix->codes, host.size());
    sync();
    std::vector<uint8_t> h = {'O', 'V', 'R', 'Q', 1, (uint8_t)ix->bits};
    put_u32(h, ix->dim);
    put_u32(h, (uint32_t)ix->n);
    write_file(path, h, host.data(), host.size());
}

// src/rank_io.rs:574-640: header, exact payload, every row a permutation of
// [0, dim) (bound check plus a per-row stamp for repeats).
// src/rank_io.rs:600-640: the reference's message for the first value of a
// row that is >= dim or repeats (empty when the row is a permutation).
std::string ovr_row_error(const uint8_t* rp, uint64_t row, uint32_t dim, std::vector<uint32_t>& seen,
                          uint32_t stamp) {
    for (uint32_t p = 0; p < dim; p++) {
        const uint32_t v = (uint32_t)rp[2 * p] | ((uint32_t)rp[2 * p + 1] << 8);
        if (v >= dim)
            return "OVR1 rank value " + std::to_string(v) + " >= dim (" + std::to_string(dim) +
                   "); ranks must be a permutation of [0, dim)";
        if (seen[v] == stamp)
            return "OVR1 row " + std::to_string(row) + " is not a permutation of [0, dim): value " + std::to_string(v) +
                   " repeats";
        seen[v] = stamp;
    }
    return "";
}

// src/rank_io.rs:574-640: header, exact payload, every row a permutation of
// [0, dim) (bound check plus a per-row stamp for repeats).
RkIndex* load_ovr(const char* path) {
    Reader r;
    r.buf = read_file(path);
    if (!magic_is(r.buf, "OVR1", "TVR1")) fail(kCorruptIndex, "not an OVR1/TVR1 (Rank) file: wrong magic");
    r.label = "OVR1";
    r.pos = 4;
    r.version();
    const uint32_t dim = r.u32("dim");
    if (dim < 2 || dim > kMaxDim) fail(kCorruptIndex, "dim " + std::to_string(dim) + " out of range [2, 65535]");
    const uint32_t n = r.u32("n_vectors");
    if (n > kMaxVectors) fail(kCorruptIndex, "n_vectors " + std::to_string(n) + " exceeds MAX_VECTORS");
    const uint64_t payload = (uint64_t)n * dim * 2;
    r.payload_exact(payload);
    const uint8_t* pay = r.buf.data() + r.pos;
    const bool on_device = payload >= kDeviceCheckBytes;
    std::vector<uint32_t> seen(dim, 0);
    if (!on_device)
        for (uint64_t row = 0; row < n; row++) {
            std::string e = ovr_row_error(pay + row * dim * 2, row, dim, seen, (uint32_t)row + 1);
            if (!e.empty()) fail(kCorruptIndex, e);
        }
    RkIndex* ix = new RkIndex();
    ix->dim = dim;
    try {
        eng().init();
        ix->reserve(n);
        h2d(ix->ranks, pay, payload);  // u16 little-endian on both sides
        if (on_device) {
            unsigned long long* fb = first_bad_slot();
            launch_check_rank_rows(ix->ranks, n, dim, fb, cur().stream);
            const uint64_t bad = read_first_bad(fb);
            if (bad < n) fail(kCorruptIndex, ovr_row_error(pay + bad * dim * 2, bad, dim, seen, 1));
        }
        sync();
    } catch (...) {
        delete ix;
        throw;
    }
    ix->n = n;
    return ix;
}

// src/rank_io.rs:544-560: "OVR1" | version 1 | dim u32 | n u32 | u16 LE ranks.
void write_ovr(const RkIndex* ix, const char* path) {
    const uint64_t bytes = ix->n * ix->row_bytes();
    if (bytes > kMaxPayload) fail(kBadArgument, "payload exceeds MAX_PAYLOAD");
    std::vector<uint8_t> host(bytes);
    d2h(host.data(), ix->ranks, bytes);
    sync();
    std::vector<uint8_t> h = {'O', 'V', 'R', '1', 1};
    put_u32(h, ix->dim);
    put_u32(h, (uint32_t)ix->n);
    write_file(path, h, host.data(), bytes);
}

// swap_remove on row-major storage (src/rank.rs:506-517, src/quant.rs:869-881):
// the last row moves into `idx`; returns the moved row's old index.
uint64_t swap_remove_rows(void* buf, uint64_t& n, uint64_t row_bytes, uint64_t idx) {
    if (idx >= n)
        fail(kBadArgument, "index " + std::to_string(idx) + " out of range (index holds " + std::to_string(n) +
                               " vectors)");
    const uint64_t last = n - 1;
    if (idx != last) {
        char* b = static_cast<char*>(buf);
        d2d(b + idx * row_bytes, b + last * row_bytes, row_bytes);
        sync();
    }
    n = last;
    return last;
}

// Row-major host words <-> interleaved device layout.
// Row-major host bitmaps -> interleaved device layout: rows are staged in
// chunks and scattered by a kernel.  ntop > 0 also checks every row's
// popcount on the device; the lowest bad row is returned (n when none).
uint64_t sb_upload_rowmajor(SbIndex* ix, const uint64_t* rows, uint64_t n, uint32_t ntop = 0) {
    ix->reserve(n);
    if (n == 0) return 0;
    const uint64_t row_bytes = (uint64_t)ix->qpv * 8;
    const uint64_t chunk = std::max<uint64_t>(1, kChunkBytes / row_bytes);
    unsigned long long* fb = ntop ? first_bad_slot() : nullptr;
    uint64_t* stage = static_cast<uint64_t*>(cur().workspace(std::min(n, chunk) * row_bytes));
    for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
        const uint64_t cnt = std::min(chunk, n - r0);
        h2d(stage, rows + r0 * ix->qpv, cnt * row_bytes);
        launch_interleave_rows(stage, cnt, r0, ix->qpv, ix->words, ntop, fb, cur().stream);
        sync();  // the staging buffer is reused by the next chunk
    }
    ix->n = n;
    return ntop ? std::min<uint64_t>(read_first_bad(fb), n) : n;
}

std::vector<uint64_t> sb_download_rowmajor(const SbIndex* ix) {
    const uint64_t nblk = (ix->n + 63) / 64, bw = ix->block_words();
    std::vector<uint64_t> inter(nblk * bw);
    d2h(inter.data(), ix->words, inter.size() * 8);
    sync();
    std::vector<uint64_t> rows(ix->n * ix->qpv);
    for (uint64_t doc = 0; doc < ix->n; doc++)
        for (uint32_t w = 0; w < ix->qpv; w++) rows[doc * ix->qpv + w] = inter[(doc / 64) * bw + (uint64_t)w * 64 + (doc & 63)];
    return rows;
}

// src/rank_io.rs:1034-1075
SbIndex* load_ovsb(const char* path) {
    Reader r;
    r.buf = read_file(path);
    if (!magic_is(r.buf, "OVSB", "TVSB")) fail(kCorruptIndex, "not an OVSB/TVSB (SignBitmap) file: wrong magic");
    r.label = "OVSB";
    r.pos = 4;
    r.version();
    uint32_t dim = r.u32("dim");
    if (dim < 64 || dim > kMaxSignDim) fail(kCorruptIndex, "OVSB dim " + std::to_string(dim) + " out of range [64, 16777216]");
    if (dim % 64) fail(kCorruptIndex, "OVSB dim " + std::to_string(dim) + " is not a multiple of 64");
    uint32_t n = r.u32("n_vectors");
    if (n > kMaxVectors) fail(kCorruptIndex, "n_vectors exceeds MAX_VECTORS");
    r.payload_exact((uint64_t)n * (dim / 64) * 8);
    std::vector<uint64_t> rows((uint64_t)n * (dim / 64));
    std::memcpy(rows.data(), r.buf.data() + r.pos, rows.size() * 8);  // little-endian host
    eng().init();
    SbIndex* ix = new SbIndex();
    ix->dim = dim;
    ix->qpv = dim / 64;
    try {
        sb_upload_rowmajor(ix, rows.data(), n);
    } catch (...) {
        delete ix;
        throw;
    }
    return ix;
}

void write_ovsb(const SbIndex* ix, const char* path) {
    std::vector<uint64_t> rows = sb_download_rowmajor(ix);
    std::vector<uint8_t> h = {'O', 'V', 'S', 'B', 1};
    put_u32(h, ix->dim);
    put_u32(h, (uint32_t)ix->n);
    write_file(path, h, rows.data(), rows.size() * 8);
}

// src/rank_io.rs:826-957 (.ovbm): "OVBM"/"TVBM" | 1 | dim u32 | n_top u32 |
// n u32 | u64 LE words row-major; every row must have exactly n_top bits set.
SbIndex* load_ovbm(const char* path) {
    Reader r;
    r.buf = read_file(path);
    if (!magic_is(r.buf, "OVBM", "TVBM")) fail(kCorruptIndex, "not an OVBM/TVBM (Bitmap) file: wrong magic");
    r.label = "OVBM";
    r.pos = 4;
    r.version();
    uint32_t dim = r.u32("dim");
    if (dim < 2 || dim > kMaxDim) fail(kCorruptIndex, "dim out of range [2, 65535]");
    if (dim % 64) fail(kCorruptIndex, "OVBM dim " + std::to_string(dim) + " is not a multiple of 64");
    uint32_t n_top = r.u32("n_top");
    if (n_top == 0 || n_top >= dim)
        fail(kCorruptIndex, "OVBM n_top " + std::to_string(n_top) + " must satisfy 0 < n_top < dim (" + std::to_string(dim) + ")");
    uint32_t n = r.u32("n_vectors");
    if (n > kMaxVectors) fail(kCorruptIndex, "n_vectors exceeds MAX_VECTORS");
    const uint32_t qpv = dim / 64;
    r.payload_exact((uint64_t)n * qpv * 8);
    std::vector<uint64_t> rows((uint64_t)n * qpv);
    std::memcpy(rows.data(), r.buf.data() + r.pos, rows.size() * 8);  // little-endian host
    // src/rank_io.rs:938-951: every row holds exactly n_top bits.
    auto row_error = [&](uint64_t i) {
        uint32_t pop = 0;
        for (uint32_t w = 0; w < qpv; w++) pop += (uint32_t)__builtin_popcountll(rows[i * qpv + w]);
        return pop == n_top ? std::string()
                            : "OVBM row " + std::to_string(i) + " has " + std::to_string(pop) +
                                  " bits set, expected n_top = " + std::to_string(n_top);
    };
    const bool on_device = (uint64_t)n * qpv * 8 >= kDeviceCheckBytes;
    if (!on_device)
        for (uint64_t i = 0; i < n; i++) {
            std::string e = row_error(i);
            if (!e.empty()) fail(kCorruptIndex, e);
        }
    eng().init();
    SbIndex* ix = new SbIndex();
    ix->dim = dim;
    ix->qpv = qpv;
    ix->ntop = n_top;
    try {
        const uint64_t bad = sb_upload_rowmajor(ix, rows.data(), n, on_device ? n_top : 0);
        if (bad < n) fail(kCorruptIndex, row_error(bad));
    } catch (...) {
        delete ix;
        throw;
    }
    return ix;
}

void write_ovbm(const SbIndex* ix, const char* path) {
    std::vector<uint64_t> rows = sb_download_rowmajor(ix);
    std::vector<uint8_t> h = {'O', 'V', 'B', 'M', 1};
    put_u32(h, ix->dim);
    put_u32(h, ix->ntop);
    put_u32(h, (uint32_t)ix->n);
    write_file(path, h, rows.data(), rows.size() * 8);
}

// src/rank_io.rs:1077-1250 (.ovfs): header, nibble domain, b=2 composition,
// zero tail padding.
FsIndex* load_ovfs(const char* path) {
    Reader r;
    r.buf = read_file(path);
    if (!magic_is(r.buf, "OVFS", nullptr)) fail(kCorruptIndex, "not an OVFS (RankQuantFastscan) file: wrong magic");
    r.label = "OVFS";
    r.pos = 4;
    r.version();
    uint32_t dim = r.u32("dim");
    if (dim < 2 || dim > kMaxDim) fail(kCorruptIndex, "OVFS dim out of range");
    if (dim % 4) fail(kCorruptIndex, "OVFS dim " + std::to_string(dim) + " is not a multiple of 4");
    uint32_t n = r.u32("n_vectors");
    if (n > kMaxVectors) fail(kCorruptIndex, "n_vectors exceeds MAX_VECTORS");
    const uint64_t pairs = dim / 2, nblk = ((uint64_t)n + 31) / 32, bytes = nblk * pairs * 32;
    r.payload_exact(bytes);
    const uint8_t* pay = r.buf.data() + r.pos;
    for (uint64_t b = 0; b < nblk; b++) {
        uint64_t inb = std::min<uint64_t>(32, n - b * 32);
        for (uint64_t l = 0; l < 32; l++) {
            uint32_t hist[4] = {0, 0, 0, 0};
            for (uint64_t p = 0; p < pairs; p++) {
                uint8_t v = pay[b * pairs * 32 + p * 32 + l];
                if (l >= inb) {
                    if (v) fail(kCorruptIndex, "OVFS block-tail padding must be zero");
                    continue;
                }
                if (v > 15) fail(kCorruptIndex, "OVFS byte is not a 4-bit FastScan code");
                hist[v >> 2]++;
                hist[v & 3]++;
            }
            if (l < inb)
                for (int k = 0; k < 4; k++)
                    if (hist[k] != dim / 4)
                        fail(kCorruptIndex, "OVFS row " + std::to_string(b * 32 + l) + " violates b=2 constant composition");
        }
    }
    eng().init();
    FsIndex* ix = new FsIndex();
    ix->dim = dim;
    ix->n = n;
    if (bytes) {
        if (hipMalloc(&ix->packed, bytes) != hipSuccess) {
            (void)hipGetLastError();
            delete ix;
            fail(kInternal, "HBM allocation for FastScan payload failed");
        }
        h2d(ix->packed, pay, bytes);
        sync();
    }
    return ix;
}

void write_ovfs(const FsIndex* ix, const char* path) {
    std::vector<uint8_t> host(ix->payload());
    d2h(host.data(), ix->packed, host.size());
    sync();
    std::vector<uint8_t> h = {'O', 'V', 'F', 'S', 1};
    put_u32(h, ix->dim);
    put_u32(h, (uint32_t)ix->n);
    write_file(path, h, host.data(), host.size());
}

// ---------------------------------------------------------------------------
// ABI v1 helpers (ordvec-ffi/src/lib.rs:246-598)
// ---------------------------------------------------------------------------
ordvec_index_info_t default_info() {
    ordvec_index_info_t i;
    std::memset(&i, 0, sizeof(i));
    i.struct_size = sizeof(ordvec_index_info_t);
    return i;
}

ordvec_search_params_t default_params() {
    ordvec_search_params_t p;
    std::memset(&p, 0, sizeof(p));
    p.struct_size = sizeof(ordvec_search_params_t);
    return p;
}

ordvec_search_stats_t default_stats() {
    ordvec_search_stats_t s;
    std::memset(&s, 0, sizeof(s));
    s.struct_size = sizeof(ordvec_search_stats_t);
    s.abi_version = ORDVEC_ABI_VERSION;
    return s;
}

void check_exact_size(uint64_t got, size_t expected, const char* what) {
    if (got != expected)
        fail(kBadStructSize, std::string(what) + ".struct_size must be exactly " + std::to_string(expected) + ", got " +
                                 std::to_string(got));
}

constexpr uint64_t kAllCaps =
    ORDVEC_CAP_FULL_SEARCH | ORDVEC_CAP_SUBSET_SEARCH | ORDVEC_CAP_STATS | ORDVEC_CAP_ID_EQUALS_ROW_ID;

// ---------------------------------------------------------------------------
// Optional event timing of kernel families (engine mutex held by callers)
// ---------------------------------------------------------------------------
struct KTimes {
    uint32_t mask = 0;  // families recorded (bit f = family f)
    struct Pending {
        hipEvent_t a, b;
        int fam;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t open[kFamCount] = {};
    uint64_t launches[kFamCount] = {};
    double total_ms[kFamCount] = {};

    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        OVX_HIP(hipEventCreate(&e));
        return e;
    }
    void resolve() {
        if (pending.empty()) return;
        OVX_HIP(hipEventSynchronize(pending.back().b));
        for (auto& p : pending) {
            float ms = 0.f;
            OVX_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            launches[p.fam]++;
            total_ms[p.fam] += ms;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
};

KTimes& kt() {
    static KTimes* t = new KTimes();
    return *t;
}

}  // namespace

void ktime_begin(int family, hipStream_t s) {
    KTimes& t = kt();
    if (!((t.mask >> family) & 1u) || (t_ctx && !t_ctx->timed) || t_no_events) return;  // serving slots are not
                                                                                         // event-timed
    hipEvent_t e = t.get();
    OVX_HIP(hipEventRecord(e, s));
    t.open[family] = e;
}

void ktime_end(int family, hipStream_t s) {
    KTimes& t = kt();
    if (!((t.mask >> family) & 1u) || (t_ctx && !t_ctx->timed) || t_no_events || !t.open[family]) return;
    hipEvent_t e = t.get();
    OVX_HIP(hipEventRecord(e, s));
    t.pending.push_back({t.open[family], e, family});
    t.open[family] = nullptr;
    if (t.pending.size() > 4096) t.resolve();
}

}  // namespace ovx

using namespace ovx;

struct ordvec_index_t {};
struct ordvec_rankquant_t {};
struct ordvec_sign_bitmap_t {};
struct ordvec_fastscan_t {};
struct ordvec_server_t {};
struct ordvec_bitmap_t {};
struct ordvec_rank_t {};

static IndexHandle* H(const ordvec_index_t* p) { return reinterpret_cast<IndexHandle*>(const_cast<ordvec_index_t*>(p)); }
static RqIndex* RQ(const ordvec_rankquant_t* p) { return reinterpret_cast<RqIndex*>(const_cast<ordvec_rankquant_t*>(p)); }
static SbIndex* SB(const ordvec_sign_bitmap_t* p) { return reinterpret_cast<SbIndex*>(const_cast<ordvec_sign_bitmap_t*>(p)); }
static RkIndex* RK(const ordvec_rank_t* p) { return reinterpret_cast<RkIndex*>(const_cast<ordvec_rank_t*>(p)); }
static FsIndex* FS(const ordvec_fastscan_t* p) { return reinterpret_cast<FsIndex*>(const_cast<ordvec_fastscan_t*>(p)); }

#define LOCK std::lock_guard<std::recursive_mutex> _lk(eng().mu)

extern "C" {

uint32_t ordvec_abi_version(void) { return ORDVEC_ABI_VERSION; }

const char* ordvec_version_string(void) { return "0.6.0"; }

const char* ordvec_status_name(ordvec_status_t s) {
    switch (s) {
        case 0: return "OK";
        case 1: return "NULL_POINTER";
        case 2: return "BAD_ARGUMENT";
        case 3: return "BAD_STRUCT_SIZE";
        case 4: return "UNSUPPORTED_FORMAT";
        case 5: return "CORRUPT_INDEX";
        case 6: return "IO";
        case 7: return "DIM_MISMATCH";
        case 8: return "NONFINITE_QUERY";
        case 9: return "ROW_ID_OUT_OF_RANGE";
        case 10: return "BUFFER_TOO_SMALL";
        case 11: return "UNSUPPORTED_OPERATION";
        case 12: return "PANIC";
        case 13: return "INTERNAL";
        default: return "UNKNOWN";
    }
}

const char* ordvec_last_error(void) { return g_last_error.c_str(); }

ordvec_status_t ordvec_init(void) {
    return boundary([&] {
        LOCK;
        eng().init();
        (void)noalloc_scratch();  // the allocation-free entry point's scratch address on this device
    });
}

void ordvec_index_info_init(ordvec_index_info_t* info) {
    if (info) *info = default_info();
}
void ordvec_search_params_init(ordvec_search_params_t* p) {
    if (p) *p = default_params();
}
void ordvec_search_stats_init(ordvec_search_stats_t* s) {
    if (s) *s = default_stats();
}

ordvec_status_t ordvec_index_load(const char* path, uint64_t flags, ordvec_index_t** out) {
    return boundary([&] {
        NONNULL(path, "path");
        NONNULL(out, "out");
        *out = nullptr;
        if (flags != 0) fail(kBadArgument, "unknown load flags: " + std::to_string(flags));
        std::vector<uint8_t> head = read_file(path, 4);
        if (head.size() < 4) fail(kCorruptIndex, "read index magic: failed to fill whole buffer");
        uint64_t fsz = file_size(path);
        if (magic_is(head, "OVRQ", "TVRQ")) {
            LOCK;
            IndexHandle* h = new IndexHandle();
            try {
                h->rq = load_ovrq(path);
            } catch (...) {
                delete h;
                throw;
            }
            h->source_file_size_bytes = fsz;
            *out = reinterpret_cast<ordvec_index_t*>(h);
        } else if (magic_is(head, "OVBM", "TVBM")) {
            LOCK;
            IndexHandle* h = new IndexHandle();
            try {
                h->bm = load_ovbm(path);
            } catch (...) {
                delete h;
                throw;
            }
            h->source_file_size_bytes = fsz;
            *out = reinterpret_cast<ordvec_index_t*>(h);
        } else if (magic_is(head, "OVR1", "TVR1") || magic_is(head, "OVSB", "TVSB")) {
            fail(kUnsupportedFormat, "ABI v1 supports only RankQuant and Bitmap indexes");
        } else if (magic_is(head, "OVFS", nullptr)) {
            fail(kUnsupportedFormat, "ABI v1 does not support RankQuantFastscan indexes (tracked in #232)");
        } else {
            fail(kCorruptIndex, "unrecognized ordvec index magic");
        }
    });
}

ordvec_status_t ordvec_index_probe(const char* path, uint64_t flags, ordvec_index_info_t* info_out) {
    return boundary([&] {
        NONNULL(path, "path");
        NONNULL(info_out, "info_out");
        if (flags != 0) fail(kBadArgument, "unknown probe flags: " + std::to_string(flags));
        check_exact_size(info_out->struct_size, sizeof(ordvec_index_info_t), "ordvec_index_info_t");
        uint64_t fsz = file_size(path);
        Reader r;
        r.buf = read_file(path, 64);
        if (r.buf.size() < 4) fail(kCorruptIndex, "probe index metadata: ordvec index header truncated while reading magic");
        ordvec_index_info_t info = default_info();
        info.format_version = 1;
        info.source_file_size_bytes = fsz;
        info.capabilities = kAllCaps;
        // Header fields are parsed from the head; the exact-length rule uses the real size.
        auto exact = [&](uint64_t header, uint64_t payload, const char* label) {
            if (payload > kMaxPayload) fail(kCorruptIndex, "payload exceeds MAX_PAYLOAD");
            uint64_t remaining = fsz >= header ? fsz - header : 0;
            if (payload > remaining)
                fail(kCorruptIndex, std::string("probe index metadata: ") + label + " payload truncated: header declares " +
                                        std::to_string(payload) + " B but file has " + std::to_string(remaining) + " B remaining");
            if (payload < remaining)
                fail(kCorruptIndex, std::string("probe index metadata: ") + label + " payload has trailing bytes: header declares " +
                                        std::to_string(payload) + " B but file has " + std::to_string(remaining) + " B remaining");
        };
        if (magic_is(r.buf, "OVRQ", "TVRQ")) {
            OvrqParsed h = parse_ovrq_header(r, false);
            exact(14, h.n * (uint64_t)(h.dim * h.bits / 8), "OVRQ");
            info.kind = ORDVEC_INDEX_KIND_RANK_QUANT;
            info.dim = h.dim;
            info.bit_width = h.bits;
            info.vector_count = h.n;
            info.bytes_per_vec = h.dim * h.bits / 8;
        } else if (magic_is(r.buf, "OVBM", "TVBM")) {
            r.label = "OVBM";
            r.pos = 4;
            r.version();
            uint32_t dim = r.u32("dim");
            if (dim < 2 || dim > kMaxDim) fail(kCorruptIndex, "dim out of range [2, 65535]");
            if (dim % 64) fail(kCorruptIndex, "OVBM dim " + std::to_string(dim) + " is not a multiple of 64");
            uint32_t n_top = r.u32("n_top");
            if (n_top == 0 || n_top >= dim) fail(kCorruptIndex, "OVBM n_top must satisfy 0 < n_top < dim");
            uint32_t n = r.u32("n_vectors");
            if (n > kMaxVectors) fail(kCorruptIndex, "n_vectors exceeds MAX_VECTORS");
            exact(17, (uint64_t)n * (dim / 64) * 8, "OVBM");
            info.kind = ORDVEC_INDEX_KIND_BITMAP;
            info.dim = dim;
            info.n_top = n_top;
            info.vector_count = n;
            info.bytes_per_vec = dim / 8;
        } else if (magic_is(r.buf, "OVR1", "TVR1") || magic_is(r.buf, "OVSB", "TVSB")) {
            fail(kUnsupportedFormat, "ABI v1 supports metadata probes only for RankQuant and Bitmap indexes");
        } else if (magic_is(r.buf, "OVFS", nullptr)) {
            fail(kCorruptIndex, "probe index metadata: OVFS (RankQuantFastscan) metadata probing is not supported in this version");
        } else {
            fail(kCorruptIndex, "probe index metadata: unknown ordvec index magic");
        }
        *info_out = info;
    });
}

ordvec_status_t ordvec_index_info(const ordvec_index_t* index, ordvec_index_info_t* info_out) {
    return boundary([&] {
        NONNULL(index, "index");
        NONNULL(info_out, "info_out");
        check_exact_size(info_out->struct_size, sizeof(ordvec_index_info_t), "ordvec_index_info_t");
        const IndexHandle* h = H(index);
        ordvec_index_info_t info = default_info();
        info.format_version = 1;
        if (h->bm) {
            info.kind = ORDVEC_INDEX_KIND_BITMAP;
            info.dim = h->bm->dim;
            info.n_top = h->bm->ntop;
            info.vector_count = h->bm->n;
            info.bytes_per_vec = h->bm->dim / 8;
        } else {
            info.kind = ORDVEC_INDEX_KIND_RANK_QUANT;
            info.dim = h->rq->dim;
            info.bit_width = h->rq->bits;
            info.vector_count = h->rq->n;
            info.bytes_per_vec = h->rq->bpv;
        }
        info.source_file_size_bytes = h->source_file_size_bytes;
        info.capabilities = kAllCaps;
        *info_out = info;
    });
}

void ordvec_index_free(ordvec_index_t* index) {
    if (!index) return;
    LOCK;
    delete H(index);
}

ordvec_status_t ordvec_index_search(const ordvec_index_t* index, const ordvec_search_params_t* params,
                                    ordvec_hit_t* hits_out, uint64_t hits_capacity, uint64_t* returned_out,
                                    ordvec_search_stats_t* stats_out) {
    return boundary([&] {
        const uint64_t t0 = now_ns();
        NONNULL(index, "index");
        NONNULL(params, "params");
        const IndexHandle* h = H(index);
        const RqIndex* ix = h->rq;
        const SbIndex* bm = h->bm;
        const uint64_t idx_dim = bm ? bm->dim : ix->dim, idx_n = bm ? bm->n : ix->n;
        // Validation order: ordvec-ffi/src/lib.rs:444-568.
        check_exact_size(params->struct_size, sizeof(ordvec_search_params_t), "ordvec_search_params_t");
        if (params->flags != ORDVEC_SEARCH_FLAG_NONE) fail(kBadArgument, "unknown search flags: " + std::to_string(params->flags));
        for (uint64_t r : params->reserved)
            if (r) fail(kBadArgument, "ordvec_search_params_t.reserved fields must be zero");
        NONNULL(returned_out, "returned_out");
        if (stats_out) check_exact_size(stats_out->struct_size, sizeof(ordvec_search_stats_t), "ordvec_search_stats_t");
        NONNULL(params->query, "query");
        if (params->dim != idx_dim)
            fail(kDimMismatch, "query dim " + std::to_string(params->dim) + " does not match index dim " + std::to_string(idx_dim));
        if (!host_all_finite(params->query, idx_dim)) fail(kNonfiniteQuery, "query contains NaN or infinity");
        const uint32_t* rows = nullptr;
        uint64_t count = params->candidate_count;
        if (count == 0 && params->candidate_rows) fail(kBadArgument, "candidate_count is zero but candidate_rows is non-NULL");
        if (count != 0 && !params->candidate_rows) fail(kNullPointer, "candidate_rows pointer is NULL but candidate_count is nonzero");
        if (count) {
            rows = params->candidate_rows;
            for (uint64_t i = 0; i < count; i++)
                if (rows[i] >= idx_n)
                    fail(kRowIdOutOfRange, "candidate row " + std::to_string(rows[i]) + " is out of range for " +
                                               std::to_string(idx_n) + " vectors");
        }
        const uint64_t space = rows ? count : idx_n;
        const uint64_t required = std::min<uint64_t>(params->k, space);
        if (required > 0) {
            NONNULL(hits_out, "hits_out");
            if (hits_capacity < required)
                fail(kBufferTooSmall, "hits_capacity " + std::to_string(hits_capacity) + " is smaller than required " +
                                          std::to_string(required));
        }
        uint64_t t_prep = now_ns(), t_score = t_prep, t_sel = t_prep;
        if (required > 0) {
            // search_asymmetric_subset asserts rows <= n_vectors (src/quant.rs:1057-1063).
            if (!bm && rows && count > ix->n)
                throw std::runtime_error("search_asymmetric_subset: candidate list length " + std::to_string(count) +
                                         " exceeds n_vectors " + std::to_string(ix->n) + "; deduplicate repeated ids before calling");
            std::vector<float> sc(required);
            std::vector<int64_t> id(required);
            {
                LOCK;
                eng().init();
                t_prep = now_ns();
                uint64_t off[2] = {0, count};
                if (bm && rows) {
                    // Bitmap::search_subset (ordvec-ffi/src/lib.rs:978-980).
                    bm_subset(bm, params->query, 1, off, rows, count, required, sc.data(), id.data(), 0);
                } else if (bm) {
                    // Bitmap::search (ordvec-ffi/src/lib.rs:974-977).
                    bm_search(bm, params->query, 1, required, sc.data(), id.data(), 0);
                } else if (rows) {
                    rq_subset_search(ix, params->query, 1, off, rows, count, 0, required, sc.data(), id.data(), 0);
                } else {
                    rq_full_search(ix, params->query, 1, required, sc.data(), id.data(), 0);
                }
                t_score = now_ns();
            }
            for (uint64_t j = 0; j < required; j++) {
                if (id[j] < 0) fail(kInternal, "engine returned a negative row id");
                hits_out[j].row_id = (uint64_t)id[j];
                hits_out[j].id = (uint64_t)id[j];
                hits_out[j].score = sc[j];
                hits_out[j].reserved = 0;
            }
            t_sel = now_ns();
        }
        *returned_out = required;
        if (stats_out) {
            ordvec_search_stats_t st = default_stats();
            st.kind = bm ? ORDVEC_INDEX_KIND_BITMAP : ORDVEC_INDEX_KIND_RANK_QUANT;
            st.dim = idx_dim;
            st.bit_width = bm ? 0 : ix->bits;
            st.n_top = bm ? bm->ntop : 0;
            st.k = params->k;
            st.user_tag = params->user_tag;
            st.vector_count = idx_n;
            st.candidate_count = space;
            st.returned_count = required;
            st.total_ns = now_ns() - t0;
            st.prepare_ns = t_prep - t0;
            st.score_ns = t_score - t_prep;
            st.select_ns = t_sel - t_score;
            st.vectors_scored = space;
            st.bytes_read = required ? space * (bm ? bm->dim / 8 : ix->bpv) + idx_dim * 4 : 0;
            *stats_out = st;
        }
    });
}

// ---- Part 2 -----------------------------------------------------------------

ordvec_status_t ordvec_set_stream(void* hip_stream) {
    return boundary([&] {
        LOCK;
        eng().init();
        if (cur().stream == static_cast<hipStream_t>(hip_stream)) return;
        OVX_HIP(hipStreamSynchronize(cur().stream));
        cur().stream = static_cast<hipStream_t>(hip_stream);
    });
}

ordvec_status_t ordvec_synchronize(void) {
    return boundary([&] {
        LOCK;
        eng().init();
        sync();
    });
}

ordvec_status_t ordvec_stage1_exact_queries(uint64_t* count) {
    return boundary([&] {
        NONNULL(count, "count");
        LOCK;
        *count = cur().exact_queries;
    });
}

ordvec_status_t ordvec_kernel_timing_enable_families(uint32_t mask) {
    return boundary([&] {
        LOCK;
        eng().init();
        KTimes& t = kt();
        t.resolve();
        for (int f = 0; f < kFamCount; f++) {
            t.launches[f] = 0;
            t.total_ms[f] = 0.0;
            t.open[f] = nullptr;
        }
        t.mask = mask & ((1u << kFamCount) - 1u);
    });
}

ordvec_status_t ordvec_kernel_timing_enable(int enable) {
    return ordvec_kernel_timing_enable_families(enable ? ~0u : 0u);
}

ordvec_status_t ordvec_engine_option(uint32_t option, int64_t value, int64_t* previous) {
    return boundary([&] {
        if (value < -1 || value > 1) fail(kBadArgument, "option value must be -1 (read), 0 or 1");
        int64_t prev = 0;
        if (option == ORDVEC_OPTION_ONE_QUERY_PATH) {
            prev = g_plain_path.load() ? 1 : 0;
            if (value >= 0) g_plain_path.store(value != 0);
        } else if (option == ORDVEC_OPTION_MFMA_SCAN) {
            prev = set_mfma_scan((int)value) ? 1 : 0;
        } else if (option == ORDVEC_OPTION_BATCH_TAIL) {
            prev = g_batch_tail.load() ? 1 : 0;
            if (value >= 0) g_batch_tail.store(value != 0);
        } else if (option == ORDVEC_OPTION_SIDE_PREP) {
            prev = g_side_prep.load() ? 1 : 0;
            if (value >= 0) g_side_prep.store(value != 0);
        } else {
            fail(kBadArgument, "unknown engine option " + std::to_string(option));
        }
        if (previous) *previous = prev;
    });
}

ordvec_status_t ordvec_kernel_timing_read(uint32_t family, uint64_t* launches, double* total_ms) {
    return boundary([&] {
        if (family >= (uint32_t)kFamCount) fail(kBadArgument, "unknown kernel family " + std::to_string(family));
        LOCK;
        KTimes& t = kt();
        t.resolve();
        if (launches) *launches = t.launches[family];
        if (total_ms) *total_ms = t.total_ms[family];
    });
}

ordvec_status_t ordvec_rankquant_new(uint64_t dim, uint32_t bits, ordvec_rankquant_t** out) {
    return boundary([&] {
        NONNULL(out, "out");
        *out = nullptr;
        validate_rq_params(dim, bits);
        LOCK;
        eng().init();
        RqIndex* ix = new RqIndex();
        ix->dim = (uint32_t)dim;
        ix->bits = bits;
        ix->bpv = (uint32_t)(dim * bits / 8);
        *out = reinterpret_cast<ordvec_rankquant_t*>(ix);
    });
}

ordvec_status_t ordvec_rankquant_add(ordvec_rankquant_t* index, const float* vectors, uint64_t n, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        if (n == 0) return;
        NONNULL(vectors, "vectors");
        LOCK;
        check_not_serving(*RQ(index), "add");
        rq_add(RQ(index), vectors, n, flags);
    });
}

ordvec_status_t ordvec_rankquant_info(const ordvec_rankquant_t* index, uint64_t* dim, uint32_t* bits, uint64_t* n) {
    return boundary([&] {
        NONNULL(index, "index");
        if (dim) *dim = RQ(index)->dim;
        if (bits) *bits = RQ(index)->bits;
        if (n) *n = RQ(index)->n;
    });
}

ordvec_status_t ordvec_rankquant_search_asymmetric(const ordvec_rankquant_t* index, const float* queries, uint64_t nq,
                                                   uint64_t k, float* out_scores, int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const RqIndex* ix = RQ(index);
        if (nq == 0 || std::min<uint64_t>(k, ix->n) == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        LOCK;

        rq_full_search(ix, queries, nq, k, out_scores, out_ids, flags);
    });
}

ordvec_status_t ordvec_rankquant_search_subset(const ordvec_rankquant_t* index, const float* query, const uint32_t* rows,
                                               uint64_t n_rows, uint64_t k, float* out_scores, int64_t* out_ids,
                                               uint64_t* returned, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const RqIndex* ix = RQ(index);
        NONNULL(query, "query");
        if (n_rows > ix->n)
            fail(kBadArgument, "search_asymmetric_subset: candidate list length " + std::to_string(n_rows) +
                                   " exceeds n_vectors " + std::to_string(ix->n) + "; deduplicate repeated ids before calling");
        const uint64_t out_k = std::min(k, n_rows);
        if (returned) *returned = out_k;
        if (out_k == 0) return;
        NONNULL(rows, "rows");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        LOCK;

        if (!(flags & ORDVEC_FLAG_DEVICE_PTRS))
            for (uint64_t i = 0; i < n_rows; i++)
                if (rows[i] >= ix->n) fail(kRowIdOutOfRange, "search_asymmetric_subset: candidate id out of range");
        uint64_t off[2] = {0, n_rows};
        if (flags & ORDVEC_FLAG_DEVICE_PTRS) {
            rq_subset_search(ix, query, 1, nullptr, rows, n_rows, n_rows, out_k, out_scores, out_ids, flags);
        } else {
            rq_subset_search(ix, query, 1, off, rows, n_rows, 0, out_k, out_scores, out_ids, flags);
        }
    });
}

ordvec_status_t ordvec_rankquant_search_subset_batched(const ordvec_rankquant_t* index, const float* queries,
                                                       uint64_t nq, const uint64_t* offsets, const uint32_t* candidates,
                                                       uint64_t n_candidates, uint64_t k, float* out_scores,
                                                       int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const RqIndex* ix = RQ(index);
        NONNULL(offsets, "offsets");
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        if (!dev) {
            // validate_csr_batch (src/quant.rs:1088-1119)
            if (offsets[0] != 0) fail(kBadArgument, "candidate_offsets[0] must be 0");
            if (offsets[nq] != n_candidates) fail(kBadArgument, "candidate_offsets[nq] must equal candidates.len()");
            for (uint64_t q = 0; q < nq; q++) {
                if (offsets[q + 1] < offsets[q]) fail(kBadArgument, "candidate_offsets must be monotonic non-decreasing");
                if (offsets[q + 1] - offsets[q] > ix->n)
                    fail(kBadArgument, "per-row candidate count exceeds n_vectors " + std::to_string(ix->n));
            }
            for (uint64_t i = 0; i < n_candidates; i++)
                if (candidates[i] >= ix->n)
                    fail(kRowIdOutOfRange, "candidate id out of range (n_vectors " + std::to_string(ix->n) + ")");
        }
        const uint64_t out_k = std::min<uint64_t>(k, ix->n);
        if (out_k == 0 || nq == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        if (n_candidates) NONNULL(candidates, "candidates");
        LOCK;

        rq_subset_search(ix, queries, nq, offsets, candidates, n_candidates, 0, out_k, out_scores, out_ids, flags);
    });
}

// Shared checks of the score-by-id pair; ids in [-1, n) (-1 = no row).
void check_score_ids(const RqIndex* ix, const int64_t* ids, uint64_t nq, uint64_t k, uint64_t flags) {
    check_search_dim(ix->dim);
    if (ix->n == 0) fail(kBadArgument, "score_ids needs a non-empty index");
    if (k > kStage1MaxM) fail(kUnsupportedOperation, "score_ids supports k <= " + std::to_string(kStage1MaxM));
    if (!(flags & ORDVEC_FLAG_DEVICE_PTRS))
        for (uint64_t i = 0; i < nq * k; i++)
            if (ids[i] < -1 || ids[i] >= (int64_t)ix->n)
                fail(kRowIdOutOfRange, "id out of range (n_vectors " + std::to_string(ix->n) + ", -1 = no row)");
}

ordvec_status_t ordvec_rankquant_score_ids(const ordvec_rankquant_t* index, const float* queries, uint64_t nq,
                                           const int64_t* ids, uint64_t k, float* out_scores, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const RqIndex* ix = RQ(index);
        if (nq == 0 || k == 0) return;
        NONNULL(queries, "queries");
        NONNULL(ids, "ids");
        NONNULL(out_scores, "out_scores");
        check_score_ids(ix, ids, nq, k, flags);
        LOCK;
        eng().init();
        check_finite(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        const uint32_t d = ix->dim;
        size_t bytes = (dev ? 0 : al(nq * d * 4) + al(nq * k * 8) + al(nq * k * 4)) + al(nq * d * 4) + al(nq * 4) + 256;
        Carver cv{static_cast<char*>(cur().workspace(bytes))};
        const float* dq = stage_queries(cv, queries, nq * d, flags);
        const int64_t* did = ids;
        if (!dev) {
            int64_t* t = cv.take<int64_t>(nq * k);
            h2d(t, ids, nq * k * 8);
            did = t;
        }
        float* ds = dev ? out_scores : cv.take<float>(nq * k);
        float* qhat = cv.take<float>(nq * d);
        float* off = cv.take<float>(nq);
        uint32_t* bad = cv.take<uint32_t>(1);
        hipStream_t s = cur().stream;
        OVX_HIP(hipMemsetAsync(bad, 0, 4, s));
        const int tier = select_tier(d, ix->bits);
        const float inv_norm = rankquant_inv_norm(d, ix->bits);
        launch_prepare_queries(dq, nq, d, ix->bits, tier, inv_norm, qhat, off, nullptr, s);
        launch_score_ids(ix->codes, ix->n, d, ix->bits, tier, inv_norm, qhat, off, nq, did, (uint32_t)k, ds, bad, s);
        if (!dev) d2h(out_scores, ds, nq * k * 4);
        if (read_qstat(bad, 1)[0]) fail(kRowIdOutOfRange, "id out of range (n_vectors " + std::to_string(ix->n) + ")");
    });
}

ordvec_status_t ordvec_rankquant_score_ids_backward(const ordvec_rankquant_t* index, const float* queries, uint64_t nq,
                                                    const int64_t* ids, uint64_t k, const float* grad_scores,
                                                    float* grad_queries, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const RqIndex* ix = RQ(index);
        if (nq == 0) return;
        NONNULL(queries, "queries");
        NONNULL(grad_queries, "grad_queries");
        if (k) {
            NONNULL(ids, "ids");
            NONNULL(grad_scores, "grad_scores");
        }
        check_score_ids(ix, ids, nq, k, flags);
        LOCK;
        eng().init();
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        const uint32_t d = ix->dim;
        size_t bytes = (dev ? 0 : 2 * al(nq * d * 4) + al(nq * k * 8) + al(nq * k * 4)) + 256;
        Carver cv{static_cast<char*>(cur().workspace(bytes))};
        const float* dq = stage_queries(cv, queries, nq * d, flags);
        const int64_t* did = ids;
        const float* dg = grad_scores;
        if (!dev && k) {
            int64_t* t = cv.take<int64_t>(nq * k);
            h2d(t, ids, nq * k * 8);
            did = t;
            float* g = cv.take<float>(nq * k);
            h2d(g, grad_scores, nq * k * 4);
            dg = g;
        }
        float* dx = dev ? grad_queries : cv.take<float>(nq * d);
        uint32_t* bad = cv.take<uint32_t>(1);
        hipStream_t s = cur().stream;
        OVX_HIP(hipMemsetAsync(bad, 0, 4, s));
        launch_score_ids_backward(ix->codes, ix->n, d, ix->bits, rankquant_inv_norm(d, ix->bits), dq, nq, did,
                                  (uint32_t)k, dg, dx, bad, s);
        if (!dev) d2h(grad_queries, dx, nq * d * 4);
        if (read_qstat(bad, 1)[0]) fail(kRowIdOutOfRange, "id out of range (n_vectors " + std::to_string(ix->n) + ")");
    });
}

ordvec_status_t ordvec_rankquant_codes(const ordvec_rankquant_t* index, uint8_t* out, uint64_t capacity, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const RqIndex* ix = RQ(index);
        uint64_t bytes = ix->n * ix->bpv;
        if (capacity < bytes) fail(kBufferTooSmall, "codes buffer too small");
        if (!bytes) return;
        NONNULL(out, "out");
        LOCK;
        if (flags & ORDVEC_FLAG_DEVICE_PTRS) {
            OVX_HIP(hipMemcpyAsync(out, ix->codes, bytes, hipMemcpyDeviceToDevice, cur().stream));
        } else {
            d2h(out, ix->codes, bytes);
            sync();
        }
    });
}

ordvec_status_t ordvec_rankquant_write(const ordvec_rankquant_t* index, const char* path) {
    return boundary([&] {
        NONNULL(index, "index");
        NONNULL(path, "path");
        LOCK;
        write_ovrq(RQ(index), path);
    });
}

ordvec_status_t ordvec_rankquant_load(const char* path, ordvec_rankquant_t** out) {
    return boundary([&] {
        NONNULL(path, "path");
        NONNULL(out, "out");
        *out = nullptr;
        LOCK;  // the loader checks the file before any device set-up
        *out = reinterpret_cast<ordvec_rankquant_t*>(load_ovrq(path));
    });
}

void ordvec_rankquant_free(ordvec_rankquant_t* index) {
    if (!index) return;
    LOCK;
    if (RQ(index)->serving) RQ(index)->free_pending = true;
    else delete RQ(index);
}

ordvec_status_t ordvec_rankquant_search(const ordvec_rankquant_t* index, const float* queries, uint64_t nq, uint64_t k,
                                       float* out_scores, int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const RqIndex* ix = RQ(index);
        if (nq == 0 || std::min<uint64_t>(k, ix->n) == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        LOCK;
        rq_symmetric_search(ix, queries, nq, k, out_scores, out_ids, flags);
    });
}

ordvec_status_t ordvec_rankquant_swap_remove(ordvec_rankquant_t* index, uint64_t idx, uint64_t* moved_from) {
    return boundary([&] {
        NONNULL(index, "index");
        LOCK;
        RqIndex* ix = RQ(index);
        check_not_serving(*ix, "swap_remove");
        uint64_t last = swap_remove_rows(ix->codes, ix->n, ix->bpv, idx);
        if (moved_from) *moved_from = last;
    });
}

// ---------------------------------------------------------------------------
// Rank (src/rank.rs:291-584)
// ---------------------------------------------------------------------------
ordvec_status_t ordvec_rank_new(uint64_t dim, ordvec_rank_t** out) {
    return boundary([&] {
        NONNULL(out, "out");
        *out = nullptr;
        if (dim < 2) fail(kBadArgument, "dim must be >= 2");
        if (dim > kMaxDim) fail(kBadArgument, "dim must fit in u16");
        LOCK;
        eng().init();
        RkIndex* ix = new RkIndex();
        ix->dim = (uint32_t)dim;
        *out = reinterpret_cast<ordvec_rank_t*>(ix);
    });
}

ordvec_status_t ordvec_rank_add(ordvec_rank_t* index, const float* vectors, uint64_t n, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        if (n == 0) return;
        NONNULL(vectors, "vectors");
        LOCK;
        rank_add(RK(index), vectors, n, flags);
    });
}

ordvec_status_t ordvec_rank_info(const ordvec_rank_t* index, uint64_t* dim, uint64_t* n) {
    return boundary([&] {
        NONNULL(index, "index");
        if (dim) *dim = RK(index)->dim;
        if (n) *n = RK(index)->n;
    });
}

static ordvec_status_t rank_search_api(const ordvec_rank_t* index, const float* queries, uint64_t nq, uint64_t k,
                                       float* out_scores, int64_t* out_ids, uint64_t flags, bool symmetric) {
    return boundary([&] {
        NONNULL(index, "index");
        const RkIndex* ix = RK(index);
        if (nq == 0 || std::min<uint64_t>(k, ix->n) == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        LOCK;
        rank_search(ix, queries, nq, k, symmetric, out_scores, out_ids, flags);
    });
}

ordvec_status_t ordvec_rank_search(const ordvec_rank_t* index, const float* queries, uint64_t nq, uint64_t k,
                                   float* out_scores, int64_t* out_ids, uint64_t flags) {
    return rank_search_api(index, queries, nq, k, out_scores, out_ids, flags, true);
}

ordvec_status_t ordvec_rank_search_asymmetric(const ordvec_rank_t* index, const float* queries, uint64_t nq,
                                              uint64_t k, float* out_scores, int64_t* out_ids, uint64_t flags) {
    return rank_search_api(index, queries, nq, k, out_scores, out_ids, flags, false);
}

ordvec_status_t ordvec_rank_ranks(const ordvec_rank_t* index, uint16_t* out, uint64_t capacity, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const RkIndex* ix = RK(index);
        const uint64_t count = ix->n * ix->dim;
        if (capacity < count) fail(kBufferTooSmall, "ranks buffer too small");
        if (!count) return;
        NONNULL(out, "out");
        LOCK;
        if (flags & ORDVEC_FLAG_DEVICE_PTRS) {
            d2d(out, ix->ranks, count * 2);
        } else {
            d2h(out, ix->ranks, count * 2);
            sync();
        }
    });
}

ordvec_status_t ordvec_rank_swap_remove(ordvec_rank_t* index, uint64_t idx, uint64_t* moved_from) {
    return boundary([&] {
        NONNULL(index, "index");
        LOCK;
        RkIndex* ix = RK(index);
        uint64_t last = swap_remove_rows(ix->ranks, ix->n, ix->row_bytes(), idx);
        if (moved_from) *moved_from = last;
    });
}

ordvec_status_t ordvec_rank_write(const ordvec_rank_t* index, const char* path) {
    return boundary([&] {
        NONNULL(index, "index");
        NONNULL(path, "path");
        LOCK;
        write_ovr(RK(index), path);
    });
}

ordvec_status_t ordvec_rank_load(const char* path, ordvec_rank_t** out) {
    return boundary([&] {
        NONNULL(path, "path");
        NONNULL(out, "out");
        *out = nullptr;
        LOCK;
        *out = reinterpret_cast<ordvec_rank_t*>(load_ovr(path));
    });
}

void ordvec_rank_free(ordvec_rank_t* index) {
    if (!index) return;
    LOCK;
    delete RK(index);
}

ordvec_status_t ordvec_sign_bitmap_new(uint64_t dim, ordvec_sign_bitmap_t** out) {
    return boundary([&] {
        NONNULL(out, "out");
        *out = nullptr;
        validate_sign_dim(dim);
        LOCK;
        eng().init();
        SbIndex* ix = new SbIndex();
        ix->dim = (uint32_t)dim;
        ix->qpv = (uint32_t)(dim / 64);
        *out = reinterpret_cast<ordvec_sign_bitmap_t*>(ix);
    });
}

ordvec_status_t ordvec_sign_bitmap_add(ordvec_sign_bitmap_t* index, const float* vectors, uint64_t n, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        if (n == 0) return;
        NONNULL(vectors, "vectors");
        LOCK;
        check_not_serving(*SB(index), "add");
        sb_add(SB(index), vectors, n, flags);
    });
}

ordvec_status_t ordvec_two_stage_add(ordvec_sign_bitmap_t* sign, ordvec_rankquant_t* rq, const float* vectors,
                                     uint64_t n, uint64_t flags) {
    return boundary([&] {
        NONNULL(sign, "sign");
        NONNULL(rq, "rq");
        SbIndex* sb = SB(sign);
        RqIndex* ix = RQ(rq);
        if (sb->dim != ix->dim) fail(kBadArgument, "sign_probe.dim must match RankQuant dim " + std::to_string(ix->dim));
        if (n == 0) return;
        NONNULL(vectors, "vectors");
        LOCK;
        check_not_serving(*sb, "add");
        check_not_serving(*ix, "add");
        if (one_pass_ok(ix->dim, ix->bits, true, true, vectors, flags)) {
            eng().init();
            add_one_pass(ix, sb, vectors, n, flags);
            return;
        }
        // no one-pass form for this (dim, bits): the two adds, vectors checked once
        check_finite(vectors, n * ix->dim, flags, kBadArgument, "vectors");
        check_new_count(ix->n, n);
        check_new_count(sb->n, n);
        sb_add(sb, vectors, n, flags | ORDVEC_FLAG_ASSUME_FINITE);
        rq_add(ix, vectors, n, flags | ORDVEC_FLAG_ASSUME_FINITE);
    });
}

ordvec_status_t ordvec_sign_bitmap_info(const ordvec_sign_bitmap_t* index, uint64_t* dim, uint64_t* n) {
    return boundary([&] {
        NONNULL(index, "index");
        if (dim) *dim = SB(index)->dim;
        if (n) *n = SB(index)->n;
    });
}

ordvec_status_t ordvec_sign_bitmap_top_m(const ordvec_sign_bitmap_t* index, const float* queries, uint64_t nq,
                                         uint64_t m, uint32_t* out_candidates, uint32_t* out_hamming, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = SB(index);
        const uint64_t m_eff = std::min<uint64_t>(m, ix->n);
        if (nq == 0 || m_eff == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_candidates, "out_candidates");
        LOCK;
        sb_top_m(ix, queries, nq, m_eff, out_candidates, out_hamming, flags);
    });
}

ordvec_status_t ordvec_rank_transform(const float* vectors, uint64_t n, uint64_t dim, uint16_t* out_ranks,
                                      uint64_t flags) {
    return boundary([&] {
        if (dim > kMaxDim) fail(kBadArgument, "length " + std::to_string(dim) + " exceeds u16::MAX (65535)");
        if (n == 0 || dim == 0) return;
        NONNULL(vectors, "vectors");
        NONNULL(out_ranks, "out_ranks");
        check_finite(vectors, n * dim, flags, kBadArgument, "vectors");
        LOCK;
        eng().init();
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        const uint32_t d = (uint32_t)dim;
        if (d == 1) {  // a single value has rank 0
            if (dev) OVX_HIP(hipMemsetAsync(out_ranks, 0, n * 2, cur().stream));
            else std::memset(out_ranks, 0, n * 2);
            return;
        }
        const uint64_t rows = dev ? n : std::max<uint64_t>(1, (256ull << 20) / (4ull * d));
        for (uint64_t r0 = 0; r0 < n; r0 += rows) {
            const uint64_t c = std::min(rows, n - r0);
            Carver cv{static_cast<char*>(cur().workspace(dev ? 256 : al(c * d * 4) + al(c * d * 2)))};
            const float* dv = stage_queries(cv, vectors + r0 * d, c * d, flags);
            uint16_t* rk = dev ? out_ranks + r0 * d : cv.take<uint16_t>(c * d);
            rank_rows(dv, c, d, rk, 0, nullptr);
            if (!dev) {
                d2h(out_ranks + r0 * d, rk, c * d * 2);
                sync();
            }
        }
    });
}

ordvec_status_t ordvec_rankquant_eval_search(const float* corpus, uint64_t n, const float* queries, uint64_t nq,
                                             uint64_t dim, uint32_t bits, uint64_t k, float* out_scores,
                                             int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        // src/quant.rs:96-117,1541-1575 parameter checks
        if (bits < 1 || bits > 8) fail(kBadArgument, "bits must be in 1..=8");
        if (dim < 2) fail(kBadArgument, "dim must be >= 2");
        if (dim > kMaxDim) fail(kBadArgument, "dim must fit in u16");
        if (dim > 8192) fail(kUnsupportedOperation, "GPU eval search stages the query in LDS and supports dim <= 8192");
        const uint64_t kk = std::min<uint64_t>(k, n);
        if (nq == 0 || kk == 0) return;
        NONNULL(corpus, "corpus");
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        check_finite(corpus, n * dim, flags, kBadArgument, "corpus");
        check_finite(queries, nq * dim, flags, kBadArgument, "queries");
        LOCK;
        eng().init();
        const uint32_t d = (uint32_t)dim;
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        hipStream_t s = cur().stream;
        // norm: sqrt of sum_rank centre(bucket(rank))^2 in f64 (src/quant.rs:106-117)
        double acc = 0.0;
        const double half = (double)((1u << bits) - 1) / 2.0;
        for (uint64_t r = 0; r < d; r++) {
            const double c = (double)(float)((double)((r << bits) / d) - half);
            acc += c * c;
        }
        const float norm = (float)std::sqrt(acc);
        const float inv_norm_sq = 1.0f / (norm * norm);
        // corpus buckets, row chunks through the rank transform
        DevBuf buckets;
        struct FreeOnExit {  // the bucket matrix lives for this call only, also on the error paths
            DevBuf& b;
            hipStream_t s;
            ~FreeOnExit() {
                if (b.p) {
                    (void)hipStreamSynchronize(s);
                    (void)hipFree(b.p);
                }
            }
        } release{buckets, s};
        uint8_t* db = static_cast<uint8_t*>(buckets.get(std::max<uint64_t>(256, n * d), s));
        const uint64_t rows = std::max<uint64_t>(1, (256ull << 20) / (4ull * d));
        for (uint64_t r0 = 0; r0 < n; r0 += rows) {
            const uint64_t c = std::min(rows, n - r0);
            Carver cv{static_cast<char*>(cur().workspace(al(c * d * 2) + (dev ? 0 : al(c * d * 4))))};
            const float* dv = stage_queries(cv, corpus + r0 * d, c * d, flags);
            uint16_t* rk = cv.take<uint16_t>(c * d);
            rank_rows(dv, c, d, rk, 0, nullptr);
            launch_eval_buckets(rk, c * d, d, bits, db + r0 * d, s);
            sync();
        }
        // queries in chunks whose score keys stay within 1 GiB
        const uint64_t qc = std::max<uint64_t>(1, std::min<uint64_t>({nq, 65535, (1ull << 30) / (16 * n + 1)}));
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint64_t cn = std::min(qc, nq - q0);
            const size_t wb = eval_search_work_bytes(n, cn, d, s);
            Carver cv{static_cast<char*>(cur().workspace(wb + (dev ? 0 : al(cn * d * 4) + al(cn * kk * 12))))};
            char* work = cv.take<char>(wb);
            const float* dq = stage_queries(cv, queries + q0 * d, cn * d, flags);
            float* ds = dev ? out_scores + q0 * kk : cv.take<float>(cn * kk);
            int64_t* di = dev ? out_ids + q0 * kk : cv.take<int64_t>(cn * kk);
            uint16_t* qr = reinterpret_cast<uint16_t*>(work);
            rank_rows(dq, cn, d, qr, 0, nullptr);
            launch_eval_search(db, n, d, bits, inv_norm_sq, qr, cn, (uint32_t)kk, ds, di, work, s);
            if (!dev) {
                d2h(out_scores + q0 * kk, ds, cn * kk * 4);
                d2h(out_ids + q0 * kk, di, cn * kk * 8);
            }
            sync();
        }
        sync();
    });
}

ordvec_status_t ordvec_sign_bitmap_query_bits(const ordvec_sign_bitmap_t* index, const float* queries, uint64_t nq,
                                              uint64_t* out_words, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = SB(index);
        if (nq == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_words, "out_words");
        host_finite_or_fail(queries, nq * ix->dim, flags, kBadArgument, "query");
        LOCK;
        eng().init();
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        const uint64_t qpv = ix->qpv;
        Carver cv{static_cast<char*>(cur().workspace(dev ? 256 : al(nq * ix->dim * 4) + al(nq * qpv * 8)))};
        const float* dq = stage_queries(cv, queries, nq * ix->dim, flags);
        uint64_t* dw = dev ? out_words : cv.take<uint64_t>(nq * qpv);
        if (dev && !(flags & ORDVEC_FLAG_ASSUME_FINITE)) check_finite(queries, nq * ix->dim, flags, kBadArgument, "query");
        launch_sign_encode(dq, nq, ix->dim, dw, 0, false, cur().stream);
        if (!dev) {
            d2h(out_words, dw, nq * qpv * 8);
            sync();
        }
    });
}

ordvec_status_t ordvec_sign_bitmap_score_all(const ordvec_sign_bitmap_t* index, const float* queries, uint64_t nq,
                                             uint32_t* out_scores, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = SB(index);
        if (nq == 0 || ix->n == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        if (ix->ntop) fail(kBadArgument, "score_all is a SignBitmap method");
        host_finite_or_fail(queries, nq * ix->dim, flags, kBadArgument, "query");
        LOCK;
        eng().init();
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        if (dev && !(flags & ORDVEC_FLAG_ASSUME_FINITE)) check_finite(queries, nq * ix->dim, flags, kBadArgument, "query");
        const uint64_t qpv = ix->qpv;
        // host outputs go back in query chunks whose scores fit 1 GiB of workspace
        const uint64_t qc = dev ? nq : std::max<uint64_t>(1, std::min<uint64_t>(nq, (1ull << 30) / (ix->n * 4)));
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint64_t cn = std::min(qc, nq - q0);
            Carver cv{static_cast<char*>(cur().workspace(al(cn * qpv * 8) + (dev ? 0 : al(cn * ix->dim * 4) +
                                                                                        al(cn * ix->n * 4))))};
            const float* dq = stage_queries(cv, queries + q0 * ix->dim, cn * ix->dim, flags);
            uint64_t* qb = cv.take<uint64_t>(cn * qpv);
            uint32_t* ds = dev ? out_scores + q0 * ix->n : cv.take<uint32_t>(cn * ix->n);
            launch_sign_encode(dq, cn, ix->dim, qb, 0, false, cur().stream);
            launch_sign_score_all(ix->words, ix->n, ix->dim, qb, cn, ds, cur().stream);
            if (!dev) {
                d2h(out_scores + q0 * ix->n, ds, cn * ix->n * 4);
                sync();
            }
        }
    });
}

ordvec_status_t ordvec_sign_bitmap_words(const ordvec_sign_bitmap_t* index, uint64_t* out, uint64_t capacity,
                                         uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = SB(index);
        uint64_t cnt = ix->n * ix->qpv;
        if (capacity < cnt) fail(kBufferTooSmall, "words buffer too small");
        if (!cnt) return;
        NONNULL(out, "out");
        if (flags & ORDVEC_FLAG_DEVICE_PTRS) fail(kUnsupportedOperation, "row-major export is host-only");
        LOCK;
        std::vector<uint64_t> rows = sb_download_rowmajor(ix);
        std::memcpy(out, rows.data(), cnt * 8);
    });
}

ordvec_status_t ordvec_sign_bitmap_write(const ordvec_sign_bitmap_t* index, const char* path) {
    return boundary([&] {
        NONNULL(index, "index");
        NONNULL(path, "path");
        LOCK;
        write_ovsb(SB(index), path);
    });
}

ordvec_status_t ordvec_sign_bitmap_load(const char* path, ordvec_sign_bitmap_t** out) {
    return boundary([&] {
        NONNULL(path, "path");
        NONNULL(out, "out");
        *out = nullptr;
        LOCK;  // the loader checks the file before any device set-up
        *out = reinterpret_cast<ordvec_sign_bitmap_t*>(load_ovsb(path));
    });
}

// src/sign_bitmap.rs:578-590 swap_remove on the interleaved layout; the
// vacated slot is cleared so padding stays zero.
static uint64_t sb_swap_remove(SbIndex* ix, uint64_t idx) {
    if (idx >= ix->n)
        fail(kBadArgument, "index " + std::to_string(idx) + " out of range (index holds " + std::to_string(ix->n) +
                               " vectors)");
    const uint64_t last = ix->n - 1;
    launch_move_doc_interleaved(ix->words, ix->qpv, last, idx, cur().stream);
    sync();
    ix->n = last;
    return last;
}

ordvec_status_t ordvec_sign_bitmap_swap_remove(ordvec_sign_bitmap_t* index, uint64_t idx, uint64_t* moved_from) {
    return boundary([&] {
        NONNULL(index, "index");
        LOCK;
        check_not_serving(*SB(index), "swap_remove");
        uint64_t last = sb_swap_remove(SB(index), idx);
        if (moved_from) *moved_from = last;
    });
}

void ordvec_sign_bitmap_free(ordvec_sign_bitmap_t* index) {
    if (!index) return;
    LOCK;
    if (SB(index)->serving) SB(index)->free_pending = true;
    else delete SB(index);
}

// ---- Bitmap (top-bucket, AND-popcount) ------------------------------------------

static SbIndex* BM(const ordvec_bitmap_t* p) { return reinterpret_cast<SbIndex*>(const_cast<ordvec_bitmap_t*>(p)); }

ordvec_status_t ordvec_bitmap_new(uint64_t dim, uint64_t n_top, ordvec_bitmap_t** out) {
    return boundary([&] {
        NONNULL(out, "out");
        *out = nullptr;
        // Bitmap::validate_params (src/bitmap.rs:71-103)
        if (dim == 0) fail(kBadArgument, "dim must be > 0");
        if (dim % 64) fail(kBadArgument, "dim must be a multiple of 64");
        if (dim > kMaxDim) fail(kBadArgument, "dim must be <= 65535 (u16 rank invariant)");
        if (!(n_top > 0 && n_top < dim)) fail(kBadArgument, "n_top must satisfy 0 < n_top < dim");
        LOCK;
        eng().init();
        SbIndex* ix = new SbIndex();
        ix->dim = (uint32_t)dim;
        ix->qpv = (uint32_t)(dim / 64);
        ix->ntop = (uint32_t)n_top;
        *out = reinterpret_cast<ordvec_bitmap_t*>(ix);
    });
}

ordvec_status_t ordvec_bitmap_add(ordvec_bitmap_t* index, const float* vectors, uint64_t n, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        if (n == 0) return;
        NONNULL(vectors, "vectors");
        LOCK;
        bm_add(BM(index), vectors, n, flags);
    });
}

ordvec_status_t ordvec_bitmap_info(const ordvec_bitmap_t* index, uint64_t* dim, uint64_t* n_top, uint64_t* n) {
    return boundary([&] {
        NONNULL(index, "index");
        if (dim) *dim = BM(index)->dim;
        if (n_top) *n_top = BM(index)->ntop;
        if (n) *n = BM(index)->n;
    });
}

ordvec_status_t ordvec_bitmap_search(const ordvec_bitmap_t* index, const float* queries, uint64_t nq, uint64_t k,
                                     float* out_scores, int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = BM(index);
        const uint64_t k_eff = std::min<uint64_t>(k, ix->n);
        if (nq == 0 || k_eff == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        LOCK;
        bm_search(ix, queries, nq, k_eff, out_scores, out_ids, flags);
    });
}

ordvec_status_t ordvec_bitmap_top_m(const ordvec_bitmap_t* index, const float* queries, uint64_t nq, uint64_t m,
                                    uint32_t* out_candidates, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = BM(index);
        const uint64_t m_eff = std::min<uint64_t>(m, ix->n);
        if (nq == 0 || m_eff == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_candidates, "out_candidates");
        LOCK;
        sb_top_m(ix, queries, nq, m_eff, out_candidates, nullptr, flags);
    });
}

ordvec_status_t ordvec_bitmap_search_subset_batched(const ordvec_bitmap_t* index, const float* queries, uint64_t nq,
                                                    const uint64_t* offsets, const uint32_t* rows, uint64_t n_rows,
                                                    uint64_t k, float* out_scores, int64_t* out_ids,
                                                    uint64_t* out_counts, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = BM(index);
        NONNULL(offsets, "offsets");
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        if (dev) fail(kUnsupportedOperation, "Bitmap subset search takes host CSR arrays");
        if (offsets[0] != 0) fail(kBadArgument, "offsets[0] must be 0");
        if (offsets[nq] != n_rows) fail(kBadArgument, "offsets[nq] must equal the number of rows");
        for (uint64_t q = 0; q < nq; q++)
            if (offsets[q + 1] < offsets[q]) fail(kBadArgument, "offsets must be monotonic non-decreasing");
        for (uint64_t i = 0; i < n_rows; i++)
            if (rows[i] >= ix->n) fail(kRowIdOutOfRange, "doc_id out of range [0, " + std::to_string(ix->n) + ")");
        // Per query k_eff = min(k, row count); rows are padded to k with -inf / -1.
        if (out_counts)
            for (uint64_t q = 0; q < nq; q++) out_counts[q] = std::min<uint64_t>(k, offsets[q + 1] - offsets[q]);
        if (nq == 0 || k == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        if (n_rows) NONNULL(rows, "rows");
        LOCK;
        bm_subset(ix, queries, nq, offsets, rows, n_rows, k, out_scores, out_ids, flags);
    });
}

ordvec_status_t ordvec_bitmap_query_bits(const ordvec_bitmap_t* index, const float* queries, uint64_t nq,
                                         uint64_t* out_words, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = BM(index);
        if (nq == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_words, "out_words");
        check_bitmap_query_dim(ix);
        host_finite_or_fail(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
        LOCK;
        eng().init();
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        size_t bytes = al(nq * 4) + (dev ? 0 : al(nq * ix->dim * 4) + al(nq * ix->qpv * 8));
        Carver cv{static_cast<char*>(cur().workspace(bytes))};
        const float* dq = stage_queries(cv, queries, nq * ix->dim, flags);
        uint32_t* qstat = cv.take<uint32_t>(nq);
        uint64_t* qb = dev ? out_words : cv.take<uint64_t>(nq * ix->qpv);
        launch_bitmap_query_bits(dq, nq, ix->dim, ix->ntop, qb, qstat, cur().stream);
        if (!dev) d2h(out_words, qb, nq * ix->qpv * 8);
        finish_status(qstat, nq, flags, kNonfiniteQuery, "queries");
    });
}

ordvec_status_t ordvec_bitmap_overlap_subset(const ordvec_bitmap_t* index, const uint64_t* q_words,
                                             const uint32_t* rows, uint64_t n_rows, uint32_t* out) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = BM(index);
        if (n_rows == 0) return;
        NONNULL(q_words, "q_words");
        NONNULL(rows, "rows");
        NONNULL(out, "out");
        for (uint64_t i = 0; i < n_rows; i++)
            if (rows[i] >= ix->n)
                fail(kRowIdOutOfRange, "body_overlap_scores_subset: doc_id out of range [0, " + std::to_string(ix->n) + ")");
        LOCK;
        eng().init();
        Carver cv{static_cast<char*>(cur().workspace(al(ix->qpv * 8) + 2 * al(n_rows * 4)))};
        uint64_t* dq = cv.take<uint64_t>(ix->qpv);
        uint32_t* dr = cv.take<uint32_t>(n_rows);
        uint32_t* dout = cv.take<uint32_t>(n_rows);
        h2d(dq, q_words, ix->qpv * 8);
        h2d(dr, rows, n_rows * 4);
        launch_bm_overlap_rows(ix->words, ix->dim, dq, dr, n_rows, dout, cur().stream);
        d2h(out, dout, n_rows * 4);
        sync();
    });
}

ordvec_status_t ordvec_bitmap_words(const ordvec_bitmap_t* index, uint64_t* out, uint64_t capacity) {
    return boundary([&] {
        NONNULL(index, "index");
        const SbIndex* ix = BM(index);
        uint64_t cnt = ix->n * ix->qpv;
        if (capacity < cnt) fail(kBufferTooSmall, "words buffer too small");
        if (!cnt) return;
        NONNULL(out, "out");
        LOCK;
        std::vector<uint64_t> rows = sb_download_rowmajor(ix);
        std::memcpy(out, rows.data(), cnt * 8);
    });
}

ordvec_status_t ordvec_bitmap_write(const ordvec_bitmap_t* index, const char* path) {
    return boundary([&] {
        NONNULL(index, "index");
        NONNULL(path, "path");
        LOCK;
        write_ovbm(BM(index), path);
    });
}

ordvec_status_t ordvec_bitmap_load(const char* path, ordvec_bitmap_t** out) {
    return boundary([&] {
        NONNULL(path, "path");
        NONNULL(out, "out");
        *out = nullptr;
        LOCK;
        *out = reinterpret_cast<ordvec_bitmap_t*>(load_ovbm(path));
    });
}

ordvec_status_t ordvec_bitmap_swap_remove(ordvec_bitmap_t* index, uint64_t idx, uint64_t* moved_from) {
    return boundary([&] {
        NONNULL(index, "index");
        LOCK;
        uint64_t last = sb_swap_remove(BM(index), idx);  // src/bitmap.rs:515-527
        if (moved_from) *moved_from = last;
    });
}

void ordvec_bitmap_free(ordvec_bitmap_t* index) {
    if (!index) return;
    LOCK;
    delete BM(index);
}

ordvec_status_t ordvec_two_stage_search(const ordvec_sign_bitmap_t* sign, const ordvec_rankquant_t* rq,
                                        const float* queries, uint64_t nq, uint64_t m, uint64_t k, float* out_scores,
                                        int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(sign, "sign");
        NONNULL(rq, "rq");
        LOCK;
        two_stage_run(SB(sign), RQ(rq), queries, nq, m, k, out_scores, out_ids, flags);
    });
}

void ordvec_two_stage_policy_default(ordvec_two_stage_policy_t* policy) {
    if (!policy) return;
    *policy = ordvec_two_stage_policy_t{};
    policy->min_candidates = 256;
    policy->k_multiplier = 32;
}

uint64_t ordvec_two_stage_policy_candidate_count(const ordvec_two_stage_policy_t* policy, uint64_t k,
                                                 uint64_t search_space) {
    ordvec_two_stage_policy_t def;
    ordvec_two_stage_policy_default(&def);
    const ordvec_two_stage_policy_t& p = policy ? *policy : def;
    if (k == 0 || search_space == 0) return 0;
    uint64_t prod = 0;
    if (__builtin_mul_overflow(k, p.k_multiplier, &prod)) prod = UINT64_MAX;  // saturating_mul
    uint64_t count = std::max(p.min_candidates, prod);
    if (p.has_max_candidates) count = std::min(count, p.max_candidates);
    return std::min(count, search_space);
}

ordvec_status_t ordvec_rankquant_search_with_sign_probe(const ordvec_rankquant_t* rq, const ordvec_sign_bitmap_t* sign,
                                                        const float* query, uint64_t dim, uint64_t k,
                                                        const ordvec_two_stage_policy_t* policy, float* out_scores,
                                                        int64_t* out_ids, uint64_t* returned, uint64_t flags) {
    return boundary([&] {
        NONNULL(rq, "rq");
        NONNULL(sign, "sign");
        NONNULL(returned, "returned");
        *returned = 0;
        const RqIndex* ix = RQ(rq);
        const SbIndex* sb = SB(sign);
        // src/quant.rs:1474-1492: parameter checks in the reference's order.
        if (sb->dim != ix->dim) fail(kBadArgument, "sign_probe.dim must match RankQuant dim " + std::to_string(ix->dim));
        if (sb->n != ix->n) fail(kBadArgument, "sign_probe.len must match RankQuant len " + std::to_string(ix->n));
        if (dim != ix->dim)
            fail(kDimMismatch, "query has length " + std::to_string(dim) + ", expected " + std::to_string(ix->dim));
        NONNULL(query, "query");
        if (!(flags & (ORDVEC_FLAG_DEVICE_PTRS | ORDVEC_FLAG_ASSUME_FINITE)) && !host_all_finite(query, dim))
            fail(kNonfiniteQuery, "query must contain only finite values");
        const uint64_t m = ordvec_two_stage_policy_candidate_count(policy, k, ix->n);
        const uint64_t ret = std::min(k, m);
        if (ret == 0) return;
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        LOCK;
        // The top min(k, m) of the m candidates: a two-stage row of width
        // min(k, m) (= search_asymmetric_subset's output length).
        two_stage_run(sb, ix, query, 1, m, ret, out_scores, out_ids, flags);
        *returned = ret;
    });
}

ordvec_status_t ordvec_two_stage_search_policy(const ordvec_sign_bitmap_t* sign, const ordvec_rankquant_t* rq,
                                               const float* queries, uint64_t nq, uint64_t k,
                                               const ordvec_two_stage_policy_t* policy, float* out_scores,
                                               int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(sign, "sign");
        NONNULL(rq, "rq");
        const uint64_t m = ordvec_two_stage_policy_candidate_count(policy, k, RQ(rq)->n);
        LOCK;
        two_stage_run(SB(sign), RQ(rq), queries, nq, m, k, out_scores, out_ids, flags);
    });
}

ordvec_status_t ordvec_two_stage_search_noalloc(const ordvec_sign_bitmap_t* sign, const ordvec_rankquant_t* rq,
                                                const float* queries, uint64_t nq, uint64_t m, uint64_t k,
                                                float* out_scores, int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(sign, "sign");
        NONNULL(rq, "rq");
        LOCK;
        two_stage_noalloc(SB(sign), RQ(rq), queries, nq, m, k, out_scores, out_ids, flags);
    });
}

ordvec_status_t ordvec_fastscan_new(uint64_t dim, ordvec_fastscan_t** out) {
    return boundary([&] {
        NONNULL(out, "out");
        *out = nullptr;
        if (dim < 2) fail(kBadArgument, "FastScan b=2: dim must be >= 2");
        if (dim > kMaxDim) fail(kBadArgument, "FastScan b=2: dim must fit in u16");
        if (dim % 4) fail(kBadArgument, "FastScan b=2: dim " + std::to_string(dim) + " must be divisible by 4");
        LOCK;
        eng().init();
        FsIndex* ix = new FsIndex();
        ix->dim = (uint32_t)dim;
        *out = reinterpret_cast<ordvec_fastscan_t*>(ix);
    });
}

ordvec_status_t ordvec_fastscan_new_bits(uint64_t dim, uint32_t bits, ordvec_fastscan_t** out) {
    if (bits == 2) return ordvec_fastscan_new(dim, out);
    return boundary([&] {
        NONNULL(out, "out");
        *out = nullptr;
        if (bits != 4) fail(kBadArgument, "FastScan bits must be 2 or 4");
        if (dim < 8 || dim % 8) fail(kBadArgument, "FastScan b=4: dim " + std::to_string(dim) + " must be a positive multiple of 8");
        if (dim > kFs4MaxDim) fail(kUnsupportedOperation, "FastScan b=4 supports dim <= " + std::to_string(kFs4MaxDim));
        LOCK;
        eng().init();
        FsIndex* ix = new FsIndex();
        ix->dim = (uint32_t)dim;
        ix->bits = 4;
        *out = reinterpret_cast<ordvec_fastscan_t*>(ix);
    });
}

ordvec_status_t ordvec_fastscan_bits(const ordvec_fastscan_t* index, uint32_t* bits) {
    return boundary([&] {
        NONNULL(index, "index");
        NONNULL(bits, "bits");
        *bits = FS(index)->bits;
    });
}

ordvec_status_t ordvec_fastscan_add(ordvec_fastscan_t* index, const float* vectors, uint64_t n, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        FsIndex* ix = FS(index);
        if (n == 0) return;
        NONNULL(vectors, "vectors");
        if (ix->n != 0)
            fail(kBadArgument, "FastScan v1: incremental add() not supported (block-32 layout has tail-padding semantics "
                               "that don't compose); construct a new RankQuantFastscan instead");
        LOCK;
        check_finite(vectors, n * ix->dim, flags, kBadArgument, "vectors");
        check_new_count(0, n);
        RqIndex tmp;
        tmp.dim = ix->dim;
        tmp.bits = ix->bits;
        tmp.bpv = ix->dim * ix->bits / 8;
        rq_add(&tmp, vectors, n, flags | ORDVEC_FLAG_ASSUME_FINITE);
        ix->n = n;
        if (ix->bits == 4) {
            const uint64_t bytes4 = ix->payload4();
            if (hipMalloc(&ix->packed4, bytes4) != hipSuccess) {
                (void)hipGetLastError();
                ix->n = 0;
                fail(kInternal, "HBM allocation for FastScan payload failed");
            }
            OVX_HIP(hipMemsetAsync(ix->packed4, 0, bytes4, cur().stream));
            launch_fs4_pack(tmp.codes, n, ix->dim, ix->packed4, cur().stream);
            sync();
            return;
        }
        uint64_t bytes = ix->payload();
        if (hipMalloc(&ix->packed, bytes) != hipSuccess) {
            (void)hipGetLastError();
            ix->n = 0;
            fail(kInternal, "HBM allocation for FastScan payload failed");
        }
        OVX_HIP(hipMemsetAsync(ix->packed, 0, bytes, cur().stream));
        launch_fastscan_pack(tmp.codes, n, ix->dim, ix->packed, cur().stream);
        sync();
    });
}

ordvec_status_t ordvec_fastscan_search(const ordvec_fastscan_t* index, const float* queries, uint64_t nq, uint64_t k,
                                       float* out_scores, int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(index, "index");
        const FsIndex* ix = FS(index);
        const uint64_t kk = std::min<uint64_t>(k, ix->n);
        if (nq == 0 || kk == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        check_k(kk);
        if (ix->dim > 4096) fail(kUnsupportedOperation, "FastScan LUT staging supports dim <= 4096");
        host_finite_or_fail(queries, nq * ix->dim, flags, kNonfiniteQuery, "queries");
        LOCK;
        if (ix->bits == 4) {
            if (kk > kFs4MaxK) fail(kUnsupportedOperation, "FastScan b=4 supports k <= " + std::to_string(kFs4MaxK));
            eng().init();
            fs4_plain_search(ix, queries, nq, kk, out_scores, out_ids, flags);
            return;
        }
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        const uint32_t d = ix->dim;
        size_t wk = fullscan_work_bytes(ix->n, nq, (uint32_t)kk);
        size_t bytes = al(nq * d * 4) * 2 + 2 * al(nq * 4) + al(nq * (d / 2) * 16) + al(nq * 8) + wk +
                       (dev ? 0 : al(nq * kk * 4) + al(nq * kk * 8));
        Carver cv{static_cast<char*>(cur().workspace(bytes))};
        const float* dq = stage_queries(cv, queries, nq * d, flags);
        float* qhat = cv.take<float>(nq * d);
        float* off = cv.take<float>(nq);
        uint32_t* qstat = cv.take<uint32_t>(nq);
        uint8_t* lut = cv.take<uint8_t>(nq * (d / 2) * 16);
        float* bq = cv.take<float>(nq * 2);
        void* work = cv.take<char>(wk);
        float* ds = dev ? out_scores : cv.take<float>(nq * kk);
        int64_t* di = dev ? out_ids : cv.take<int64_t>(nq * kk);
        const float inv_norm = rankquant_inv_norm(d, 2);
        hipStream_t s = cur().stream;
        launch_prepare_queries(dq, nq, d, 2, 0, inv_norm, qhat, off, qstat, s);
        launch_fastscan_prepare(qhat, nq, d, lut, bq, s);
        launch_fullscan_fs(ix->packed, ix->n, d, lut, bq, inv_norm, nq, (uint32_t)kk, work, ds, di, s);
        if (!dev) {
            d2h(out_scores, ds, nq * kk * 4);
            d2h(out_ids, di, nq * kk * 8);
        }
        finish_status(qstat, nq, flags, kNonfiniteQuery, "queries");
    });
}

ordvec_status_t ordvec_fastscan_info(const ordvec_fastscan_t* index, uint64_t* dim, uint64_t* n) {
    return boundary([&] {
        NONNULL(index, "index");
        if (dim) *dim = FS(index)->dim;
        if (n) *n = FS(index)->n;
    });
}

ordvec_status_t ordvec_fastscan_write(const ordvec_fastscan_t* index, const char* path) {
    return boundary([&] {
        NONNULL(index, "index");
        NONNULL(path, "path");
        if (FS(index)->bits != 2)
            fail(kUnsupportedOperation, "the .ovfs format holds FastScan b=2 indexes only");
        LOCK;
        write_ovfs(FS(index), path);
    });
}

ordvec_status_t ordvec_fastscan_load(const char* path, ordvec_fastscan_t** out) {
    return boundary([&] {
        NONNULL(path, "path");
        NONNULL(out, "out");
        *out = nullptr;
        LOCK;  // the loader checks the file before any device set-up
        *out = reinterpret_cast<ordvec_fastscan_t*>(load_ovfs(path));
    });
}

// ---- batched multi-index FastScan b=4 -------------------------------------------

struct ordvec_fastscan_batch_t {};
static Fs4Batch* FB(ordvec_fastscan_batch_t* p) { return reinterpret_cast<Fs4Batch*>(p); }

ordvec_status_t ordvec_fastscan_batch_new(const ordvec_fastscan_t* const* indexes, uint64_t n_indexes,
                                          const uint32_t* queries_per_index, const uint32_t* k_per_index,
                                          ordvec_fastscan_batch_t** out) {
    return boundary([&] {
        NONNULL(out, "out");
        *out = nullptr;
        if (n_indexes == 0) fail(kBadArgument, "a batch needs at least one index");
        NONNULL(indexes, "indexes");
        NONNULL(queries_per_index, "queries_per_index");
        NONNULL(k_per_index, "k_per_index");
        std::unique_ptr<Fs4Batch> b(new Fs4Batch());
        uint64_t docs = 0, nq = 0;
        for (uint64_t i = 0; i < n_indexes; i++) {
            NONNULL(indexes[i], "indexes[i]");
            const FsIndex* ix = FS(indexes[i]);
            if (ix->bits != 4) fail(kBadArgument, "batched FastScan takes b=4 indexes");
            if (i == 0) b->d = ix->dim;
            if (ix->dim != b->d) fail(kDimMismatch, "every index of a batch must have the same dim");
            if (k_per_index[i] == 0 || k_per_index[i] > kFs4MaxK)
                fail(kBadArgument, "k must be in [1, " + std::to_string(kFs4MaxK) + "]");
            b->ixs.push_back(ix);
            b->nq_per.push_back(queries_per_index[i]);
            b->k_per.push_back(k_per_index[i]);
            docs += ix->n;
            nq += queries_per_index[i];
        }
        if (nq == 0) fail(kBadArgument, "a batch needs at least one query");
        if (nq > (1u << 26)) fail(kBadArgument, "too many queries in one batch");
        LOCK;
        eng().init();
        OVX_HIP(hipStreamCreateWithFlags(&b->ctx.stream, hipStreamNonBlocking));
        CtxScope scope(&b->ctx);
        uint32_t kcap = 1;
        for (size_t i = 0; i < b->ixs.size(); i++) kcap = std::max<uint32_t>(kcap, (uint32_t)std::min<uint64_t>(b->k_per[i], b->ixs[i]->n));
        b->key = Fs4TuneKey{b->d, kcap, ilog2(std::max<uint64_t>(docs, 1)), ilog2(nq)};
        Fs4Config c;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = tune_cache().find(b->key);
            if (it != tune_cache().end()) {
                c = it->second;
            } else {
                const bool two = fs4_config_fits(2, b->d) && nq >= 8 * n_indexes;
                c.qg = two ? 2 : 1;
                c.seg_blocks = 64;
                c.grid = device_cus() * (two ? 1 : 2);
            }
        }
        b->build(c);
        *out = reinterpret_cast<ordvec_fastscan_batch_t*>(b.release());
    });
}

// Queries of the batch (rows ordered by index, queries_per_index[i] each) into
// the batch's buffer on its stream.
static void fs4_stage(Fs4Batch* b, const float* queries, uint64_t flags) {
    const uint64_t count = (uint64_t)b->dv.nq * b->d;
    if (flags & ORDVEC_FLAG_DEVICE_PTRS) d2d(b->dv.dq, queries, count * 4);
    else h2d(b->dv.dq, queries, count * 4);
}

ordvec_status_t ordvec_fastscan_batch_autotune(ordvec_fastscan_batch_t* batch, const float* queries, uint32_t reps,
                                               uint64_t flags) {
    return boundary([&] {
        NONNULL(batch, "batch");
        NONNULL(queries, "queries");
        Fs4Batch* b = FB(batch);
        CtxScope scope(&b->ctx);
        host_finite_or_fail(queries, (uint64_t)b->dv.nq * b->d, flags, kNonfiniteQuery, "queries");
        if (reps == 0) reps = 2;
        const uint32_t cus = device_cus();
        std::vector<Fs4Config> cands;
        for (uint32_t qg : {1u, 2u}) {
            if (!fs4_config_fits(qg, b->d)) continue;
            for (uint32_t sb : {16u, 32u, 64u, 128u, 256u})
                for (uint32_t gm : {1u, 2u, 4u}) cands.push_back(fs4_config_of(4 * qg, sb, cus * gm));
        }
        if (fs4_config_fits(4, b->d)) {  // matrix-core tiles: one 512-thread workgroup per CU
            for (uint32_t sb : {16u, 32u, 64u, 128u, 256u})
                for (uint32_t gm : {1u, 2u}) cands.push_back(fs4_config_of(16, sb, cus * gm));
            // mixed: whole 16-query sets on the matrix cores, the rest on LUT tiles
            for (uint32_t qpt : {20u, 24u}) {
                if (!fs4_config_fits(fs4_config_of(qpt, 64, cus), b->d)) continue;
                for (uint32_t sb : {64u, 128u, 256u})
                    for (uint32_t gm : {1u, 2u}) cands.push_back(fs4_config_of(qpt, sb, cus * gm));
            }
        }
        size_t bytes = 0;
        for (const Fs4Config& c : cands) bytes = std::max(bytes, fs4_dev_bytes(fs4_tables(b->ixs, b->nq_per, b->k_per, c), b->d, c.qg, 0));
        b->drop_graph();
        if (bytes > b->dv.buf.size && b->dv.buf.p) {
            sync();
            (void)hipFree(b->dv.buf.p);
            b->dv.buf.p = nullptr;
            b->dv.buf.size = 0;
        }
        void* base = b->dv.buf.get(bytes, cur().stream);
        hipEvent_t e0, e1;
        OVX_HIP(hipEventCreate(&e0));
        OVX_HIP(hipEventCreate(&e1));
        b->tuning.clear();
        try {
            for (const Fs4Config& c : cands) {
                Fs4Tables t = fs4_tables(b->ixs, b->nq_per, b->k_per, c);
                fs4_upload(b->dv, t, b->d, c.qg, 0, base);
                fs4_stage(b, queries, flags);
                fs4_enqueue(b->dv, b->d, c, true);  // warm
                double best = 1e30;
                for (uint32_t r = 0; r < reps; r++) {
                    OVX_HIP(hipEventRecord(e0, cur().stream));
                    fs4_enqueue(b->dv, b->d, c, true);
                    OVX_HIP(hipEventRecord(e1, cur().stream));
                    OVX_HIP(hipEventSynchronize(e1));
                    float ms = 0;
                    OVX_HIP(hipEventElapsedTime(&ms, e0, e1));
                    best = std::min(best, (double)ms);
                }
                b->tuning.push_back(Fs4TuneRow{c, best});
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        size_t bi = 0;
        for (size_t i = 1; i < b->tuning.size(); i++)
            if (b->tuning[i].ms < b->tuning[bi].ms) bi = i;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            tune_cache()[b->key] = b->tuning[bi].c;
        }
        b->build(b->tuning[bi].c);
    });
}

ordvec_status_t ordvec_fastscan_batch_run(ordvec_fastscan_batch_t* batch, const float* queries, float* out_scores,
                                          int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(batch, "batch");
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        Fs4Batch* b = FB(batch);
        CtxScope scope(&b->ctx);
        const uint64_t nq = b->dv.nq, kout = b->dv.kcap;
        host_finite_or_fail(queries, nq * b->d, flags, kNonfiniteQuery, "queries");
        fs4_stage(b, queries, flags);
        if (flags & ORDVEC_FLAG_NO_GRAPH) fs4_enqueue(b->dv, b->d, b->cfg, true);
        else OVX_HIP(hipGraphLaunch(b->exec, cur().stream));
        const bool dev = flags & ORDVEC_FLAG_DEVICE_PTRS;
        if (dev) {
            d2d(out_scores, b->dv.out_s, nq * kout * 4);
            d2d(out_ids, b->dv.out_i, nq * kout * 8);
        } else {
            d2h(out_scores, b->dv.out_s, nq * kout * 4);
            d2h(out_ids, b->dv.out_i, nq * kout * 8);
        }
        std::vector<uint32_t> st = read_qstat(b->dv.qstat, nq);
        if (dev && !(flags & ORDVEC_FLAG_ASSUME_FINITE) && any_bit(st, kStatNonfinite))
            fail(kNonfiniteQuery, "queries contain NaN or infinity");
    });
}

ordvec_status_t ordvec_fastscan_batch_config(ordvec_fastscan_batch_t* batch, uint64_t* total_queries, uint32_t* kout,
                                             uint32_t* queries_per_tile, uint32_t* blocks_per_tile,
                                             uint32_t* grid, uint64_t* n_tiles, uint64_t* n_tuned) {
    return boundary([&] {
        NONNULL(batch, "batch");
        Fs4Batch* b = FB(batch);
        if (total_queries) *total_queries = b->dv.nq;
        if (kout) *kout = b->dv.kcap;
        if (queries_per_tile) *queries_per_tile = b->cfg.qpt();
        if (blocks_per_tile) *blocks_per_tile = b->cfg.seg_blocks;
        if (grid) *grid = b->cfg.grid;
        if (n_tiles) *n_tiles = b->dv.ntiles;
        if (n_tuned) *n_tuned = b->tuning.size();
    });
}

ordvec_status_t ordvec_fastscan_batch_set_config(ordvec_fastscan_batch_t* batch, uint32_t queries_per_tile,
                                                 uint32_t blocks_per_tile, uint32_t grid) {
    return boundary([&] {
        NONNULL(batch, "batch");
        Fs4Batch* b = FB(batch);
        if (queries_per_tile != 4 && queries_per_tile != 8 && queries_per_tile != 16 && queries_per_tile != 20 &&
            queries_per_tile != 24)
            fail(kBadArgument, "queries_per_tile must be 4, 8, 16, 20 or 24");
        if (blocks_per_tile == 0 || grid == 0) fail(kBadArgument, "blocks_per_tile and grid must be >= 1");
        const Fs4Config c = fs4_config_of(queries_per_tile, blocks_per_tile, grid);
        if (!fs4_config_fits(c, b->d))
            fail(kBadArgument, std::to_string(queries_per_tile) + " queries per tile do not fit dim " + std::to_string(b->d));
        CtxScope scope(&b->ctx);
        b->build(c);
    });
}

ordvec_status_t ordvec_fastscan_batch_tuning(ordvec_fastscan_batch_t* batch, uint64_t i, uint32_t* queries_per_tile,
                                             uint32_t* blocks_per_tile, uint32_t* grid, double* ms) {
    return boundary([&] {
        NONNULL(batch, "batch");
        Fs4Batch* b = FB(batch);
        if (i >= b->tuning.size()) fail(kBadArgument, "tuning row out of range");
        const Fs4TuneRow& r = b->tuning[i];
        if (queries_per_tile) *queries_per_tile = r.c.qpt();
        if (blocks_per_tile) *blocks_per_tile = r.c.seg_blocks;
        if (grid) *grid = r.c.grid;
        if (ms) *ms = r.ms;
    });
}

void ordvec_fastscan_batch_free(ordvec_fastscan_batch_t* batch) {
    if (!batch) return;
    delete FB(batch);
}

// ---- two-stage batch plans -----------------------------------------------------
static TsBatch* TB(const ordvec_two_stage_batch_t* p) {
    return reinterpret_cast<TsBatch*>(const_cast<ordvec_two_stage_batch_t*>(p));
}

static void shape_out(const TsShape& sh, ordvec_two_stage_shape_t* o) {
    *o = ordvec_two_stage_shape_t{};
    o->scan_tiles = sh.t.scan_nqt;
    o->scan_pairs = sh.t.scan_pairs;
    o->scan_flat = sh.t.scan_flat;
    o->sample_wgs = sh.t.sample_wgs;
    o->side_prep = sh.side_prep ? 1u : 0u;
}

constexpr uint64_t kTsMaxQueries = 65536;

ordvec_status_t ordvec_two_stage_batch_new(const ordvec_sign_bitmap_t* sign, const ordvec_rankquant_t* rq,
                                           uint64_t max_queries, uint64_t m, uint64_t k,
                                           ordvec_two_stage_batch_t** out) {
    return boundary([&] {
        NONNULL(out, "out");
        *out = nullptr;
        NONNULL(sign, "sign");
        NONNULL(rq, "rq");
        LOCK;
        SbIndex* sb = SB(sign);
        RqIndex* ix = RQ(rq);
        if (sb->dim != ix->dim) fail(kBadArgument, "sign_probe.dim must match RankQuant dim " + std::to_string(ix->dim));
        if (sb->n != ix->n) fail(kBadArgument, "sign_probe.len must match RankQuant len " + std::to_string(ix->n));
        if (ix->n == 0) fail(kBadArgument, "a two-stage batch plan needs a non-empty index");
        if (max_queries == 0 || max_queries > kTsMaxQueries)
            fail(kBadArgument, "max_queries must be in [1, " + std::to_string(kTsMaxQueries) + "]");
        if (m == 0 || k == 0) fail(kBadArgument, "m and k must be >= 1");
        const uint64_t m_eff = std::min<uint64_t>(m, ix->n), out_k = std::min<uint64_t>(k, ix->n);
        if (m_eff > kStage1MaxM || out_k > kTopKMax)
            fail(kUnsupportedOperation, "two-stage batch plans take m <= " + std::to_string(kStage1MaxM) +
                                            " and k <= " + std::to_string(kTopKMax) +
                                            " (ordvec_two_stage_search takes larger ones)");
        check_search_dim(ix->dim);
        if (ix->dim > kSignMaxDim) fail(kUnsupportedOperation, "GPU sign scan supports dim <= 8192");
        eng().init();
        std::unique_ptr<TsBatch> b(new TsBatch());
        b->sb = sb;
        b->rq = ix;
        b->max_nq = max_queries;
        b->m_eff = m_eff;
        b->out_k = out_k;
        b->ctx.timed = true;
        OVX_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
        OVX_HIP(hipEventCreateWithFlags(&b->st_ev, hipEventDisableTiming));
        size_t need = b->call_bytes(max_queries);
        for (uint64_t q = 1; q < std::min<uint64_t>(max_queries, 128); q++) need = std::max(need, b->call_bytes(q));
        b->ws_bytes = b->stage_bytes() + need;
        if (hipMalloc(&b->ws, b->ws_bytes) != hipSuccess) {
            (void)hipGetLastError();
            b->ws = nullptr;
            fail(kInternal, "device allocation of the plan's workspace (" + std::to_string(b->ws_bytes) +
                                " bytes) failed");
        }
        Carver cv{b->ws};
        b->hq = cv.take<float>(max_queries * ix->dim);
        b->hs = cv.take<float>(max_queries * out_k);
        b->hi = cv.take<int64_t>(max_queries * out_k);
        {
            b->ctx.stream = b->cap;
            CtxScope scope(&b->ctx);
            b->hstat = cur().pinned_u32(max_queries);
            b->ctx.stream = nullptr;
        }
        sb->serving++;
        ix->serving++;
        *out = reinterpret_cast<ordvec_two_stage_batch_t*>(b.release());
    });
}

ordvec_status_t ordvec_two_stage_batch_run(ordvec_two_stage_batch_t* batch, const float* queries, uint64_t nq,
                                           float* out_scores, int64_t* out_ids, uint64_t flags) {
    return boundary([&] {
        NONNULL(batch, "batch");
        TsBatch* b = TB(batch);
        if (nq > b->max_nq)
            fail(kBadArgument, "the plan takes at most " + std::to_string(b->max_nq) + " queries per call");
        if (nq == 0) return;
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        LOCK;
        ts_run(b, queries, nq, out_scores, out_ids, flags);
    });
}

ordvec_status_t ordvec_two_stage_batch_autotune(ordvec_two_stage_batch_t* batch, const float* queries, uint64_t nq,
                                                uint32_t reps, uint64_t flags) {
    return boundary([&] {
        NONNULL(batch, "batch");
        NONNULL(queries, "queries");
        TsBatch* b = TB(batch);
        if (nq == 0 || nq > b->max_nq)
            fail(kBadArgument, "autotune takes 1 to " + std::to_string(b->max_nq) + " queries");
        if (reps == 0) reps = 3;
        LOCK;
        const uint32_t d = b->rq->dim;
        host_finite_or_fail(queries, nq * d, flags, kNonfiniteQuery, "queries");
        const float* dq = queries;
        if (!(flags & ORDVEC_FLAG_DEVICE_PTRS)) {
            b->ctx.stream = eng().def.stream;
            CtxScope scope(&b->ctx);
            h2d(b->hq, queries, nq * d * 4);
            dq = b->hq;
        }
        ts_tune(b, dq, nq, reps);
    });
}

ordvec_status_t ordvec_two_stage_batch_shape(const ordvec_two_stage_batch_t* batch, ordvec_two_stage_shape_t* shape) {
    return boundary([&] {
        NONNULL(batch, "batch");
        NONNULL(shape, "shape");
        shape_out(TB(batch)->shape, shape);
    });
}

ordvec_status_t ordvec_two_stage_batch_set_shape(ordvec_two_stage_batch_t* batch,
                                                 const ordvec_two_stage_shape_t* shape) {
    return boundary([&] {
        NONNULL(batch, "batch");
        NONNULL(shape, "shape");
        if (shape->scan_tiles != 0 && shape->scan_tiles != 1 && shape->scan_tiles != 2 && shape->scan_tiles != 4)
            fail(kBadArgument, "scan_tiles must be 0, 1, 2 or 4");
        if (shape->scan_flat > 1 || shape->side_prep > 1) fail(kBadArgument, "scan_flat and side_prep must be 0 or 1");
        if (shape->reserved[0] || shape->reserved[1] || shape->reserved[2])
            fail(kBadArgument, "reserved fields must be zero");
        if (shape->scan_pairs > (1u << 20) || shape->sample_wgs > (1u << 16))
            fail(kBadArgument, "scan_pairs must be <= 2^20 and sample_wgs <= 65536");
        TsBatch* b = TB(batch);
        LOCK;
        TsShape sh;
        sh.t.scan_nqt = shape->scan_tiles;
        sh.t.scan_pairs = shape->scan_pairs;
        sh.t.scan_flat = shape->scan_flat;
        sh.t.sample_wgs = shape->sample_wgs;
        sh.side_prep = shape->side_prep != 0;
        if (eng().def.stream) OVX_HIP(hipStreamSynchronize(eng().def.stream));
        else OVX_HIP(hipDeviceSynchronize());
        b->drop_graphs();
        b->shape = sh;
    });
}

ordvec_status_t ordvec_two_stage_batch_tuning(const ordvec_two_stage_batch_t* batch, uint64_t i,
                                              ordvec_two_stage_shape_t* shape, double* ms) {
    return boundary([&] {
        NONNULL(batch, "batch");
        const TsBatch* b = TB(batch);
        if (i >= b->tuning.size()) fail(kBadArgument, "tuning row out of range");
        if (shape) shape_out(b->tuning[i].c, shape);
        if (ms) *ms = b->tuning[i].ms;
    });
}

ordvec_status_t ordvec_two_stage_batch_stats(const ordvec_two_stage_batch_t* batch, uint64_t* graph_captures,
                                             uint64_t* graph_calls, uint64_t* eager_calls, uint64_t* exact_queries,
                                             uint64_t* tuned_shapes) {
    return boundary([&] {
        NONNULL(batch, "batch");
        const TsBatch* b = TB(batch);
        if (graph_captures) *graph_captures = b->captures;
        if (graph_calls) *graph_calls = b->replays;
        if (eager_calls) *eager_calls = b->eager;
        if (exact_queries) *exact_queries = b->ctx.exact_queries;
        if (tuned_shapes) *tuned_shapes = b->tuning.size();
    });
}

void ordvec_two_stage_batch_free(ordvec_two_stage_batch_t* batch) {
    if (!batch) return;
    TsBatch* b = TB(batch);
    LOCK;
    if (eng().def.stream) (void)hipStreamSynchronize(eng().def.stream);
    else (void)hipDeviceSynchronize();
    SbIndex* sb = b->sb;
    RqIndex* rq = b->rq;
    delete b;
    if (--sb->serving == 0 && sb->free_pending) delete sb;
    if (--rq->serving == 0 && rq->free_pending) delete rq;
}

void ordvec_fastscan_free(ordvec_fastscan_t* index) {
    if (!index) return;
    LOCK;
    delete FS(index);
}


// ---- serving -------------------------------------------------------------------

static Server* SV(ordvec_server_t* p) { return reinterpret_cast<Server*>(p); }

ordvec_status_t ordvec_server_start(const ordvec_sign_bitmap_t* sign, const ordvec_rankquant_t* rq, uint64_t m,
                                    uint64_t max_k, uint32_t slots, uint64_t max_batch_queries,
                                    uint64_t memory_budget_bytes, ordvec_server_t** out) {
    return boundary([&] {
        NONNULL(sign, "sign");
        NONNULL(rq, "rq");
        NONNULL(out, "out");
        *out = nullptr;
        const SbIndex* sb = SB(sign);
        const RqIndex* ix = RQ(rq);
        if (sb->dim != ix->dim || sb->n != ix->n) fail(kBadArgument, "sign and RankQuant indexes must hold the same documents");
        if (slots == 0 || slots > 16) fail(kBadArgument, "slots must be in [1, 16]");
        if (max_k == 0 || max_k > kTopKMax) fail(kBadArgument, "max_k must be in [1, 1024]");
        check_m(std::min<uint64_t>(m, ix->n));
        LOCK;
        eng().init();
        // Admission by device memory: the budget is one pool; the largest
        // batch is the largest whose workspace fits all of it (batches that
        // run together share it, each holding its own need until it ends).
        uint64_t cap = std::max<uint64_t>(1, max_batch_queries);
        while (cap > 1 && serve_ws_bytes(sb, ix, cap, m, max_k) > memory_budget_bytes) cap = cap * 3 / 4;
        if (serve_ws_bytes(sb, ix, cap, m, max_k) > memory_budget_bytes)
            fail(kBadArgument, "memory budget is too small for a single query");
        std::unique_ptr<Server> sv(new Server());
        sv->sb = sb;
        sv->rq = ix;
        sv->m = m;
        sv->max_batch = cap;
        sv->kcap = max_k;
        sv->ws_need.assign(cap + 1, 0);
        for (uint64_t q = 1; q <= cap; q++)
            sv->ws_need[q] = std::max(sv->ws_need[q - 1], two_stage_bytes(sb, ix, q, m, max_k));
        sv->pool.init(std::max<size_t>(sv->ws_need[cap], memory_budget_bytes));
        for (uint32_t i = 0; i < slots; i++) {
            std::unique_ptr<ServeSlot> sl(new ServeSlot());
            OVX_HIP(hipStreamCreateWithFlags(&sl->ctx.stream, hipStreamNonBlocking));
            OVX_HIP(hipHostMalloc((void**)&sl->q_stage, cap * ix->dim * 4, hipHostMallocDefault));
            OVX_HIP(hipHostMalloc((void**)&sl->s_stage, cap * max_k * 4, hipHostMallocDefault));
            OVX_HIP(hipHostMalloc((void**)&sl->i_stage, cap * max_k * 8, hipHostMallocDefault));
            {
                CtxScope scope(&sl->ctx);
                (void)cur().pinned_u32(cap);
                (void)cur().zero_words(kZwTailDone + cap);
            }
            sv->slots.push_back(std::move(sl));
        }
        Server* raw = sv.get();
        const_cast<SbIndex*>(sb)->serving++;
        const_cast<RqIndex*>(ix)->serving++;
        for (auto& sl : raw->slots) sl->th = std::thread(&Server::worker, raw, sl.get());
        *out = reinterpret_cast<ordvec_server_t*>(sv.release());
    });
}

ordvec_status_t ordvec_server_info(ordvec_server_t* server, uint64_t* max_batch_queries, uint64_t* batches,
                                   uint64_t* batched_queries) {
    return boundary([&] {
        NONNULL(server, "server");
        Server* sv = SV(server);
        std::lock_guard<std::mutex> lk(sv->mu);
        if (max_batch_queries) *max_batch_queries = sv->max_batch;
        if (batches) *batches = sv->batches;
        if (batched_queries) *batched_queries = sv->batched_queries;
    });
}

ordvec_status_t ordvec_server_submit(ordvec_server_t* server, const float* queries, uint64_t nq, uint64_t k,
                                     float* out_scores, int64_t* out_ids, uint64_t* ticket) {
    return boundary([&] {
        NONNULL(server, "server");
        NONNULL(ticket, "ticket");
        Server* sv = SV(server);
        if (nq == 0 || nq > sv->max_batch)
            fail(kBadArgument, "request size must be in [1, " + std::to_string(sv->max_batch) + "] queries");
        if (k == 0 || k > sv->kcap) fail(kBadArgument, "k must be in [1, max_k]");
        NONNULL(queries, "queries");
        NONNULL(out_scores, "out_scores");
        NONNULL(out_ids, "out_ids");
        host_finite_or_fail(queries, nq * sv->rq->dim, 0, kNonfiniteQuery, "queries");
        ServeRequest* r = new ServeRequest();
        r->queries = queries;
        r->nq = nq;
        r->k = k;
        r->out_s = out_scores;
        r->out_i = out_ids;
        r->t_arrival_ns = now_ns();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            *ticket = sv->next_ticket++;
            sv->tickets[*ticket] = r;
        }
        sv->enqueue(r);
    });
}

ordvec_status_t ordvec_server_wait(ordvec_server_t* server, uint64_t ticket, uint64_t* latency_ns,
                                   uint64_t* batch_queries) {
    ServeRequest* r = nullptr;
    ordvec_status_t st = boundary([&] {
        NONNULL(server, "server");
        Server* sv = SV(server);
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->tickets.find(ticket);
        if (it == sv->tickets.end()) fail(kBadArgument, "unknown ticket " + std::to_string(ticket));
        r = it->second;
        sv->cv_done.wait(lk, [&] { return r->done; });
        sv->tickets.erase(it);
        if (latency_ns) *latency_ns = r->t_done_ns - r->t_arrival_ns;
        if (batch_queries) *batch_queries = r->batch_queries;
        if (r->status) {
            std::string msg = r->error;
            uint32_t code = r->status;
            delete r;
            r = nullptr;
            throw Error(code, msg);
        }
    });
    delete r;
    return st;
}

ordvec_status_t ordvec_server_open_loop(ordvec_server_t* server, const float* query_pool, uint64_t pool_n,
                                        double offered_qps, double duration_s, uint32_t min_nq, uint32_t max_nq,
                                        uint32_t min_k, uint32_t max_k, uint64_t seed, ordvec_serve_report_t* report) {
    return boundary([&] {
        NONNULL(server, "server");
        NONNULL(query_pool, "query_pool");
        NONNULL(report, "report");
        check_exact_size(report->struct_size, sizeof(ordvec_serve_report_t), "ordvec_serve_report_t");
        Server* sv = SV(server);
        if (min_nq == 0 || max_nq < min_nq || max_nq > sv->max_batch || max_nq > pool_n)
            fail(kBadArgument, "request sizes must satisfy 1 <= min_nq <= max_nq <= min(max batch, pool)");
        if (min_k == 0 || max_k < min_k || max_k > sv->kcap) fail(kBadArgument, "k range must lie in [1, max_k]");
        if (!(offered_qps > 0) || !(duration_s > 0)) fail(kBadArgument, "rate and duration must be positive");
        const uint64_t d = sv->rq->dim;
        host_finite_or_fail(query_pool, pool_n * d, 0, kNonfiniteQuery, "query_pool");
        std::mt19937_64 rng(seed);
        std::uniform_int_distribution<uint32_t> unq(min_nq, max_nq), uk(min_k, max_k);
        const double mean_nq = 0.5 * (min_nq + max_nq);
        std::exponential_distribution<double> gap(offered_qps / mean_nq);  // request arrivals per second
        std::vector<std::unique_ptr<ServeRequest>> reqs;
        uint64_t b0 = 0, exact0 = 0, ov0 = 0, of0 = 0;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            b0 = sv->batches;
            ov0 = sv->overlapped;
            sv->max_in_flight = sv->in_flight;
            sv->pool.peak = sv->pool.in_use;
            for (auto& sl : sv->slots) {
                exact0 += sl->ctx.exact_queries;
                of0 += sl->ctx.lent_overflows;
            }
        }
        const uint64_t t0 = now_ns() + 2000000;  // start 2 ms from now
        double t = 0.0;
        for (;;) {
            t += gap(rng);
            if (t >= duration_s) break;
            std::unique_ptr<ServeRequest> r(new ServeRequest());
            r->nq = unq(rng);
            r->k = uk(rng);
            const uint64_t start = std::uniform_int_distribution<uint64_t>(0, pool_n - r->nq)(rng);
            r->queries = query_pool + start * d;
            r->t_arrival_ns = t0 + (uint64_t)(t * 1e9);
            reqs.push_back(std::move(r));
        }
        for (auto& r : reqs) {  // allocate outputs before the clock starts
            const uint64_t rk = std::min<uint64_t>(r->k, sv->rq->n);
            r->own_s.resize(r->nq * rk);
            r->own_i.resize(r->nq * rk);
            r->out_s = r->own_s.data();
            r->out_i = r->own_i.data();
        }
        // Open loop: submit at the scheduled instant regardless of completions.
        for (auto& r : reqs) {
            while (now_ns() < r->t_arrival_ns) {
                if (r->t_arrival_ns - now_ns() > 200000) std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            sv->enqueue(r.get());
        }
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            for (auto& r : reqs) sv->cv_done.wait(lk, [&] { return r->done; });
        }
        ordvec_serve_report_t rep;
        std::memset(&rep, 0, sizeof(rep));
        rep.struct_size = sizeof(rep);
        std::vector<double> lat, queue, service;
        uint64_t last_done = t0, queries = 0;
        for (auto& r : reqs) {
            if (r->status) {
                rep.errors++;
                continue;
            }
            lat.push_back((double)(r->t_done_ns - r->t_arrival_ns) * 1e-6);
            queue.push_back((double)(r->t_start_ns > r->t_arrival_ns ? r->t_start_ns - r->t_arrival_ns : 0) * 1e-6);
            service.push_back((double)(r->t_done_ns - r->t_start_ns) * 1e-6);
            last_done = std::max(last_done, r->t_done_ns);
            queries += r->nq;
        }
        rep.requests = reqs.size();
        rep.queries = queries;
        rep.duration_s = duration_s;
        rep.offered_qps = offered_qps;
        const double span = (double)(last_done - t0) * 1e-9;
        rep.achieved_qps = span > 0 ? queries / span : 0;
        rep.achieved_rps = span > 0 ? (double)lat.size() / span : 0;
        if (!lat.empty()) {
            std::sort(lat.begin(), lat.end());
            auto pct = [&](double p) { return lat[std::min(lat.size() - 1, (size_t)(p * (lat.size() - 1) + 0.5))]; };
            rep.p50_ms = pct(0.50);
            rep.p90_ms = pct(0.90);
            rep.p99_ms = pct(0.99);
            rep.p999_ms = pct(0.999);
            rep.max_ms = lat.back();
            double sum = 0;
            for (double v : lat) sum += v;
            rep.mean_ms = sum / lat.size();
            std::sort(queue.begin(), queue.end());
            std::sort(service.begin(), service.end());
            const size_t i99 = std::min(lat.size() - 1, (size_t)(0.99 * (lat.size() - 1) + 0.5));
            rep.p99_queue_ms = queue[i99];
            rep.p99_service_ms = service[i99];
        }
        std::lock_guard<std::mutex> lk(sv->mu);
        uint64_t exact1 = 0, of1 = 0;
        for (auto& sl : sv->slots) {
            exact1 += sl->ctx.exact_queries;
            of1 += sl->ctx.lent_overflows;
        }
        rep.exact_path_queries = exact1 - exact0;
        rep.pool_bytes = sv->pool.size;
        rep.peak_pool_bytes = sv->pool.peak;
        rep.overlapped_batches = sv->overlapped - ov0;
        rep.max_batches_in_flight = sv->max_in_flight;
        rep.workspace_overflows = of1 - of0;
        rep.batches = (double)(sv->batches - b0);
        rep.mean_batch_queries = rep.batches > 0 ? (double)queries / rep.batches : 0;
        *report = rep;
    });
}

void ordvec_server_stop(ordvec_server_t* server) {
    if (!server) return;
    Server* sv = SV(server);
    sv->stop();
    {
        LOCK;
        SbIndex* sb = const_cast<SbIndex*>(sv->sb);
        RqIndex* rq = const_cast<RqIndex*>(sv->rq);
        if (--sb->serving == 0 && sb->free_pending) delete sb;
        if (--rq->serving == 0 && rq->free_pending) delete rq;
    }
    delete sv;
}

}  // extern "C"
