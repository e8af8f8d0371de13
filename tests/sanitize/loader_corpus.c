/*
 * Forged-file corpus for the index loaders, built and run under
 * AddressSanitizer + UBSan on the CPU (tests/sanitize/Makefile).
 *
 * Restates the reference's malformed-file contract (tests/index/main.rs:
 * 100-273: every loader returns an error, never a panic, on empty files,
 * garbage, wrong magic, oversized or inconsistent headers and truncated
 * payloads) and its fuzz targets (fuzz/fuzz_targets/load_*.rs: arbitrary
 * bytes into each loader) as a deterministic corpus:
 *   - the reference's fourteen cases, under the legacy TV* and current OV*
 *     magics;
 *   - one small valid file per format, every prefix of it, every header
 *     byte replaced by 0x00 / 0xFF / itself + 1, and one trailing byte;
 *   - seeded random mutations and random garbage;
 *   - a directory, a missing path and an empty path.
 * Each file goes through every loader (ABI v1 load and probe, and the
 * per-kind loads of ABI v2).  A status outside the enum, ORDVEC_STATUS_PANIC
 * (an exception the boundary did not expect), ORDVEC_STATUS_OK on a file
 * that must be rejected, or any sanitizer report fails the run.  Without a
 * GPU a well-formed file stops at device set-up; the header and payload
 * checks before it are what this run covers.
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "ordvec.h"

static char g_dir[512];
static int g_failures = 0;
static unsigned long g_calls = 0;
static int g_verbose = 0;
static const char *g_accepting = NULL; /* loader that must accept the current file */

typedef struct {
    unsigned char *p;
    size_t n, cap;
} buf_t;

static void put(buf_t *b, const void *src, size_t n) {
    if (b->n + n > b->cap) {
        b->cap = (b->n + n) * 2 + 64;
        b->p = (unsigned char *)realloc(b->p, b->cap);
        if (!b->p) abort();
    }
    memcpy(b->p + b->n, src, n);
    b->n += n;
}
static void put_u8(buf_t *b, unsigned v) {
    unsigned char c = (unsigned char)v;
    put(b, &c, 1);
}
static void put_u32(buf_t *b, uint32_t v) {
    unsigned char c[4] = {(unsigned char)v, (unsigned char)(v >> 8), (unsigned char)(v >> 16), (unsigned char)(v >> 24)};
    put(b, c, 4);
}
static void put_u16(buf_t *b, unsigned v) {
    put_u8(b, v & 0xff);
    put_u8(b, v >> 8);
}
static void put_u64(buf_t *b, uint64_t v) {
    put_u32(b, (uint32_t)v);
    put_u32(b, (uint32_t)(v >> 32));
}

static void write_bytes(const char *path, const unsigned char *p, size_t n) {
    FILE *f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "cannot create %s: %s\n", path, strerror(errno));
        exit(2);
    }
    if (n && fwrite(p, 1, n, f) != n) abort();
    fclose(f);
}

static void check(const char *label, const char *loader, ordvec_status_t st, int must_fail) {
    g_calls++;
    if (g_verbose) printf("%-20s %-18s %-20s %s\n", label, loader, ordvec_status_name(st), ordvec_last_error());
    if (st > ORDVEC_STATUS_INTERNAL) {
        fprintf(stderr, "FAIL %s / %s: status %d outside the enum\n", label, loader, (int)st);
        g_failures++;
    } else if (st == ORDVEC_STATUS_PANIC) {
        fprintf(stderr, "FAIL %s / %s: PANIC (%s)\n", label, loader, ordvec_last_error());
        g_failures++;
    } else if (must_fail && st == ORDVEC_STATUS_OK) {
        fprintf(stderr, "FAIL %s / %s: accepted a malformed file\n", label, loader);
        g_failures++;
    }
    if (g_accepting && strcmp(loader, g_accepting) == 0 && st != ORDVEC_STATUS_OK &&
        !(st == ORDVEC_STATUS_INTERNAL && strstr(ordvec_last_error(), "no HIP device"))) {
        fprintf(stderr, "FAIL %s / %s: a well-formed file was rejected: %s\n", label, loader, ordvec_last_error());
        g_failures++;
    }
    if (st != ORDVEC_STATUS_OK && strlen(ordvec_last_error()) == 0) {
        fprintf(stderr, "FAIL %s / %s: error status %s with an empty message\n", label, loader,
                ordvec_status_name(st));
        g_failures++;
    }
}

/* Every loader on one path. */
static void run_loaders(const char *path, const char *label, int must_fail) {
    ordvec_index_t *ix = NULL;
    ordvec_status_t st = ordvec_index_load(path, 0, &ix);
    check(label, "index_load", st, must_fail);
    if (st == ORDVEC_STATUS_OK) ordvec_index_free(ix);

    ordvec_index_info_t info;
    ordvec_index_info_init(&info);
    check(label, "index_probe", ordvec_index_probe(path, 0, &info), must_fail);

    ordvec_rankquant_t *rq = NULL;
    st = ordvec_rankquant_load(path, &rq);
    check(label, "rankquant_load", st, must_fail);
    if (st == ORDVEC_STATUS_OK) ordvec_rankquant_free(rq);

    ordvec_rank_t *rk = NULL;
    st = ordvec_rank_load(path, &rk);
    check(label, "rank_load", st, must_fail);
    if (st == ORDVEC_STATUS_OK) ordvec_rank_free(rk);

    ordvec_sign_bitmap_t *sb = NULL;
    st = ordvec_sign_bitmap_load(path, &sb);
    check(label, "sign_bitmap_load", st, must_fail);
    if (st == ORDVEC_STATUS_OK) ordvec_sign_bitmap_free(sb);

    ordvec_bitmap_t *bm = NULL;
    st = ordvec_bitmap_load(path, &bm);
    check(label, "bitmap_load", st, must_fail);
    if (st == ORDVEC_STATUS_OK) ordvec_bitmap_free(bm);

    ordvec_fastscan_t *fs = NULL;
    st = ordvec_fastscan_load(path, &fs);
    check(label, "fastscan_load", st, must_fail);
    if (st == ORDVEC_STATUS_OK) ordvec_fastscan_free(fs);
}

static void run_bytes(const char *label, const unsigned char *p, size_t n, int must_fail) {
    char path[640];
    snprintf(path, sizeof path, "%s/case.bin", g_dir);
    write_bytes(path, p, n);
    run_loaders(path, label, must_fail);
}

/* The reference's fourteen malformed files (tests/index/main.rs:114-239),
 * with the given magic family ('T' legacy or 'O' current). */
static void reference_cases(char fam) {
    char m_r1[5] = {fam, 'V', 'R', '1', 0}, m_rq[5] = {fam, 'V', 'R', 'Q', 0}, m_bm[5] = {fam, 'V', 'B', 'M', 0},
         m_sb[5] = {fam, 'V', 'S', 'B', 0};
    buf_t b = {0};
    char label[64];
#define CASE(name)                                                \
    do {                                                          \
        snprintf(label, sizeof label, "%c:%s", fam, name);        \
        run_bytes(label, b.p, b.n, 1);                            \
        b.n = 0;                                                  \
    } while (0)
    CASE("empty");
    for (int i = 0; i < 16; i++) put_u8(&b, 0xAB);
    CASE("garbage_16");
    for (int i = 0; i < 4096; i++) put_u8(&b, 0xCC);
    CASE("garbage_4k");
    put(&b, "XXXX\x01", 5);
    for (int i = 0; i < 8; i++) put_u8(&b, 0);
    CASE("wrong_magic");
    put(&b, m_r1, 4), put_u8(&b, 1), put_u32(&b, UINT32_MAX), put_u32(&b, UINT32_MAX);
    CASE("ovr_oversize");
    put(&b, m_rq, 4), put_u8(&b, 1), put_u8(&b, 255), put_u32(&b, 1024), put_u32(&b, 10);
    CASE("ovrq_bad_bits");
    put(&b, m_rq, 4), put_u8(&b, 1), put_u8(&b, 2), put_u32(&b, 13), put_u32(&b, 0);
    CASE("ovrq_bad_dim");
    put(&b, m_rq, 4), put_u8(&b, 1), put_u8(&b, 4), put_u32(&b, UINT32_MAX), put_u32(&b, UINT32_MAX);
    CASE("ovrq_oversize");
    put(&b, m_bm, 4), put_u8(&b, 1), put_u32(&b, 100), put_u32(&b, 25), put_u32(&b, 5);
    CASE("ovbm_dim_not_64");
    put(&b, m_bm, 4), put_u8(&b, 1), put_u32(&b, 128), put_u32(&b, 128), put_u32(&b, 5);
    CASE("ovbm_bad_n_top");
    put(&b, m_bm, 4), put_u8(&b, 1), put_u32(&b, UINT32_MAX), put_u32(&b, 1), put_u32(&b, UINT32_MAX);
    CASE("ovbm_oversize");
    put(&b, m_r1, 4), put_u8(&b, 1), put_u32(&b, 64), put_u32(&b, 100);
    for (int i = 0; i < 100; i++) put_u8(&b, 0);
    CASE("ovr_truncated");
    put(&b, m_sb, 4), put_u8(&b, 1), put_u32(&b, 100), put_u32(&b, 5);
    CASE("ovsb_dim_not_64");
    put(&b, m_sb, 4), put_u8(&b, 1), put_u32(&b, UINT32_MAX), put_u32(&b, UINT32_MAX);
    CASE("ovsb_oversize");
    put(&b, m_sb, 4), put_u8(&b, 1), put_u32(&b, 128), put_u32(&b, 8);
    CASE("ovsb_truncated");
#undef CASE
    free(b.p);
}

/* One small well-formed file per format and its header length. */
typedef struct {
    const char *name, *loader;
    buf_t b;
    size_t header;
} valid_t;

static void make_valid(valid_t v[6]) {
    memset(v, 0, sizeof(valid_t) * 6);
    /* .ovrq: dim 8, bits 2, rows 00 00 01 01 10 10 11 11 and a permutation */
    v[0].name = "ovrq";
    v[0].loader = "rankquant_load";
    put(&v[0].b, "OVRQ", 4), put_u8(&v[0].b, 1), put_u8(&v[0].b, 2), put_u32(&v[0].b, 8), put_u32(&v[0].b, 2);
    put_u8(&v[0].b, 0x05), put_u8(&v[0].b, 0xAF), put_u8(&v[0].b, 0xE4), put_u8(&v[0].b, 0x1B);
    v[0].header = 14;
    /* .ovr: dim 4, two permutations */
    v[1].name = "ovr";
    v[1].loader = "rank_load";
    put(&v[1].b, "OVR1", 4), put_u8(&v[1].b, 1), put_u32(&v[1].b, 4), put_u32(&v[1].b, 2);
    for (unsigned r = 0; r < 4; r++) put_u16(&v[1].b, r);
    for (unsigned r = 0; r < 4; r++) put_u16(&v[1].b, 3 - r);
    v[1].header = 13;
    /* .ovsb: dim 128, two rows */
    v[2].name = "ovsb";
    v[2].loader = "sign_bitmap_load";
    put(&v[2].b, "OVSB", 4), put_u8(&v[2].b, 1), put_u32(&v[2].b, 128), put_u32(&v[2].b, 2);
    put_u64(&v[2].b, 0x0123456789abcdefull), put_u64(&v[2].b, ~0ull), put_u64(&v[2].b, 0), put_u64(&v[2].b, 5);
    v[2].header = 13;
    /* .ovbm: dim 64, n_top 3, two rows of three set bits */
    v[3].name = "ovbm";
    v[3].loader = "bitmap_load";
    put(&v[3].b, "OVBM", 4), put_u8(&v[3].b, 1), put_u32(&v[3].b, 64), put_u32(&v[3].b, 3), put_u32(&v[3].b, 2);
    put_u64(&v[3].b, 0x7ull), put_u64(&v[3].b, 0x8000000000000101ull);
    v[3].header = 17;
    /* .ovfs: dim 4 (two pairs), one row with buckets 0 1 2 3, 31 zero tail lanes */
    v[4].name = "ovfs";
    v[4].loader = "fastscan_load";
    put(&v[4].b, "OVFS", 4), put_u8(&v[4].b, 1), put_u32(&v[4].b, 4), put_u32(&v[4].b, 1);
    for (int p = 0; p < 2; p++)
        for (int l = 0; l < 32; l++) put_u8(&v[4].b, l ? 0 : (p ? (2 << 2 | 3) : (0 << 2 | 1)));
    v[4].header = 13;
    /* legacy .tvrq of the first file */
    v[5].name = "tvrq";
    v[5].loader = "rankquant_load";
    put(&v[5].b, v[0].b.p, v[0].b.n);
    memcpy(v[5].b.p, "TVRQ", 4);
    v[5].header = 14;
}

static uint64_t g_rng = 0x9E3779B97F4A7C15ull;
static uint64_t rnd(void) {
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return g_rng;
}

int main(int argc, char **argv) {
    const int mutations = argc > 1 ? atoi(argv[1]) : 400;
    g_verbose = argc > 2 && strcmp(argv[2], "-v") == 0;
    snprintf(g_dir, sizeof g_dir, "/tmp/ordvec_corpus_%d", (int)getpid());
    if (mkdir(g_dir, 0700) != 0 && errno != EEXIST) {
        perror("mkdir");
        return 2;
    }
    reference_cases('T');
    reference_cases('O');

    valid_t v[6];
    make_valid(v);
    char label[96];
    unsigned char *tmp = (unsigned char *)malloc(1 << 16);
    for (int f = 0; f < 6; f++) {
        const buf_t *b = &v[f].b;
        snprintf(label, sizeof label, "%s:valid", v[f].name);
        g_accepting = v[f].loader;
        run_bytes(label, b->p, b->n, 0);
        g_accepting = NULL;
        for (size_t len = 0; len < b->n; len++) {  /* every strict prefix is truncated */
            snprintf(label, sizeof label, "%s:prefix%zu", v[f].name, len);
            run_bytes(label, b->p, len, 1);
        }
        memcpy(tmp, b->p, b->n);
        tmp[b->n] = 0;
        snprintf(label, sizeof label, "%s:trailing", v[f].name);
        run_bytes(label, tmp, b->n + 1, 1);
        for (size_t i = 0; i < v[f].header; i++) {
            const unsigned char repl[3] = {0x00, 0xFF, (unsigned char)(b->p[i] + 1)};
            for (int k = 0; k < 3; k++) {
                if (repl[k] == b->p[i]) continue;
                memcpy(tmp, b->p, b->n);
                tmp[i] = repl[k];
                snprintf(label, sizeof label, "%s:hdr%zu=%02x", v[f].name, i, repl[k]);
                run_bytes(label, tmp, b->n, 0);
            }
        }
        for (int m = 0; m < mutations; m++) {  /* 1-4 random byte edits, random length changes */
            size_t n = b->n;
            memcpy(tmp, b->p, n);
            const int edits = 1 + (int)(rnd() % 4);
            for (int e = 0; e < edits; e++) {
                const uint64_t r = rnd();
                switch (r % 4) {
                    case 0: tmp[r / 4 % n] = (unsigned char)(r >> 32); break;
                    case 1: tmp[r / 4 % n] ^= (unsigned char)(1u << (r >> 40) % 8); break;
                    case 2: if (n > 1) n -= 1 + (size_t)(r / 4 % (n - 1)); break;
                    default: {
                        const size_t add = (size_t)(r / 4 % 64);
                        for (size_t j = 0; j < add && n < 4096; j++) tmp[n++] = (unsigned char)rnd();
                    }
                }
            }
            snprintf(label, sizeof label, "%s:mut%d", v[f].name, m);
            run_bytes(label, tmp, n, 0);
        }
    }
    for (int g = 0; g < mutations; g++) {  /* garbage behind a valid magic */
        static const char *magics[] = {"OVRQ", "TVRQ", "OVR1", "TVR1", "OVSB", "TVSB", "OVBM", "TVBM", "OVFS"};
        const size_t n = 4 + (size_t)(rnd() % 200);
        memcpy(tmp, magics[g % 9], 4);
        for (size_t j = 4; j < n; j++) tmp[j] = (unsigned char)rnd();
        if (n > 4) tmp[4] = (rnd() & 3) ? 1 : tmp[4];  /* mostly version 1, to reach the fields */
        snprintf(label, sizeof label, "garbage%d", g);
        run_bytes(label, tmp, n, 0);
    }
    run_loaders(g_dir, "directory", 1);
    char missing[640];
    snprintf(missing, sizeof missing, "%s/does-not-exist", g_dir);
    run_loaders(missing, "missing", 1);
    run_loaders("", "empty-path", 1);

    free(tmp);
    for (int f = 0; f < 6; f++) free(v[f].b.p);
    char path[640];
    snprintf(path, sizeof path, "%s/case.bin", g_dir);
    unlink(path);
    rmdir(g_dir);
    printf("loader corpus: %lu loader calls, %d failures\n", g_calls, g_failures);
    return g_failures ? 1 : 0;
}
