/*
 * A C caller of the ordvec ABI v1, compiled with `cc` against include/ordvec.h
 * and linked with libordvec_amd.so (as ordvec-ffi/tests/c_link_smoke.rs links
 * a C program against the reference cdylib).  It writes the reference FFI
 * fixture (RankQuant(16, 2), four copies of the document 0..15; ordvec-ffi/src/
 * lib.rs:1022-1033) as a .ovrq file, then loads, probes and searches it and
 * walks the status paths of ordvec-ffi/src/lib.rs:1274-1500.
 *
 * usage: c_link_smoke <scratch directory>; exit status 0 = every check passed,
 * otherwise the number of the first failed check (printed to stderr).
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "ordvec.h"

static int g_check = 0;

#define CHECK(cond)                                                                      \
    do {                                                                                 \
        g_check++;                                                                       \
        if (!(cond)) {                                                                   \
            fprintf(stderr, "check %d failed: %s (line %d; last error: %s)\n", g_check, \
                    #cond, __LINE__, ordvec_last_error());                               \
            return g_check;                                                              \
        }                                                                                \
    } while (0)

/* "OVRQ" | version 1 | bits 2 | dim 16 | n 4 | four rows of buckets
 * 0,0,0,0, 1,1,1,1, 2,2,2,2, 3,3,3,3 packed MSB-first: 00 55 AA FF. */
static int write_fixture(const char *path) {
    static const unsigned char hdr[14] = {'O', 'V', 'R', 'Q', 1, 2, 16, 0, 0, 0, 4, 0, 0, 0};
    static const unsigned char row[4] = {0x00, 0x55, 0xAA, 0xFF};
    FILE *f = fopen(path, "wb");
    if (!f) return 0;
    fwrite(hdr, 1, sizeof hdr, f);
    for (int i = 0; i < 4; i++) fwrite(row, 1, sizeof row, f);
    return fclose(f) == 0;
}

static ordvec_status_t search(const ordvec_index_t *idx, const ordvec_search_params_t *p, ordvec_hit_t *hits,
                              uint64_t cap, uint64_t *ret, ordvec_search_stats_t *st) {
    return ordvec_index_search(idx, p, hits, cap, ret, st);
}

int main(int argc, char **argv) {
    char path[4096];
    snprintf(path, sizeof path, "%s/fixture.ovrq", argc > 1 ? argv[1] : ".");
    CHECK(write_fixture(path));
    CHECK(ordvec_abi_version() == ORDVEC_ABI_VERSION);
    CHECK(strcmp(ordvec_status_name(ORDVEC_STATUS_BUFFER_TOO_SMALL), "BUFFER_TOO_SMALL") == 0);
    CHECK(ordvec_init() == ORDVEC_STATUS_OK);

    /* probe without loading (lib.rs:1104-1133) */
    ordvec_index_info_t info;
    ordvec_index_info_init(&info);
    CHECK(ordvec_index_probe(path, 0, &info) == ORDVEC_STATUS_OK);
    CHECK(info.kind == ORDVEC_INDEX_KIND_RANK_QUANT && info.format_version == 1);
    CHECK(info.dim == 16 && info.bit_width == 2 && info.n_top == 0 && info.vector_count == 4);
    CHECK(info.bytes_per_vec == 4 && info.source_file_size_bytes == 30);
    CHECK((info.capabilities & ORDVEC_CAP_SUBSET_SEARCH) == ORDVEC_CAP_SUBSET_SEARCH);

    /* load, info, free (lib.rs:1079-1102) */
    ordvec_index_t *idx = NULL;
    CHECK(ordvec_index_load(path, 0, &idx) == ORDVEC_STATUS_OK && idx != NULL);
    ordvec_index_info_init(&info);
    CHECK(ordvec_index_info(idx, &info) == ORDVEC_STATUS_OK);
    CHECK(info.kind == ORDVEC_INDEX_KIND_RANK_QUANT && info.dim == 16 && info.bit_width == 2 && info.vector_count == 4);

    /* full and subset search (lib.rs:1135-1224) */
    float query[16];
    for (int j = 0; j < 16; j++) query[j] = (float)j;
    ordvec_search_params_t p;
    ordvec_search_params_init(&p);
    p.query = query;
    p.dim = 16;
    p.k = 2;
    p.user_tag = 77;
    ordvec_hit_t hits[3];
    uint64_t returned = 0;
    ordvec_search_stats_t stats;
    ordvec_search_stats_init(&stats);
    CHECK(search(idx, &p, hits, 2, &returned, &stats) == ORDVEC_STATUS_OK);
    CHECK(returned == 2 && hits[0].row_id == 0 && hits[0].id == hits[0].row_id && hits[1].row_id == 1);
    CHECK(stats.user_tag == 77 && stats.candidate_count == 4 && stats.vectors_scored == 4 && stats.returned_count == 2);
    /* every doc equals the query: score = <q_hat, centres> / R, positive and equal */
    CHECK(hits[0].score > 0.0f && hits[0].score == hits[1].score && isfinite(hits[0].score));

    const uint32_t cands[3] = {3, 1, 2};
    p.candidate_rows = cands;
    p.candidate_count = 3;
    CHECK(search(idx, &p, hits, 2, &returned, NULL) == ORDVEC_STATUS_OK);
    CHECK(returned == 2 && hits[0].row_id == 1 && hits[1].row_id == 2);

    const uint32_t dup[4] = {3, 1, 1, 2};
    p.k = 3;
    p.candidate_rows = dup;
    p.candidate_count = 4;
    ordvec_search_stats_init(&stats);
    CHECK(search(idx, &p, hits, 3, &returned, &stats) == ORDVEC_STATUS_OK);
    CHECK(returned == 3 && hits[0].row_id == 1 && hits[1].row_id == 1 && hits[2].row_id == 2);
    CHECK(stats.candidate_count == 4 && stats.vectors_scored == 4);

    /* validation failures in the reference's order (lib.rs:1274-1472) */
    float ones[16];
    for (int j = 0; j < 16; j++) ones[j] = 1.0f;
    ordvec_search_params_t good;
    ordvec_search_params_init(&good);
    good.query = ones;
    good.dim = 16;
    good.k = 1;
    ordvec_search_params_t bad;
    CHECK(search(NULL, &good, hits, 1, &returned, NULL) == ORDVEC_STATUS_NULL_POINTER);
    CHECK(search(idx, NULL, hits, 1, &returned, NULL) == ORDVEC_STATUS_NULL_POINTER);
    bad = good;
    bad.struct_size -= 1;
    CHECK(search(idx, &bad, hits, 1, &returned, NULL) == ORDVEC_STATUS_BAD_STRUCT_SIZE);
    bad = good;
    bad.flags = 1;
    CHECK(search(idx, &bad, hits, 1, &returned, NULL) == ORDVEC_STATUS_BAD_ARGUMENT);
    bad = good;
    bad.reserved[0] = 1;
    CHECK(search(idx, &bad, hits, 1, &returned, NULL) == ORDVEC_STATUS_BAD_ARGUMENT);
    bad = good;
    bad.query = NULL;
    CHECK(search(idx, &bad, hits, 1, &returned, NULL) == ORDVEC_STATUS_NULL_POINTER);
    bad = good;
    bad.dim = 15;
    CHECK(search(idx, &bad, hits, 1, &returned, NULL) == ORDVEC_STATUS_DIM_MISMATCH);
    float nonfinite[16];
    memcpy(nonfinite, ones, sizeof ones);
    nonfinite[0] = NAN;
    bad = good;
    bad.query = nonfinite;
    CHECK(search(idx, &bad, hits, 1, &returned, NULL) == ORDVEC_STATUS_NONFINITE_QUERY);
    CHECK(strlen(ordvec_last_error()) > 0);
    CHECK(search(idx, &good, hits, 1, NULL, NULL) == ORDVEC_STATUS_NULL_POINTER);
    CHECK(search(idx, &good, NULL, 1, &returned, NULL) == ORDVEC_STATUS_NULL_POINTER);
    CHECK(search(idx, &good, hits, 0, &returned, NULL) == ORDVEC_STATUS_BUFFER_TOO_SMALL);
    const uint32_t far[1] = {99};
    bad = good;
    bad.candidate_rows = far;
    bad.candidate_count = 1;
    CHECK(search(idx, &bad, hits, 1, &returned, NULL) == ORDVEC_STATUS_ROW_ID_OUT_OF_RANGE);
    bad = good;
    bad.candidate_rows = far;
    bad.candidate_count = 0;
    CHECK(search(idx, &bad, NULL, 0, &returned, NULL) == ORDVEC_STATUS_BAD_ARGUMENT);
    bad = good;
    bad.candidate_rows = NULL;
    bad.candidate_count = 1;
    CHECK(search(idx, &bad, NULL, 0, &returned, NULL) == ORDVEC_STATUS_NULL_POINTER);
    ordvec_search_stats_t short_stats;
    ordvec_search_stats_init(&short_stats);
    short_stats.struct_size = 8;
    CHECK(search(idx, &good, hits, 1, &returned, &short_stats) == ORDVEC_STATUS_BAD_STRUCT_SIZE);
    /* a subset longer than the index: the reference's assert, reported as PANIC */
    const uint32_t five[5] = {0, 1, 2, 3, 0};
    bad = good;
    bad.candidate_rows = five;
    bad.candidate_count = 5;
    CHECK(search(idx, &bad, hits, 1, &returned, NULL) == ORDVEC_STATUS_PANIC);

    /* zero required hits allows NULL hits (lib.rs:1474-1500) */
    ordvec_search_params_t zero = good;
    zero.k = 0;
    returned = 42;
    CHECK(search(idx, &zero, NULL, 0, &returned, NULL) == ORDVEC_STATUS_OK && returned == 0);
    CHECK(strlen(ordvec_last_error()) == 0);

    ordvec_index_free(idx);
    ordvec_index_free(NULL);

    /* a truncated file is CORRUPT_INDEX and leaves *out NULL (lib.rs:1503-1553) */
    char corrupt[4096];
    snprintf(corrupt, sizeof corrupt, "%s/corrupt.ovrq", argc > 1 ? argv[1] : ".");
    FILE *f = fopen(corrupt, "wb");
    CHECK(f != NULL);
    fwrite("OVRQ\x01", 1, 5, f);
    fclose(f);
    ordvec_index_t *none = (ordvec_index_t *)&info;
    CHECK(ordvec_index_load(corrupt, 0, &none) == ORDVEC_STATUS_CORRUPT_INDEX && none == NULL);
    printf("c_link_smoke: %d checks passed\n", g_check);
    return 0;
}
