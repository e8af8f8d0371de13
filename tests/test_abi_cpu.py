"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, struct layouts match ordvec-ffi, and argument validation
that happens before any device work returns the reference's status codes
(ordvec-ffi/src/lib.rs:1061-1066, 1274-1472)."""
import ctypes as C
import os
import re

import pytest

from ordvec_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ordvec.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ordvec_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = N.lib()
    names = header_functions()
    assert len(names) >= 40
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.SIGNATURES), set(names) ^ set(N.SIGNATURES)


def test_struct_sizes_match_abi_v1():
    assert C.sizeof(N.IndexInfo) == 128
    assert C.sizeof(N.SearchParams) == 128
    assert C.sizeof(N.Hit) == 24
    assert C.sizeof(N.SearchStats) == 184
    assert C.sizeof(N.ServeReport) == 192


def test_versions_and_status_names():
    lib = N.lib()
    assert lib.ordvec_abi_version() == 1
    assert lib.ordvec_version_string() == b"0.6.0"
    names = [lib.ordvec_status_name(i).decode() for i in range(15)]
    assert names == ["OK", "NULL_POINTER", "BAD_ARGUMENT", "BAD_STRUCT_SIZE", "UNSUPPORTED_FORMAT",
                     "CORRUPT_INDEX", "IO", "DIM_MISMATCH", "NONFINITE_QUERY", "ROW_ID_OUT_OF_RANGE",
                     "BUFFER_TOO_SMALL", "UNSUPPORTED_OPERATION", "PANIC", "INTERNAL", "UNKNOWN"]


def test_init_defaults():
    lib = N.lib()
    info, params, stats = N.IndexInfo(), N.SearchParams(), N.SearchStats()
    lib.ordvec_index_info_init(C.byref(info))
    lib.ordvec_search_params_init(C.byref(params))
    lib.ordvec_search_stats_init(C.byref(stats))
    assert info.struct_size == 128 and params.struct_size == 128 and stats.struct_size == 184
    assert stats.abi_version == 1 and params.flags == 0 and params.k == 0
    lib.ordvec_index_info_init(None)  # NULL is a no-op


def test_load_and_probe_status_paths(tmp_path):
    lib = N.lib()
    out = C.c_void_p()
    assert lib.ordvec_index_load(None, 0, C.byref(out)) == N.NULL_POINTER
    assert "path" in N.last_error()
    p = tmp_path / "x.ovrq"
    p.write_bytes(b"OVRQ\x01\x02" + (8).to_bytes(4, "little") + (1).to_bytes(4, "little") + b"\x05\xaf")
    assert lib.ordvec_index_load(str(p).encode(), 1, C.byref(out)) == N.BAD_ARGUMENT
    assert lib.ordvec_index_load(str(tmp_path / "missing").encode(), 0, C.byref(out)) == N.IO
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"ZZZZ\x01")
    assert lib.ordvec_index_load(str(bad).encode(), 0, C.byref(out)) == N.CORRUPT_INDEX
    sb = tmp_path / "a.ovsb"
    sb.write_bytes(b"OVSB\x01" + (64).to_bytes(4, "little") + (1).to_bytes(4, "little") + bytes(8))
    assert lib.ordvec_index_load(str(sb).encode(), 0, C.byref(out)) == N.UNSUPPORTED_FORMAT
    # probe: metadata only, no device work
    info = N.IndexInfo()
    lib.ordvec_index_info_init(C.byref(info))
    assert lib.ordvec_index_probe(str(p).encode(), 0, C.byref(info)) == N.OK
    assert (info.kind, info.dim, info.bit_width, info.vector_count, info.bytes_per_vec) == (1, 8, 2, 1, 2)
    assert info.source_file_size_bytes == 16 and info.capabilities == 15
    info.struct_size = 127
    assert lib.ordvec_index_probe(str(p).encode(), 0, C.byref(info)) == N.BAD_STRUCT_SIZE
    info.struct_size = 128
    trailing = tmp_path / "t.ovrq"
    trailing.write_bytes(p.read_bytes() + b"\x00")
    assert lib.ordvec_index_probe(str(trailing).encode(), 0, C.byref(info)) == N.CORRUPT_INDEX
    assert "trailing" in N.last_error()
    assert lib.ordvec_index_probe(str(sb).encode(), 0, C.byref(info)) == N.UNSUPPORTED_FORMAT
    bm = tmp_path / "a.ovbm"
    bm.write_bytes(b"OVBM\x01" + (64).to_bytes(4, "little") + (2).to_bytes(4, "little") +
                   (1).to_bytes(4, "little") + bytes(7) + b"\xc0")
    assert lib.ordvec_index_probe(str(bm).encode(), 0, C.byref(info)) == N.OK
    assert (info.kind, info.n_top, info.bytes_per_vec) == (2, 2, 8)
    # .ovbm loads through ABI v1; a row without exactly n_top bits is corrupt
    # (src/rank_io.rs:938-951), found before any device work.
    forged = tmp_path / "forged.ovbm"
    forged.write_bytes(bm.read_bytes()[:-1] + b"\xe0")
    assert lib.ordvec_index_load(str(forged).encode(), 0, C.byref(out)) == N.CORRUPT_INDEX
    assert "bits set" in N.last_error()
    assert lib.ordvec_bitmap_load(str(forged).encode(), C.byref(out)) == N.CORRUPT_INDEX
    zero_top = tmp_path / "z.ovbm"
    zero_top.write_bytes(b"OVBM\x01" + (64).to_bytes(4, "little") + (0).to_bytes(4, "little") +
                         (0).to_bytes(4, "little"))
    assert lib.ordvec_index_load(str(zero_top).encode(), 0, C.byref(out)) == N.CORRUPT_INDEX
    v2 = tmp_path / "v2.ovrq"
    v2.write_bytes(b"OVRQ\x02" + p.read_bytes()[5:])
    assert lib.ordvec_index_probe(str(v2).encode(), 0, C.byref(info)) == N.CORRUPT_INDEX
    assert "version" in N.last_error()


def test_parameter_validation_needs_no_device():
    lib = N.lib()
    h = C.c_void_p()
    assert lib.ordvec_rankquant_new(3, 2, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_rankquant_new(64, 3, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_rankquant_new(1, 1, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_rankquant_new(70000, 1, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_sign_bitmap_new(100, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_fastscan_new(6, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_bitmap_new(100, 4, C.byref(h)) == N.BAD_ARGUMENT  # dim % 64
    assert lib.ordvec_bitmap_new(64, 0, C.byref(h)) == N.BAD_ARGUMENT   # 0 < n_top < dim
    assert lib.ordvec_bitmap_new(64, 64, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_rankquant_new(64, 2, None) == N.NULL_POINTER
    assert lib.ordvec_index_search(None, None, None, 0, None, None) == N.NULL_POINTER
    lib.ordvec_index_free(None)
    lib.ordvec_rankquant_free(None)
    assert lib.ordvec_server_start(None, None, 1000, 10, 1, 64, 1 << 30, C.byref(h)) == N.NULL_POINTER
    assert lib.ordvec_server_submit(None, None, 1, 10, None, None, None) == N.NULL_POINTER
    assert lib.ordvec_server_wait(None, 1, None, None) == N.NULL_POINTER
    lib.ordvec_server_stop(None)


def test_python_api_validation():
    import ordvec_amd
    from ordvec_amd.index import _float_rows

    with pytest.raises(ValueError):
        ordvec_amd.RankQuant(3, 2)
    with pytest.raises(ValueError):
        _float_rows(__import__("numpy").zeros((2, 5), "float32"), 4, "vectors")
    with pytest.raises(ValueError):
        _float_rows(__import__("numpy").zeros((2, 4), "int32"), 4, "vectors")


def test_rank_validation_and_forged_ovr_need_no_device(tmp_path):
    """Rank::new bounds (src/rank.rs:317-324) and the .ovr loader's checks
    (src/rank_io.rs:580-640) run before any device work."""
    lib = N.lib()
    h = C.c_void_p()
    assert lib.ordvec_rank_new(1, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_rank_new(65536, C.byref(h)) == N.BAD_ARGUMENT
    assert lib.ordvec_rank_new(64, None) == N.NULL_POINTER
    assert lib.ordvec_rank_swap_remove(None, 0, None) == N.NULL_POINTER
    assert lib.ordvec_rankquant_swap_remove(None, 0, None) == N.NULL_POINTER
    assert lib.ordvec_rank_search(None, None, 1, 1, None, None, 0) == N.NULL_POINTER
    lib.ordvec_rank_free(None)

    def hdr(dim, n, magic=b"OVR1", version=1):
        return magic + bytes([version]) + dim.to_bytes(4, "little") + n.to_bytes(4, "little")

    def rows(*rs):
        return b"".join(int(v).to_bytes(2, "little") for r in rs for v in r)

    cases = [
        (hdr(4, 1) + rows([0, 1, 2, 4]), "4 >= dim"),
        (hdr(4, 2) + rows([0, 1, 2, 3], [3, 1, 1, 0]), "row 1 is not a permutation"),
        (hdr(4, 1) + rows([0, 1, 2, 3])[:-1], "truncated"),
        (hdr(4, 1) + rows([0, 1, 2, 3]) + b"\x00", "trailing"),
        (hdr(1, 0), "out of range"),
        (hdr(4, 1, version=2) + rows([0, 1, 2, 3]), "version"),
        (hdr(4, 1, magic=b"OVRQ") + rows([0, 1, 2, 3]), "wrong magic"),
        (b"OVR1\x01\x04\x00", "truncated"),
    ]
    for j, (body, msg) in enumerate(cases):
        p = tmp_path / f"f{j}.ovr"
        p.write_bytes(body)
        assert lib.ordvec_rank_load(str(p).encode(), C.byref(h)) == N.CORRUPT_INDEX, j
        assert msg in N.last_error(), (j, N.last_error())
    assert lib.ordvec_rank_load(str(tmp_path / "missing.ovr").encode(), C.byref(h)) == N.IO


def test_two_stage_candidate_policy_kats():
    """tests/index/two_stage.rs:76-89 (N = 256), through the C ABI."""
    import ordvec_amd as ov

    n = 256
    default = ov.TwoStageCandidatePolicy()
    assert default.candidate_count(0, n) == 0
    assert default.candidate_count(1, n) == min(256, n)
    assert default.candidate_count(10, n) == n
    assert default.candidate_count(10, 10**6) == 320
    assert default.candidate_count(5, 0) == 0
    capped = ov.TwoStageCandidatePolicy(min_candidates=4, k_multiplier=2**64 - 1, max_candidates=37)
    assert capped.candidate_count(2**64 - 1, n) == 37  # k * multiplier saturates
    assert capped.candidate_count(10, 9) == 9
    assert ov.TwoStageCandidatePolicy(0, 4, 37).candidate_count(5, n) == 20
    assert ov.TwoStageCandidatePolicy(0, 0).candidate_count(3, 10) == 0
    assert ov.TwoStageCandidatePolicy(7, 0, 0).candidate_count(3, 10) == 0  # Some(0) caps at zero
    # NULL policy = the default
    assert N.lib().ordvec_two_stage_policy_candidate_count(None, 10, 10**6) == 320
    p = N.TwoStagePolicy()
    N.lib().ordvec_two_stage_policy_default(C.byref(p))
    assert (p.min_candidates, p.k_multiplier, p.has_max_candidates) == (256, 32, 0)
