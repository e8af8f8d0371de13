"""The library's host code under AddressSanitizer + UBSan, on the CPU: the
index loaders against a forged-file corpus (tests/sanitize/loader_corpus.c,
after the reference's tests/index/main.rs:100-273 and its load_* fuzz
targets), and the Python-level ABI checks run against the sanitizer build.
The reference runs its loaders under ASAN the same way
(.github/workflows/sanitizers.yml)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OUT = os.path.join(ROOT, "_build", "asan")
CLANG = "/opt/rocm/llvm/bin/clang"


@pytest.fixture(scope="module")
def asan_runtime():
    if not os.path.exists(CLANG):
        pytest.skip("ROCm clang (sanitizer runtime) not installed")
    if not all(os.path.exists(os.path.join(ROOT, "_build", f + ".o"))
               for f in ("kernels.hip", "fastscan4.hip", "segsort.hip")):
        from ordvec_amd import _build

        _build.build(force=False)
    r = subprocess.run(["make", "-s", "-C", SAN, "-j4"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    rt = subprocess.run(["make", "-s", "-C", SAN, "runtime"], capture_output=True, text=True).stdout.strip()
    assert os.path.exists(rt), rt
    return rt


def test_loader_corpus_under_asan_ubsan(asan_runtime):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(OUT, "loader_corpus"), "400"], cwd="/tmp", env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:] + r.stderr[-4000:])
    assert " 0 failures" in r.stdout and "ERROR: AddressSanitizer" not in r.stderr


def test_abi_checks_under_asan_ubsan(asan_runtime):
    so = os.path.join(OUT, "libordvec_amd_asan.so")
    env = dict(os.environ, LD_PRELOAD=asan_runtime, ORDVEC_AMD_LIB=so,
               ASAN_OPTIONS="detect_leaks=0:alloc_dealloc_mismatch=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    probe = ("from ordvec_amd import _native as N; N.lib(); "
             "print(any(l.rstrip().endswith('libordvec_amd_asan.so') for l in open('/proc/self/maps')))")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("True"), (r.stdout + r.stderr)[-3000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "tests/test_abi_cpu.py", "tests/test_primitives_cpu.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:] + r.stderr[-3000:])
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in r.stderr
