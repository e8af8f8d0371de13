"""A C program built with `cc -std=c11` against include/ordvec.h and linked
with libordvec_amd.so (ordvec-ffi/tests/c_link_smoke.rs:158-170): compiling
and linking needs no GPU; running it (the reference FFI fixture, full and
subset search, every status path) does."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ordvec_amd", "lib", "libordvec_amd.so")
SRC = os.path.join(ROOT, "tests", "c_link_smoke.c")


def _build(tmp_path):
    cc = shutil.which("cc") or shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "c_link_smoke")
    cmd = [cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC, LIB,
           "-Wl,-rpath," + os.path.dirname(LIB), "-lm", "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


def test_c_program_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_c_program_runs_against_the_library(tmp_path):
    exe = _build(tmp_path)
    res = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, (res.returncode, res.stdout, res.stderr)
    assert "checks passed" in res.stdout
