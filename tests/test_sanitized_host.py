"""The host C of the library and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

Both are rebuilt with -fsanitize=address,undefined (opengraph_amd/Makefile.san
over the ordinary kernel objects, `make -C oracle san`) and every other CPU
test runs again against those builds in a child process with libasan
preloaded: the .csr reader and writer and their validation, the edge-list
generator, the oracle's CSR build, PageRank variants, SpMV, relabelling,
rank statistics and MLP, and the library's exported symbols.  Any
sanitizer report fails the child and this test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "opengraph_amd", "_build_san", "libopengraph_amd_san.so")
SAN_ORACLE = os.path.join(ROOT, "oracle", "_ref", "libogoracle_san.so")


def _libasan():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                           timeout=30).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_cpu_suite_under_asan_and_ubsan():
    asan = _libasan()
    if asan is None:
        pytest.skip("gcc has no AddressSanitizer runtime here")
    kernels = os.path.join(ROOT, "opengraph_amd", "_build")
    if not os.path.isdir(kernels) or not any(f.endswith(".hip.o") for f in os.listdir(kernels)):
        pytest.skip("kernel objects not built (make -C opengraph_amd first)")
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-C", os.path.join(ROOT, "opengraph_amd"), "-f", "Makefile.san", "-j", jobs],
                   check=True, capture_output=True, timeout=600)
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "san"], check=True, capture_output=True,
                   timeout=600)
    env = dict(os.environ, LD_PRELOAD=asan, OGPU_LIB=SAN_LIB, OG_ORACLE_LIB=SAN_ORACLE,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        "--ignore", os.path.abspath(__file__), os.path.join(ROOT, "tests")],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert r.returncode == 0, out[-3000:]
    assert "AddressSanitizer" not in out and "runtime error:" not in out
    assert " passed" in out
