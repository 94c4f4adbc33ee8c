"""Host sanitizer run (CPU): the library's host code (table builder, init
self tests, C ABI, batch and server argument handling) and the oracle,
built with -fsanitize=address,undefined (`make asan` in the package and in
oracle/), run the host-side CPU tests with gcc's ASan runtime preloaded; any
sanitizer report fails the test.  The reference's CPU code is the analogue:
its tables and CRT constants are built on the host (solve.cu:1150-1487), and
solve.cu:966-970 against 1003 shows the kind of latent memory error this
run is for."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

from conftest import PKG, ROOT


def _tool_path(compiler: str, name: str) -> str:
    return subprocess.run([compiler, f"-print-file-name={name}"], capture_output=True, text=True,
                          check=True).stdout.strip()


def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    if not (PKG / "_obj" / "nc_kernels.o").exists():
        pytest.skip("the gfx950 kernel object is not built (run __graft_entry__.build())")
    for d in (PKG, ROOT / "oracle"):
        subprocess.run(["make", "-C", str(d), "asan"], check=True, capture_output=True, timeout=600)
    asan = _tool_path("gcc", "libasan.so")
    stdcxx = _tool_path("g++", "libstdc++.so.6")  # preloaded too, so the runtime's C++ interceptors resolve
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{asan} {stdcxx}",
        "ASAN_OPTIONS": f"detect_leaks=0:halt_on_error=1:log_path={tmp_path / 'asan'}",
        "UBSAN_OPTIONS": f"halt_on_error=1:print_stacktrace=1:log_path={tmp_path / 'ubsan'}",
        "NC_LIB_PATH": str(PKG / "lib" / "libnegacyclic_asan.so"),
        "NC_ORACLE_LIB": str(ROOT / "oracle" / "libncoracle_asan.so"),
    })
    files = [str(ROOT / "tests" / f) for f in ("test_abi.py", "test_oracle.py", "test_twiddles.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", *files, "-q", "-m", "not gpu", "-p", "no:cacheprovider"],
                       env=env, capture_output=True, text=True, timeout=900, cwd=str(ROOT))
    reports = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith(("asan", "ubsan")))
    detail = "".join((tmp_path / n).read_text()[:3000] for n in reports)
    assert not reports, detail
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert " passed" in r.stdout
