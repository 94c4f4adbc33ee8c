"""The drop-in boundary: fherma_init / fherma_run / fherma_free.

CPU: the shim library exports exactly the three C++ symbols the platform
harness calls (fherma.h:45-47, mangled as the reference's main.cpp expects),
and the point-directory generator writes the harness layout.
GPU: the repository's runner (bin/nc_solution) and the reference's own
unchanged main.cpp linked against the shim (oracle/_ref/fherma_main, built
from /root/reference by `make -C oracle refharness`) both produce c.bin
bit-identical to the CPU oracle and to the closed forms of the edge cases.
"""
from __future__ import annotations

import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import make_point
import ncmul
import oracle as O
from conftest import ROOT, modulus

SHIM = ncmul.PKG_DIR / "lib" / "libfherma_negacyclic.so"
RUNNER = ncmul.PKG_DIR / "bin" / "nc_solution"
REF_MAIN = ROOT / "oracle" / "_ref" / "fherma_main"


def test_shim_exports_fherma_symbols():
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", str(SHIM)], capture_output=True, text=True,
                         check=True).stdout
    assert re.search(r"\bfherma_init\(fherma::Point const&\)", out)
    assert re.search(r"\bfherma_run\(void\*, fherma::Inputs const&\)", out)
    assert re.search(r"\bfherma_free\(void\*\)", out)


def test_make_point_layout(tmp_path):
    n = make_point.write_point(tmp_path, 16, (1 << 868) - 467, 2, 5, True)
    assert n == 5
    man = json.loads((tmp_path / "manifest.json").read_text())
    assert man == {"N": 16, "W": 868, "L": 28, "cases": 5}
    assert (tmp_path / "point" / "q.bin").stat().st_size == 28 * 4
    for i in range(5):
        for f in ("a.bin", "b.bin"):
            assert (tmp_path / "cases" / f"{i:06d}" / f).stat().st_size == 16 * 28 * 4
    q = modulus("q868")
    a3 = np.fromfile(tmp_path / "cases" / "000003" / "a.bin", dtype="<u4").reshape(16, 28)
    assert np.array_equal(a3, O.gen_poly(10, 16, q))


def run_point(binary: Path, point: Path, N: int, ncases: int):
    r = subprocess.run([str(binary), str(point)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    res = json.loads((point / "out" / "results.json").read_text())
    assert [c["status"] for c in res["cases"]] == ["ok"] * ncases
    assert res["init_s"] >= 0
    q = modulus("q868")
    plan = O.Plan(N, q)
    for i in range(ncases):
        d = point / "cases" / f"{i:06d}"
        a = np.fromfile(d / "a.bin", dtype="<u4").reshape(N, 28)
        b = np.fromfile(d / "b.bin", dtype="<u4").reshape(N, 28)
        c = np.fromfile(point / "out" / f"{i:06d}" / "c.bin", dtype="<u4").reshape(N, 28)
        assert np.array_equal(c, plan.multiply(a, b)), i
    return res


@pytest.mark.gpu
@pytest.mark.parametrize("N", [1024, 16384])
def test_runner_point_directory(gpu, tmp_path, N):
    n = make_point.write_point(tmp_path, N, (1 << 868) - 467, 2, 1, True)
    print("runner", N, flush=True)
    run_point(RUNNER, tmp_path, N, n)
    qi = (1 << 868) - 467
    c0 = np.fromfile(tmp_path / "out" / "000000" / "c.bin", dtype="<u4").reshape(N, 28)
    assert O.limbs_to_int(c0[0]) == (2 - N) % qi  # all q-1: c_k = 2k + 2 - N


@pytest.mark.gpu
def test_reference_main_links_and_runs(gpu, tmp_path):
    """The reference's generated harness, unchanged, over this library."""
    if not REF_MAIN.exists():
        pytest.skip("oracle/_ref/fherma_main not built (needs /root/reference at build time)")
    n = make_point.write_point(tmp_path, 4096, (1 << 868) - 467, 2, 7, True)
    print("reference main.cpp harness", flush=True)
    run_point(REF_MAIN, tmp_path, 4096, n)


@pytest.mark.gpu
def test_runner_records_crashed_case(gpu, tmp_path):
    n = make_point.write_point(tmp_path, 64, (1 << 868) - 467, 2, 3, False)
    bad = tmp_path / "cases" / "000001" / "a.bin"
    bad.write_bytes(bad.read_bytes()[:-4])  # wrong size: the case fails, the run goes on
    r = subprocess.run([str(RUNNER), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    res = json.loads((tmp_path / "out" / "results.json").read_text())
    assert [c["status"] for c in res["cases"]] == ["ok", "crashed"]
    assert res["cases"][1]["seconds"] is None
