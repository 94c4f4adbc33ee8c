"""Register and scratch budgets of the built gfx950 kernels, read from the
code object's metadata notes (no device needed).

The kernels of the batched step (decomposition, inner NTT, fused CRT) must
not spill.  The persistent batched form must stay at its two-waves-per-SIMD
budget: built for three waves per SIMD (168 VGPRs) it spilled about 104
VGPRs and 64 SGPRs to 400-424 bytes of scratch per lane and returned one
wave's share of wrong output in 2-7 % of its launches; at two waves per SIMD
it spills 16 VGPRs once at entry (68 bytes) and returned none in thousands
of checked launches.  A change that brings the heavy spilling back fails
here, before any GPU run.
"""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
OBJ = ROOT / "engel-cuda-negacyclic-multiply_amd" / "_obj" / "nc_kernels.o"
LLVM = Path("/opt/rocm/llvm/bin")


def kernel_resources(tmp_path: Path) -> dict:
    """kernel symbol -> {vgpr, vgpr_spill, sgpr_spill, scratch} from the gfx950 code object."""
    objdump, readelf = LLVM / "llvm-objdump", LLVM / "llvm-readelf"
    if not OBJ.exists() or not objdump.exists() or not readelf.exists():
        pytest.skip("kernel object or LLVM tools not present (build first)")
    obj = tmp_path / "nc_kernels.o"
    shutil.copy(OBJ, obj)  # the extraction writes next to its input
    subprocess.run([str(objdump), "--offloading", str(obj)], check=True, capture_output=True, cwd=tmp_path)
    code = [p for p in tmp_path.iterdir() if "amdgcn-amd-amdhsa--gfx950" in p.name]
    assert len(code) == 1, [p.name for p in tmp_path.iterdir()]
    notes = subprocess.run([str(readelf), "--notes", str(code[0])], check=True, capture_output=True,
                           text=True).stdout
    out = {}
    # one metadata map per kernel; the fields of interest follow its .name in
    # alphabetical order, so each block runs from one "- .agpr_count" to the next
    for block in notes.split("- .agpr_count")[1:]:
        name = re.search(r"\.name:\s+(\S+)", block)
        if not name:
            continue

        def field(key: str) -> int:
            m = re.search(rf"\.{key}:\s+(\d+)", block)
            return int(m.group(1)) if m else 0

        out[name.group(1)] = {"vgpr": field("vgpr_count"), "vgpr_spill": field("vgpr_spill_count"),
                              "sgpr_spill": field("sgpr_spill_count"), "scratch": field("private_segment_fixed_size")}
    return out


def test_batched_step_kernels_do_not_spill(tmp_path):
    res = kernel_resources(tmp_path)
    headline = {
        "decomposition": r"decompose30_mfma_kernelILi2ELi1E",
        "inner NTT": r"ntt_inner30_kernelILi12E",
        "fused CRT": r"crt30_mfma_kernelILi4ELi8ELi28ELi2E",
    }
    for what, pat in headline.items():
        hits = {k: v for k, v in res.items() if re.search(pat, k)}
        assert hits, what
        for k, v in hits.items():
            assert v["vgpr_spill"] == 0 and v["scratch"] == 0, (what, k, v)


def test_persistent_kernel_stays_at_two_waves_budget(tmp_path):
    res = kernel_resources(tmp_path)
    hits = {k: v for k, v in res.items() if "persist30_kernel" in k}
    assert len(hits) == 5, sorted(hits)  # blocks of 2^8 .. 2^12
    for k, v in hits.items():
        assert v["vgpr"] > 168, (k, v)  # built for two waves per SIMD, not three
        assert v["vgpr_spill"] <= 24 and v["scratch"] <= 128, (k, v)
