"""CPU tests of the C ABI library: it loads, exports every symbol that
include/negacyclic.h declares, validates arguments before touching a device,
and its synthetic-input generator matches the oracle's and the golden inputs.
No compute call needs a GPU here."""
from __future__ import annotations

import ctypes
import hashlib
import re
import subprocess

import numpy as np
import pytest

import ncmul
import oracle as O
from conftest import ROOT, load_golden, modulus


def declared(header):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared(ncmul.HEADER)
    assert "nc_multiply" in names and "nc_plan_create" in names
    lib = ncmul.lib()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", str(ncmul.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_every_header_declares_only_what_is_exported():
    for h in (ROOT / "include").glob("*.h"):
        for n in declared(h):
            assert hasattr(ncmul.lib(), n), f"{h.name}: {n}"


def test_status_strings():
    lib = ncmul.lib()
    for code in ncmul.STATUS:
        assert lib.nc_status_string(code)
    assert lib.nc_status_string(0) == b"ok"


@pytest.mark.parametrize("N,L,qbits,err", [
    (3, 28, 868, -1), (0, 28, 868, -1), (1, 28, 868, -1), (1 << 17, 28, 868, -1),
    (1024, 0, 868, -1), (1024, 33, 868, -1),
])
def test_plan_create_rejects_bad_shapes(N, L, qbits, err):
    q = np.zeros(max(L, 1), dtype=np.uint32)
    q[0] = 3
    h = ctypes.c_void_p()
    rc = ncmul.lib().nc_plan_create(N, L, q.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(h))
    assert rc == err
    assert not h.value
    assert ncmul.lib().nc_last_error()


def test_plan_create_rejects_even_q_and_wide_q():
    h = ctypes.c_void_p()
    q = modulus("q868").copy()
    q[0] &= ~np.uint32(1)
    p32 = ctypes.POINTER(ctypes.c_uint32)
    assert ncmul.lib().nc_plan_create(1024, 28, q.ctypes.data_as(p32), ctypes.byref(h)) == -1
    wide = np.full(32, 0xFFFFFFFF, dtype=np.uint32)  # 1024-bit q: beyond bits(q)+70 <= 1088
    assert ncmul.lib().nc_plan_create(1024, 32, wide.ctypes.data_as(p32), ctypes.byref(h)) == -1
    assert b"1018" in ncmul.lib().nc_last_error()
    assert ncmul.lib().nc_plan_create(1024, 28, None, ctypes.byref(h)) == -1


def test_destroy_null_is_safe():
    ncmul.lib().nc_plan_destroy(None)


@pytest.mark.parametrize("name", ["q868", "q868b", "q200"])
def test_generator_matches_oracle_and_golden(name):
    q = modulus(name)
    for N, seed in [(1, 0), (7, 123), (1024, 2)]:
        assert np.array_equal(ncmul.gen_uniform(seed, N, q), O.gen_poly(seed, N, q))
    for case in load_golden()["cases"]:
        if case["modulus"] == name and case["N"] <= 16384:
            a = ncmul.gen_uniform(2 * case["seed"], case["N"], q)
            assert hashlib.sha256(a.astype("<u4").tobytes()).hexdigest() == case["sha256_a"]


@pytest.mark.parametrize("N,L,bits", [
    (16384, 28, 868), (1024, 7, 200), (64, 1, 13), (64, 4, 120), (64, 13, 414), (32, 27, 855), (64, 32, 1017),
    (2, 28, 868), (65536, 28, 868),
])
def test_check_tables_host_self_tests(N, L, bits):
    """Host tables of both prime sets and the init known-answer tests,
    including the host emulation of the int8 matrix-core CRT contraction,
    with no device."""
    import random
    r = random.Random(N * 131 + L)
    qi = (1 << 868) - 467 if bits == 868 else (1 << (bits - 1)) | r.getrandbits(bits - 1) | 1
    ncmul.check_tables(N, O.int_to_limbs(qi, L))


def test_check_tables_rejects():
    with pytest.raises(ncmul.NcError) as e:
        ncmul.check_tables(1000, modulus("q868"))
    assert e.value.status == -1
    with pytest.raises(ncmul.NcError):
        ncmul.check_tables(1024, np.full(32, 0xFFFFFFFF, dtype=np.uint32))


def test_batch_and_server_argument_checks_need_no_device():
    """Argument validation of the batch / server ABI fails before any device call."""
    lib = ncmul.lib()
    h = ctypes.c_void_p()
    q = modulus("q868")
    p32 = ctypes.POINTER(ctypes.c_uint32)
    assert lib.nc_batch_create(28, None, ctypes.byref(h)) == -1
    assert lib.nc_batch_create(0, q.ctypes.data_as(p32), ctypes.byref(h)) == -1
    assert lib.nc_batch_create(28, q.ctypes.data_as(p32), None) == -1
    assert lib.nc_server_create(None, ctypes.byref(h)) == -1
    cfg = ncmul.ServerConfig(28, q.ctypes.data_as(p32), None, 0, 8, 2, 1 << 30, 0)
    assert lib.nc_server_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert not h.value
    for f in (lib.nc_batch_destroy, lib.nc_graph_destroy, lib.nc_server_destroy, lib.nc_host_free):
        f(None)  # null safe
    assert lib.nc_graph_launch(None, None) == -1
    assert lib.nc_server_submit(None, 1024, None, None, None, None) == -1
