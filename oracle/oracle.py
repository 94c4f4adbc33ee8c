"""ctypes front end of the CPU oracle (oracle/nc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() (as the
checker) and bench.py (the cpu_baseline leg).  The shipped library never
imports it.

Two independent formulations are reachable from here:
  * ``schoolbook`` / ``schoolbook_coeffs``: the exact big-integer definition
    c_k = sum_{i+j=k} a_i b_j - sum_{i+j=k+N} a_i b_j (mod q).
  * ``Plan``: the C restatement of the reference's RNS path (solve.cu:
    prime rule 1174-1204, tables 1224-1369, run 1489-1528, fast CRT 752-836,
    Garner 838-906).
plus pure-Python helpers (``py_*``) on Python integers that restate neither.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# NC_ORACLE_LIB: another build of the oracle (the sanitizer build,
# libncoracle_asan.so, that tests/test_sanitize.py loads)
LIB_PATH = Path(os.environ["NC_ORACLE_LIB"]) if os.environ.get("NC_ORACLE_LIB") else HERE / "libncoracle.so"

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)

_lib = None


def build() -> None:
    """Compile the oracle library in place (gcc + OpenMP)."""
    import subprocess

    subprocess.run(["make", "-B", "-C", str(HERE)], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.orc_schoolbook.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p, ctypes.c_int]
        L.orc_schoolbook_coeffs.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p,
                                            ctypes.c_uint32, _u32p, ctypes.c_int]
        L.orc_gen_poly.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]
        L.orc_gen_poly.restype = None
        L.orc_plan_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.POINTER(ctypes.c_void_p)]
        L.orc_plan_destroy.argtypes = [ctypes.c_void_p]
        L.orc_plan_destroy.restype = None
        L.orc_plan_channels.argtypes = [ctypes.c_void_p]
        L.orc_plan_channels.restype = ctypes.c_uint32
        L.orc_plan_primes.argtypes = [ctypes.c_void_p, _u64p]
        L.orc_plan_primes.restype = None
        L.orc_residues.argtypes = [ctypes.c_void_p, _u32p, _u64p]
        L.orc_product_residues.argtypes = [ctypes.c_void_p, _u32p, _u32p, _u64p, ctypes.c_int]
        L.orc_crt_fast.argtypes = [ctypes.c_void_p, _u64p, _u32p, ctypes.c_int]
        L.orc_crt_garner.argtypes = [ctypes.c_void_p, _u64p, _u32p, ctypes.c_int]
        L.orc_multiply.argtypes = [ctypes.c_void_p, _u32p, _u32p, _u32p, ctypes.c_int]
        L.orc_crt_decisions.argtypes = [ctypes.c_void_p, _u64p, _u32p, _u8p]
        L.orc_selftest.argtypes = [ctypes.c_void_p]
        for name in ("orc_mul_mod", "orc_pow_mod"):
            getattr(L, name).argtypes = [ctypes.c_uint64] * 3
            getattr(L, name).restype = ctypes.c_uint64
        L.orc_mont_mul.argtypes = [ctypes.c_uint64] * 4
        L.orc_mont_mul.restype = ctypes.c_uint64
        L.orc_shoup_mul.argtypes = [ctypes.c_uint64] * 4
        L.orc_shoup_mul.restype = ctypes.c_uint64
        L.orc_reduce_128.argtypes = [ctypes.c_uint64] * 5
        L.orc_reduce_128.restype = ctypes.c_uint64
        L.orc_barrett_reduce.argtypes = [ctypes.c_uint64] * 4
        L.orc_barrett_reduce.restype = ctypes.c_uint64
        L.orc_barrett_mu.argtypes = [ctypes.c_uint64]
        L.orc_barrett_mu.restype = ctypes.c_uint64
        L.orc_inv_mod_2_64.argtypes = [ctypes.c_uint64]
        L.orc_inv_mod_2_64.restype = ctypes.c_uint64
        L.orc_is_prime_u64.argtypes = [ctypes.c_uint64]
        L.orc_find_psi.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        L.orc_find_psi.restype = ctypes.c_uint64
        L.orc_div_pow2_64.argtypes = [ctypes.c_uint64, _u64p]
        L.orc_div_pow2_64.restype = ctypes.c_uint64
        L.orc_mod_q_barrett.argtypes = [_u32p, _u32p, _u32p]
        L.orc_mod_q_barrett.restype = None
        L.orc_mod_q.argtypes = [_u32p, _u32p, ctypes.c_int, ctypes.c_int]
        L.orc_mod_q.restype = None
        L.orc_barrett_mu_996.argtypes = [_u32p, ctypes.c_int, _u32p]
        L.orc_barrett_mu_996.restype = None
        _lib = L
    return _lib


def _p32(x: np.ndarray):
    assert x.dtype == np.uint32 and x.flags.c_contiguous
    return x.ctypes.data_as(_u32p)


def _p64(x: np.ndarray):
    assert x.dtype == np.uint64 and x.flags.c_contiguous
    return x.ctypes.data_as(_u64p)


# ---------------------------------------------------------------------------
# limb <-> Python int helpers (pure Python, independent of the C code)
# ---------------------------------------------------------------------------

def int_to_limbs(x: int, L: int) -> np.ndarray:
    return np.frombuffer(int(x).to_bytes(4 * L, "little"), dtype="<u4").astype(np.uint32)


def limbs_to_int(v) -> int:
    return int.from_bytes(np.ascontiguousarray(v, dtype="<u4").tobytes(), "little")


def poly_to_ints(poly: np.ndarray) -> list:
    N, L = poly.shape
    raw = np.ascontiguousarray(poly, dtype="<u4").tobytes()
    return [int.from_bytes(raw[4 * L * i: 4 * L * (i + 1)], "little") for i in range(N)]


def ints_to_poly(xs, L: int) -> np.ndarray:
    raw = b"".join(int(x).to_bytes(4 * L, "little") for x in xs)
    return np.frombuffer(raw, dtype="<u4").astype(np.uint32).reshape(len(xs), L)


def py_negacyclic(a_ints, b_ints, q: int) -> list:
    """Exact negacyclic product mod q on Python ints by Kronecker substitution:
    one big-integer multiplication of the packed polynomials, then the
    X^N = -1 fold.  Independent of every C routine in this directory."""
    N = len(a_ints)
    assert len(b_ints) == N
    bound = N * q * q
    w = ((bound.bit_length() + 8) // 8) * 8  # slot width in bits, byte aligned
    wb = w // 8
    A = int.from_bytes(b"".join(int(x).to_bytes(wb, "little") for x in a_ints), "little")
    B = int.from_bytes(b"".join(int(x).to_bytes(wb, "little") for x in b_ints), "little")
    raw = (A * B).to_bytes(2 * N * wb, "little")
    lin = [int.from_bytes(raw[wb * k: wb * (k + 1)], "little") for k in range(2 * N)]
    return [(lin[k] - lin[k + N]) % q for k in range(N)]


def py_negacyclic_exact(a_ints, b_ints) -> list:
    """The exact signed integer negacyclic product c_int (no reduction mod q),
    by the same Kronecker substitution as py_negacyclic."""
    N = len(a_ints)
    bound = N * max(max(a_ints), 1) * max(max(b_ints), 1)
    w = ((bound.bit_length() + 8) // 8) * 8
    wb = w // 8
    A = int.from_bytes(b"".join(int(x).to_bytes(wb, "little") for x in a_ints), "little")
    B = int.from_bytes(b"".join(int(x).to_bytes(wb, "little") for x in b_ints), "little")
    raw = (A * B).to_bytes(2 * N * wb, "little")
    lin = [int.from_bytes(raw[wb * k: wb * (k + 1)], "little") for k in range(2 * N)]
    return [lin[k] - lin[k + N] for k in range(N)]


def py_negacyclic_loops(a_ints, b_ints, q: int) -> list:
    """Direct double loop over the definition (tiny N only)."""
    N = len(a_ints)
    out = []
    for k in range(N):
        s = 0
        for i in range(N):
            j = k - i
            if j >= 0:
                s += a_ints[i] * b_ints[j]
            else:
                s -= a_ints[i] * b_ints[j + N]
        out.append(s % q)
    return out


# ---------------------------------------------------------------------------
# C oracle entry points
# ---------------------------------------------------------------------------

def gen_poly(seed: int, N: int, q: np.ndarray) -> np.ndarray:
    L = q.shape[0]
    out = np.empty((N, L), dtype=np.uint32)
    lib().orc_gen_poly(ctypes.c_uint64(seed), N, L, _p32(q), _p32(out))
    return out


def schoolbook(a: np.ndarray, b: np.ndarray, q: np.ndarray, threads: int = 0) -> np.ndarray:
    N, L = a.shape
    c = np.empty_like(a)
    rc = lib().orc_schoolbook(N, L, _p32(q), _p32(a), _p32(b), _p32(c), threads)
    if rc:
        raise RuntimeError(f"orc_schoolbook: {rc}")
    return c


def schoolbook_coeffs(a, b, q, idx, threads: int = 0) -> np.ndarray:
    N, L = a.shape
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    out = np.empty((idx.shape[0], L), dtype=np.uint32)
    rc = lib().orc_schoolbook_coeffs(N, L, _p32(q), _p32(a), _p32(b), _p32(idx), idx.shape[0], _p32(out), threads)
    if rc:
        raise RuntimeError(f"orc_schoolbook_coeffs: {rc}")
    return out


class Plan:
    """Restated reference plan (prime rule + tables) for one (N, q)."""

    def __init__(self, N: int, q: np.ndarray):
        self.N = int(N)
        self.L = int(q.shape[0])
        self.q = np.ascontiguousarray(q, dtype=np.uint32)
        h = ctypes.c_void_p()
        rc = lib().orc_plan_create(self.N, self.L, _p32(self.q), ctypes.byref(h))
        if rc:
            raise ValueError(f"orc_plan_create failed: {rc}")
        self._h = h
        self.k = int(lib().orc_plan_channels(h))

    def close(self):
        if self._h:
            lib().orc_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def primes(self) -> np.ndarray:
        out = np.empty(self.k, dtype=np.uint64)
        lib().orc_plan_primes(self._h, _p64(out))
        return out

    def selftest(self) -> int:
        return int(lib().orc_selftest(self._h))

    def residues(self, poly: np.ndarray) -> np.ndarray:
        out = np.empty((self.k, self.N), dtype=np.uint64)
        lib().orc_residues(self._h, _p32(poly), _p64(out))
        return out

    def product_residues(self, a, b, threads: int = 0) -> np.ndarray:
        out = np.empty((self.k, self.N), dtype=np.uint64)
        rc = lib().orc_product_residues(self._h, _p32(a), _p32(b), _p64(out), threads)
        if rc:
            raise RuntimeError(rc)
        return out

    def crt_fast(self, res: np.ndarray, threads: int = 0) -> np.ndarray:
        c = np.empty((self.N, self.L), dtype=np.uint32)
        lib().orc_crt_fast(self._h, _p64(np.ascontiguousarray(res)), _p32(c), threads)
        return c

    def crt_garner(self, res: np.ndarray, threads: int = 0) -> np.ndarray:
        c = np.empty((self.N, self.L), dtype=np.uint32)
        lib().orc_crt_garner(self._h, _p64(np.ascontiguousarray(res)), _p32(c), threads)
        return c

    def crt_decisions(self, res: np.ndarray):
        hi = np.empty(self.N, dtype=np.uint32)
        neg = np.empty(self.N, dtype=np.uint8)
        lib().orc_crt_decisions(self._h, _p64(np.ascontiguousarray(res)), _p32(hi), neg.ctypes.data_as(_u8p))
        return hi, neg

    def multiply(self, a, b, threads: int = 0) -> np.ndarray:
        c = np.empty((self.N, self.L), dtype=np.uint32)
        rc = lib().orc_multiply(self._h, _p32(a), _p32(b), _p32(c), threads)
        if rc:
            raise RuntimeError(rc)
        return c


def cpu_count() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1
