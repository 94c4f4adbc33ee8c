"""Python front end (ctypes) of lib/libnegacyclic.so, the gfx950 library.

Used by the tests, bench.py and __graft_entry__; the product itself is the
C ABI in include/negacyclic.h.  Device buffers are torch tensors (torch is
only the allocator and stream provider here).  Every call goes to the HIP
library; if the library is missing this module raises instead of falling
back to anything else.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
# NC_LIB_PATH: another build of the same library (the sanitizer build,
# lib/libnegacyclic_asan.so, that tests/test_sanitize.py loads)
LIB_PATH = Path(os.environ["NC_LIB_PATH"]) if os.environ.get("NC_LIB_PATH") else PKG_DIR / "lib" / "libnegacyclic.so"
HEADER = PKG_DIR.parent / "include" / "negacyclic.h"

NC_OK = 0
NC_ERR_UNSUPPORTED = -9  # a form switched off in this build (nc_plan_set_persist mode 1)
STATUS = {
    0: "NC_OK", -1: "NC_ERR_INVALID", -2: "NC_ERR_SIZE", -3: "NC_ERR_DEVICE", -4: "NC_ERR_NOMEM",
    -5: "NC_ERR_PRIMES", -6: "NC_ERR_SELFTEST", -7: "NC_ERR_CAPACITY", -8: "NC_ERR_STATE",
    -9: "NC_ERR_UNSUPPORTED",
}

MFMA_DECOMPOSE = 1  # int8 matrix-core paths of the 30-bit prime set (nc_plan_set_mfma mask)
MFMA_CRT = 2
MFMA_ALL = 3

NOALLOC_SCRATCH_BYTES = 32 << 20  # NC_NOALLOC_SCRATCH_BYTES (negacyclic.h)

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p


class NcError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {detail}")
        self.status = status


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_uint32), ("L", ctypes.c_uint32), ("q_bits", ctypes.c_uint32), ("k62", ctypes.c_uint32),
        ("k30", ctypes.c_uint32), ("width", ctypes.c_uint32), ("log_blocks", ctypes.c_uint32),
        ("mfma_paths", ctypes.c_uint32), ("table_bytes", ctypes.c_uint64), ("bytes_per_product", ctypes.c_uint64),
    ]


class Group(ctypes.Structure):
    """nc_group: `count` products of degree N, contiguous device operands."""
    _fields_ = [("N", ctypes.c_uint32), ("count", ctypes.c_uint32), ("d_a", _vp), ("d_b", _vp), ("d_c", _vp)]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("L", ctypes.c_uint32), ("q", _u32p), ("degrees", _u32p), ("n_degrees", ctypes.c_uint32),
        ("max_batch", ctypes.c_uint32), ("lanes_per_degree", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
        ("autotune", ctypes.c_uint32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "launches", "max_batch_seen", "peak_bytes",
                                                "admission_waits", "device_bytes", "max_turn_gap_us",
                                                "max_enqueue_us", "max_batch_us", "mapped_requests")]


_lib = None


def lib():
    """Load the HIP library (raises FileNotFoundError if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}`")
        L = ctypes.CDLL(str(LIB_PATH))
        L.nc_status_string.restype = ctypes.c_char_p
        L.nc_status_string.argtypes = [ctypes.c_int]
        L.nc_last_error.restype = ctypes.c_char_p
        L.nc_last_error.argtypes = []
        L.nc_plan_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.POINTER(_vp)]
        L.nc_check_tables.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p]
        L.nc_host_twiddles.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_uint32, ctypes.c_int, _vp,
                                       ctypes.POINTER(ctypes.c_size_t)]
        L.nc_plan_copy_twiddles.argtypes = [_vp, ctypes.c_uint32, ctypes.c_int, _vp, ctypes.c_size_t]
        L.nc_tableset_create.argtypes = [ctypes.c_uint32, _u32p, ctypes.POINTER(_vp)]
        L.nc_tableset_destroy.argtypes = [_vp]
        L.nc_tableset_destroy.restype = None
        L.nc_tableset_get_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_float)]
        L.nc_plan_create_shared.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.nc_plan_destroy.argtypes = [_vp]
        L.nc_plan_destroy.restype = None
        L.nc_plan_get_info.argtypes = [_vp, ctypes.POINTER(PlanInfo)]
        L.nc_plan_primes.argtypes = [_vp, ctypes.c_uint32, _u64p]
        L.nc_plan_set_tuning.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32]
        L.nc_plan_set_mfma.argtypes = [_vp, ctypes.c_int]
        L.nc_plan_get_choice.argtypes = [_vp, ctypes.c_uint32] + [ctypes.POINTER(ctypes.c_uint32)] * 4
        L.nc_plan_set_chunks.argtypes = [_vp, ctypes.c_uint32]
        L.nc_plan_set_persist.argtypes = [_vp, ctypes.c_int, ctypes.c_uint32]
        L.nc_plan_get_persist.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.nc_plan_persist_status.argtypes = [_vp, _vp, ctypes.POINTER(ctypes.c_uint32)]
        L.nc_persist_emulate.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
        L.nc_multiply.argtypes = [_vp, _u32p, _u32p, _u32p, ctypes.c_size_t]
        L.nc_multiply_device.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint32, _vp]
        L.nc_multiply_device_noalloc.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint32, _vp]
        L.nc_stage_residues.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, ctypes.c_uint32, _vp]
        L.nc_stage_product_residues.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_uint32, _vp]
        L.nc_stage_crt.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, ctypes.c_uint32, _vp]
        L.nc_gen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]
        L.nc_plan_autotune.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.POINTER(ctypes.c_uint32)]
        L.nc_set_tune_cache.argtypes = [ctypes.c_char_p]
        L.nc_plan_get_tune_info.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.nc_profile_stages.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp,
                                        ctypes.POINTER(ctypes.c_float)]
        L.nc_measure_mul32_peak.argtypes = [ctypes.POINTER(ctypes.c_double)]
        # multi-degree batches and serving (include/negacyclic_batch.h)
        L.nc_batch_create.argtypes = [ctypes.c_uint32, _u32p, ctypes.POINTER(_vp)]
        L.nc_batch_destroy.argtypes = [_vp]
        L.nc_batch_destroy.restype = None
        L.nc_batch_plan.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.nc_batch_autotune.argtypes = [_vp, ctypes.POINTER(Group), ctypes.c_uint32, _vp]
        L.nc_batch_run.argtypes = [_vp, ctypes.POINTER(Group), ctypes.c_uint32, _vp]
        L.nc_batch_capture.argtypes = [_vp, ctypes.POINTER(Group), ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.nc_graph_launch.argtypes = [_vp, _vp]
        L.nc_graph_destroy.argtypes = [_vp]
        L.nc_graph_destroy.restype = None
        L.nc_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
        L.nc_server_destroy.argtypes = [_vp]
        L.nc_server_destroy.restype = None
        L.nc_server_footprint.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]
        L.nc_server_submit.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_uint64)]
        L.nc_server_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.nc_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
        L.nc_server_reset_stats.argtypes = [_vp]
        L.nc_server_replay.argtypes = [_vp, ctypes.c_uint32, _u32p, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.nc_host_alloc.argtypes = [ctypes.c_size_t]
        L.nc_host_alloc.restype = _vp
        L.nc_host_free.argtypes = [_vp]
        L.nc_host_free.restype = None
        _lib = L
    return _lib


def _check(rc: int, where: str):
    if rc != NC_OK:
        raise NcError(rc, where, lib().nc_last_error().decode())


def _p32(x: np.ndarray):
    assert x.dtype == np.uint32 and x.flags.c_contiguous
    return x.ctypes.data_as(_u32p)


def gen_uniform(seed: int, N: int, q: np.ndarray) -> np.ndarray:
    """Seeded uniform polynomial in [0, q) (SplitMix64, see negacyclic.h)."""
    q = np.ascontiguousarray(q, dtype=np.uint32)
    out = np.empty((N, q.shape[0]), dtype=np.uint32)
    _check(lib().nc_gen_uniform(ctypes.c_uint64(seed), N, q.shape[0], _p32(q), _p32(out)), "nc_gen_uniform")
    return out


def check_tables(N: int, q: np.ndarray):
    """Host tables and init self tests of (N, q) without a device (raises NcError)."""
    q = np.ascontiguousarray(q, dtype=np.uint32)
    _check(lib().nc_check_tables(int(N), q.shape[0], _p32(q)), "nc_check_tables")


def host_twiddles(N: int, q: np.ndarray, width: int, inverse: bool) -> np.ndarray:
    """Twiddle table [k][N][2] ({w, Shoup quotient}) built on the host, no device."""
    q = np.ascontiguousarray(q, dtype=np.uint32)
    n = ctypes.c_size_t(0)
    args = (int(N), q.shape[0], _p32(q), int(width), int(bool(inverse)))
    _check(lib().nc_host_twiddles(*args, None, ctypes.byref(n)), "nc_host_twiddles")
    dt = np.uint64 if width == 62 else np.uint32
    out = np.empty(n.value // np.dtype(dt).itemsize, dtype=dt)
    _check(lib().nc_host_twiddles(*args, _vp(out.ctypes.data), ctypes.byref(n)), "nc_host_twiddles")
    return out.reshape(-1, N, 2)


def set_tune_cache(path):
    """Process-wide autotuner cache file (None disables; NC_TUNE_CACHE otherwise)."""
    _check(lib().nc_set_tune_cache(None if path is None else str(path).encode()), "nc_set_tune_cache")


def persist_emulate(N: int, k: int, batch: int, xcds: int, workgroups: int, seed: int = 1) -> dict:
    """CPU emulation of the persistent form's queues (raises NcError on a failed check)."""
    mk, wf = ctypes.c_double(), ctypes.c_double()
    _check(lib().nc_persist_emulate(N, k, batch, xcds, workgroups, seed, ctypes.byref(mk), ctypes.byref(wf)),
           "nc_persist_emulate")
    return {"makespan": float(mk.value), "wait_fraction": float(wf.value)}


def measure_mul32_peak() -> float:
    """Measured 32x32->64 integer multiplies per second of the current device."""
    out = ctypes.c_double()
    _check(lib().nc_measure_mul32_peak(ctypes.byref(out)), "nc_measure_mul32_peak")
    return out.value


def _stream_handle(stream):
    if stream is None:
        import torch

        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class TableSet:
    """One shared table set for every degree N <= 65536 of one modulus
    (nc_tableset_create); plans on it come from Plan(N, q, tables=set)."""

    def __init__(self, q: np.ndarray):
        self.q = np.ascontiguousarray(q, dtype=np.uint32)
        h = _vp()
        _check(lib().nc_tableset_create(self.q.shape[0], _p32(self.q), ctypes.byref(h)), "nc_tableset_create")
        self._h = h

    def info(self) -> dict:
        k62, k30, nb, ms = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_float()
        _check(lib().nc_tableset_get_info(self._h, ctypes.byref(k62), ctypes.byref(k30), ctypes.byref(nb),
                                          ctypes.byref(ms)), "nc_tableset_get_info")
        return {"k62": k62.value, "k30": k30.value, "table_bytes": nb.value, "create_ms": ms.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().nc_tableset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Plan:
    """One (N, q) plan: resident prime tables on the current device -- its
    own (nc_plan_create), or a shared set's (tables=TableSet,
    nc_plan_create_shared)."""

    def __init__(self, N: int, q: np.ndarray, tables: TableSet = None):
        self.q = np.ascontiguousarray(q, dtype=np.uint32)
        self.N = int(N)
        self.L = int(self.q.shape[0])
        h = _vp()
        if tables is None:
            _check(lib().nc_plan_create(self.N, self.L, _p32(self.q), ctypes.byref(h)), "nc_plan_create")
        else:
            _check(lib().nc_plan_create_shared(tables._h, self.N, ctypes.byref(h)), "nc_plan_create_shared")
        self._h = h
        self.info = self.get_info()
        self.k62, self.k30 = self.info.k62, self.info.k30

    def close(self):
        if getattr(self, "_h", None):
            lib().nc_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_info(self) -> PlanInfo:
        info = PlanInfo()
        _check(lib().nc_plan_get_info(self._h, ctypes.byref(info)), "nc_plan_get_info")
        return info

    def k(self, width: int) -> int:
        return self.k62 if width == 62 else self.k30

    def primes(self, width: int = 62) -> np.ndarray:
        out = np.empty(self.k(width), dtype=np.uint64)
        _check(lib().nc_plan_primes(self._h, width, out.ctypes.data_as(_u64p)), "nc_plan_primes")
        return out

    def twiddles(self, width: int, inverse: bool) -> np.ndarray:
        """The plan's device-built twiddle table [k][N][2], copied to the host."""
        dt = np.uint64 if width == 62 else np.uint32
        out = np.empty((self.k(width), self.N, 2), dtype=dt)
        _check(lib().nc_plan_copy_twiddles(self._h, int(width), int(bool(inverse)), _vp(out.ctypes.data),
                                           out.nbytes), "nc_plan_copy_twiddles")
        return out

    def set_tuning(self, width: int, log_blocks=None):
        """Pin prime width (62 / 30) and NTT split; width 0 unpins."""
        lb = 0xFFFFFFFF if log_blocks is None else int(log_blocks)
        _check(lib().nc_plan_set_tuning(self._h, int(width), lb), "nc_plan_set_tuning")
        self.info = self.get_info()

    def choice(self, batch: int):
        """(prime width, NTT split, matrix-core paths mask, chunks) used for `batch` products."""
        w, lb, m, ch = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().nc_plan_get_choice(self._h, batch, ctypes.byref(w), ctypes.byref(lb), ctypes.byref(m),
                                        ctypes.byref(ch)), "nc_plan_get_choice")
        return int(w.value), int(lb.value), int(m.value), int(ch.value)

    def set_chunks(self, chunks: int):
        """Batch slices on two staggered streams: 0 tuned, else forced."""
        _check(lib().nc_plan_set_chunks(self._h, int(chunks)), "nc_plan_set_chunks")

    def set_persist(self, mode: int, wait_bound_us: int = 0):
        """Persistent batched form: -1 tuned, 0 never, 1 wherever available; bound of one in-launch wait."""
        _check(lib().nc_plan_set_persist(self._h, int(mode), int(wait_bound_us)), "nc_plan_set_persist")

    def persistent(self, batch: int) -> bool:
        """Whether calls with `batch` products run the persistent form."""
        v = ctypes.c_uint32()
        _check(lib().nc_plan_get_persist(self._h, batch, ctypes.byref(v)), "nc_plan_get_persist")
        return bool(v.value)

    def persist_status(self, stream=None) -> int:
        """Error bits of the persistent launches since the last call (0: all drained); clears them."""
        v = ctypes.c_uint32()
        _check(lib().nc_plan_persist_status(self._h, _stream_handle(stream), ctypes.byref(v)),
               "nc_plan_persist_status")
        return int(v.value)

    def set_mfma(self, mode: int):
        """Matrix-core paths: a mask of MFMA_DECOMPOSE | MFMA_CRT (0: VALU only), or -1 tuned."""
        _check(lib().nc_plan_set_mfma(self._h, int(mode)), "nc_plan_set_mfma")
        self.info = self.get_info()

    def autotune(self, batch: int, stream=None):
        """Returns the (prime width, NTT split) chosen for this batch size."""
        w, lb = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().nc_plan_autotune(self._h, batch, _stream_handle(stream), ctypes.byref(w), ctypes.byref(lb)),
               "nc_plan_autotune")
        self.info = self.get_info()
        return int(w.value), int(lb.value)

    def tune_info(self, batch: int) -> dict:
        """How bucket `batch` was tuned: from_cache, tune_ms, median_ms."""
        fc, tm, md = ctypes.c_uint32(), ctypes.c_float(), ctypes.c_float()
        _check(lib().nc_plan_get_tune_info(self._h, batch, ctypes.byref(fc), ctypes.byref(tm), ctypes.byref(md)),
               "nc_plan_get_tune_info")
        return {"from_cache": bool(fc.value), "tune_ms": float(tm.value), "median_ms": float(md.value)}

    def profile_stages(self, d_a, d_b, d_c, batch: int, reps: int = 5, stream=None):
        """Mean per-launch milliseconds: decompose, ntt_outer_fwd, ntt_inner, ntt_outer_inv, crt."""
        ms = (ctypes.c_float * 5)()
        _check(lib().nc_profile_stages(self._h, _vp(d_a.data_ptr()), _vp(d_b.data_ptr()), _vp(d_c.data_ptr()),
                                       batch, reps, _stream_handle(stream), ms), "nc_profile_stages")
        return dict(zip(("decompose", "ntt_outer_fwd", "ntt_inner", "ntt_outer_inv", "crt"), list(ms)))

    # -- host buffers (fherma_run equivalent) --
    def multiply(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        c = np.empty((self.N, self.L), dtype=np.uint32)
        _check(lib().nc_multiply(self._h, _p32(a), _p32(b), _p32(c), a.size), "nc_multiply")
        return c

    # -- device buffers (torch tensors, int32 storage of u32 limbs) --
    def multiply_device(self, d_a, d_b, d_c, batch: int, stream=None):
        _check(lib().nc_multiply_device(self._h, _vp(d_a.data_ptr()), _vp(d_b.data_ptr()), _vp(d_c.data_ptr()),
                                        batch, _stream_handle(stream)), "nc_multiply_device")

    def multiply_device_noalloc(self, d_a, d_b, d_c, batch: int, stream=None):
        """nc_multiply_device with the library's per-device scratch array as
        workspace (no device allocation; batch * bytes_per_product must fit
        NOALLOC_SCRATCH_BYTES)."""
        _check(lib().nc_multiply_device_noalloc(self._h, _vp(d_a.data_ptr()), _vp(d_b.data_ptr()),
                                                _vp(d_c.data_ptr()), batch, _stream_handle(stream)),
               "nc_multiply_device_noalloc")

    def stage_residues(self, width: int, d_poly, d_res, batch: int, stream=None):
        _check(lib().nc_stage_residues(self._h, width, _vp(d_poly.data_ptr()), _vp(d_res.data_ptr()), batch,
                                       _stream_handle(stream)), "nc_stage_residues")

    def stage_product_residues(self, width: int, d_a, d_b, d_res, batch: int, stream=None):
        _check(lib().nc_stage_product_residues(self._h, width, _vp(d_a.data_ptr()), _vp(d_b.data_ptr()),
                                               _vp(d_res.data_ptr()), batch, _stream_handle(stream)),
               "nc_stage_product_residues")

    def stage_crt(self, width: int, d_res, d_c, batch: int, stream=None):
        _check(lib().nc_stage_crt(self._h, width, _vp(d_res.data_ptr()), _vp(d_c.data_ptr()), batch,
                                  _stream_handle(stream)), "nc_stage_crt")


# ---------------------------------------------------------------------------
# numpy <-> torch helpers for u32 / u64 data (torch has no unsigned 32/64-bit
# arithmetic we need; the bits travel in int32 / int64 storage)
# ---------------------------------------------------------------------------

def to_device_u32(x: np.ndarray, device="cuda"):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.uint32).view(np.int32)).to(device)


def to_device_u64(x: np.ndarray, device="cuda"):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.uint64).view(np.int64)).to(device)


def from_device_u32(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint32)


def from_device_u64(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint64)


def empty_device_u32(shape, device="cuda"):
    import torch

    return torch.empty(shape, dtype=torch.int32, device=device)


def empty_device_u64(shape, device="cuda"):
    import torch

    return torch.empty(shape, dtype=torch.int64, device=device)


# ---------------------------------------------------------------------------
# pinned host arrays, multi-degree batches, serving
# ---------------------------------------------------------------------------
def host_array(shape, dtype=np.uint32) -> np.ndarray:
    """numpy array in pinned host memory (nc_host_alloc), freed with the array."""
    import weakref

    count = int(np.prod(shape))
    nbytes = max(1, count * np.dtype(dtype).itemsize)
    ptr = lib().nc_host_alloc(nbytes)
    if not ptr:
        raise MemoryError(f"nc_host_alloc({nbytes}) failed")
    buf = (ctypes.c_char * nbytes).from_address(ptr)
    arr = np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)
    weakref.finalize(buf, lib().nc_host_free, ctypes.c_void_p(ptr))
    return arr


def _groups(groups):
    """[(N, count, d_a, d_b, d_c)] with torch device tensors -> nc_group array."""
    arr = (Group * len(groups))()
    for i, (N, count, a, b, c) in enumerate(groups):
        arr[i] = Group(N, count, a.data_ptr(), b.data_ptr(), c.data_ptr())
    return arr


class Graph:
    """An executable HIP graph of one multi-degree batch (nc_batch_capture)."""

    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep  # the bound device buffers stay alive with the graph

    def launch(self, stream=None):
        _check(lib().nc_graph_launch(self._h, _stream_handle(stream)), "nc_graph_launch")

    def close(self):
        if self._h:
            lib().nc_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Resident per-degree plans for one modulus; runs or captures groups of
    products of mixed degree.  Groups are (N, count, d_a, d_b, d_c) with
    torch device tensors of shape (count, N, L)."""

    def __init__(self, q: np.ndarray):
        self.q = np.ascontiguousarray(q, dtype=np.uint32)
        h = ctypes.c_void_p()
        _check(lib().nc_batch_create(self.q.shape[0], _p32(self.q), ctypes.byref(h)), "nc_batch_create")
        self._h = h
        self._graphs = []

    def autotune(self, groups, stream=None):
        g = _groups(groups)
        _check(lib().nc_batch_autotune(self._h, g, len(groups), _stream_handle(stream)), "nc_batch_autotune")

    def run(self, groups, stream=None):
        g = _groups(groups)
        _check(lib().nc_batch_run(self._h, g, len(groups), _stream_handle(stream)), "nc_batch_run")

    def set_persist(self, degrees, mode: int, wait_bound_us: int = 0):
        """nc_plan_set_persist on the batch's plan of each degree."""
        for N in degrees:
            h = _vp()
            _check(lib().nc_batch_plan(self._h, int(N), ctypes.byref(h)), "nc_batch_plan")
            _check(lib().nc_plan_set_persist(h, int(mode), int(wait_bound_us)), "nc_plan_set_persist")

    def persist_status(self, degrees) -> int:
        """OR of nc_plan_persist_status over the batch's plans of these degrees."""
        flags = 0
        for N in degrees:
            h, v = _vp(), ctypes.c_uint32()
            _check(lib().nc_batch_plan(self._h, int(N), ctypes.byref(h)), "nc_batch_plan")
            _check(lib().nc_plan_persist_status(h, None, ctypes.byref(v)), "nc_plan_persist_status")
            flags |= int(v.value)
        return flags

    def capture(self, groups) -> Graph:
        g = _groups(groups)
        h = ctypes.c_void_p()
        _check(lib().nc_batch_capture(self._h, g, len(groups), ctypes.byref(h)), "nc_batch_capture")
        graph = Graph(h, list(groups))
        self._graphs.append(graph)
        return graph

    def close(self):
        for g in self._graphs:
            g.close()
        self._graphs = []
        if self._h:
            lib().nc_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Continuous-batching server over resident per-degree plans (nc_server_*).
    Request buffers should come from host_array (pinned)."""

    def __init__(self, q: np.ndarray, degrees, max_batch: int = 32, lanes: int = 2, hbm_budget: int = 8 << 30,
                 autotune: bool = True):
        self.q = np.ascontiguousarray(q, dtype=np.uint32)
        self.L = self.q.shape[0]
        self.degrees = np.ascontiguousarray(sorted(set(int(n) for n in degrees)), dtype=np.uint32)
        cfg = ServerConfig(self.L, _p32(self.q), _p32(self.degrees), len(self.degrees), max_batch, lanes,
                           hbm_budget, 1 if autotune else 0)
        h = ctypes.c_void_p()
        _check(lib().nc_server_create(ctypes.byref(cfg), ctypes.byref(h)), "nc_server_create")
        self._h = h
        self._live = {}  # ticket -> buffers kept alive until waited for

    def footprint(self, N: int) -> int:
        out = ctypes.c_uint64()
        _check(lib().nc_server_footprint(self._h, N, ctypes.byref(out)), "nc_server_footprint")
        return int(out.value)

    def submit(self, a: np.ndarray, b: np.ndarray, c: np.ndarray) -> int:
        N = a.shape[0]
        for x in (a, b, c):
            assert x.dtype == np.uint32 and x.flags.c_contiguous and x.shape == (N, self.L)
        t = ctypes.c_uint64()
        _check(lib().nc_server_submit(self._h, N, a.ctypes.data, b.ctypes.data, c.ctypes.data, ctypes.byref(t)),
               "nc_server_submit")
        self._live[t.value] = (a, b, c)
        return int(t.value)

    def wait(self, ticket: int, timeout_ms: int = -1) -> float:
        lat = ctypes.c_double()
        rc = lib().nc_server_wait(self._h, ticket, timeout_ms, ctypes.byref(lat))
        self._live.pop(ticket, None)
        _check(rc, "nc_server_wait")
        return lat.value

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().nc_server_get_stats(self._h, ctypes.byref(st)), "nc_server_get_stats")
        return {n: int(getattr(st, n)) for n, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(lib().nc_server_reset_stats(self._h), "nc_server_reset_stats")

    def replay(self, degrees, arrivals, a_list, b_list, c_list):
        """Open-loop replay; returns (latencies seconds, makespan seconds)."""
        n = len(degrees)
        deg = np.ascontiguousarray(degrees, dtype=np.uint32)
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        pa = (_vp * n)(*[x.ctypes.data for x in a_list])
        pb = (_vp * n)(*[x.ctypes.data for x in b_list])
        pc = (_vp * n)(*[x.ctypes.data for x in c_list])
        lat = np.zeros(n, dtype=np.float64)
        mk = ctypes.c_double()
        _check(lib().nc_server_replay(self._h, n, _p32(deg), arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      pa, pb, pc, lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      ctypes.byref(mk)), "nc_server_replay")
        return lat, mk.value

    def close(self):
        if self._h:
            lib().nc_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
