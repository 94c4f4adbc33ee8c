"""Twiddle tables: the host construction against the definitions, and the
device builder (what nc_plan_create uses) against the host construction.

The reference builds these tables on the host in fherma_init
(solve.cu:1224-1257: psi by the smallest-base rule 495-502, Shoup
quotients 1250-1253).  Here entry x of prime j is psi_j^(+-bitrev(x)) with
floor(w 2^b / p_j), {0, 0} at x = 0, the 30-bit forward value stored as
2^32 - w.  The CPU test restates the prime rule, the psi rule and the entry
definition in Python integers and also checks identities that hold for any
primitive 2N-th root (tw[1]^2 = -1, tw[2x]^2 = tw[x],
tw[2x+1] = tw[2x] tw[1], forward times inverse = 1).
"""
from __future__ import annotations

import numpy as np
import pytest
import sympy

import ncmul


def bitrev(x: int, bits: int) -> int:
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def rule_primes(N: int, bits: int, k: int) -> list:
    """p = 1 mod 2N, descending from 2^bits - 1, the first k primes."""
    step = 2 * N
    cand = (1 << bits) - 1
    cand -= (cand - 1) % step
    out = []
    while len(out) < k:
        if sympy.isprime(cand):
            out.append(cand)
        cand -= step
    return out


def rule_psi(p: int, N: int) -> int:
    """Smallest base g >= 2 with (g^((p-1)/2N))^N = -1."""
    e = (p - 1) // (2 * N)
    g = 2
    while True:
        psi = pow(g, e, p)
        if pow(psi, N, p) == p - 1:
            return psi
        g += 1


def values(table: np.ndarray, width: int, inverse: bool) -> np.ndarray:
    w = table[..., 0].astype(np.uint64)
    if width == 30 and not inverse:
        w = (np.uint64(1 << 32) - w) & np.uint64(0xFFFFFFFF)
    return w


@pytest.mark.parametrize("N", [2, 16, 512])
@pytest.mark.parametrize("width", [62, 30])
def test_host_twiddles_match_definitions(N, width, q868):
    fwd = ncmul.host_twiddles(N, q868, width, False)
    inv = ncmul.host_twiddles(N, q868, width, True)
    k = fwd.shape[0]
    assert fwd.shape == inv.shape == (k, N, 2)
    log_n = N.bit_length() - 1
    b = 64 if width == 62 else 32
    for j, p in enumerate(rule_primes(N, width, k)):
        wf = [int(v) for v in values(fwd[j], width, False)]
        wi = [int(v) for v in values(inv[j], width, True)]
        assert wf[0] == wi[0] == 0 and int(fwd[j, 0, 1]) == int(inv[j, 0, 1]) == 0
        psi = rule_psi(p, N)
        psi_inv = pow(psi, -1, p)
        for x in range(1, N):
            e = bitrev(x, log_n)
            assert wf[x] == pow(psi, e, p), (j, x)
            assert wi[x] == pow(psi_inv, e, p), (j, x)
            assert int(fwd[j, x, 1]) == (wf[x] << b) // p
            assert int(inv[j, x, 1]) == (wi[x] << b) // p
            assert wf[x] * wi[x] % p == 1
        # identities of any primitive 2N-th root in this layout
        assert wf[1] * wf[1] % p == p - 1
        for x in range(1, N // 2):
            assert wf[2 * x] * wf[2 * x] % p == wf[x]
            assert wf[2 * x + 1] == wf[2 * x] * wf[1] % p


def test_host_twiddles_size_query_and_errors(q868):
    import ctypes

    n = ctypes.c_size_t(0)
    q = np.ascontiguousarray(q868)
    qp = q.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    assert ncmul.lib().nc_host_twiddles(64, 28, qp, 30, 0, None, ctypes.byref(n)) == 0
    k = n.value // (64 * 8)
    assert n.value == k * 64 * 8 and k > 0
    small = (ctypes.c_uint8 * 8)()
    n2 = ctypes.c_size_t(8)
    assert ncmul.lib().nc_host_twiddles(64, 28, qp, 30, 0, ctypes.cast(small, ctypes.c_void_p),
                                        ctypes.byref(n2)) == -2  # NC_ERR_SIZE
    assert ncmul.lib().nc_host_twiddles(64, 28, qp, 31, 0, None, ctypes.byref(n)) == -1  # bad width
    assert ncmul.lib().nc_host_twiddles(3, 28, qp, 30, 0, None, ctypes.byref(n)) == -1  # N not a power of two


@pytest.mark.gpu
@pytest.mark.parametrize("N", [2, 64, 1024, 16384, 65536])
def test_device_twiddles_equal_host(gpu, N, q868):
    print(f"device twiddles N={N}", flush=True)
    with ncmul.Plan(N, q868) as plan:
        for width in (62, 30):
            for inverse in (False, True):
                dev = plan.twiddles(width, inverse)
                host = ncmul.host_twiddles(N, q868, width, inverse)
                assert dev.shape == host.shape
                assert np.array_equal(dev, host), (width, inverse)
