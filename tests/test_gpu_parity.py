"""GPU parity: the HIP library against the golden vectors and the oracle,
bit for bit, through the C ABI (include/negacyclic.h), for both prime sets
(62-bit: the reference's prime rule; 30-bit: the same rule below 2^30).

Sizes: every golden N (2 .. 65536) end to end; stage-level checks at
N = 1024 and 16384 (62-bit, against the CPU restatement of the reference)
and N = 1024 (30-bit, against exact Python integers); batch and tuning
invariance; the reference's edge inputs by closed form; size-independent
properties (commutativity, bilinearity, X^N = -1) at N = 65536.
"""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

import ncmul
import oracle as O
from conftest import load_golden, modulus

pytestmark = pytest.mark.gpu

WIDTHS = (62, 30)


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u4").tobytes()).hexdigest()


def say(*a):
    print(*a, flush=True)


def crt_weights(primes):
    """y_j = (M / p_j)^-1 mod p_j: the weight the inverse transform applies."""
    M = 1
    for p in primes:
        M *= p
    return [pow(M // p, -1, p) for p in primes]


def crt_exact(v_cols, primes, qi):
    """c mod q from weighted residues v_j (one list per coefficient): the
    centred value of sum_j v_j M / p_j mod M, reduced mod q."""
    M = 1
    for p in primes:
        M *= p
    out = []
    for v in v_cols:
        x = sum(vj * (M // p) for vj, p in zip(v, primes)) % M
        if x > M // 2:
            x -= M
        out.append(x % qi)
    return out


CASES = load_golden()["cases"]


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['modulus']}-N{c['N']}-s{c['seed']}")
def test_golden_host_path(gpu, case, width):
    """nc_multiply (the fherma_run replacement) against the Python-int vectors."""
    q = modulus(case["modulus"])
    N, s = case["N"], case["seed"]
    say("golden", case["modulus"], N, s, "width", width)
    a, b = ncmul.gen_uniform(2 * s, N, q), ncmul.gen_uniform(2 * s + 1, N, q)
    with ncmul.Plan(N, q) as plan:
        plan.set_tuning(width)
        c = plan.multiply(a, b)
    assert sha(c) == case["sha256_c"]
    for idx, hx in case["samples"].items():
        assert O.limbs_to_int(c[int(idx)]) == int(hx, 16)


@pytest.mark.parametrize("N", [1024, 16384])
def test_stage_residues_62(gpu, N, q868):
    say("stage residues 62", N)
    a = ncmul.gen_uniform(21, N, q868)
    with ncmul.Plan(N, q868) as plan:
        assert np.array_equal(plan.primes(62), O.Plan(N, q868).primes())
        d_a = ncmul.to_device_u32(np.stack([a, a[::-1].copy()]))
        d_res = ncmul.empty_device_u64((2, plan.k62, N))
        plan.stage_residues(62, d_a, d_res, 2)
        got = ncmul.from_device_u64(d_res)
    orc = O.Plan(N, q868)
    assert np.array_equal(got[0], orc.residues(a))
    assert np.array_equal(got[1], orc.residues(a[::-1].copy()))


@pytest.mark.parametrize("mfma", [1, 0])
def test_stage_residues_30(gpu, q868, mfma):
    """30-bit residues on the matrix cores (int8 byte split) and on the VALU."""
    N = 1024
    say("stage residues 30 mfma", mfma, N)
    a = ncmul.gen_uniform(22, N, q868)
    a[5] = O.int_to_limbs(O.limbs_to_int(q868) - 1, 28)  # all-ones bytes where q allows
    a[6] = 0
    ai = O.poly_to_ints(a)
    with ncmul.Plan(N, q868) as plan:
        plan.set_mfma(mfma)
        pr = [int(p) for p in plan.primes(30)]
        assert all(p < (1 << 30) and p % (2 * N) == 1 for p in pr)
        d_a = ncmul.to_device_u32(a[None])
        d_res = ncmul.empty_device_u32((1, plan.k30, N))
        plan.stage_residues(30, d_a, d_res, 1)
        got = ncmul.from_device_u32(d_res)[0]
    for j, p in enumerate(pr):
        assert [int(x) for x in got[j]] == [x % p for x in ai], j


@pytest.mark.parametrize("N", [32, 512, 16384, 65536])
def test_residues_30_mfma_equals_valu(gpu, N, q868):
    say("mfma vs valu residues", N)
    B = 3
    a = np.stack([ncmul.gen_uniform(70 + i, N, q868) for i in range(B)])
    with ncmul.Plan(N, q868) as plan:
        d_a = ncmul.to_device_u32(a)
        outs = []
        for mfma in (1, 0):
            plan.set_mfma(mfma)
            d_res = ncmul.empty_device_u32((B, plan.k30, N))
            plan.stage_residues(30, d_a, d_res, B)
            outs.append(ncmul.from_device_u32(d_res))
        assert np.array_equal(outs[0], outs[1])
        p0 = int(plan.primes(30)[0])
        assert [int(x) for x in outs[0][1, 0, :8]] == [x % p0 for x in O.poly_to_ints(a[1][:8])]


@pytest.mark.parametrize("N", [1024, 16384])
def test_stage_product_residues_and_crt_62(gpu, N, q868):
    say("stage product residues 62", N)
    a, b = ncmul.gen_uniform(31, N, q868), ncmul.gen_uniform(32, N, q868)
    orc = O.Plan(N, q868)
    pres = orc.product_residues(a, b)
    pr = [int(p) for p in orc.primes()]
    y = crt_weights(pr)
    want = np.array([[int(r) * y[j] % pr[j] for r in pres[j]] for j in range(len(pr))], dtype=np.uint64)
    with ncmul.Plan(N, q868) as plan:
        d_a, d_b = ncmul.to_device_u32(a[None]), ncmul.to_device_u32(b[None])
        d_res = ncmul.empty_device_u64((1, plan.k62, N))
        plan.stage_product_residues(62, d_a, d_b, d_res, 1)
        assert np.array_equal(ncmul.from_device_u64(d_res)[0], want)
        d_c = ncmul.empty_device_u32((1, N, 28))
        plan.stage_crt(62, ncmul.to_device_u64(want[None]), d_c, 1)
        assert np.array_equal(ncmul.from_device_u32(d_c)[0], orc.crt_fast(pres))


@pytest.mark.parametrize("mfma", [ncmul.MFMA_ALL, 0])
def test_stage_product_residues_and_crt_30(gpu, q868, mfma):
    N = 1024
    say("stage product residues 30 mfma", mfma, N)
    qi = O.limbs_to_int(q868)
    a, b = ncmul.gen_uniform(33, N, q868), ncmul.gen_uniform(34, N, q868)
    cint = O.py_negacyclic_exact(O.poly_to_ints(a), O.poly_to_ints(b))
    with ncmul.Plan(N, q868) as plan:
        plan.set_mfma(mfma)
        pr = [int(p) for p in plan.primes(30)]
        y = crt_weights(pr)
        want = np.array([[c % p * y[j] % p for c in cint] for j, p in enumerate(pr)], dtype=np.uint32)
        d_a, d_b = ncmul.to_device_u32(a[None]), ncmul.to_device_u32(b[None])
        d_res = ncmul.empty_device_u32((1, plan.k30, N))
        plan.stage_product_residues(30, d_a, d_b, d_res, 1)
        assert np.array_equal(ncmul.from_device_u32(d_res)[0], want)
        d_c = ncmul.empty_device_u32((1, N, 28))
        plan.stage_crt(30, ncmul.to_device_u32(want[None]), d_c, 1)
        assert O.poly_to_ints(ncmul.from_device_u32(d_c)[0]) == [c % qi for c in cint]


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("N", [16, 64, 256, 1024, 2048, 8192, 16384, 32768, 65536])
def test_tuning_invariance(gpu, N, width, q868):
    """Every accepted NTT split gives the same bits as the oracle (split 2
    with the matrix-core CRT runs the last inverse stages inside the CRT;
    N = 64 is its smallest case, one wave per product)."""
    a, b = ncmul.gen_uniform(41, N, q868), ncmul.gen_uniform(42, N, q868)
    want = O.Plan(N, q868).multiply(a, b)
    with ncmul.Plan(N, q868) as plan:
        logn = N.bit_length() - 1
        lo, hi = max(0, logn - 14), min(4, logn - 4)
        for lb in range(lo, hi + 1):
            for mfma in ((ncmul.MFMA_ALL, 0) if width == 30 else (0,)):
                say("tuning", N, width, lb, "mfma", mfma)
                plan.set_mfma(mfma)
                plan.set_tuning(width, lb)
                assert np.array_equal(plan.multiply(a, b), want), (lb, mfma)


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("N", [1024, 16384])
def test_device_batch(gpu, N, width, q868):
    say("batch", N, width)
    B = 5
    a = np.stack([ncmul.gen_uniform(100 + 2 * i, N, q868) for i in range(B)])
    b = np.stack([ncmul.gen_uniform(101 + 2 * i, N, q868) for i in range(B)])
    orc = O.Plan(N, q868)
    with ncmul.Plan(N, q868) as plan:
        plan.set_tuning(width)
        d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
        d_c = ncmul.empty_device_u32((B, N, 28))
        plan.multiply_device(d_a, d_b, d_c, B)
        got = ncmul.from_device_u32(d_c)
        for i in range(B):
            assert np.array_equal(got[i], orc.multiply(a[i], b[i])), i
        plan.multiply_device(d_b, d_a, d_c, 2)  # smaller batch reuses the workspace
        got2 = ncmul.from_device_u32(d_c)
        for i in range(2):
            assert np.array_equal(got2[i], got[i])


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("chunks", [2, 3, 8])
def test_chunked_batch(gpu, width, chunks, q868):
    """Batches split into chunks on two staggered streams (uneven last chunk,
    more chunks than products) give the same bits as the oracle."""
    N, B = 1024, 7
    say("chunked", width, chunks)
    a = np.stack([ncmul.gen_uniform(300 + 2 * i, N, q868) for i in range(B)])
    b = np.stack([ncmul.gen_uniform(301 + 2 * i, N, q868) for i in range(B)])
    orc = O.Plan(N, q868)
    with ncmul.Plan(N, q868) as plan:
        # 30-bit: split 2, so chunks also run the CRT with the inverse stages inside
        plan.set_tuning(width, 2 if width == 30 else None)
        plan.set_chunks(chunks)
        assert plan.choice(B)[3] == min(chunks, B)
        d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
        d_c = ncmul.empty_device_u32((B, N, 28))
        for rep in range(2):
            d_c.zero_()
            plan.multiply_device(d_a, d_b, d_c, B)
            got = ncmul.from_device_u32(d_c)
            for i in range(B):
                assert np.array_equal(got[i], orc.multiply(a[i], b[i])), (rep, i)


def test_decompose_wave_split(gpu, q868):
    """Batch sizes on both sides of the decomposition's small-launch
    thresholds (at N = 2^14: 1 and 3 products share each chunk's prime tiles
    over 4 waves, 5 over 2, 9 over 1), for each fused-stage count (split 0,
    1, 2), give the oracle's bits."""
    N = 16384
    a = np.stack([ncmul.gen_uniform(700 + 2 * i, N, q868) for i in range(9)])
    b = np.stack([ncmul.gen_uniform(701 + 2 * i, N, q868) for i in range(9)])
    orc = O.Plan(N, q868)
    want = [orc.multiply(a[i], b[i]) for i in range(9)]
    d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
    d_c = ncmul.empty_device_u32((9, N, 28))
    with ncmul.Plan(N, q868) as plan:
        plan.set_mfma(ncmul.MFMA_ALL)
        for split in (0, 1, 2):
            plan.set_tuning(30, split)
            for B in (1, 3, 5, 9):
                say("wave split", split, B)
                d_c.zero_()
                plan.multiply_device(d_a, d_b, d_c, B)
                got = ncmul.from_device_u32(d_c)
                for i in range(B):
                    assert np.array_equal(got[i], want[i]), (split, B, i)


@pytest.mark.parametrize("width,split", [(62, None), (30, 2), (30, 0)])
def test_profile_stages(gpu, width, split, q868):
    """The stage profiler runs the real pipeline back to back: its last
    repetition leaves the products in c, every stage it reports is positive,
    and stages that run inside a neighbour report 0."""
    N, B = 2048, 3
    say("profile", width, split)
    a = np.stack([ncmul.gen_uniform(500 + 2 * i, N, q868) for i in range(B)])
    b = np.stack([ncmul.gen_uniform(501 + 2 * i, N, q868) for i in range(B)])
    orc = O.Plan(N, q868)
    with ncmul.Plan(N, q868) as plan:
        plan.set_tuning(width, split)
        d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
        d_c = ncmul.empty_device_u32((B, N, 28))
        ms = plan.profile_stages(d_a, d_b, d_c, B, reps=7)
        got = ncmul.from_device_u32(d_c)
        for i in range(B):
            assert np.array_equal(got[i], orc.multiply(a[i], b[i])), i
        assert set(ms) == {"decompose", "ntt_outer_fwd", "ntt_inner", "ntt_outer_inv", "crt"}
        assert ms["decompose"] > 0 and ms["ntt_inner"] > 0 and ms["crt"] > 0
        assert all(v >= 0 for v in ms.values())
        if split == 0:  # one block per prime: no outer stages at all
            assert ms["ntt_outer_fwd"] == 0 and ms["ntt_outer_inv"] == 0
        with pytest.raises(ncmul.NcError):
            plan.profile_stages(d_a, d_b, d_c, B, reps=0)
        with pytest.raises(ncmul.NcError):
            plan.profile_stages(d_a, d_b, d_c, B, reps=4097)


def test_autotune_keeps_bits(gpu, q868):
    N = 4096
    a, b = ncmul.gen_uniform(61, N, q868), ncmul.gen_uniform(62, N, q868)
    want = O.Plan(N, q868).multiply(a, b)
    with ncmul.Plan(N, q868) as plan:
        for batch in (1, 8):
            w, lb = plan.autotune(batch)
            say("autotune", batch, w, lb, "choice", plan.choice(batch))
            assert w in WIDTHS
        assert np.array_equal(plan.multiply(a, b), want)
        # the batch-8 choice (possibly chunked) through the device entry
        # point, every product against the oracle
        A = np.stack([ncmul.gen_uniform(63 + 2 * i, N, q868) for i in range(8)])
        B = np.stack([ncmul.gen_uniform(64 + 2 * i, N, q868) for i in range(8)])
        d_c = ncmul.empty_device_u32((8, N, 28))
        d_c.fill_(-1)
        plan.multiply_device(ncmul.to_device_u32(A), ncmul.to_device_u32(B), d_c, 8)
        got = ncmul.from_device_u32(d_c)
        oplan = O.Plan(N, q868)
        for i in range(8):
            assert np.array_equal(got[i], oplan.multiply(A[i], B[i])), i


def test_one_plan_two_streams(gpu, q868):
    """Calls on one plan from different streams share its residue
    workspace; the plan orders them on the device.  A long batch on a side
    stream, then the host-path call (the plan's own stream) and a device call
    on a second side stream, with no synchronisation in between: every
    product equals the oracle's."""
    import torch

    N, n = 4096, 48
    A = np.stack([ncmul.gen_uniform(700 + 2 * i, N, q868) for i in range(n)])
    B = np.stack([ncmul.gen_uniform(701 + 2 * i, N, q868) for i in range(n)])
    oplan = O.Plan(N, q868)
    want = [oplan.multiply(A[i], B[i]) for i in range(n)]
    d_a, d_b = ncmul.to_device_u32(A), ncmul.to_device_u32(B)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with ncmul.Plan(N, q868) as plan:
        plan.set_chunks(1)
        for rep in range(3):
            c1 = ncmul.empty_device_u32((n, N, 28))
            c2 = ncmul.empty_device_u32((4, N, 28))
            c1.fill_(-1)
            c2.fill_(-1)
            torch.cuda.synchronize()
            plan.multiply_device(d_a, d_b, c1, n, stream=s1)      # long, asynchronous
            host = plan.multiply(A[5], B[5])                        # plan stream, blocks
            plan.multiply_device(d_a[8:12], d_b[8:12], c2, 4, stream=s2)
            plan.multiply_device(d_b, d_a, c1, n, stream=s1)        # commuted: same products
            torch.cuda.synchronize()
            say("two streams rep", rep)
            assert np.array_equal(host, want[5])
            got1, got2 = ncmul.from_device_u32(c1), ncmul.from_device_u32(c2)
            for i in range(n):
                assert np.array_equal(got1[i], want[i]), (rep, i)
            for i in range(4):
                assert np.array_equal(got2[i], want[8 + i]), (rep, i)


@pytest.mark.parametrize("N", [2, 1024, 16384, 65536])
def test_edges_closed_form(gpu, N, q868):
    say("edges", N)
    qi = O.limbs_to_int(q868)
    with ncmul.Plan(N, q868) as plan:
        for width in WIDTHS:
            plan.set_tuning(width)
            full = np.tile(O.int_to_limbs(qi - 1, 28), (N, 1))
            c = plan.multiply(full, full)
            for k in sorted({0, 1, N // 2, N - 1}):
                assert O.limbs_to_int(c[k]) == (2 * k + 2 - N) % qi
            zero = np.zeros((N, 28), dtype=np.uint32)
            assert not plan.multiply(zero, full).any()
            a, b = zero.copy(), zero.copy()
            v = (qi * 5) // 11
            a[N - 1] = O.int_to_limbs(v, 28)
            b[1] = O.int_to_limbs(1, 28)
            c = plan.multiply(a, b)
            assert O.limbs_to_int(c[0]) == qi - v
            assert not c[1:].any()
            x = ncmul.gen_uniform(7, N, q868)
            one = zero.copy()
            one[0] = O.int_to_limbs(1, 28)
            assert np.array_equal(plan.multiply(x, one), x)
            neg = zero.copy()
            neg[0] = O.int_to_limbs(qi - 1, 28)
            assert O.poly_to_ints(plan.multiply(x, neg)) == [(qi - y) % qi for y in O.poly_to_ints(x)]


def test_properties_max_size(gpu, q868):
    """N = 65536 (beyond the reference's cap): commutativity, bilinearity and
    the X^N = -1 shift, all exact, on the default configuration."""
    N = 65536
    qi = O.limbs_to_int(q868)
    say("properties", N)
    a, a2, b = (ncmul.gen_uniform(s, N, q868) for s in (51, 52, 53))
    with ncmul.Plan(N, q868) as plan:
        ab = plan.multiply(a, b)
        assert np.array_equal(plan.multiply(b, a), ab)
        s = O.ints_to_poly([(x + y) % qi for x, y in zip(O.poly_to_ints(a), O.poly_to_ints(a2))], 28)
        lhs = O.poly_to_ints(plan.multiply(s, b))
        rhs = [(x + y) % qi for x, y in zip(O.poly_to_ints(ab), O.poly_to_ints(plan.multiply(a2, b)))]
        assert lhs == rhs
        x1 = np.zeros_like(a)
        x1[1] = O.int_to_limbs(1, 28)
        sh = O.poly_to_ints(plan.multiply(a, x1))
        ai = O.poly_to_ints(a)
        assert sh[0] == (qi - ai[-1]) % qi and sh[1:] == ai[:-1]


def _wide_modulus(L, bits, seed):
    import random
    r = random.Random(seed)
    return (1 << (bits - 1)) | r.getrandbits(bits - 1) | 1


@pytest.mark.parametrize("N,B,mod", [
    (32, 3, "q868"), (64, 2, "q200"), (1024, 1, "q868b"), (16384, 2, "q868"), (64, 3, "w1017"), (32, 1, "w500"),
])
def test_crt_mfma_equals_exact(gpu, N, B, mod):
    """The int8 matrix-core CRT and the VALU CRT against exact Python
    integers: weighted residues of integers c with |c| <= N (q-1)^2 (the
    range of negacyclic product coefficients), including both ends, 0 and -1;
    odd tile counts (B N / 32 odd) leave half a wave idle."""
    if mod == "w1017":
        qi, L = _wide_modulus(32, 1017, 9), 32
    elif mod == "w500":
        qi, L = _wide_modulus(16, 500, 10), 16
    else:
        q = modulus(mod)
        qi, L = O.limbs_to_int(q), q.shape[0]
    q = O.int_to_limbs(qi, L)
    import random
    r = random.Random(N * 7 + B)
    bound = N * (qi - 1) ** 2
    cs = [[r.randrange(-bound, bound + 1) for _ in range(N)] for _ in range(B)]
    cs[0][:4] = [bound, -bound, 0, -1][: N]
    with ncmul.Plan(N, q) as plan:
        assert plan.info.mfma_paths & ncmul.MFMA_CRT
        pr = [int(p) for p in plan.primes(30)]
        y = crt_weights(pr)
        v = np.array([[[c % p * yj % p for c in row] for p, yj in zip(pr, y)] for row in cs], dtype=np.uint32)
        d_v = ncmul.to_device_u32(v)
        outs = {}
        for mode in (ncmul.MFMA_CRT, 0):
            say("crt", mod, N, B, "mode", mode)
            plan.set_mfma(mode)
            d_c = ncmul.empty_device_u32((B, N, L))
            plan.stage_crt(30, d_v, d_c, B)
            outs[mode] = ncmul.from_device_u32(d_c)
    for mode, got in outs.items():
        for bi in range(B):
            assert O.poly_to_ints(got[bi]) == [c % qi for c in cs[bi]], (mode, bi)


def test_errors(gpu, q868):
    with ncmul.Plan(1024, q868) as plan:
        a = ncmul.gen_uniform(1, 1024, q868)
        with pytest.raises(ncmul.NcError) as e:
            plan.multiply(a[:-1], a[:-1])
        assert e.value.status == -2
        with pytest.raises(ncmul.NcError):
            plan.set_tuning(62, 9)
        with pytest.raises(ncmul.NcError):
            plan.set_tuning(31)


def test_autotune_cache_second_process(gpu, q868, tmp_path):
    """A choice tuned in one process is appended to the cache file; a second
    process with the same device and shape takes it without timing anything
    (from_cache, a tuning call of a few milliseconds), and its products are
    the oracle's."""
    import json
    import os
    import subprocess
    import sys

    N, batch = 2048, 8
    cache = tmp_path / "tune.txt"
    ncmul.set_tune_cache(cache)
    try:
        with ncmul.Plan(N, q868) as plan:
            plan.autotune(batch)
            first = plan.tune_info(batch)
            choice = plan.choice(batch)
    finally:
        ncmul.set_tune_cache(None)
    say("first tuning", first, choice, cache.read_text())
    assert not first["from_cache"] and cache.read_text().count("\n") == 1
    child = r'''
import json, sys
import numpy as np
sys.path[:0] = sys.argv[1:3]
import torch
torch.cuda.init()  # torch's device runtime first, as in every other entry point
import ncmul, oracle as O
q = np.frombuffer(bytes.fromhex(sys.argv[3]), dtype="<u4").astype(np.uint32)
N, batch = 2048, 8
with ncmul.Plan(N, q) as plan:
    plan.autotune(batch)
    info, choice = plan.tune_info(batch), plan.choice(batch)
    A = np.stack([ncmul.gen_uniform(900 + 2 * i, N, q) for i in range(batch)])
    B = np.stack([ncmul.gen_uniform(901 + 2 * i, N, q) for i in range(batch)])
    d_c = ncmul.empty_device_u32((batch, N, 28))
    plan.multiply_device(ncmul.to_device_u32(A), ncmul.to_device_u32(B), d_c, batch)
    got = ncmul.from_device_u32(d_c)
    op = O.Plan(N, q)
    exact = all(np.array_equal(got[i], op.multiply(A[i], B[i])) for i in range(batch))
print(json.dumps({"info": info, "choice": list(choice), "exact": exact}))
'''
    from conftest import PKG, ROOT
    env = dict(os.environ, NC_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", child, str(PKG), str(ROOT / "oracle"), q868.astype("<u4").tobytes().hex()],
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = json.loads(r.stdout.strip().splitlines()[-1])
    say("second process", got)
    assert got["info"]["from_cache"] and got["info"]["tune_ms"] < 50.0
    assert tuple(got["choice"]) == choice and got["exact"]


# ---- shared table sets (nc_tableset_create / nc_plan_create_shared) -------

_SETS = {}


def _table_set(name):
    if name not in _SETS:
        _SETS[name] = ncmul.TableSet(modulus(name))
    return _SETS[name]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['modulus']}-N{c['N']}-s{c['seed']}")
def test_golden_shared_tables(gpu, case):
    """Every golden vector through a plan on its modulus's one shared table
    set (primes p = 1 mod 2^17, degree N reading the prefix of 65536-entry
    twiddle rows), for both prime widths."""
    q = modulus(case["modulus"])
    N, s = case["N"], case["seed"]
    ts = _table_set(case["modulus"])
    a, b = ncmul.gen_uniform(2 * s, N, q), ncmul.gen_uniform(2 * s + 1, N, q)
    with ncmul.Plan(N, q, tables=ts) as plan:
        for width in WIDTHS:
            say("golden shared", case["modulus"], N, s, "width", width)
            plan.set_tuning(width)
            c = plan.multiply(a, b)
            assert sha(c) == case["sha256_c"], width
            for idx, hx in case["samples"].items():
                assert O.limbs_to_int(c[int(idx)]) == int(hx, 16)


def test_shared_tables_are_prefixes_of_the_master(gpu, q868):
    """The set's 65536-entry rows equal the host construction at N = 65536
    (the prime rule and psi rule there are the per-N ones), and every degree's
    table is the prefix of every row; a shared plan's primes are the set's."""
    ts = _table_set("q868")
    info = ts.info()
    say("table set", info)
    with ncmul.Plan(65536, q868, tables=ts) as top:
        full = {(w, inv): top.twiddles(w, inv) for w in WIDTHS for inv in (False, True)}
        for (w, inv), t in full.items():
            assert np.array_equal(t, ncmul.host_twiddles(65536, q868, w, inv)), (w, inv)
        for N in (2, 16, 1024, 16384):
            with ncmul.Plan(N, q868, tables=ts) as p:
                assert np.array_equal(p.primes(30), top.primes(30)) and np.array_equal(p.primes(62), top.primes(62))
                assert all(int(x) % (1 << 17) == 1 for x in p.primes(30))
                for (w, inv), t in full.items():
                    assert np.array_equal(p.twiddles(w, inv), t[:, :N]), (N, w, inv)


@pytest.mark.parametrize("mfma", [3, 0])
def test_shared_plan_stage_values(gpu, q868, mfma):
    """Stage level on a shared plan: CRT-weighted product residues over the
    set's primes equal exact Python integers, and the CRT maps them back."""
    N = 2048
    qi = O.limbs_to_int(q868)
    a, b = ncmul.gen_uniform(35, N, q868), ncmul.gen_uniform(36, N, q868)
    cint = O.py_negacyclic_exact(O.poly_to_ints(a), O.poly_to_ints(b))
    with ncmul.Plan(N, q868, tables=_table_set("q868")) as plan:
        plan.set_mfma(mfma)
        pr = [int(p) for p in plan.primes(30)]
        y = crt_weights(pr)
        want = np.array([[c % p * y[j] % p for c in cint] for j, p in enumerate(pr)], dtype=np.uint32)
        d_a, d_b = ncmul.to_device_u32(a[None]), ncmul.to_device_u32(b[None])
        d_res = ncmul.empty_device_u32((1, plan.k30, N))
        plan.stage_product_residues(30, d_a, d_b, d_res, 1)
        assert np.array_equal(ncmul.from_device_u32(d_res)[0], want)
        d_c = ncmul.empty_device_u32((1, N, 28))
        plan.stage_crt(30, ncmul.to_device_u32(want[None]), d_c, 1)
        assert O.poly_to_ints(ncmul.from_device_u32(d_c)[0]) == [c % qi for c in cint]


@pytest.mark.parametrize("N", [64, 2048, 16384, 65536])
def test_shared_plan_every_split(gpu, q868, N):
    """Every accepted split and width on a shared plan, batched, equals the oracle."""
    a = np.stack([ncmul.gen_uniform(51 + 2 * i, N, q868) for i in range(3)])
    b = np.stack([ncmul.gen_uniform(52 + 2 * i, N, q868) for i in range(3)])
    op = O.Plan(N, q868)
    want = [op.multiply(a[i], b[i]) for i in range(3)]
    logn = N.bit_length() - 1
    with ncmul.Plan(N, q868, tables=_table_set("q868")) as plan:
        for width in WIDTHS:
            for s1 in range(max(0, logn - 14), min(4, logn - 4) + 1):
                plan.set_tuning(width, s1)
                d_c = ncmul.empty_device_u32((3, N, 28))
                d_c.fill_(-1)
                plan.multiply_device(ncmul.to_device_u32(a), ncmul.to_device_u32(b), d_c, 3)
                got = ncmul.from_device_u32(d_c)
                say("shared split", N, width, s1)
                for i in range(3):
                    assert np.array_equal(got[i], want[i]), (width, s1, i)
