"""GPU parity of the persistent batched form (nc_plan_set_persist): one
launch per batch, every XCD draining a queue of decomposition, NTT and CRT
items, against the oracle's restatement of the reference (solve.cu) and the
plain launch sequence, bit for bit, through the C ABI.

Sizes: every degree the form covers (N = 2^10 .. 2^14) at batch 8 (one
product per XCD), uneven batches, the 128- and 256-product batches at
N = 2^14 with every product checked, a second 868-bit modulus; edge inputs
(all q - 1, zeros); the mixed-degree graph and the allocation-free entry
point; the error path forced from the host with a 1 us readiness bound;
hundreds of launches in a row against the launch sequence.
"""
from __future__ import annotations

import numpy as np
import pytest

import ncmul
import oracle as O

pytestmark = pytest.mark.gpu

L = 28


def say(*a):
    print(*a, flush=True)


def batch_inputs(N, B, q, seed):
    a = np.stack([ncmul.gen_uniform(seed + 2 * i, N, q) for i in range(B)])
    b = np.stack([ncmul.gen_uniform(seed + 2 * i + 1, N, q) for i in range(B)])
    return a, b


def run_device(plan, a, b):
    B, N = a.shape[0], a.shape[1]
    d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
    d_c = ncmul.empty_device_u32((B, N, L))
    d_c.fill_(-1)  # products the launch leaves out would show as all ones
    plan.multiply_device(d_a, d_b, d_c, B)
    return ncmul.from_device_u32(d_c)


@pytest.mark.parametrize("N,B", [(1024, 8), (1024, 19), (2048, 8), (4096, 11), (8192, 8), (8192, 24),
                                 (16384, 8), (16384, 13)])
def test_persist_matches_oracle(gpu, N, B, q868):
    say("persist", N, B)
    a, b = batch_inputs(N, B, q868, 7000 + N + B)
    orc = O.Plan(N, q868)
    with ncmul.Plan(N, q868) as plan:
        plan.set_persist(1)
        if not plan.persistent(B):
            pytest.skip("the device does not report its XCD count")
        got = run_device(plan, a, b)
        assert plan.persist_status() == 0
        for i in range(B):
            assert np.array_equal(got[i], orc.multiply(a[i], b[i], threads=0)), i
        # a smaller batch on the same workspace: the control block starts clean
        B2 = max(8, B // 2)
        got2 = run_device(plan, b[:B2], a[:B2])
        assert plan.persist_status() == 0
        for i in range(B2):
            assert np.array_equal(got2[i], got[i]), i


def test_persist_headline_batch(gpu, q868):
    """128 products at N = 2^14 (the headline shape), every one against the
    oracle, then twice more on the same buffers (launch-to-launch reuse)."""
    N, B = 16384, 128
    say("persist headline", N, B)
    a, b = batch_inputs(N, B, q868, 9100)
    orc = O.Plan(N, q868)
    want = [orc.multiply(a[i], b[i], threads=0) for i in range(B)]
    with ncmul.Plan(N, q868) as plan:
        plan.set_persist(1)
        if not plan.persistent(B):
            pytest.skip("the device does not report its XCD count")
        d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
        d_c = ncmul.empty_device_u32((B, N, L))
        for rep in range(3):
            d_c.fill_(-1)
            plan.multiply_device(d_a, d_b, d_c, B)
            got = ncmul.from_device_u32(d_c)
            assert plan.persist_status() == 0, rep
            bad = [i for i in range(B) if not np.array_equal(got[i], want[i])]
            assert not bad, (rep, bad[:8])


def test_persist_equals_launch_sequence(gpu, q868):
    """The persistent form and the plain launch sequence give the same bits,
    on random inputs and on the reference's edge inputs (all coefficients
    q - 1, all zero)."""
    N, B = 16384, 16
    say("persist vs sequence", N, B)
    a, b = batch_inputs(N, B, q868, 9300)
    qm1 = np.tile(q868 - np.array([1] + [0] * (L - 1), dtype=np.uint32), (N, 1))
    a[0], b[0] = qm1, qm1
    a[1] = 0
    b[2] = qm1
    with ncmul.Plan(N, q868) as plan:
        plan.set_persist(0)
        plan.set_tuning(30, 2)
        want = run_device(plan, a, b)
        plan.set_persist(1)
        if not plan.persistent(B):
            pytest.skip("the device does not report its XCD count")
        got = run_device(plan, a, b)
        assert plan.persist_status() == 0
    assert np.array_equal(got, want)
    # closed forms: (q-1)^2 summed negacyclically gives c_k = 2k + 2 - N mod q
    qi = O.limbs_to_int(q868)
    for k_ in (0, 1, N // 2, N - 1):
        assert O.limbs_to_int(got[0][k_]) == (2 * k_ + 2 - N) % qi
    assert not got[1].any()


def test_persist_wait_bound_ends_launch(gpu, q868):
    """A readiness bound of 1 us, forced from the host, makes the NTT items
    of a one-product-per-XCD batch give up while the decompositions run: the
    launch still ends, the plan's error word says so, and the next launch
    with the normal bound is exact."""
    N, B = 16384, 8
    say("persist wait bound", N, B)
    a, b = batch_inputs(N, B, q868, 9500)
    orc = O.Plan(N, q868)
    with ncmul.Plan(N, q868) as plan:
        plan.set_persist(1, wait_bound_us=1)
        if not plan.persistent(B):
            pytest.skip("the device does not report its XCD count")
        run_device(plan, a, b)
        assert plan.persist_status() & 1
        assert plan.persist_status() == 0  # read and cleared
        plan.set_persist(1)
        got = run_device(plan, a, b)
        assert plan.persist_status() == 0
        for i in (0, B - 1):
            assert np.array_equal(got[i], orc.multiply(a[i], b[i], threads=0)), i


def test_persist_autotune_candidate(gpu, q868):
    """The autotuner times the persistent form at a headline-sized batch and
    keeps it only if its runs drained every queue; whatever it picks, the
    output stays exact."""
    N, B = 16384, 64
    say("persist autotune", N, B)
    a, b = batch_inputs(N, B, q868, 9700)
    orc = O.Plan(N, q868)
    with ncmul.Plan(N, q868) as plan:
        plan.autotune(B)
        say("persistent chosen:", plan.persistent(B))
        got = run_device(plan, a, b)
        assert plan.persist_status() == 0
        for i in (0, 17, B - 1):
            assert np.array_equal(got[i], orc.multiply(a[i], b[i], threads=0)), i


def test_persist_in_mixed_degree_graph(gpu, q868):
    """The mixed-degree graph batch with the persistent form on for every
    degree: one persistent launch per group of 2^10 .. 2^14 (groups below
    one product per XCD and 2^15 keep the launch sequence), captured in one
    HIP graph, three replays with eager work between them, every product
    equal to the plain path's; then 20 more replays, each compared on the
    device with the first."""
    import torch
    from test_gpu_batch_serve import make_groups, mixed_degrees, plain_products

    degrees = mixed_degrees(256, 10, 15, seed=14)
    say("persist graph degrees", sorted(set(degrees)))
    groups, host = make_groups(degrees, q868, 4000)
    want = plain_products(q868, host)
    all_n = sorted(set(degrees))
    with ncmul.Batch(q868) as batch:
        batch.set_persist(all_n, 1)
        graph = batch.capture(groups)
        for rep in range(3):
            for _, _, _, _, c in groups:
                c.fill_(-1)
            graph.launch()
            torch.cuda.synchronize()
            assert batch.persist_status(all_n) == 0, rep
            for N, n, a, b, c in groups:
                assert np.array_equal(ncmul.from_device_u32(c), want[N]), (rep, N)
            batch.run(groups[:2])
            torch.cuda.synchronize()
        first = [c.clone() for _, _, _, _, c in groups]
        for rep in range(20):
            for _, _, _, _, c in groups:
                c.fill_(-1)
            graph.launch()
            torch.cuda.synchronize()
            for (N, _, _, _, c), f in zip(groups, first):
                assert torch.equal(c, f), (rep, N)
        assert batch.persist_status(all_n) == 0


def test_persist_noalloc_scratch(gpu, q868):
    """The allocation-free entry point with the persistent form on: its
    control block sits in the scratch array's per-product trailer and is
    cleared by the form's reset kernel; two calls in a row, the second
    smaller, each equal to the plain launch sequence through
    nc_multiply_device."""
    N = 1024
    say("persist noalloc", N)
    with ncmul.Plan(N, q868) as plan:
        for B in (16, 9):
            a, b = batch_inputs(N, B, q868, 9900 + B)
            plan.set_persist(0)
            want = run_device(plan, a, b)
            plan.set_persist(1)
            if not plan.persistent(B):
                pytest.skip("the device does not report its XCD count")
            d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
            d_c = ncmul.empty_device_u32((B, N, L))
            d_c.fill_(-1)
            plan.multiply_device_noalloc(d_a, d_b, d_c, B)
            got = ncmul.from_device_u32(d_c)
            assert plan.persist_status() == 0, B
            assert np.array_equal(got, want), B


def test_persist_random_prime_modulus(gpu):
    """A second 868-bit modulus (a seeded random prime, whose reduction the
    reference's own Barrett step gets wrong): every product of an uneven
    batch against the oracle's exact reduction."""
    from conftest import modulus

    q = modulus("q868b")
    N, B = 16384, 11
    say("persist q868b", N, B)
    a, b = batch_inputs(N, B, q, 9800)
    orc = O.Plan(N, q)
    with ncmul.Plan(N, q) as plan:
        plan.set_persist(1)
        if not plan.persistent(B):
            pytest.skip("the device does not report its XCD count")
        got = run_device(plan, a, b)
        assert plan.persist_status() == 0
        for i in range(B):
            assert np.array_equal(got[i], orc.multiply(a[i], b[i], threads=0)), i


def test_persist_batch_256(gpu, q868):
    """The bench's headline batch, 256 products at N = 2^14 (32 per XCD),
    every product against the oracle."""
    N, B = 16384, 256
    say("persist batch 256", N, B)
    a, b = batch_inputs(N, B, q868, 12000)
    orc = O.Plan(N, q868)
    with ncmul.Plan(N, q868) as plan:
        plan.set_persist(1)
        if not plan.persistent(B):
            pytest.skip("the device does not report its XCD count")
        got = run_device(plan, a, b)
        assert plan.persist_status() == 0
    bad = [i for i in range(B) if not np.array_equal(got[i], orc.multiply(a[i], b[i], threads=0))]
    assert not bad, bad[:8]


def test_persist_repeated_launches(gpu, q868):
    """Hundreds of launches in a row, two operand sets alternating and the
    output cleared before each, every launch equal to the launch sequence's
    products (two of which are checked against the oracle).  Built for three
    waves per SIMD (168 VGPRs, about 104 VGPRs and 64 SGPRs spilled), the
    kernel returned one wave's share of wrong output in 2-7 % of such
    launches; the two-wave build (16 VGPRs spilled once at entry) returned
    none in 1200."""
    import torch

    for N, B, launches in ((4096, 64, 300), (16384, 128, 150), (16384, 256, 60)):
        say("persist repeated", N, B, launches)
        sets = [batch_inputs(N, B, q868, 15000 + 1000 * s_) for s_ in range(2)]
        dev = [(ncmul.to_device_u32(a), ncmul.to_device_u32(b)) for a, b in sets]
        d_c = ncmul.empty_device_u32((B, N, L))
        with ncmul.Plan(N, q868) as plan:
            plan.set_tuning(30, 2)
            plan.set_chunks(1)
            plan.set_persist(0)
            refs = []
            for d_a, d_b in dev:
                d_c.fill_(-1)
                plan.multiply_device(d_a, d_b, d_c, B)
                torch.cuda.synchronize()
                refs.append(d_c.clone())
            if N == 4096:
                orc = O.Plan(N, q868)
                got = ncmul.from_device_u32(refs[0])
                for i in (0, B - 1):
                    assert np.array_equal(got[i], orc.multiply(sets[0][0][i], sets[0][1][i], threads=0)), i
            plan.set_persist(1)
            if not plan.persistent(B):
                pytest.skip("the device does not report its XCD count")
            bad = []
            for it in range(launches):
                d_a, d_b = dev[it & 1]
                d_c.fill_(-1)
                plan.multiply_device(d_a, d_b, d_c, B)
                torch.cuda.synchronize()
                if not torch.equal(d_c, refs[it & 1]):
                    wrong = torch.nonzero((d_c != refs[it & 1]).reshape(B, -1).any(dim=1)).flatten().tolist()
                    bad.append((it, wrong[:8]))
            assert plan.persist_status() == 0
            assert not bad, (N, B, len(bad), bad[:4])
