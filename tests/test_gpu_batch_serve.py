"""GPU tests of the batched and served forms (include/negacyclic_batch.h):
every product they return must equal the plain path's product of the same
operands, bit for bit, and the plain path is itself pinned to the oracle and
the golden vectors (test_gpu_parity.py); here a few products are also checked
against the oracle directly.

  * mixed-degree batch: 256 products, N log-uniform in {2^10, ..., 2^15},
    captured as one HIP graph; every replay (with eager work and
    synchronisation between replays, and with new operands written into the
    bound buffers) equals the plain path;
  * serving: requests of mixed degree submitted together (continuous
    batching, overlapping lanes) return what each returns alone; a tight HBM
    budget forces admission waits without changing results; an open-loop
    Poisson replay completes every request correctly.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import ncmul
import oracle as O
from conftest import modulus

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def mixed_degrees(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return [1 << int(x) for x in rng.integers(lo, hi + 1, size=n)]


def make_groups(degrees, q, seed0):
    """Group products by degree; returns (groups, host operands per product)."""
    L = q.shape[0]
    by = {}
    for i, N in enumerate(degrees):
        by.setdefault(N, []).append(i)
    groups, host = [], {}
    for N, idx in sorted(by.items()):
        a = np.stack([ncmul.gen_uniform(2 * (seed0 + i), N, q) for i in idx])
        b = np.stack([ncmul.gen_uniform(2 * (seed0 + i) + 1, N, q) for i in idx])
        d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
        d_c = ncmul.empty_device_u32((len(idx), N, L))
        groups.append((N, len(idx), d_a, d_b, d_c))
        host[N] = (idx, a, b)
    return groups, host


def plain_products(q, host):
    """The plain path: one product per call (nc_multiply), per degree plan."""
    out = {}
    for N, (idx, a, b) in host.items():
        with ncmul.Plan(N, q) as p:
            out[N] = np.stack([p.multiply(a[i], b[i]) for i in range(len(idx))])
    return out


def test_mixed_degree_graph_matches_plain(gpu, q868):
    degrees = mixed_degrees(256, 10, 15, seed=4)
    say("mixed batch degrees", sorted(set(degrees)))
    groups, host = make_groups(degrees, q868, 1000)
    want = plain_products(q868, host)
    with ncmul.Batch(q868) as batch:
        batch.autotune([(N, n, a, b, c) for N, n, a, b, c in groups])
        graph = batch.capture(groups)
        for rep in range(3):
            for _, _, _, _, c in groups:
                c.zero_()
            graph.launch()
            torch.cuda.synchronize()
            for N, n, a, b, c in groups:
                assert np.array_equal(ncmul.from_device_u32(c), want[N]), (rep, N)
            # eager work on the same plans between replays
            batch.run(groups[:2])
            torch.cuda.synchronize()
        # the graph is bound to the buffers: new operands give new products
        N, n, d_a, d_b, d_c = groups[-1]
        d_a.copy_(torch.flip(d_a, dims=[0]))
        graph.launch()
        torch.cuda.synchronize()
        got, a_new, b_host = ncmul.from_device_u32(d_c), ncmul.from_device_u32(d_a), host[N][2]
        with ncmul.Plan(N, q868) as p:
            for i in range(n):
                assert np.array_equal(got[i], p.multiply(a_new[i], b_host[i])), i
        assert np.array_equal(got[0], O.Plan(N, q868).multiply(a_new[0], b_host[0]))


def test_batch_run_matches_graph(gpu, q868):
    degrees = mixed_degrees(24, 10, 13, seed=6)
    groups, host = make_groups(degrees, q868, 3000)
    with ncmul.Batch(q868) as batch:
        batch.run(groups)
        torch.cuda.synchronize()
        eager = {N: ncmul.from_device_u32(c) for N, _, _, _, c in groups}
        graph = batch.capture(groups)
        for _, _, _, _, c in groups:
            c.zero_()
        graph.launch()
        torch.cuda.synchronize()
        for N, _, _, _, c in groups:
            assert np.array_equal(ncmul.from_device_u32(c), eager[N]), N
        N0 = groups[0][0]
        idx, a, b = host[N0]
        assert np.array_equal(eager[N0][0], O.Plan(N0, q868).multiply(a[0], b[0]))


def test_batch_rejects_bad_groups(gpu, q868):
    with ncmul.Batch(q868) as batch:
        with pytest.raises(ncmul.NcError):
            batch.run([])
        d = ncmul.empty_device_u32((1, 1000, 28))
        with pytest.raises(ncmul.NcError):
            batch.run([(1000, 1, d, d, d)])  # not a power of two


def _requests(degrees, q, seed0):
    L = q.shape[0]
    reqs = []
    for i, N in enumerate(degrees):
        a, b, c = ncmul.host_array((N, L)), ncmul.host_array((N, L)), ncmul.host_array((N, L))
        a[:] = ncmul.gen_uniform(2 * (seed0 + i), N, q)
        b[:] = ncmul.gen_uniform(2 * (seed0 + i) + 1, N, q)
        c[:] = 0
        reqs.append((N, a, b, c))
    return reqs


def _alone(q, reqs):
    plans, out = {}, []
    try:
        for N, a, b, _ in reqs:
            if N not in plans:
                plans[N] = ncmul.Plan(N, q)
            out.append(plans[N].multiply(a, b))
    finally:
        for p in plans.values():
            p.close()
    return out


def test_server_mix_equals_alone(gpu, q868):
    degrees = mixed_degrees(80, 11, 14, seed=8)
    reqs = _requests(degrees, q868, 5000)
    alone = _alone(q868, reqs)
    with ncmul.Server(q868, sorted(set(degrees)), max_batch=8, lanes=2, autotune=False) as srv:
        tickets = [srv.submit(a, b, c) for _, a, b, c in reqs]  # all in flight together
        lat = [srv.wait(t, timeout_ms=60000) for t in tickets]
        st = srv.stats()
    say("server stats", st)
    assert all(x > 0 for x in lat)
    assert st["completed"] == len(reqs) and st["launches"] < len(reqs)  # batched, not one per request
    for i, (N, a, b, c) in enumerate(reqs):
        assert np.array_equal(c, alone[i]), (i, N)
    N, a, b, c = reqs[0]
    assert np.array_equal(c, O.Plan(N, q868).multiply(np.asarray(a), np.asarray(b)))


def test_server_transfer_paths(gpu, q868):
    """Requests (below 1 MiB per buffer) whose operands and result all lie
    in mapped pinned memory move by the transfer kernels, the others by
    copies (pageable memory,
    pinned memory at an address that is not 8-byte aligned); both kinds in
    one batch, pinned views at offsets inside a larger allocation included,
    and every result exact."""
    N, L = 4096, q868.shape[0]
    reqs = _requests([N] * 6, q868, 13000)  # all pinned
    n0, a0, b0, c0 = reqs[1]
    reqs[1] = (N, a0, b0, np.zeros((N, L), np.uint32))  # pageable result
    n0, a0, b0, c0 = reqs[2]
    reqs[2] = (N, np.array(a0), np.array(b0), c0)  # pageable operands
    odd = ncmul.host_array((N * L + 1,))  # pinned, 4 bytes past 8-byte alignment
    n0, a0, b0, c0 = reqs[3]
    odd_c = odd[1:].reshape(N, L)
    odd_c[:] = 0
    reqs[3] = (N, a0, b0, odd_c)
    rows = ncmul.host_array((3, N, L))  # pinned rows at offsets inside one allocation
    n0, a0, b0, c0 = reqs[4]
    rows[0], rows[2] = a0, b0
    reqs[4] = (N, rows[0], rows[2], rows[1])
    alone = _alone(q868, reqs)
    with ncmul.Server(q868, [N], max_batch=8, lanes=1, autotune=False) as srv:
        tickets = [srv.submit(a, b, c) for _, a, b, c in reqs]
        for t in tickets:
            srv.wait(t, timeout_ms=60000)
        st = srv.stats()
    say("transfer paths", st)
    assert st["mapped_requests"] == 3  # requests 0, 4 and 5
    for i, (_, a, b, c) in enumerate(reqs):
        assert np.array_equal(c, alone[i]), i


def test_server_admission_budget(gpu, q868):
    degrees = [4096] * 24 + [8192] * 8
    reqs = _requests(degrees, q868, 7000)
    alone = _alone(q868, reqs)
    with ncmul.Server(q868, [4096, 8192], max_batch=8, lanes=2, hbm_budget=1 << 40, autotune=False) as probe:
        f4, f8 = probe.footprint(4096), probe.footprint(8192)
    budget = 2 * f8 + f4  # room for about two requests at a time
    with ncmul.Server(q868, [4096, 8192], max_batch=8, lanes=2, hbm_budget=budget, autotune=False) as srv:
        tickets = [srv.submit(a, b, c) for _, a, b, c in reqs]
        for t in tickets:
            srv.wait(t, timeout_ms=60000)
        st = srv.stats()
    say("admission stats", st, "budget", budget)
    # the arena (all the device memory requests get) stays within the budget
    assert st["device_bytes"] <= budget
    assert st["peak_bytes"] <= st["device_bytes"] and st["admission_waits"] > 0
    for i, (N, a, b, c) in enumerate(reqs):
        assert np.array_equal(c, alone[i]), (i, N)


def test_server_rejects(gpu, q868):
    with pytest.raises(ncmul.NcError):  # budget below one request
        ncmul.Server(q868, [1024], hbm_budget=1024, autotune=False)
    with ncmul.Server(q868, [1024], autotune=False) as srv:
        a = ncmul.host_array((2048, 28))
        with pytest.raises(ncmul.NcError):  # degree not served
            srv.submit(a, a, a)
        with pytest.raises(ncmul.NcError):
            srv.wait(12345, timeout_ms=0)


def test_server_poisson_replay(gpu, q868):
    n = 200
    degrees = mixed_degrees(n, 11, 16, seed=10)
    reqs = _requests(degrees, q868, 9000)
    rng = np.random.default_rng(11)
    arrivals = np.cumsum(rng.exponential(1.0 / 2000.0, size=n))  # 2000 requests/s
    with ncmul.Server(q868, sorted(set(degrees)), max_batch=16, lanes=2, autotune=False) as srv:
        lat, makespan = srv.replay([r[0] for r in reqs], arrivals, [r[1] for r in reqs], [r[2] for r in reqs],
                                   [r[3] for r in reqs])
        st = srv.stats()
    say("replay", n, "makespan", makespan, "p50", np.percentile(lat, 50), "p99", np.percentile(lat, 99), st)
    assert st["completed"] == n and (lat > 0).all() and makespan >= arrivals[-1]
    # every buffer pinned: the requests below 1 MiB per buffer moved by the transfer kernels
    assert st["mapped_requests"] == sum(N * 28 * 4 < (1 << 20) for N in degrees)
    # every response, each with its own operands and result buffer
    alone = _alone(q868, reqs)
    for i in range(n):
        assert np.array_equal(reqs[i][3], alone[i]), i


def test_server_arena_bounds_device_memory(gpu, q868):
    """The request arena is min(hbm_budget, every lane full) bytes; with a
    budget below what max_batch x lanes would take, batches shrink to what
    fits and every result stays exact."""
    degrees = [16384] * 12
    reqs = _requests(degrees, q868, 11000)
    alone = _alone(q868, reqs)
    with ncmul.Server(q868, [16384], max_batch=8, lanes=2, hbm_budget=1 << 40, autotune=False) as probe:
        f = probe.footprint(16384)
        full = probe.stats()["device_bytes"]
    assert full <= 2 * 8 * f  # no more than every lane holding a full batch
    budget = 3 * f + 1000  # three requests at a time, not 16
    with ncmul.Server(q868, [16384], max_batch=8, lanes=2, hbm_budget=budget, autotune=False) as srv:
        tickets = [srv.submit(a, b, c) for _, a, b, c in reqs]
        for t in tickets:
            srv.wait(t, timeout_ms=60000)
        st = srv.stats()
    say("arena stats", st, "budget", budget, "footprint", f)
    assert st["device_bytes"] <= budget and st["peak_bytes"] <= budget
    assert st["max_batch_seen"] <= 3  # max_batch is 8: the arena, not the lanes, limited the batches
    for i in range(len(reqs)):
        assert np.array_equal(reqs[i][3], alone[i]), i
