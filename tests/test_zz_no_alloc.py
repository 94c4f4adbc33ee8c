"""GPU test of nc_multiply_device_noalloc (include/negacyclic.h), the
allocation-free variant of nc_multiply_device: its workspace is the
library's per-device scratch array.  Every product must equal what
nc_multiply_device returns for the same operands, bit for bit, on calls in
a row whose sizes shrink (the array keeps the previous call's contents), and
one product per shape is also checked against the oracle."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import ncmul
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _operands(N, batch, q, seed0):
    a = np.stack([ncmul.gen_uniform(2 * (seed0 + i), N, q) for i in range(batch)])
    b = np.stack([ncmul.gen_uniform(2 * (seed0 + i) + 1, N, q) for i in range(batch)])
    return a, b


def _both(plan, a, b):
    """(noalloc result, nc_multiply_device result) of the same operands; the
    outputs start poisoned so nothing stale can pass."""
    batch, N, L = a.shape
    d_a, d_b = ncmul.to_device_u32(a), ncmul.to_device_u32(b)
    d_c0 = ncmul.empty_device_u32((batch, N, L))
    d_c1 = ncmul.empty_device_u32((batch, N, L))
    d_c0.fill_(-1)
    d_c1.fill_(-1)
    plan.multiply_device_noalloc(d_a, d_b, d_c0, batch)
    torch.cuda.synchronize()
    got = ncmul.from_device_u32(d_c0)
    plan.multiply_device(d_a, d_b, d_c1, batch)
    torch.cuda.synchronize()
    return got, ncmul.from_device_u32(d_c1)


def test_noalloc_equals_device_entry_point(gpu, q868):
    N = 16384
    with ncmul.Plan(N, q868) as plan:
        per = plan.info.bytes_per_product
        cap = ncmul.NOALLOC_SCRATCH_BYTES // per
        say("N", N, "bytes_per_product", per, "fits", cap)
        assert cap >= 4
        # two calls in a row, the second smaller
        for batch, seed in ((cap, 100), (2, 200)):
            say("noalloc N", N, "batch", batch)
            a, b = _operands(N, batch, q868, seed)
            got, want = _both(plan, a, b)
            for i in range(batch):
                assert np.array_equal(got[i], want[i]), (batch, i)
        assert np.array_equal(got[0], O.Plan(N, q868).multiply(a[0], b[0]))
        # beyond the array: a clear error from the host check, nothing launched
        d = ncmul.empty_device_u32((cap + 1, N, 28))
        with pytest.raises(ncmul.NcError) as ei:
            plan.multiply_device_noalloc(d, d, d, cap + 1)
        assert ei.value.status == -7 and "scratch" in str(ei.value)


def test_noalloc_shapes_in_a_row(gpu, q868):
    # the largest degree (one product fills most of the array), then smaller
    # degrees and batches on other plans sharing the same array
    for N, batch, seed in ((65536, 1, 300), (1024, 8, 400), (4096, 3, 500)):
        say("noalloc N", N, "batch", batch)
        with ncmul.Plan(N, q868) as plan:
            a, b = _operands(N, batch, q868, seed)
            got, want = _both(plan, a, b)
            for i in range(batch):
                assert np.array_equal(got[i], want[i]), (N, batch, i)
            if N <= 4096:
                assert np.array_equal(got[-1], O.Plan(N, q868).multiply(a[-1], b[-1]))
