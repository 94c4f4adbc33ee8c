"""The persistent batched form is switched off in this library build: forcing
it is refused with NC_ERR_UNSUPPORTED, no batch size reports it, the
autotuner leaves it out, and calls run the launch sequence, exact against the
oracle (the reference's solve.cu restatement)."""
from __future__ import annotations

import numpy as np
import pytest

import ncmul
import oracle as O

pytestmark = pytest.mark.gpu

L = 28


def test_persist_mode_refused(gpu, q868):
    N, B = 16384, 16
    print("persist off", N, B, flush=True)
    a = np.stack([ncmul.gen_uniform(13100 + 2 * i, N, q868) for i in range(B)])
    b = np.stack([ncmul.gen_uniform(13101 + 2 * i, N, q868) for i in range(B)])
    orc = O.Plan(N, q868)
    with ncmul.Plan(N, q868) as plan:
        with pytest.raises(ncmul.NcError) as err:
            plan.set_persist(1)
        assert err.value.status == ncmul.NC_ERR_UNSUPPORTED
        plan.set_persist(0)
        plan.set_persist(-1)
        plan.autotune(B)
        assert not plan.persistent(B)
        d_c = ncmul.empty_device_u32((B, N, L))
        d_c.fill_(-1)
        plan.multiply_device(ncmul.to_device_u32(a), ncmul.to_device_u32(b), d_c, B)
        got = ncmul.from_device_u32(d_c)
        assert plan.persist_status() == 0
    for i in (0, 7, B - 1):
        assert np.array_equal(got[i], orc.multiply(a[i], b[i], threads=0)), i


def test_batch_persist_mode_refused(gpu, q868):
    with ncmul.Batch(q868) as batch:
        with pytest.raises(ncmul.NcError) as err:
            batch.set_persist([1024, 16384], 1)
        assert err.value.status == ncmul.NC_ERR_UNSUPPORTED
