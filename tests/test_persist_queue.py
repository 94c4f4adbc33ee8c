"""CPU checks of the persistent form's work queue (nc_persist_emulate): the
schedule the kernel decodes, run by an emulation of workgroups taking queue
positions in order -- each fetching its next position before running the
current item, as the kernel does -- with random item durations.  Every item
must run exactly once and only after its inputs, and every XCD's queue must
drain, for any number of workgroups (one workgroup per XCD is the case that
deadlocks if an item's inputs ever came after it in the queue).
"""
from __future__ import annotations

import pytest

import ncmul


@pytest.mark.parametrize("N", [1024, 2048, 4096, 8192, 16384])
@pytest.mark.parametrize("batch,xcds,wgs", [(8, 8, 96), (128, 8, 96), (256, 8, 96), (9, 8, 1), (17, 8, 2),
                                            (128, 8, 1), (64, 4, 96), (3, 1, 5), (1, 1, 1), (40, 8, 7)])
def test_queue_drains_in_dependency_order(N, batch, xcds, wgs):
    for seed in (1, 2):
        st = ncmul.persist_emulate(N, 59, batch, xcds, wgs, seed)
        assert st["makespan"] > 0
        assert 0.0 <= st["wait_fraction"] < 1.0


def test_queue_other_prime_counts():
    for k in (57, 60, 64):
        ncmul.persist_emulate(16384, k, 128, 8, 96, 3)


def test_headline_queue_rarely_waits():
    """At the headline shape (N = 2^14, 128 products, 96 workgroups per XCD)
    an item's inputs sit a whole round earlier in the queue, so workgroups
    spend almost no time waiting."""
    st = ncmul.persist_emulate(16384, 59, 128, 8, 96, 11)
    assert st["wait_fraction"] < 0.05


def test_queue_rejects_shapes_outside_the_form():
    with pytest.raises(ncmul.NcError):
        ncmul.persist_emulate(32768, 59, 128, 8, 96, 1)
    with pytest.raises(ncmul.NcError):
        ncmul.persist_emulate(16384, 59, 128, 0, 96, 1)
