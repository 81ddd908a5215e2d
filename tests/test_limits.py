"""Host-side limits (no GPU): an instance over the kernels' limits is
refused when it is added to a batch, before any device work."""
import numpy as np
import pytest

from workloads import instance


def test_batch_refuses_instance_over_node_limit(aw):
    # 130 haplotypes, one removed: 129 leaves, 257 nodes per tree (limit 256)
    tr, seqs, _, mdl = instance(aw, 130, 300, 81)
    b = aw.Batch()
    with pytest.raises(RuntimeError):
        b.add(tr, seqs, *mdl, 129, np.random.RandomState(1).uniform(size=300))
    aw.delete_local_trees(tr)


def test_batch_takes_instance_at_node_limit(aw):
    # 129 haplotypes, one removed: 255 nodes per tree
    tr, seqs, _, mdl = instance(aw, 129, 200, 82)
    b = aw.Batch()
    try:
        assert b.add(tr, seqs, *mdl, 128, None) == 0
    except RuntimeError:
        # over 2048 states in a block is the other limit
        pytest.skip("instance exceeds the states-per-block limit")
    aw.delete_local_trees(tr)


def test_traceback_options_out_of_range(aw):
    # options per traceback tile: 1 .. 4; warm-up columns >= 0
    b = aw.Batch()
    b.set_traceback(0, 1)
    b.set_traceback(4096, 4)
    for bad in ((256, 0), (256, 5), (-1, 3)):
        with pytest.raises(ValueError):
            b.set_traceback(*bad)


def test_tail_tiles_out_of_range(aw):
    # the re-run instances' traceback tile length: >= 0 (0: the run's)
    b = aw.Batch()
    b.set_tail_tiles(0)
    b.set_tail_tiles(4000)
    with pytest.raises(ValueError):
        b.set_tail_tiles(-1)
