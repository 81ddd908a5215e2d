"""The no-allocation entry point (awamd_thread_forward_noalloc) against
awamd_thread_forward, through the C ABI: two calls in a row, the second on
a smaller instance, each bit-identical to the allocating entry point (same
kernels, same schedule), in FP64 and FP32; an instance too large for the
fixed workspace is refused with an error before any kernel runs on it."""
import numpy as np
import pytest

from workloads import instance

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_allocating_entry_point(aw, prec):
    for k, L, seed in [(10, 6000, 601), (6, 1500, 602)]:   # second call: smaller input
        print("noalloc prec=%d k=%d L=%d seed=%d" % (prec, k, L, seed), flush=True)
        tr, seqs, _, (times, pops, rho, mu) = instance(aw, k, L, seed)
        u = np.random.RandomState(seed).uniform(size=L)
        ref = aw.thread_forward(tr, seqs, times, pops, rho, mu, k - 1, uniforms=u,
                                precision=prec)
        got = aw.thread_forward(tr, seqs, times, pops, rho, mu, k - 1, uniforms=u,
                                precision=prec, noalloc=True)
        assert np.array_equal(got["alpha"], ref["alpha"])
        assert np.array_equal(got["path"], ref["path"])
        assert got["lnl"] == ref["lnl"]
        aw.delete_local_trees(tr)


def test_noalloc_refuses_instance_over_workspace(aw):
    # 40 haplotypes over 60000 columns: the FP64 forward table alone is
    # 60000 * 64R * 8 bytes, over the 64 MiB workspace
    k, L, seed = 40, 60000, 603
    print("noalloc too large k=%d L=%d" % (k, L), flush=True)
    tr, seqs, _, (times, pops, rho, mu) = instance(aw, k, L, seed)
    u = np.random.RandomState(seed).uniform(size=L)
    with pytest.raises(RuntimeError):
        aw.thread_forward(tr, seqs, times, pops, rho, mu, k - 1, uniforms=u, noalloc=True)
    # the workspace is still usable after the refusal
    k, L, seed = 6, 1500, 602
    tr2, seqs2, _, m2 = instance(aw, k, L, seed)
    u2 = np.random.RandomState(seed).uniform(size=L)
    ref = aw.thread_forward(tr2, seqs2, *m2, k - 1, uniforms=u2)
    got = aw.thread_forward(tr2, seqs2, *m2, k - 1, uniforms=u2, noalloc=True)
    assert np.array_equal(got["alpha"], ref["alpha"]) and np.array_equal(got["path"], ref["path"])
    aw.delete_local_trees(tr)
    aw.delete_local_trees(tr2)
