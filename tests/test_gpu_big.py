"""Blocks of more than 1024 states (32 states per lane of the wave): 128
haplotypes under a population size of 3e4 give local trees with 1050-1110
states.  FP64 table, log-likelihood and thread against the oracle; the
drop-in sample_thread against the compiled reference; a batch that mixes
such an instance with small ones, FP32 forward tiles and traceback tiles,
against each instance run alone."""
import numpy as np
import pytest

import oracle_bind as O
from workloads import instance

pytestmark = pytest.mark.gpu

BIG = dict(k=128, L=1500, popsizes=np.full(20, 3e4))


def _big(aw, seed):
    tr, seqs, inst, mdl = instance(aw, BIG["k"], BIG["L"], seed, popsizes=BIG["popsizes"])
    _, S = aw.num_column_states(tr, 20)
    assert S.max() > 1024, S.max()
    return tr, seqs, inst, mdl


@pytest.mark.parametrize("seed", [1, 2])
def test_big_blocks_match_oracle(aw, seed):
    print("big blocks seed=%d" % seed, flush=True)
    tr, seqs, inst, (times, pops, rho, mu) = _big(aw, seed)
    L = BIG["L"]
    fw, lnl = inst.forward()
    u = np.random.RandomState(seed).uniform(size=L)
    out = aw.thread_forward(tr, seqs, times, pops, rho, mu, BIG["k"] - 1, uniforms=u)
    rel = np.abs(out["alpha"] - fw) / np.maximum(np.abs(fw), 1e-300)
    assert rel.max() <= 1e-12, rel.max()
    assert abs(out["lnl"] - lnl) <= 1e-12 * abs(lnl)
    assert np.array_equal(out["path"], inst.traceback(fw, u))
    o32 = aw.thread_forward(tr, seqs, times, pops, rho, mu, BIG["k"] - 1, uniforms=u,
                            precision=1)
    assert abs(o32["lnl"] - lnl) <= 1e-6 * abs(lnl)
    assert np.mean(o32["path"] == out["path"]) >= 0.98
    aw.delete_local_trees(tr)


@pytest.mark.skipif(not O.ref_available(), reason="reference library oracle/_ref not built")
def test_big_blocks_sample_thread_matches_reference(aw):
    """Same srand seed: the threaded ARG is identical to the compiled
    reference's, node numbering included."""
    print("big blocks sample_thread", flush=True)
    from workloads import model
    times, _, rho, mu = model()
    pops = BIG["popsizes"]
    k, L = BIG["k"], BIG["L"]
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, 3)
    P, Ag, S, bl = aw.trees_arrays(tr)
    rt = O.ref_new_trees(P, Ag, S, bl)
    aw.remove_thread(tr, k - 1)
    O.ref_remove_thread(rt, k - 1)
    assert aw.num_column_states(tr, 20)[1].max() > 1024
    O.srand(703)
    aw.sample_thread(tr, seqs, times, pops, rho, mu)
    O.srand(703)
    O.ref_sample_thread(rt, seqs, times, pops, rho, mu)
    mine, ref = aw.trees_arrays(tr), O.ref_trees_arrays(rt)
    assert all(np.array_equal(a, b) for a, b in zip(mine, ref))
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


def test_big_block_in_mixed_batch(aw):
    """A big instance beside small ones in one FP32 batch (forward tiles with
    re-runs, traceback tiles with options): each instance's log-likelihood,
    thread and table equal the instance run alone."""
    print("big block in a mixed batch", flush=True)
    items = []
    tr, seqs, _, mdl = _big(aw, 4)
    items.append((tr, seqs, mdl, BIG["k"] - 1, BIG["L"]))
    for j, (k, L) in enumerate([(8, 1200), (20, 1500), (40, 900)]):
        t2, s2, _, m2 = instance(aw, k, L, 60 + j)
        items.append((t2, s2, m2, k - 1, L))

    def batch(sel):
        b = aw.Batch()
        for i in sel:
            t, s, m, nc, L = items[i]
            b.add(t, s, *m, nc, np.random.RandomState(100 + i).uniform(size=L))
        b.upload()
        return b

    kw = dict(fw_tile=400, fw_warm=60, fw_tol=1e-5)
    b = batch(range(len(items)))
    b.run(1, 200, **kw)
    for i in range(len(items)):
        one = batch([i])
        one.run(1, 200, **kw)
        assert one.lnl(0) == b.lnl(i), i
        assert np.array_equal(one.path(0), b.path(i)), i
        assert np.array_equal(one.alpha(0), b.alpha(i)), i
    for t, *_ in items:
        aw.delete_local_trees(t)
