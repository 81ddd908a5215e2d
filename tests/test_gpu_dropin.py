"""The reference's C entry points, called through this library on the GPU,
against the same calls into the reference library compiled from its sources
(oracle/_ref): forward tables, and the ARGs that sample_thread / resample_arg
/ sample_arg_seq produce from the same glibc rand() stream."""
import numpy as np
import pytest

import oracle_bind as O
from workloads import model

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.ref_available(),
                                 reason="reference library oracle/_ref not built")]


def _pair(aw, k, L, seed, remove=True):
    """The same ARG in both libraries: simulated here, rebuilt in the
    reference from the full arrays, then the same leaf removed in each."""
    times, pops, rho, mu = model()
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    P, Ag, S, bl = aw.trees_arrays(tr)
    rt = O.ref_new_trees(P, Ag, S, bl)
    if remove:
        aw.remove_thread(tr, k - 1)
        O.ref_remove_thread(rt, k - 1)
    return tr, rt, seqs, (times, pops, rho, mu)


def _same_arg(aw, tr, rt):
    mine = aw.trees_arrays(tr)
    ref = O.ref_trees_arrays(rt)
    return all(np.array_equal(a, b) for a, b in zip(mine, ref))


@pytest.mark.parametrize("k,L,seed", [(6, 800, 31), (10, 2500, 32)])
def test_arghmm_forward_alg(aw, k, L, seed):
    print("forward_alg k=%d L=%d seed=%d" % (k, L, seed), flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, k, L, seed)
    rows = aw.forward_algorithm(tr, seqs, times, pops, rho, mu)
    nst = aw.get_nstates(tr, len(times), L)
    ref = O.ref_forward(rt, seqs, times, pops, rho, mu, nst)
    for c in range(L):
        a, b = rows[c], ref[c]
        assert a.shape == b.shape, c
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        assert rel.max() <= 1e-12, (c, rel.max())
    O.ref_delete(rt)


@pytest.mark.parametrize("k,L,seed", [(6, 800, 41), (12, 2000, 42)])
def test_arghmm_sample_thread(aw, k, L, seed):
    """Same srand seed: the threaded ARG is identical, node numbering included."""
    print("sample_thread k=%d L=%d seed=%d" % (k, L, seed), flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, k, L, seed)
    O.srand(500 + seed)
    aw.sample_thread(tr, seqs, times, pops, rho, mu)
    O.srand(500 + seed)
    O.ref_sample_thread(rt, seqs, times, pops, rho, mu)
    assert _same_arg(aw, tr, rt)
    O.ref_delete(rt)


def test_arghmm_resample_arg(aw):
    print("resample_arg k=6 L=600", flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, 6, 600, 51, remove=False)
    O.srand(77)
    aw.resample_arg(tr, seqs, times, pops, rho, mu, niters=2)
    O.srand(77)
    O.ref_resample_arg(rt, seqs, times, pops, rho, mu, 2)
    assert _same_arg(aw, tr, rt)
    O.ref_delete(rt)


def test_arghmm_sample_arg_seq(aw):
    print("sample_arg_seq k=4 L=400", flush=True)
    times, pops, rho, mu = model()
    _, seqs = aw.simulate(4, 400, times, pops, rho, mu, 61)
    O.srand(99)
    tr = aw.sample_arg_seq(seqs, times, pops, rho, mu)
    O.srand(99)
    rt = O.ref_sample_arg_seq(seqs, times, pops, rho, mu)
    assert _same_arg(aw, tr, rt)
    O.ref_delete(rt)
