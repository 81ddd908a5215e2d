"""ARG probability exports (host code, no GPU) against the compiled
reference: the log-likelihood of the sequences given the ARG, the prior of
the trees and SPRs, their sum and the first tree's prior, on seeded ARGs,
including masked ('N') sites and the reference's reuse of a tree's first
invariant-site likelihood for its later invariant sites."""
import ctypes

import numpy as np
import pytest

import oracle_bind as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _ref_probs(tr_arrays, seqs, times, pops, rho, mu):
    r = O.ref()
    for f in ("arghmm_likelihood", "arghmm_prior_prob", "arghmm_tree_prior_prob",
              "arghmm_joint_prob"):
        getattr(r, f).restype = ctypes.c_double
    rt = O.ref_new_trees(*tr_arrays)
    s = O._Seqs(seqs)
    k, L = s.arr.shape
    t = np.ascontiguousarray(times, np.float64)
    p = np.ascontiguousarray(pops, np.float64)
    vt = ctypes.c_void_p(rt)
    out = (r.arghmm_likelihood(vt, O._dp(t), len(t), ctypes.c_double(mu), s.ptr, k, L),
           r.arghmm_prior_prob(vt, O._dp(t), len(t), O._dp(p), ctypes.c_double(rho)),
           r.arghmm_tree_prior_prob(vt, O._dp(t), len(t), O._dp(p)),
           r.arghmm_joint_prob(vt, O._dp(t), len(t), O._dp(p), ctypes.c_double(mu),
                               ctypes.c_double(rho), s.ptr, k, L))
    O.ref_delete(rt)
    return out


@needs_ref
@pytest.mark.parametrize("k,L,seed,masked", [(6, 3000, 1, False), (12, 8000, 2, False),
                                              (20, 5000, 3, True)])
def test_arg_probabilities_match_reference(aw, k, L, seed, masked):
    times = np.array(aw.get_time_points(20, 200e3))
    rng = np.random.RandomState(seed)
    pops = np.exp(rng.uniform(np.log(5e3), np.log(5e4), size=20))
    rho, mu = 1.6e-7, 1.8e-7
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    if masked:
        seqs = seqs.copy()
        seqs[rng.uniform(size=seqs.shape) < 0.02] = ord("N")
        seqs[:, ::97] = ord("N")   # all-masked columns: invariant, likelihood 1
    ref = _ref_probs(aw.trees_arrays(tr), seqs, times, pops, rho, mu)
    mine = (aw.arg_likelihood(tr, seqs, times, mu), aw.arg_prior(tr, times, pops, rho),
            aw.tree_prior(tr, times, pops), aw.arg_joint_prob(tr, seqs, times, pops, mu, rho))
    # same terms summed in the same order as the reference: bit-identical
    for name, a, b in zip(("likelihood", "prior", "tree prior", "joint"), mine, ref):
        assert a == b, (name, a, b)
    assert mine[3] == mine[0] + mine[1]
    aw.delete_local_trees(tr)


def test_likelihood_of_identical_sequences(aw):
    """Reference-free: with every sequence equal, every column is invariant
    and the log-likelihood is the number of columns times the log of the
    first column's likelihood; masking a leaf's whole sequence leaves a
    likelihood no smaller (a masked leaf sums over its base)."""
    times = np.array(aw.get_time_points(20, 200e3))
    pops = np.full(20, 1e4)
    tr, seqs = aw.simulate(8, 2000, times, pops, 1.6e-7, 1.8e-7, 9)
    same = np.repeat(seqs[:1], 8, axis=0)
    lk = aw.arg_likelihood(tr, same, times, 1.8e-7)
    assert lk < 0
    one = aw.arg_likelihood(tr, same[:, :1].repeat(2000, axis=1), times, 1.8e-7)
    # per-tree reuse of the first invariant column: equal bases give equal
    # site likelihoods (JC69 is symmetric in the base), so both sums agree
    assert abs(lk - one) <= 1e-9 * abs(one)
    masked = same.copy()
    masked[3] = ord("N")
    assert aw.arg_likelihood(tr, masked, times, 1.8e-7) >= lk
    aw.delete_local_trees(tr)
