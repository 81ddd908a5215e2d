"""log P(sequences | ARG) and log P(ARG) on the device (awamd_arg_likelihood,
awamd_arg_prior) against the host statistics arghmm_likelihood and
arghmm_prior_prob, which tests/test_arg_stats.py pins bit for bit to the
compiled reference.  The device sums the same terms in another order, so
the bound is relative 1e-11 (a few thousand to 120000 terms of magnitude
about 1 to 30)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,L,seed,masked", [(8, 3000, 1, False), (20, 20000, 2, True),
                                             (60, 120000, 3, False), (128, 1500, 4, True)])
def test_device_statistics_match_host(aw, k, L, seed, masked):
    print("device likelihood k=%d L=%d seed=%d masked=%s" % (k, L, seed, masked), flush=True)
    times = np.array(aw.get_time_points(20, 200e3))
    rng = np.random.RandomState(seed)
    pops = np.exp(rng.uniform(np.log(5e3), np.log(5e4), size=20))
    rho, mu = 1.6e-7, 1.8e-7
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    if masked:
        seqs = seqs.copy()
        seqs[rng.uniform(size=seqs.shape) < 0.02] = ord("N")
        seqs[:, ::97] = ord("N")   # all-masked columns count as invariant
    host = aw.arg_likelihood(tr, seqs, times, mu)
    dev = aw.arg_likelihood_device(tr, seqs, times, mu)
    assert np.isfinite(dev)
    assert abs(dev - host) <= 1e-11 * abs(host), (dev, host)
    hp = aw.arg_prior(tr, times, pops, rho)
    dp = aw.arg_prior_device(tr, times, pops, rho)
    assert abs(dp - hp) <= 1e-11 * abs(hp), (dp, hp)
    # after a leaf is removed, the rows beyond the tree's leaves still take
    # part in the invariant test, as in the reference
    aw.remove_thread(tr, k - 1)
    host = aw.arg_likelihood(tr, seqs, times, mu)
    dev = aw.arg_likelihood_device(tr, seqs, times, mu)
    assert abs(dev - host) <= 1e-11 * abs(host), (dev, host)
    hp = aw.arg_prior(tr, times, pops, rho)
    dp = aw.arg_prior_device(tr, times, pops, rho)
    assert abs(dp - hp) <= 1e-11 * abs(hp), (dp, hp)
    aw.delete_local_trees(tr)
