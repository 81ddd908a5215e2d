"""HIP path against the oracle: FP64 forward table, log-likelihood and the
sampled thread for the same uniform stream."""
import numpy as np
import pytest

import oracle_bind as O
from workloads import instance

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,L,seed", [(4, 500, 1), (8, 3000, 2), (8, 3000, 3), (12, 4000, 4)])
def test_forward_fp64_matches_oracle(aw, k, L, seed):
    tr, seqs, inst, (times, pops, rho, mu) = instance(aw, k, L, seed)
    print("case k=%d L=%d seed=%d" % (k, L, seed), flush=True)
    fw, lnl = inst.forward()
    rng = np.random.RandomState(seed)
    u = rng.uniform(size=L)
    out = aw.thread_forward(tr, seqs, times, pops, rho, mu, k - 1, uniforms=u)
    rel = np.abs(out["alpha"] - fw) / np.maximum(np.abs(fw), 1e-300)
    assert rel.max() <= 1e-12, rel.max()
    assert abs(out["lnl"] - lnl) <= 1e-12 * abs(lnl)
    path = inst.traceback(fw, u)
    assert np.array_equal(out["path"], path)


@pytest.mark.parametrize("ntimes,k,L,seed", [(12, 6, 800, 5), (28, 6, 800, 6), (32, 5, 600, 7)])
def test_other_time_grids(aw, ntimes, k, L, seed):
    """Time grids other than 20 points (the kernels are instantiated for up
    to 19 and up to 31 time bins): FP64 table and thread against the
    oracle, and the FP32 run within its bounds."""
    tr, seqs, inst, (times, pops, rho, mu) = instance(aw, k, L, seed, ntimes=ntimes)
    print("case ntimes=%d k=%d L=%d seed=%d" % (ntimes, k, L, seed), flush=True)
    fw, lnl = inst.forward()
    u = np.random.RandomState(seed).uniform(size=L)
    out = aw.thread_forward(tr, seqs, times, pops, rho, mu, k - 1, uniforms=u)
    rel = np.abs(out["alpha"] - fw) / np.maximum(np.abs(fw), 1e-300)
    assert rel.max() <= 1e-12, rel.max()
    assert abs(out["lnl"] - lnl) <= 1e-12 * abs(lnl)
    assert np.array_equal(out["path"], inst.traceback(fw, u))
    o32 = aw.thread_forward(tr, seqs, times, pops, rho, mu, k - 1, uniforms=u, precision=1)
    assert abs(o32["lnl"] - lnl) <= 1e-6 * abs(lnl)
    assert np.mean(o32["path"] == out["path"]) >= 0.98
