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
