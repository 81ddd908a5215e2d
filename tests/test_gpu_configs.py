"""Parity at the workload sizes and inputs the benchmark runs on the device:

- config 2 at full size (k = 50, 100000 columns): the drop-in forward table
  to 1e-12 of the compiled reference and the identical threaded ARG;
- config 5 (k = 100, 500000 columns, N(t) piecewise-constant over the 20
  intervals, log-uniform in [5e3, 5e4]): one re-threading identical to the
  reference's, and a small N(t) case against the oracle;
- masked columns (all 'N') and 'N' bases inside variant columns;
- config 3: eight full-length instances (k = 8..64, up to 200000 columns) in
  one batch, FP64 against the oracle and FP32 with the benchmark's forward
  tiles against FP64 within the FP32 contract;
- sample_posterior, sample_arg_refine and resample_arg(nremove = 3) against
  the reference.

Reference calls: arghmm_forward_alg (sample_thread.cpp:887-926),
arghmm_sample_thread (:1010-1023), arghmm_sample_posterior (:930-977),
arghmm_sample_arg_refine / arghmm_resample_arg (sample_arg.cpp:682-718).
"""
import numpy as np
import pytest

import oracle_bind as O
from workloads import instance, model

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")

# FP32 contract (forward tiles included), stated against the FP64 table:
#   |dlogL| / |logL|                     <= 1e-6
#   |alpha32 - alpha64| per column entry <= 2e-4 (columns sum to 1)
FP32_DLOGL = 1e-6
FP32_ALPHA_ABS = 2e-4


def log_uniform_popsizes(seed, ntimes=20, lo=5e3, hi=5e4):
    """N(t) of config 5: one size per time interval, log-uniform in [lo, hi]."""
    rng = np.random.RandomState(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=ntimes))


def _pair(aw, k, L, seed, pops=None, seqs_fn=None):
    """The same ARG in both libraries with leaf k-1 removed (rebuilt in the
    reference from the full arrays; the leaf removed inside each library)."""
    times, p0, rho, mu = model()
    pops = p0 if pops is None else pops
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    if seqs_fn is not None:
        seqs = seqs_fn(seqs)
    rt = O.ref_new_trees(*aw.trees_arrays(tr))
    aw.remove_thread(tr, k - 1)
    O.ref_remove_thread(rt, k - 1)
    return tr, rt, seqs, (times, pops, rho, mu)


def _same_arg(aw, tr, rt):
    return all(np.array_equal(a, b) for a, b in zip(aw.trees_arrays(tr), O.ref_trees_arrays(rt)))


def _forward_vs_ref(aw, tr, rt, seqs, m):
    times, pops, rho, mu = m
    L = seqs.shape[1]
    mine = np.concatenate(aw.forward_algorithm(tr, seqs, times, pops, rho, mu))
    ref = np.concatenate(O.ref_forward(rt, seqs, times, pops, rho, mu,
                                       aw.get_nstates(tr, len(times), L)))
    assert mine.shape == ref.shape
    rel = np.abs(mine - ref) / np.maximum(np.abs(ref), 1e-300)
    return rel.max()


def _thread_vs_ref(aw, tr, rt, seqs, m, seed):
    times, pops, rho, mu = m
    O.srand(seed)
    assert aw.sample_thread(tr, seqs, times, pops, rho, mu)
    O.srand(seed)
    O.ref_sample_thread(rt, seqs, times, pops, rho, mu)
    return _same_arg(aw, tr, rt)


@needs_ref
def test_config2_full_size(aw):
    """k = 50, 100000 columns: forward table <= 1e-12 and identical thread."""
    print("config 2: k=50 L=100000 seed=201", flush=True)
    tr, rt, seqs, m = _pair(aw, 50, 100000, 201)
    rel = _forward_vs_ref(aw, tr, rt, seqs, m)
    print("config 2 forward max rel %.3g" % rel, flush=True)
    assert rel <= 1e-12
    assert _thread_vs_ref(aw, tr, rt, seqs, m, 2201)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


@needs_ref
def test_config5_piecewise_popsizes_rethread(aw):
    """k = 100, 500000 columns, N(t) log-uniform: one re-threading step
    (remove a leaf, thread it back) gives the reference's ARG exactly."""
    print("config 5: k=100 L=500000 seed=501 N(t) log-uniform", flush=True)
    pops = log_uniform_popsizes(501)
    tr, rt, seqs, m = _pair(aw, 100, 500000, 501, pops=pops)
    assert _thread_vs_ref(aw, tr, rt, seqs, m, 5501)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


@pytest.mark.parametrize("k,L,seed", [(10, 3000, 511), (24, 5000, 512)])
def test_piecewise_popsizes_against_oracle(aw, k, L, seed):
    """N(t) enters the transition vectors and the switch sums: FP64 table,
    log-likelihood and thread against the oracle."""
    print("N(t) k=%d L=%d seed=%d" % (k, L, seed), flush=True)
    pops = log_uniform_popsizes(seed)
    tr, seqs, inst, (times, pops, rho, mu) = instance(aw, k, L, seed, popsizes=pops)
    fw, lnl = inst.forward()
    u = np.random.RandomState(seed).uniform(size=L)
    out = aw.thread_forward(tr, seqs, times, pops, rho, mu, k - 1, uniforms=u)
    rel = np.abs(out["alpha"] - fw) / np.maximum(np.abs(fw), 1e-300)
    assert rel.max() <= 1e-12, rel.max()
    assert abs(out["lnl"] - lnl) <= 1e-12 * abs(lnl)
    assert np.array_equal(out["path"], inst.traceback(fw, u))
    # a constant N gives a different table: the sizes are really read
    const = aw.thread_forward(tr, seqs, times, np.full(20, 1e4), rho, mu, k - 1)
    assert np.abs(const["alpha"] - out["alpha"]).max() > 1e-6
    aw.delete_local_trees(tr)


def _mask(seed, frac_n=0.03, every=37):
    def fn(seqs):
        rng = np.random.RandomState(seed)
        s = seqs.copy()
        s[rng.uniform(size=s.shape) < frac_n] = ord("N")   # 'N' inside variant columns
        s[:, ::every] = ord("N")                           # all-'N' columns
        s[:, 5::every * 3] = ord("N")
        s[-1, 11::every * 2] = ord("N")                    # masked in the new sequence only
        return s
    return fn


@needs_ref
@pytest.mark.parametrize("k,L,seed", [(8, 4000, 521), (20, 6000, 522)])
def test_masked_columns_against_reference(aw, k, L, seed):
    """All-'N' columns (emission 1) and 'N' leaves inside variant columns
    (emit.cpp:30-58, 815-820): forward table and thread against the reference."""
    print("masked k=%d L=%d seed=%d" % (k, L, seed), flush=True)
    tr, rt, seqs, m = _pair(aw, k, L, seed, seqs_fn=_mask(seed))
    assert (seqs == ord("N")).all(axis=0).sum() >= L // 37
    rel = _forward_vs_ref(aw, tr, rt, seqs, m)
    assert rel <= 1e-12, rel
    assert _thread_vs_ref(aw, tr, rt, seqs, m, 3000 + seed)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


def test_masked_columns_against_oracle(aw):
    """The same masking through the batched FP64 and FP32 paths against the
    oracle restatement."""
    k, L, seed = 12, 5000, 523
    print("masked batch k=%d L=%d seed=%d" % (k, L, seed), flush=True)
    tr, seqs0, _, (times, pops, rho, mu) = instance(aw, k, L, seed)
    seqs = _mask(seed)(seqs0)
    P, Ag, S, bl = aw.trees_arrays(tr)
    maps, seqids = aw.trees_internal(tr)
    inst = O.Instance(times, pops, rho, mu, P, Ag, S, bl, seqs, seqids=seqids, new_chrom=k - 1,
                      mappings=maps)
    fw, lnl = inst.forward()
    u = np.random.RandomState(seed).uniform(size=L)
    b = aw.Batch()
    b.add(tr, seqs, times, pops, rho, mu, k - 1, u)
    b.upload()
    b.run(0, 700)
    rel = np.abs(b.alpha(0) - fw) / np.maximum(np.abs(fw), 1e-300)
    assert rel.max() <= 1e-12
    assert np.array_equal(b.path(0), inst.traceback(fw, u))
    b.run(1, 700)
    assert abs(b.lnl(0) - lnl) <= FP32_DLOGL * abs(lnl)
    assert np.abs(b.alpha(0) - fw).max() <= FP32_ALPHA_ABS
    aw.delete_local_trees(tr)


# config 3: k in 8..64, 5000..200000 columns (50 kb .. 2 Mb at compression 10)
CONFIG3 = [(8, 5000, 301), (64, 200000, 302), (16, 120000, 303), (40, 20000, 304),
           (24, 200000, 305), (56, 80000, 306), (32, 50000, 307), (48, 10000, 308)]


def test_config3_full_length_batch(aw):
    """Eight config-3 instances in one batch.  FP64: the two shortest against
    the oracle (table 1e-12, identical thread).  FP32 with the benchmark's
    schedule (forward tiles of 100000 columns with 40000 warm-up columns,
    tolerance 1e-4; traceback tiles of 14000 columns): log-likelihood and
    table against FP64 within the FP32 contract, for every instance."""
    items = []
    b = aw.Batch()
    for k, L, seed in CONFIG3:
        print("config 3 instance k=%d L=%d seed=%d" % (k, L, seed), flush=True)
        tr, seqs, inst, mdl = instance(aw, k, L, seed)
        u = np.random.RandomState(seed).uniform(size=L)
        b.add(tr, seqs, *mdl, k - 1, u)
        items.append((k, L, tr, inst, u))
    b.upload()
    print("config 3: FP64 batch", flush=True)
    b.run(0, 14000)
    lnl64 = [b.lnl(i) for i in range(len(items))]
    path64 = [b.path(i) for i in range(len(items))]
    for i, (k, L, tr, inst, u) in enumerate(items):
        if L > 10000:
            continue
        fw, lnl = inst.forward()
        rel = np.abs(b.alpha(i) - fw) / np.maximum(np.abs(fw), 1e-300)
        assert rel.max() <= 1e-12, (k, L, rel.max())
        assert abs(lnl64[i] - lnl) <= 1e-12 * abs(lnl)
        assert np.array_equal(path64[i], inst.traceback(fw, u))
    alpha64 = {i: b.alpha(i) for i in (1, 2)}
    print("config 3: FP32 batch with forward tiles", flush=True)
    st = b.run(1, 14000, fw_tile=100000, fw_warm=40000, fw_tol=1e-4)
    assert st["fw_tiles"] > len(items)   # the long instances really ran as tiles
    for i, (k, L, tr, inst, u) in enumerate(items):
        d = abs(b.lnl(i) - lnl64[i]) / abs(lnl64[i])
        agree = np.mean(b.path(i) == path64[i])
        print("  k=%d L=%d dlogL %.2e path agreement %.4f" % (k, L, d, agree), flush=True)
        assert d <= FP32_DLOGL
        assert len(b.path(i)) == L
    for i, a64 in alpha64.items():
        err = np.abs(b.alpha(i) - a64).max()
        print("  instance %d max |alpha32 - alpha64| %.2e" % (i, err), flush=True)
        assert err <= FP32_ALPHA_ABS
    for it in items:
        aw.delete_local_trees(it[2])


@needs_ref
def test_sample_posterior_one_tree(aw):
    """arghmm_sample_posterior on a one-tree ARG (where the reference's block
    conversion is fully defined): the same (node, time) path as the reference
    from the same rand() stream."""
    print("sample_posterior k=9 L=3000", flush=True)
    times, pops, rho, mu = model()
    tr, seqs = aw.simulate(9, 3000, times, pops, rho, mu, 531)
    aw.remove_thread(tr, 8)
    P, Ag, S, bl = aw.trees_arrays(tr)
    one = (P[:1], Ag[:1], np.full((1, 4), -1, np.int32), np.array([3000], np.int32))
    O.srand(9531)
    mine = aw.sample_posterior(*one, seqs, times, pops, rho, mu)
    O.srand(9531)
    ref = O.ref_sample_posterior(*one, seqs, times, pops, rho, mu)
    assert np.array_equal(mine, ref)
    aw.delete_local_trees(tr)


@needs_ref
def test_sample_arg_refine(aw):
    print("sample_arg_refine k=5 L=600", flush=True)
    times, pops, rho, mu = model()
    _, seqs = aw.simulate(5, 600, times, pops, rho, mu, 541)
    O.srand(541)
    tr = aw.sample_arg_refine(seqs, times, pops, rho, mu, niters=2, nremove=2)
    O.srand(541)
    rt = O.ref_sample_arg_refine(seqs, times, pops, rho, mu, 2, 2)
    assert _same_arg(aw, tr, rt)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


@needs_ref
def test_resample_arg_nremove3(aw):
    print("resample_arg nremove=3 k=8 L=1500", flush=True)
    times, pops, rho, mu = model()
    tr, seqs = aw.simulate(8, 1500, times, pops, rho, mu, 551)
    rt = O.ref_new_trees(*aw.trees_arrays(tr))
    O.srand(551)
    aw.resample_arg(tr, seqs, times, pops, rho, mu, niters=2, nremove=3)
    O.srand(551)
    O.ref_resample_arg(rt, seqs, times, pops, rho, mu, 2, nremove=3)
    assert _same_arg(aw, tr, rt)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)
