"""The oracle restatement pinned on the CPU.

Each check follows one of the reference's own consistency tests or a
closed form, or compares against the reference library compiled from its
sources (oracle/_ref) when that was built in this environment:
  - time grid: closed form and the 20 points printed by the reference's
    example run (examples/sim1/sim1.sample/out.log:11)
  - transition matrix: rows sum to one (test/test_hmm.py:114) and the k=2
    closed form of test_trans_two (test/test_hmm.py:22-115)
  - forward: fast compressed recursion equals the dense one
    (test/test_hmm.py:314-349), column sums are one
  - traceback: CDF inversion checked against a direct implementation
  - against oracle/_ref: forward tables, ARG surgery and the whole
    sample_thread chain under the same srand seed
"""
import numpy as np
import pytest

import oracle_bind as O
from workloads import instance, model

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def test_time_points_closed_form_and_logged_grid():
    t = O.time_points(20, 200e3, 0.01)
    i = np.arange(20)
    closed = (np.exp(i / 19.0 * np.log(1 + 0.01 * 200e3)) - 1) / 0.01
    assert np.allclose(t, closed, rtol=1e-15, atol=0)
    # printed by the reference run with --ntimes 20 --maxtime 200e3
    # (examples/sim1/sim1.sample/out.log:11, 6 decimals)
    logged = [0.000000, 49.193481, 122.586947, 232.085215, 395.449492, 639.178343,
              1002.805899, 1545.314509, 2354.701987, 3562.255340, 5363.846221,
              8051.702368, 12061.808515, 18044.625462, 26970.598323, 40287.567936,
              60155.618452, 89797.454603, 134021.141756, 200000.000000]
    assert np.abs(t - np.array(logged)).max() < 5.1e-7
    assert t[0] == 0.0 and np.all(np.diff(t) > 0) and abs(t[-1] - 200e3) < 1e-6


def test_transition_rows_substochastic(aw):
    # the compressed within-block matrix leaves out part of the SPR mass
    # (rows sum to ~0.97-0.99 in these trees); it is pinned entry by entry
    # against the reference in test_transition_matrix_matches_reference
    _, _, inst, _ = instance(aw, 8, 1000, 3)
    for tree in range(min(inst.ntrees, 6)):
        P = inst.trans_dense(tree)
        s = P.sum(1)
        assert (P >= 0).all() and (s <= 1.0 + 1e-12).all() and (s > 0.9).all()


@needs_ref
@pytest.mark.parametrize("seed", [1, 2])
def test_reference_brute_force_spr_sums(aw, seed):
    """The reference's own fast-vs-brute-force checks (sum of calc_spr_prob over
    explicit SPRs, trans.cpp:856-1032; literal emissions, emit.cpp:1222-1250),
    run on the trees these tests use; the oracle equals those fast matrices
    entry by entry (test_transition_matrix_matches_reference), so this pins
    the oracle's transitions and emissions to the brute-force definitions."""
    from ctypes import c_double
    _, seqs, inst, (times, pops, rho, mu) = instance(aw, 5, 800, seed, ntimes=10)
    r = O.ref()
    T = inst.ntimes
    for tree in range(min(inst.ntrees, 4)):
        ptree = np.ascontiguousarray(inst.ptrees[tree])
        ages = np.ascontiguousarray(inst.ages[tree])
        assert r.arghmm_assert_transmat(inst.nnodes, O._ip(ptree), O._ip(ages), T, O._dp(times),
                                        O._dp(pops), c_double(rho))
        if tree > 0:
            last = np.ascontiguousarray(inst.ptrees[tree - 1])
            lages = np.ascontiguousarray(inst.ages[tree - 1])
            sp = inst.sprs[tree]
            assert r.arghmm_assert_transmat_switch(
                inst.nnodes, O._ip(last), O._ip(lages), int(sp[0]), int(sp[1]), int(sp[2]),
                int(sp[3]), T, O._dp(times), O._dp(pops), c_double(rho))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_forward_fast_equals_dense_and_columns_normalised(aw, seed):
    _, _, inst, _ = instance(aw, 6, 800, seed)
    fw, lnl = inst.forward()
    fws, lnls = inst.forward(slow=True)
    assert np.allclose(fw, fws, rtol=1e-10, atol=0)
    assert abs(lnl - lnls) <= 1e-10 * abs(lnl)
    off, S = inst.block_offsets(), inst.nstates()
    for b in range(inst.ntrees):
        blk = fw[off[b]:off[b + 1]].reshape(-1, S[b])
        start = 1 if b == 0 else 0   # column 0 is the unnormalised prior
        assert np.allclose(blk[start:].sum(1), 1.0, rtol=1e-12)


def test_forward_equals_dense_hmm_brute_force(aw):
    # tiny instance: likelihood by explicit matrix products over all states
    _, _, inst, _ = instance(aw, 3, 40, 5)
    fw, lnl = inst.forward()
    S = inst.nstates()
    prior = inst.prior()
    vec = prior.copy()
    logmass = 0.0
    col = 0
    for b in range(inst.ntrees):
        P = inst.trans_dense(b)
        E = inst.emissions(b)
        for i in range(int(inst.blocklens[b])):
            if col == 0:
                col += 1
                continue
            if i == 0:
                det, dp, rr, cr, srcs = inst.switch(b)
                M = np.zeros((S[b - 1], S[b]))
                for j in range(S[b - 1]):
                    if j == srcs[1]:
                        M[j] = cr
                    elif j == srcs[0]:
                        M[j] = rr
                    elif det[j] >= 0:
                        M[j, det[j]] = dp[j]
                vec = (vec @ M) * E[i]
            else:
                vec = (vec @ P) * E[i]
            s = vec.sum()
            logmass += np.log(s)
            vec /= s
            col += 1
    assert abs(logmass - lnl) <= 1e-10 * abs(lnl)


def test_traceback_is_cdf_inversion(aw):
    _, _, inst, _ = instance(aw, 5, 300, 9)
    fw, _ = inst.forward()
    u = np.random.RandomState(3).uniform(size=inst.seqlen)
    path = inst.traceback(fw, u)
    # restate one step directly: last column
    off, S = inst.block_offsets(), inst.nstates()
    last = fw[off[-1]:off[-1 + 1] if len(off) > 1 else None]
    last = fw[off[inst.ntrees - 1]:off[inst.ntrees]].reshape(-1, S[-1])[-1]
    c = np.cumsum(last)
    expect = int(np.searchsorted(c, u[0] * c[-1], side="left"))
    assert path[-1] == min(expect, S[-1] - 1)
    assert path.min() >= 0


@needs_ref
@pytest.mark.parametrize("k,L,seed", [(4, 600, 1), (8, 3000, 2), (8, 3000, 3), (16, 2000, 4)])
def test_oracle_forward_matches_compiled_reference(aw, k, L, seed):
    times, pops, rho, mu = model()
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    P, Ag, S, bl = aw.trees_arrays(tr)
    rt = O.ref_new_trees(P, Ag, S, bl)
    O.ref_remove_thread(rt, k - 1)
    aw.remove_thread(tr, k - 1)
    P2, A2, S2, bl2 = aw.trees_arrays(tr)
    assert all(np.array_equal(x, y) for x, y in zip((P2, A2, S2, bl2), O.ref_trees_arrays(rt)))
    maps, seqids = aw.trees_internal(tr)
    inst = O.Instance(times, pops, rho, mu, P2, A2, S2, bl2, seqs, seqids=seqids,
                      new_chrom=k - 1, mappings=maps)
    fw, _ = inst.forward()
    ref = np.concatenate(O.ref_forward(rt, seqs, times, pops, rho, mu,
                                       aw.get_nstates(tr, 20, L)))
    O.ref_delete(rt)
    rel = np.abs(fw - ref) / np.maximum(np.abs(ref), 1e-300)
    assert rel.max() <= 1e-12


@needs_ref
@pytest.mark.parametrize("seed,k,L", [(1, 8, 3000), (2, 8, 3000), (3, 8, 3000), (4, 8, 3000),
                                      (5, 24, 60000), (6, 40, 40000), (7, 12, 150000)])
def test_sample_thread_chain_matches_compiled_reference(aw, seed, k, L):
    """oracle traceback + recombination sampling + this library's add_arg_thread
    reproduce the reference's arghmm_sample_thread node for node.  The last
    case has enough local trees for the sampler's per-block rates to be
    computed on several threads."""
    times, pops, rho, mu = model()
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    if L > 3000:
        assert aw.get_local_trees_ntrees(tr) >= 256
    P, Ag, S, bl = aw.trees_arrays(tr)
    rt = O.ref_new_trees(P, Ag, S, bl)
    O.ref_remove_thread(rt, k - 1)
    aw.remove_thread(tr, k - 1)
    P2, A2, S2, bl2 = aw.trees_arrays(tr)
    maps, seqids = aw.trees_internal(tr)
    inst = O.Instance(times, pops, rho, mu, P2, A2, S2, bl2, seqs, seqids=seqids,
                      new_chrom=k - 1, mappings=maps)
    O.srand(1000 + seed)
    O.ref_sample_thread(rt, seqs, times, pops, rho, mu)
    R = O.ref_trees_arrays(rt)
    O.ref_delete(rt)
    O.srand(1000 + seed)
    u = O.draw_uniforms(L)
    fw, _ = inst.forward()
    path = inst.traceback(fw, u)
    pos, node, tm = inst.sample_recombinations(path)
    O.srand(1000 + seed)
    O.draw_uniforms(L)
    pos2, node2, tm2 = aw.sample_recombinations(tr, times, pops, rho, mu, path)
    assert np.array_equal(pos, pos2) and np.array_equal(node, node2) and np.array_equal(tm, tm2)
    aw.add_thread(tr, 20, path, k - 1, pos, node, tm)
    B = aw.trees_arrays(tr)
    assert all(np.array_equal(x, y) for x, y in zip(B, R))


def _ref_matrix(ptr, n, m):
    import ctypes
    out = np.zeros((n, m))
    for i in range(n):
        out[i] = np.ctypeslib.as_array(ptr[i], shape=(m,))
    return out


@needs_ref
@pytest.mark.parametrize("seed", [1, 2])
def test_transition_matrix_matches_reference(aw, seed):
    """a5/a6: dense within-block and switch matrices against the reference's
    new_transition_probs / new_transition_probs_switch (trans.cpp:1190-1245)."""
    import ctypes
    from ctypes import c_double, c_int
    _, _, inst, (times, pops, rho, mu) = instance(aw, 8, 1500, seed)
    r = O.ref()
    T = inst.ntimes
    steps = np.ascontiguousarray(np.append(np.diff(times), 0.0))
    zeros = np.zeros(T, np.int32)
    for tree in range(min(inst.ntrees, 8)):
        st = np.ascontiguousarray(inst.states(tree), np.int32)
        S = len(st)
        ptree = np.ascontiguousarray(inst.ptrees[tree])
        ages = np.ascontiguousarray(inst.ages[tree])
        ptr = r.new_transition_probs(
            inst.nnodes, O._ip(ptree), O._ip(ages), c_double(0.0), O._ip(st), S, T,
            O._dp(times), O._dp(steps), O._ip(zeros), O._ip(zeros), O._ip(zeros),
            O._dp(pops), c_double(rho))
        ref = np.exp(_ref_matrix(ptr, S, S))
        r.delete_transition_probs(ptr, S)
        ours = inst.trans_dense(tree)
        assert np.allclose(ours, ref, rtol=1e-12, atol=0)
        if tree == 0 or inst.mappings is not None and not np.array_equal(
                inst.mappings[tree][inst.mappings[tree] >= 0],
                np.arange(inst.nnodes)[inst.mappings[tree] >= 0]):
            continue   # the reference entry derives the default node mapping
        st1 = np.ascontiguousarray(inst.states(tree - 1), np.int32)
        sp = inst.sprs[tree]
        last = np.ascontiguousarray(inst.ptrees[tree - 1])
        lages = np.ascontiguousarray(inst.ages[tree - 1])
        ptr = r.new_transition_probs_switch(
            O._ip(ptree), O._ip(last), inst.nnodes, int(sp[0]), int(sp[1]), int(sp[2]),
            int(sp[3]), O._ip(ages), O._ip(lages), c_double(0.0), c_double(0.0),
            O._ip(st1), len(st1), O._ip(st), S, T, O._dp(times), O._dp(steps),
            O._ip(zeros), O._ip(zeros), O._ip(zeros), O._dp(pops), c_double(rho))
        refsw = np.exp(_ref_matrix(ptr, len(st1), S))
        r.delete_transition_probs(ptr, len(st1))
        det, dp, rr, cr, srcs = inst.switch(tree)
        M = np.zeros((len(st1), S))
        for j in range(len(st1)):
            if j == srcs[1]:
                M[j] = cr
            elif j == srcs[0]:
                M[j] = rr
            elif det[j] >= 0:
                M[j, det[j]] = dp[j]
        assert np.allclose(M, refsw, rtol=1e-12, atol=0)


@needs_ref
@pytest.mark.parametrize("seed", [3, 4])
def test_emissions_match_reference(aw, seed):
    """a8: per-block emissions against the reference's new_emissions (emit.cpp:1288-1303)."""
    from ctypes import POINTER, c_char_p, c_double
    _, seqs, inst, (times, pops, rho, mu) = instance(aw, 8, 2000, seed)
    r = O.ref()
    T = inst.ntimes
    start = 0
    for tree in range(min(inst.ntrees, 10)):
        L = int(inst.blocklens[tree])
        st = np.ascontiguousarray(inst.states(tree), np.int32)
        rows = [np.ascontiguousarray(seqs[inst.seqids[i], start:start + L]) for i in range(inst.nleaves)]
        rows.append(np.ascontiguousarray(seqs[inst.new_chrom, start:start + L]))
        arr = (c_char_p * len(rows))(*[c_char_p(x.tobytes()) for x in rows])
        ptree = np.ascontiguousarray(inst.ptrees[tree])
        ages = np.ascontiguousarray(inst.ages[tree])
        ptr = r.new_emissions(O._ip(st), len(st), O._ip(ptree), inst.nnodes, O._ip(ages),
                              arr, len(rows), L, O._dp(times), T, c_double(mu))
        ref = _ref_matrix(ptr, L, len(st))
        r.delete_emissions(ptr, L)
        ours = inst.emissions(tree)
        assert np.allclose(ours, ref, rtol=1e-13, atol=0)
        start += L
