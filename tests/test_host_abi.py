"""The reference's host-side exports that are not on the threading path,
against the same calls into the reference library compiled from its sources
(oracle/_ref), and against what the formats and the mathematics fix:

- new_transition_probs / _switch (trans.cpp:1190-1244) entry by entry;
- arghmm_assert_transmat / _switch / arghmm_assert_emit (the reference's own
  brute-force checks, trans.cpp:856-1032, emit.cpp:1222-1250) pass, as the
  reference's do on the same trees;
- arghmm_assert_transmat_internal / _switch_internal / arghmm_assert_emit_internal
  (the same checks for an internal branch, trans.cpp:1035-1186, 1314-1379,
  emit.cpp:1171-1280, 1340-1380) pass under the reference's rand() stream,
  and leave the ARG as the reference leaves it;
- new_emissions (emit.cpp:1288-1303) with masked columns and 'N' bases;
- forward_alg / backward_alg / sample_hmm_posterior (hmm.cpp:27-84), the
  posterior sampler under the same rand() stream;
- arghmm_est_popsizes_trees (est_popsize.cpp:68-147);
- .smc read / write (local_tree.cpp:1286-1545) both ways with the
  reference, and the reference's example file (examples/sim1, committed as
  tests/golden/sim1_out0.smc.gz) read to the arrays of tests/golden/sim1.npz;
- .sites read (sequences.cpp:173-313) of examples/sim1/sim1.sites
  (tests/golden/sim1.sites), to sequences, and the compression of the sites
  and the trees (sequences.cpp:523-626, local_tree.cpp:790-873).
"""
import ctypes
import gzip
import os
from ctypes import POINTER, c_bool, c_char_p, c_double, c_int, c_void_p

import numpy as np
import pytest

import oracle_bind as O
from workloads import model

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")

c_int_p = POINTER(c_int)
c_double_p = POINTER(c_double)
c_double_p_p = POINTER(c_double_p)
intstate_p = POINTER(c_int * 2)


def _ref_fn(name, restype, argtypes):
    # a fresh function object (indexing does not cache), so the argument
    # types set here do not leak into other tests' calls
    fn = O.ref()[name]
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


def _i(a):
    return a.ctypes.data_as(c_int_p)


def _d(a):
    return a.ctypes.data_as(c_double_p)


def _trees(aw, k, L, seed, ntimes=20):
    times, pops, rho, mu = model(ntimes=ntimes)
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    return tr, seqs, aw.trees_arrays(tr), (times, pops, rho, mu)


def _ref_matrix(m, nrows, ncols):
    out = np.array([np.ctypeslib.as_array(m[i], shape=(ncols,)) for i in range(nrows)])
    return out


@needs_ref
@pytest.mark.parametrize("k,ntimes,seed", [(6, 20, 1), (10, 20, 2), (8, 12, 3)])
def test_transition_probs_match_reference(aw, k, ntimes, seed):
    tr, _, (P, A, S, bl), (times, pops, rho, mu) = _trees(aw, k, 3000, seed, ntimes)
    pops = np.exp(np.random.RandomState(seed).uniform(np.log(5e3), np.log(5e4), ntimes))
    fn = _ref_fn("new_transition_probs", c_double_p_p,
                 [c_int, c_int_p, c_int_p, c_double, intstate_p, c_int, c_int, c_double_p,
                  c_double_p, c_int_p, c_int_p, c_int_p, c_double_p, c_double])
    z = np.zeros(ntimes, np.int32)
    for b in (0, len(bl) // 2, len(bl) - 1):
        p, a = np.ascontiguousarray(P[b]), np.ascontiguousarray(A[b])
        st = aw.tree_states(p, a, ntimes)
        mine = aw.transition_probs(p, a, times, pops, rho)
        m = fn(len(p), _i(p), _i(a), 0.0, st.ctypes.data_as(intstate_p), len(st), ntimes, _d(times),
               _d(times), _i(z), _i(z), _i(z), _d(pops), rho)
        ref = _ref_matrix(m, len(st), len(st))
        O.ref().delete_transition_probs(m, len(st))
        fin = np.isfinite(ref)
        assert np.array_equal(fin, np.isfinite(mine))
        assert np.allclose(mine[fin], ref[fin], rtol=1e-12, atol=0)
    aw.delete_local_trees(tr)


@needs_ref
@pytest.mark.parametrize("k,seed", [(6, 11), (12, 12)])
def test_transition_probs_switch_match_reference(aw, k, seed):
    tr, _, (P, A, S, bl), (times, pops, rho, mu) = _trees(aw, k, 4000, seed)
    fn = _ref_fn("new_transition_probs_switch", c_double_p_p,
                 [c_int_p, c_int_p, c_int, c_int, c_int, c_int, c_int, c_int_p, c_int_p, c_double,
                  c_double, intstate_p, c_int, intstate_p, c_int, c_int, c_double_p, c_double_p,
                  c_int_p, c_int_p, c_int_p, c_double_p, c_double])
    z = np.zeros(20, np.int32)
    assert len(bl) > 3
    for b in range(1, min(len(bl), 6)):
        p, lp = np.ascontiguousarray(P[b]), np.ascontiguousarray(P[b - 1])
        a, la = np.ascontiguousarray(A[b]), np.ascontiguousarray(A[b - 1])
        spr = S[b]
        mine = aw.transition_probs_switch(p, lp, a, la, spr, times, pops, rho)
        s1, s2 = aw.tree_states(lp, la, 20), aw.tree_states(p, a, 20)
        m = fn(_i(p), _i(lp), len(p), int(spr[0]), int(spr[1]), int(spr[2]), int(spr[3]), _i(a),
               _i(la), 0.0, 0.0, s1.ctypes.data_as(intstate_p), len(s1),
               s2.ctypes.data_as(intstate_p), len(s2), 20, _d(times), _d(times), _i(z), _i(z),
               _i(z), _d(pops), rho)
        ref = _ref_matrix(m, len(s1), len(s2))
        O.ref().delete_transition_probs(m, len(s1))
        fin = np.isfinite(ref)
        assert np.array_equal(fin, np.isfinite(mine)), b
        assert np.allclose(mine[fin], ref[fin], rtol=1e-12, atol=0), b
    aw.delete_local_trees(tr)


@needs_ref
@pytest.mark.parametrize("k,ntimes,seed", [(4, 10, 21), (5, 10, 22)])
def test_assert_transmat_passes_like_reference(aw, k, ntimes, seed):
    """The brute-force SPR sums of the reference's own check agree with the
    compressed matrix, here as in the reference."""
    tr, _, (P, A, S, bl), (times, pops, rho, mu) = _trees(aw, k, 800, seed, ntimes)
    fn = _ref_fn("arghmm_assert_transmat", c_bool,
                 [c_int, c_int_p, c_int_p, c_int, c_double_p, c_double_p, c_double])
    p, a = np.ascontiguousarray(P[0]), np.ascontiguousarray(A[0])
    assert fn(len(p), _i(p), _i(a), ntimes, _d(times), _d(pops), rho)
    assert aw.assert_transmat(p, a, times, pops, rho)
    aw.delete_local_trees(tr)


@needs_ref
def test_assert_transmat_switch_passes_like_reference(aw):
    tr, _, (P, A, S, bl), (times, pops, rho, mu) = _trees(aw, 6, 3000, 31, 12)
    fn = _ref_fn("arghmm_assert_transmat_switch", c_bool,
                 [c_int, c_int_p, c_int_p, c_int, c_int, c_int, c_int, c_int, c_double_p,
                  c_double_p, c_double])
    checked = 0
    for b in range(1, len(bl)):
        p, a = np.ascontiguousarray(P[b - 1]), np.ascontiguousarray(A[b - 1])
        spr = S[b]
        ref_ok = fn(len(p), _i(p), _i(a), int(spr[0]), int(spr[1]), int(spr[2]), int(spr[3]), 12,
                    _d(times), _d(pops), rho)
        assert aw.assert_transmat_switch(p, a, spr, times, pops, rho) == ref_ok
        assert ref_ok
        checked += 1
        if checked == 4:
            break
    assert checked > 0
    aw.delete_local_trees(tr)


@needs_ref
@pytest.mark.parametrize("masked", [False, True])
def test_emissions_match_reference(aw, masked):
    tr, seqs, (P, A, S, bl), (times, pops, rho, mu) = _trees(aw, 9, 2500, 41)
    k = seqs.shape[0]
    if masked:
        rng = np.random.RandomState(41)
        seqs = seqs.copy()
        seqs[rng.uniform(size=seqs.shape) < 0.03] = ord("N")
        seqs[:, ::29] = ord("N")
    fn = _ref_fn("new_emissions", c_double_p_p,
                 [intstate_p, c_int, c_int_p, c_int, c_int_p, POINTER(c_char_p), c_int, c_int,
                  c_double_p, c_int, c_double])
    # a tree of the first k-1 leaves: remove leaf k-1 from block 0's tree
    aw.remove_thread(tr, k - 1)
    P2, A2, _, _ = aw.trees_arrays(tr)
    p, a = np.ascontiguousarray(P2[0]), np.ascontiguousarray(A2[0])
    st = aw.tree_states(p, a, 20)
    mine = aw.emissions(p, a, seqs, times, mu)
    s = O._Seqs(seqs)
    m = fn(st.ctypes.data_as(intstate_p), len(st), _i(p), len(p), _i(a), s.ptr, k,
           seqs.shape[1], _d(times), 20, mu)
    ref = _ref_matrix(m, seqs.shape[1], len(st))
    O.ref().delete_emissions(m, seqs.shape[1])
    assert np.allclose(mine, ref, rtol=1e-13, atol=0)
    aw.delete_local_trees(tr)


def test_assert_emit_passes(aw):
    """Fast emissions against the literal slow path (the branch attached,
    then a full pruning pass), the reference's own check."""
    tr, seqs, _, (times, pops, rho, mu) = _trees(aw, 5, 300, 51)
    seqs = seqs.copy()
    seqs[1, ::17] = ord("N")
    aw.remove_thread(tr, 4)
    assert aw.assert_emit(tr, seqs, times, mu)
    aw.delete_local_trees(tr)


@needs_ref
@pytest.mark.parametrize("k,ntimes,seed", [(5, 10, 61), (6, 12, 62), (8, 10, 63)])
def test_assert_transmat_internal_passes_like_reference(aw, k, ntimes, seed):
    """Internal within-block matrix (minimum age, subtree/maintree lineage
    counts) against the brute-force SPR sums, for several drawn branches."""
    tr, _, (P, A, S, bl), (times, pops, rho, mu) = _trees(aw, k, 800, seed, ntimes)
    fn = _ref_fn("arghmm_assert_transmat_internal", c_bool,
                 [c_int, c_int_p, c_int_p, c_int, c_double_p, c_double_p, c_double])
    p, a = np.ascontiguousarray(P[0]), np.ascontiguousarray(A[0])
    for draw in range(4):
        O.srand(seed * 10 + draw)
        ref_ok = fn(len(p), _i(p), _i(a), ntimes, _d(times), _d(pops), rho)
        O.srand(seed * 10 + draw)
        assert aw.assert_transmat_internal(p, a, times, pops, rho) == ref_ok
        assert ref_ok
    aw.delete_local_trees(tr)


@needs_ref
@pytest.mark.parametrize("k,L,ntimes,seed", [(5, 3000, 10, 71), (7, 5000, 12, 72)])
def test_assert_transmat_switch_internal_passes_like_reference(aw, k, L, ntimes, seed):
    """Every internal switch matrix (recombination and re-coalescence rows,
    deterministic moves, the special cases of a tree left complete) against
    the SPR probability on the trees carrying the subtree."""
    tr, _, arrays, (times, pops, rho, mu) = _trees(aw, k, L, seed, ntimes)
    fn = _ref_fn("arghmm_assert_transmat_switch_internal", c_bool,
                 [c_void_p, c_int, c_double_p, c_double_p, c_double])
    for draw in range(3):
        rt = O.ref_new_trees(*aw.trees_arrays(tr))
        mine = aw.new_trees(*aw.trees_arrays(tr))
        O.srand(seed * 10 + draw)
        ref_ok = fn(c_void_p(rt), ntimes, _d(times), _d(pops), rho)
        O.srand(seed * 10 + draw)
        assert aw.assert_transmat_switch_internal(mine, times, pops, rho) == ref_ok
        assert ref_ok
        # the ARG is left in subtree/maintree form, node for node as the reference's
        assert all(np.array_equal(x, y) for x, y in
                   zip(aw.trees_arrays(mine), O.ref_trees_arrays(rt)))
        O.ref_delete(rt)
        aw.delete_local_trees(mine)
    aw.delete_local_trees(tr)


@needs_ref
@pytest.mark.parametrize("masked", [False, True])
def test_assert_emit_internal_passes_like_reference(aw, masked):
    """Internal emissions (subtree partials, maintree root, both tree
    lengths) against the full pruning pass with the subtree attached at each
    state, on a uniformly drawn removal path."""
    tr, seqs, _, (times, pops, rho, mu) = _trees(aw, 6, 600, 81)
    if masked:
        seqs = seqs.copy()
        seqs[2, ::13] = ord("N")
        seqs[:, ::31] = ord("N")
    fn = _ref_fn("arghmm_assert_emit_internal", c_bool,
                 [c_void_p, c_int, c_double_p, c_double, POINTER(c_char_p), c_int, c_int])
    s = O._Seqs(seqs)
    for draw in range(3):
        # a fresh copy per call: the reference's check damages the ARG it is
        # given (its LocalTrees copy shares the trees), so a second call on
        # the same one crashes it
        rt = O.ref_new_trees(*aw.trees_arrays(tr))
        O.srand(90 + draw)
        ref_ok = fn(c_void_p(rt), 20, _d(times), mu, s.ptr, seqs.shape[0], seqs.shape[1])
        O.srand(90 + draw)
        assert aw.assert_emit_internal(tr, seqs, times, mu) == ref_ok
        assert ref_ok
    aw.delete_local_trees(tr)


@needs_ref
def test_generic_hmm_matches_reference(aw):
    rng = np.random.RandomState(61)
    n, S = 40, 7
    trans = np.log(rng.dirichlet(np.ones(S), size=S))
    emit = np.log(rng.uniform(0.05, 1.0, size=(n, S)))
    first = np.log(np.full(S, 1.0 / S)) + emit[0]
    fw = aw.hmm_forward(trans, emit, first)
    bw = aw.hmm_backward(trans, emit)
    F = O.ref()
    for name in ("forward_alg", "backward_alg"):
        getattr(F, name).restype = None
    T, E = aw.DoubleMatrix(trans), aw.DoubleMatrix(emit)
    rf = np.zeros((n, S))
    rf[0] = first
    RF = aw.DoubleMatrix(rf)
    F.forward_alg(n, S, T.ptr, E.ptr, RF.ptr)
    rb = aw.DoubleMatrix(np.zeros((n, S)))
    F.backward_alg(n, S, T.ptr, E.ptr, rb.ptr)
    assert np.array_equal(fw, RF.arr)
    assert np.array_equal(bw, rb.arr)
    # forward and backward give the same total likelihood at every column
    tot = [np.logaddexp.reduce(fw[i] + bw[i]) for i in range(n)]
    assert np.allclose(tot, tot[0], rtol=1e-9)
    O.srand(7)
    path = aw.hmm_sample_posterior(trans, fw, 3)
    O.srand(7)
    rp = np.zeros(n, np.int32)
    rp[-1] = 3
    F.sample_hmm_posterior.restype = None
    F.sample_hmm_posterior(n, S, T.ptr, aw.DoubleMatrix(fw).ptr, _i(rp))
    assert np.array_equal(path, rp)


@needs_ref
def test_est_popsizes_matches_reference(aw):
    tr, _, (P, A, S, bl), (times, pops, rho, mu) = _trees(aw, 12, 20000, 71)
    mine = aw.est_popsizes_trees(tr, times, 500)
    rt = O.ref_new_trees(P, A, S, bl)
    ref = np.zeros(20)
    fn = _ref_fn("arghmm_est_popsizes_trees", None, [c_void_p, c_double_p, c_int, c_int, c_double_p])
    fn(rt, _d(times), 20, 500, _d(ref))
    assert np.array_equal(mine[:19], ref[:19])
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


def _arrays_equal(x, y):
    return all(np.array_equal(a, b) for a, b in zip(x, y))


@needs_ref
def test_smc_round_trip_with_reference(aw, tmp_path):
    tr, _, arrs, (times, pops, rho, mu) = _trees(aw, 7, 5000, 81)
    names = ["s%d" % i for i in range(7)]
    mine_path = str(tmp_path / "mine.smc")
    aw.write_trees(mine_path, tr, times, names)
    # the reference reads what this library writes
    rd = _ref_fn("read_local_trees", c_void_p, [c_char_p, c_double_p, c_int])
    rt = rd(mine_path.encode(), _d(times), 20)
    assert rt
    assert _arrays_equal(O.ref_trees_arrays(rt), arrs)
    # and this library reads what the reference writes, byte for byte the same file
    ref_path = str(tmp_path / "ref.smc")
    wr = _ref_fn("write_local_trees", None, [c_char_p, c_void_p, POINTER(c_char_p), c_double_p, c_int])
    cn = (c_char_p * 7)(*[n.encode() for n in names])
    wr(ref_path.encode(), rt, cn, _d(times), 20)
    assert open(ref_path).read() == open(mine_path).read()
    back = aw.read_trees(ref_path, times)
    assert _arrays_equal(aw.trees_arrays(back), arrs)
    O.ref_delete(rt)
    aw.delete_local_trees(back)
    aw.delete_local_trees(tr)


def test_smc_reads_reference_example(aw):
    """examples/sim1/sim1.sample/out.0.smc.gz (gzip) read to the arrays the
    fixture holds (read there by an independent Python parser)."""
    g = np.load(os.path.join(GOLD, "sim1.npz"), allow_pickle=False)
    tr = aw.read_trees(os.path.join(GOLD, "sim1_out0.smc.gz"), g["times"])
    P, A, S, bl = aw.trees_arrays(tr)
    assert np.array_equal(P, g["ptrees"]) and np.array_equal(A, g["ages"])
    assert np.array_equal(S[1:], g["sprs"][1:])
    # block lengths in bases (the fixture's columns may be compressed)
    with gzip.open(os.path.join(GOLD, "sim1_out0.smc.gz"), "rt") as f:
        lens = [int(l.split("\t")[2]) - int(l.split("\t")[1]) + 1 for l in f if l.startswith("TREE")]
    assert list(bl) == lens
    aw.delete_local_trees(tr)


def _parse_sites(path):
    names, pos, cols, region = [], [], [], None
    for line in open(path):
        f = line.rstrip("\n").split("\t")
        if f[0] == "NAMES":
            names = f[1:]
        elif f[0] == "REGION":
            region = (int(f[2]) - 1, int(f[3]))
        else:
            pos.append(int(f[0]) - 1)
            cols.append(f[1])
    return names, region, pos, cols


def test_sites_read_and_sequences(aw):
    path = os.path.join(GOLD, "sim1.sites")
    names, region, pos, cols = _parse_sites(path)
    s = aw.Sites(path)
    assert s.names == names
    nseqs, nsites, start, end = s.shape()
    assert (nseqs, nsites, start, end) == (len(names), len(pos), region[0], region[1])
    p, c = s.data()
    assert list(p) == pos
    assert [bytes(r).decode() for r in c] == cols
    seqs = s.sequences("A")
    assert seqs.shape == (nseqs, end - start)
    for x, col in zip(pos, cols):
        assert bytes(seqs[:, x - start]).decode() == col
    mask = np.ones(end - start, bool)
    mask[np.array(pos) - start] = False
    assert (seqs[:, mask] == ord("A")).all()


def test_sites_compression_and_tree_blocks(aw, tmp_path):
    """find_compress_cols keeps every variant site, one per compressed
    column, in order; the trees of the reference's example compress to the
    compressed length and survive uncompress -> compress unchanged
    (assert_uncompress_local_trees, local_tree.cpp:852-870)."""
    path = os.path.join(GOLD, "sim1.sites")
    s = aw.Sites(path)
    _, _, start, end = s.shape()
    p0, c0 = s.data()
    m = s.compress(20)
    _, nsites, cstart, cend = s.shape()
    p1, c1 = s.data()
    assert cstart == 0 and cend >= (end - start) // 20
    assert np.all(np.diff(p1) >= 1) and p1[-1] < cend
    assert np.array_equal(c0, c1)
    # a variant at old position x sits in compressed column ~ (x - start) / 20
    assert np.all(np.abs(p1 - (p0 - start) // 20) <= np.arange(len(p0)) + 1)
    g = np.load(os.path.join(GOLD, "sim1.npz"), allow_pickle=False)
    tr = aw.read_trees(os.path.join(GOLD, "sim1_out0.smc.gz"), g["times"])
    assert aw.awamd_compress_local_trees(tr, m) == 0
    _, _, _, bl1 = aw.trees_arrays(tr)
    assert bl1.sum() == cend - cstart and (bl1 > 0).all()
    assert aw.awamd_uncompress_local_trees(tr, m) == 0
    _, _, _, blu = aw.trees_arrays(tr)
    assert blu.sum() == end - start
    assert aw.awamd_compress_local_trees(tr, m) == 0
    assert np.array_equal(aw.trees_arrays(tr)[3], bl1)
    s.uncompress(m)
    assert np.array_equal(s.data()[0], p0) and s.shape()[2:] == (start, end)
    aw.awamd_delete_sites_mapping(m)
    aw.delete_local_trees(tr)


def test_compress_small_by_hand(aw, tmp_path):
    """A 100-base region with variants at 1-based positions 5, 6 and 57 and
    compression 10: columns of 10 bases, each variant taking its own column
    (find_compress_cols, sequences.cpp:523-599): 5 -> 0, 6 -> 1, 57 -> 5;
    the region becomes [0, 10)."""
    f = tmp_path / "t.sites"
    f.write_text("NAMES\ta\tb\nREGION\tchr\t1\t100\n5\tAC\n6\tGT\n57\tAA\n")
    s = aw.Sites(str(f))
    m = s.compress(10)
    assert list(s.data()[0]) == [0, 1, 5]
    assert s.shape()[2:] == (0, 10)
    aw.awamd_delete_sites_mapping(m)
