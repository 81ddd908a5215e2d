"""Internal-branch threading (the reference's default MCMC move) on the
device against the compiled reference (oracle/_ref), from the same ARG, the
same removal path and the same rand() stream:

- arghmm_forward_alg with internal=true after arghmm_remove_arg_thread_path
  (sample_thread.cpp:887-926): the table to 1e-12, trees in subtree/maintree
  form, trees left complete by the path (one placeholder state) included;
- arghmm_sample_arg_thread_internal (:981-1006): the identical thread;
- arghmm_resample_all_arg and arghmm_resample_climb_arg
  (sample_arg.cpp:722-806): the identical ARG after several re-threadings;
- arghmm_resample_arg_region and arghmm_resample_mcmc_arg (conditioned
  re-threading of a region, accept / reject, cutting and joining the ARG):
  the identical ARG.
"""
import ctypes
from ctypes import POINTER, c_bool, c_double, c_int, c_void_p

import numpy as np
import pytest

import oracle_bind as O
from workloads import model

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")]
c_int_p = POINTER(c_int)
c_double_p = POINTER(c_double)


def _i(a):
    return a.ctypes.data_as(c_int_p)


def _d(a):
    return a.ctypes.data_as(c_double_p)


def _pair(aw, k, L, seed, pops=None):
    times, p0, rho, mu = model()
    pops = p0 if pops is None else pops
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    rt = O.ref_new_trees(*aw.trees_arrays(tr))
    return tr, rt, seqs, (times, pops, rho, mu)


def _same(aw, tr, rt):
    return all(np.array_equal(a, b) for a, b in zip(aw.trees_arrays(tr), O.ref_trees_arrays(rt)))


def _remove_path(aw, tr, rt, kind, node, seed, ntimes):
    ntrees = aw.get_local_trees_ntrees(tr)
    path = np.zeros(ntrees, np.int32)
    rpath = np.zeros(ntrees, np.int32)
    name = {"leaf": "arghmm_sample_arg_removal_leaf_path",
            "node": "arghmm_sample_arg_removal_path"}[kind]
    O.srand(seed)
    aw._lib[name](c_void_p(tr), c_int(node), _i(path))
    O.srand(seed)
    O.ref()[name](c_void_p(rt), c_int(node), _i(rpath))
    assert np.array_equal(path, rpath)
    aw._lib["arghmm_remove_arg_thread_path"](c_void_p(tr), _i(path), c_int(ntimes + 1))
    O.ref()["arghmm_remove_arg_thread_path"](c_void_p(rt), _i(rpath), c_int(ntimes + 1))
    assert _same(aw, tr, rt)


def _forward_internal(lib, trees, seqs, times, pops, rho, mu):
    s = O._Seqs(seqs)
    fn = lib["arghmm_forward_alg"]
    fn.restype = POINTER(c_double_p)
    L = s.arr.shape[1]
    fw = fn(c_void_p(trees), _d(times), len(times), _d(pops), c_double(rho), c_double(mu), s.ptr,
            s.arr.shape[0], L, c_bool(False), None, c_bool(True), c_bool(False))
    assert fw
    ns = np.zeros(L, np.int32)
    lib["arghmm_get_nstates"](c_void_p(trees), c_int(len(times)), c_bool(True), _i(ns))
    rows = [np.ctypeslib.as_array(fw[i], shape=(max(int(ns[i]), 1),)).copy() for i in range(L)]
    lib["delete_forward_matrix"](fw, L)
    return rows, ns


# nodes 2k - 2 start at the root of the first tree: the path leaves some
# trees complete (no states, one placeholder)
@pytest.mark.parametrize("k,L,seed,kind,node", [(6, 3000, 1, "leaf", 2), (8, 5000, 2, "node", 9),
                                                (10, 6000, 3, "node", 14),
                                                (12, 8000, 4, "leaf", 0),
                                                (8, 5000, 5, "node", 14),
                                                (10, 8000, 6, "node", 18)])
def test_forward_alg_internal_matches_reference(aw, k, L, seed, kind, node):
    print("internal forward k=%d L=%d seed=%d %s %d" % (k, L, seed, kind, node), flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, k, L, seed)
    _remove_path(aw, tr, rt, kind, node, 100 + seed, len(times))
    mine, ns = _forward_internal(aw._lib, tr, seqs, times, pops, rho, mu)
    ref, rns = _forward_internal(O.ref(), rt, seqs, times, pops, rho, mu)
    assert np.array_equal(ns, rns)
    for c in range(L):
        rel = np.abs(mine[c] - ref[c]) / np.maximum(np.abs(ref[c]), 1e-300)
        assert rel.max() <= 1e-12, (c, int(ns[c]), rel.max(), mine[c][:4], ref[c][:4])
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


@pytest.mark.parametrize("k,L,seed,kind,node", [(6, 3000, 11, "leaf", 1), (9, 6000, 12, "node", 10)])
def test_sample_arg_thread_internal_matches_reference(aw, k, L, seed, kind, node):
    print("internal thread k=%d L=%d seed=%d" % (k, L, seed), flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, k, L, seed)
    _remove_path(aw, tr, rt, kind, node, 200 + seed, len(times))
    s = O._Seqs(seqs)
    mine = np.zeros(L, np.int32)
    ref = np.zeros(L, np.int32)
    args = (_d(times), c_int(len(times)), _d(pops), c_double(rho), c_double(mu), s.ptr,
            c_int(s.arr.shape[0]), c_int(L))
    O.srand(300 + seed)
    aw._lib["arghmm_sample_arg_thread_internal"](c_void_p(tr), *args, _i(mine))
    O.srand(300 + seed)
    O.ref()["arghmm_sample_arg_thread_internal"](c_void_p(rt), *args, _i(ref))
    assert np.array_equal(mine, ref)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


@pytest.mark.parametrize("name,extra,k,L,seed", [
    ("arghmm_resample_all_arg", c_double(0.1), 6, 3000, 21),
    ("arghmm_resample_all_arg", c_double(0.1), 10, 5000, 22),
    ("arghmm_resample_climb_arg", c_double(0.9), 8, 4000, 23)])
def test_internal_resampling_matches_reference(aw, name, extra, k, L, seed):
    """Several internal re-threadings (removal path, forward, traceback,
    recombinations, thread insertion) give the reference's ARG exactly."""
    print("%s k=%d L=%d seed=%d" % (name, k, L, seed), flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, k, L, seed)
    s = O._Seqs(seqs)
    args = (_d(times), c_int(len(times)), _d(pops), c_double(rho), c_double(mu), s.ptr,
            c_int(s.arr.shape[0]), c_int(L), c_int(4), extra)
    fm = aw._lib[name]
    fm.restype = c_void_p
    fr = O.ref()[name]
    fr.restype = c_void_p
    O.srand(400 + seed)
    assert fm(c_void_p(tr), *args)
    O.srand(400 + seed)
    fr(c_void_p(rt), *args)
    assert _same(aw, tr, rt)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


def _block_starts(aw, tr):
    bl = aw.trees_arrays(tr)[3]
    return np.concatenate([[0], np.cumsum(bl)])[:-1]


@pytest.mark.parametrize("k,L,seed,where", [(6, 3000, 31, "middle"), (8, 4000, 32, "start"),
                                            (8, 4000, 33, "end"), (10, 5000, 34, "block")])
def test_resample_arg_region_matches_reference(aw, k, L, seed, where):
    """arghmm_resample_arg_region (sample_arg.cpp:429-527, 828-842): the ARG
    cut at the region's ends (a zero-length block when the start is a block
    start), three conditioned internal re-threadings with accept / reject,
    and the pieces joined again (congruent mapping or SPR repair) give the
    reference's ARG exactly."""
    print("region k=%d L=%d seed=%d %s" % (k, L, seed, where), flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, k, L, seed)
    if where == "middle":
        r0, r1 = L // 3, 2 * L // 3
    elif where == "start":
        r0, r1 = 0, L // 2
    elif where == "end":
        r0, r1 = L // 2, L
    else:
        starts = _block_starts(aw, tr)
        r0 = int(starts[len(starts) // 3])
        r1 = int(starts[2 * len(starts) // 3])
        assert 0 < r0 < r1 < L
    s = O._Seqs(seqs)
    args = (_d(times), c_int(len(times)), _d(pops), c_double(rho), c_double(mu), s.ptr,
            c_int(s.arr.shape[0]), c_int(L), c_int(r0), c_int(r1), c_int(3))
    fm = aw._lib["arghmm_resample_arg_region"]
    fm.restype = c_void_p
    fr = O.ref()["arghmm_resample_arg_region"]
    fr.restype = c_void_p
    O.srand(500 + seed)
    assert fm(c_void_p(tr), *args)
    O.srand(500 + seed)
    fr(c_void_p(rt), *args)
    assert _same(aw, tr, rt)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


@pytest.mark.parametrize("k,L,seed,window", [(6, 3000, 41, 1000), (8, 4000, 42, 1500)])
def test_resample_mcmc_arg_matches_reference(aw, k, L, seed, window):
    """arghmm_resample_mcmc_arg (sample_arg.cpp:190-203, 532-552, 743-765):
    leaf re-threadings and windowed region re-threadings chosen at random
    give the reference's ARG exactly."""
    print("mcmc k=%d L=%d seed=%d window=%d" % (k, L, seed, window), flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, k, L, seed)
    s = O._Seqs(seqs)
    args = (_d(times), c_int(len(times)), _d(pops), c_double(rho), c_double(mu), s.ptr,
            c_int(s.arr.shape[0]), c_int(L), c_int(3), c_int(2), c_int(window))
    fm = aw._lib["arghmm_resample_mcmc_arg"]
    fm.restype = c_void_p
    fr = O.ref()["arghmm_resample_mcmc_arg"]
    fr.restype = c_void_p
    O.srand(600 + seed)
    assert fm(c_void_p(tr), *args)
    O.srand(600 + seed)
    fr(c_void_p(rt), *args)
    assert _same(aw, tr, rt)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


def test_resample_arg_region_changes_the_arg(aw):
    """The region re-threadings above are not all rejected: with more
    iterations the ARG changes (and still equals the reference's)."""
    changed = 0
    for seed in (51, 52, 53):
        tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, 8, 4000, seed)
        before = aw.trees_arrays(tr)
        s = O._Seqs(seqs)
        args = (_d(times), c_int(len(times)), _d(pops), c_double(rho), c_double(mu), s.ptr,
                c_int(s.arr.shape[0]), c_int(4000), c_int(1000), c_int(3000), c_int(6))
        fm = aw._lib["arghmm_resample_arg_region"]
        fm.restype = c_void_p
        fr = O.ref()["arghmm_resample_arg_region"]
        fr.restype = c_void_p
        O.srand(700 + seed)
        assert fm(c_void_p(tr), *args)
        O.srand(700 + seed)
        fr(c_void_p(rt), *args)
        assert _same(aw, tr, rt)
        changed += not all(np.array_equal(a, b) for a, b in zip(before, aw.trees_arrays(tr)))
        O.ref_delete(rt)
        aw.delete_local_trees(tr)
    assert changed >= 1


def test_internal_big_blocks_match_reference(aw):
    """Internal threading with blocks of more than 1024 internal states (128
    haplotypes under a population size of 3e4; 32 states per lane, the
    kernels' symbols for such instances, planes prepared on the host): the
    forward table within 1e-12 of the compiled reference's and the same
    sampled thread."""
    k, L, seed = 128, 1200, 7
    print("internal big blocks k=%d L=%d seed=%d" % (k, L, seed), flush=True)
    tr, rt, seqs, (times, pops, rho, mu) = _pair(aw, k, L, seed, pops=np.full(20, 3e4))
    _remove_path(aw, tr, rt, "leaf", 0, 400 + seed, len(times))
    mine, ns = _forward_internal(aw._lib, tr, seqs, times, pops, rho, mu)
    ref, rns = _forward_internal(O.ref(), rt, seqs, times, pops, rho, mu)
    assert np.array_equal(ns, rns)
    assert ns.max() > 1024, ns.max()
    for c in range(L):
        rel = np.abs(mine[c] - ref[c]) / np.maximum(np.abs(ref[c]), 1e-300)
        assert rel.max() <= 1e-12, (c, int(ns[c]), rel.max())
    s = O._Seqs(seqs)
    p1 = np.zeros(L, np.int32)
    p2 = np.zeros(L, np.int32)
    args = (_d(times), c_int(len(times)), _d(pops), c_double(rho), c_double(mu), s.ptr,
            c_int(s.arr.shape[0]), c_int(L))
    O.srand(500 + seed)
    aw._lib["arghmm_sample_arg_thread_internal"](c_void_p(tr), *args, _i(p1))
    O.srand(500 + seed)
    O.ref()["arghmm_sample_arg_thread_internal"](c_void_p(rt), *args, _i(p2))
    assert np.array_equal(p1, p2)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)
