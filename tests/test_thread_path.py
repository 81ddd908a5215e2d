"""Internal-branch re-threading on the host against the compiled reference
(oracle/_ref), from the same ARG and the same rand() stream:

- removal paths: the leaf path, the path forward from a node of the first
  tree, the path through a node at a position (forward and backward), the
  path weighted towards recombination-carrying branches
  (thread.cpp:672-912, 1681-1703);
- removing a branch along a path (thread.cpp:1481-1679): the ARG left in
  subtree/maintree form, node for node, and the removed branch as internal
  state indices (remove_arg_thread_path2);
- internal state counts (states.cpp:109-166, arghmm_get_nstates with
  internal) and the thread times of a path (arghmm_get_thread_times).
"""
import ctypes
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np
import pytest

import oracle_bind as O
from workloads import model

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
c_int_p = POINTER(c_int)


def _ref_fn(name, restype, argtypes):
    fn = O.ref()[name]
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


def _i(a):
    return a.ctypes.data_as(c_int_p)


def _pair(aw, k, L, seed):
    times, pops, rho, mu = model()
    tr, seqs = aw.simulate(k, L, times, pops, rho, mu, seed)
    rt = O.ref_new_trees(*aw.trees_arrays(tr))
    return tr, rt, seqs, (times, pops, rho, mu)


def _same(aw, tr, rt):
    return all(np.array_equal(a, b) for a, b in zip(aw.trees_arrays(tr), O.ref_trees_arrays(rt)))


@pytest.mark.parametrize("k,L,seed", [(6, 4000, 1), (12, 20000, 2)])
def test_removal_paths_match_reference(aw, k, L, seed):
    tr, rt, _, _ = _pair(aw, k, L, seed)
    nnodes = 2 * k - 1
    for node in (0, k - 1):
        m, r = _paths_named(aw, "arghmm_sample_arg_removal_leaf_path", tr, rt, (c_int(node),), 3)
        assert np.array_equal(m, r)
    for node in (1, k + 2, nnodes - 1):
        m, r = _paths_named(aw, "arghmm_sample_arg_removal_path", tr, rt, (c_int(node),), 4)
        assert np.array_equal(m, r)
        m, r = _paths_named(aw, "arghmm_sample_arg_removal_path2", tr, rt,
                            (c_int(node), c_int(L // 3)), 5)
        assert np.array_equal(m, r)
    for pref in (0.5, 0.9):
        m, r = _paths_named(aw, "arghmm_sample_arg_removal_path_recomb", tr, rt,
                            (c_double(pref),), 6)
        assert np.array_equal(m, r)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)


def _paths_named(aw, name, tr, rt, args, seed):
    ntrees = aw.get_local_trees_ntrees(tr)
    mine = np.full(ntrees, -7, np.int32)
    ref = np.full(ntrees, -7, np.int32)
    lib = aw._lib
    O.srand(seed)
    lib[name](c_void_p(tr), *args, _i(mine))
    O.srand(seed)
    O.ref()[name](c_void_p(rt), *args, _i(ref))
    return mine, ref


@pytest.mark.parametrize("k,L,seed,kind", [(6, 4000, 11, "leaf"), (8, 8000, 12, "node"),
                                           (12, 20000, 13, "recomb"), (10, 12000, 14, "node")])
def test_remove_thread_path_matches_reference(aw, k, L, seed, kind):
    """The branch removed along a path: the ARG in subtree/maintree form is
    the reference's, node for node, and so is the removed branch as internal
    state indices."""
    tr, rt, _, (times, _, _, _) = _pair(aw, k, L, seed)
    ntimes = len(times)
    maxtime = ntimes + 1
    if kind == "leaf":
        path, rpath = _paths_named(aw, "arghmm_sample_arg_removal_leaf_path", tr, rt,
                                   (c_int(k // 2),), 21)
    elif kind == "node":
        path, rpath = _paths_named(aw, "arghmm_sample_arg_removal_path", tr, rt,
                                   (c_int(k + 1),), 22)
    else:
        path, rpath = _paths_named(aw, "arghmm_sample_arg_removal_path_recomb", tr, rt,
                                   (c_double(0.9),), 23)
    assert np.array_equal(path, rpath)
    L = aw.seqlen_of(tr)
    orig = np.full(L, -7, np.int32)
    rorig = np.full(L, -7, np.int32)
    aw._lib["arghmm_remove_arg_thread_path2"](c_void_p(tr), _i(path), c_int(maxtime), _i(orig))
    O.ref()["arghmm_remove_arg_thread_path2"](c_void_p(rt), _i(rpath), c_int(maxtime), _i(rorig))
    assert _same(aw, tr, rt)
    assert np.array_equal(orig, rorig)
    # internal state counts per column and the times of the removed branch
    ns = np.zeros(L, np.int32)
    rns = np.zeros(L, np.int32)
    aw.argweaver_get_nstates(tr, ntimes, True, _i(ns))
    O.ref()["arghmm_get_nstates"](c_void_p(rt), c_int(ntimes), ctypes.c_bool(True), _i(rns))
    assert np.array_equal(ns, rns)
    if (ns > 0).all():
        tt = np.zeros(L, np.int32)
        rtt = np.zeros(L, np.int32)
        aw._lib["arghmm_get_thread_times"](c_void_p(tr), c_int(ntimes), _i(orig), _i(tt))
        O.ref()["arghmm_get_thread_times"](c_void_p(rt), c_int(ntimes), _i(rorig), _i(rtt))
        assert np.array_equal(tt, rtt)
    O.ref_delete(rt)
    aw.delete_local_trees(tr)
