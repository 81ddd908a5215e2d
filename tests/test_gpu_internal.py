"""Internal-branch threading (the reference's default MCMC move) on the
device against the compiled reference (oracle/_ref), from the same ARG, the
same removal path and the same rand() stream:

- arghmm_forward_alg with internal=true after arghmm_remove_arg_thread_path
  (sample_thread.cpp:887-926): the table to 1e-12, trees in subtree/maintree
  form, trees left complete by the path (one placeholder state) included;
- arghmm_sample_arg_thread_internal (:981-1006): the identical thread;
- arghmm_resample_all_arg and arghmm_resample_climb_arg
  (sample_arg.cpp:722-806): the identical ARG after several re-threadings.
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
