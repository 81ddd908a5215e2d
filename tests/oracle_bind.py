"""ctypes bindings of the checkers: the oracle restatement (oracle/liboracle.so)
and, where it was built, the compiled reference (oracle/_ref/libargweaver.so).

Test infrastructure only.  Instances are described by the arrays the C ABI
exchanges (ptrees/ages/sprs/blocklens in the ARG's internal node numbering,
plus explicit per-block mappings and the leaf -> sequence map).
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_void_p

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libargweaver.so")

c_int_p = POINTER(c_int)
c_double_p = POINTER(c_double)


class OrcInstance(ctypes.Structure):
    _fields_ = [
        ("ntimes", c_int), ("times", c_double_p), ("popsizes", c_double_p),
        ("rho", c_double), ("mu", c_double),
        ("ntrees", c_int), ("nnodes", c_int), ("start_coord", c_int),
        ("ptrees", c_int_p), ("ages", c_int_p), ("sprs", c_int_p),
        ("mappings", c_int_p), ("blocklens", c_int_p),
        ("nseqs", c_int), ("seqlen", c_int), ("seqs", POINTER(c_char_p)),
        ("seqids", c_int_p), ("new_chrom", c_int),
    ]


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle oracle)")
        _oracle = ctypes.CDLL(ORACLE_PATH)
        _oracle.orc_forward_size.restype = ctypes.c_longlong
        _oracle.orc_treelen.restype = c_double
    return _oracle


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_PATH)
        _ref.arghmm_new_trees.restype = c_void_p
        _ref.arghmm_forward_alg.restype = POINTER(c_double_p)
        _ref.arghmm_sample_thread.restype = c_void_p
        _ref.arghmm_resample_arg.restype = c_void_p
        _ref.arghmm_sample_arg_seq.restype = c_void_p
        _ref.new_transition_probs.restype = POINTER(c_double_p)
        _ref.new_transition_probs_switch.restype = POINTER(c_double_p)
        _ref.new_emissions.restype = POINTER(c_double_p)
        _ref.arghmm_assert_transmat.restype = ctypes.c_bool
        _ref.arghmm_assert_transmat_switch.restype = ctypes.c_bool
        _ref.arghmm_assert_emit.restype = ctypes.c_bool
    return _ref


def _dp(a):
    return a.ctypes.data_as(c_double_p)


def _ip(a):
    return a.ctypes.data_as(c_int_p)


class Instance(object):
    """A threading instance as flat arrays, usable by the oracle and the reference."""

    def __init__(self, times, popsizes, rho, mu, ptrees, ages, sprs, blocklens, seqs,
                 seqids=None, new_chrom=None, mappings=None, start_coord=0):
        self.times = np.ascontiguousarray(times, np.float64)
        self.popsizes = np.ascontiguousarray(popsizes, np.float64)
        self.rho, self.mu = float(rho), float(mu)
        self.ptrees = np.ascontiguousarray(ptrees, np.int32)
        self.ages = np.ascontiguousarray(ages, np.int32)
        self.sprs = np.ascontiguousarray(sprs, np.int32)
        self.blocklens = np.ascontiguousarray(blocklens, np.int32)
        self.ntrees, self.nnodes = self.ptrees.shape
        self.nleaves = (self.nnodes + 1) // 2
        self.seqs = np.ascontiguousarray(seqs, np.uint8)
        self.nseqs, self.seqlen = self.seqs.shape
        self.seqids = np.ascontiguousarray(
            np.arange(self.nleaves) if seqids is None else seqids, np.int32)
        self.new_chrom = self.nleaves if new_chrom is None else int(new_chrom)
        self.mappings = None if mappings is None else np.ascontiguousarray(mappings, np.int32)
        self.start_coord = start_coord
        self._rows = (c_char_p * self.nseqs)(
            *[ctypes.cast(self.seqs[i].ctypes.data, c_char_p) for i in range(self.nseqs)])

    @property
    def ntimes(self):
        return len(self.times)

    def c(self):
        s = OrcInstance()
        s.ntimes = self.ntimes
        s.times, s.popsizes = _dp(self.times), _dp(self.popsizes)
        s.rho, s.mu = self.rho, self.mu
        s.ntrees, s.nnodes, s.start_coord = self.ntrees, self.nnodes, self.start_coord
        s.ptrees, s.ages, s.sprs = _ip(self.ptrees), _ip(self.ages), _ip(self.sprs)
        s.mappings = _ip(self.mappings) if self.mappings is not None else None
        s.blocklens = _ip(self.blocklens)
        s.nseqs, s.seqlen = self.nseqs, self.seqlen
        s.seqs = ctypes.cast(self._rows, POINTER(c_char_p))
        s.seqids, s.new_chrom = _ip(self.seqids), self.new_chrom
        self._c = s
        return ctypes.byref(s)

    # ---- oracle ---------------------------------------------------------
    def nstates(self):
        out = np.zeros(self.ntrees, np.int32)
        oracle().orc_nstates(self.c(), _ip(out))
        return out

    def forward(self, slow=False):
        n = oracle().orc_forward_size(self.c())
        fw = np.zeros(n, np.float64)
        lnl = c_double(0)
        oracle().orc_forward(self.c(), int(slow), _dp(fw), ctypes.byref(lnl))
        return fw, lnl.value

    def traceback(self, fw, u):
        path = np.zeros(self.seqlen, np.int32)
        u = np.ascontiguousarray(u, np.float64)
        oracle().orc_traceback(self.c(), _dp(fw), _dp(u), _ip(path))
        return path

    def sample_recombinations(self, path, cap=1 << 20):
        path = np.ascontiguousarray(path, np.int32)
        pos, node, tm = (np.zeros(cap, np.int32) for _ in range(3))
        n = oracle().orc_sample_recombinations(self.c(), _ip(path), _ip(pos), _ip(node),
                                               _ip(tm), cap)
        return pos[:n], node[:n], tm[:n]

    def trans_vectors(self, tree):
        out = np.zeros(9 * self.ntimes, np.float64)
        oracle().orc_trans_vectors(self.c(), tree, _dp(out))
        return out.reshape(9, self.ntimes)

    def trans_dense(self, tree):
        S = int(self.nstates()[tree])
        out = np.zeros(S * S, np.float64)
        oracle().orc_trans_dense(self.c(), tree, _dp(out))
        return out.reshape(S, S)

    def states(self, tree):
        S = int(self.nstates()[tree])
        out = np.zeros(2 * S, np.int32)
        oracle().orc_states(self.c(), tree, _ip(out))
        return out.reshape(S, 2)

    def switch(self, tree):
        S = self.nstates()
        S1, S2 = int(S[tree - 1]), int(S[tree])
        determ = np.zeros(S1, np.int32)
        dp = np.zeros(S1, np.float64)
        rr = np.zeros(S2, np.float64)
        cr = np.zeros(S2, np.float64)
        srcs = np.zeros(2, np.int32)
        oracle().orc_switch(self.c(), tree, _ip(determ), _dp(dp), _dp(rr), _dp(cr), _ip(srcs))
        return determ, dp, rr, cr, srcs

    def prior(self):
        S = int(self.nstates()[0])
        out = np.zeros(S, np.float64)
        oracle().orc_prior(self.c(), _dp(out))
        return out

    def emissions(self, tree):
        S = int(self.nstates()[tree])
        out = np.zeros(int(self.blocklens[tree]) * S, np.float64)
        oracle().orc_emissions(self.c(), tree, _dp(out))
        return out.reshape(int(self.blocklens[tree]), S)

    def block_offsets(self):
        S = self.nstates().astype(np.int64)
        return np.concatenate([[0], np.cumsum(S * self.blocklens)])


def time_points(ntimes=20, maxtime=200e3, delta=.01):
    out = np.zeros(ntimes, np.float64)
    oracle().orc_time_points(ntimes, c_double(maxtime), c_double(delta), _dp(out))
    return out


# ---- compiled reference -------------------------------------------------

class _IntMat(object):
    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr, np.int32)
        self.rows = (c_int_p * max(self.arr.shape[0], 1))(
            *[self.arr[i].ctypes.data_as(c_int_p) for i in range(self.arr.shape[0])])

    @property
    def ptr(self):
        return ctypes.cast(self.rows, POINTER(c_int_p))


def ref_new_trees(ptrees, ages, sprs, blocklens, start_coord=0):
    P, A, S = _IntMat(ptrees), _IntMat(ages), _IntMat(sprs)
    bl = np.ascontiguousarray(blocklens, np.int32)
    return ref().arghmm_new_trees(P.ptr, A.ptr, S.ptr, _ip(bl), P.arr.shape[0],
                                  P.arr.shape[1], start_coord)


def ref_trees_arrays(trees):
    r = ref()
    ntrees = r.get_local_trees_ntrees(c_void_p(trees))
    nnodes = r.get_local_trees_nnodes(c_void_p(trees))
    P = _IntMat(np.zeros((ntrees, nnodes), np.int32))
    A = _IntMat(np.zeros((ntrees, nnodes), np.int32))
    S = _IntMat(np.zeros((ntrees, 4), np.int32))
    bl = np.zeros(ntrees, np.int32)
    r.get_local_trees_ptrees(c_void_p(trees), P.ptr, A.ptr, S.ptr, _ip(bl))
    return P.arr, A.arr, S.arr, bl


def ref_remove_thread(trees, seqid):
    fn = getattr(ref(), "_ZN9argweaver17remove_arg_threadEPNS_10LocalTreesEi")
    fn.restype = None
    fn(c_void_p(trees), c_int(seqid))


def ref_delete(trees):
    ref().delete_local_trees(c_void_p(trees))


class _Seqs(object):
    def __init__(self, seqs):
        self.arr = np.ascontiguousarray(seqs, np.uint8)
        self.rows = (c_char_p * self.arr.shape[0])(
            *[ctypes.cast(self.arr[i].ctypes.data, c_char_p) for i in range(self.arr.shape[0])])

    @property
    def ptr(self):
        return ctypes.cast(self.rows, POINTER(c_char_p))


def ref_forward(trees, seqs, times, popsizes, rho, mu, nstates_per_col, slow=False):
    """arghmm_forward_alg of the compiled reference as a list of rows."""
    s = _Seqs(seqs)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    L = s.arr.shape[1]
    fw = ref().arghmm_forward_alg(c_void_p(trees), _dp(times), len(times), _dp(popsizes),
                                  c_double(rho), c_double(mu), s.ptr, s.arr.shape[0], L,
                                  ctypes.c_bool(False), None, ctypes.c_bool(False),
                                  ctypes.c_bool(slow))
    rows = [np.ctypeslib.as_array(fw[i], shape=(max(int(nstates_per_col[i]), 1),)).copy()
            for i in range(L)]
    ref().delete_forward_matrix(fw, L)
    return rows


def ref_sample_thread(trees, seqs, times, popsizes, rho, mu):
    s = _Seqs(seqs)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    ref().arghmm_sample_thread(c_void_p(trees), _dp(times), len(times), _dp(popsizes),
                               c_double(rho), c_double(mu), s.ptr, s.arr.shape[0],
                               s.arr.shape[1])


def ref_resample_arg(trees, seqs, times, popsizes, rho, mu, niters, nremove=1):
    s = _Seqs(seqs)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    ref().arghmm_resample_arg(c_void_p(trees), _dp(times), len(times), _dp(popsizes),
                              c_double(rho), c_double(mu), s.ptr, s.arr.shape[0],
                              s.arr.shape[1], niters, nremove)


def ref_sample_arg_seq(seqs, times, popsizes, rho, mu):
    s = _Seqs(seqs)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    return ref().arghmm_sample_arg_seq(_dp(times), len(times), _dp(popsizes), c_double(rho),
                                       c_double(mu), s.ptr, s.arr.shape[0], s.arr.shape[1])


def ref_sample_arg_refine(seqs, times, popsizes, rho, mu, niters, nremove):
    s = _Seqs(seqs)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    fn = ref().arghmm_sample_arg_refine
    fn.restype = c_void_p
    return fn(_dp(times), len(times), _dp(popsizes), c_double(rho), c_double(mu), s.ptr,
              s.arr.shape[0], s.arr.shape[1], niters, nremove)


def ref_sample_posterior(ptrees, ages, sprs, blocklens, seqs, times, popsizes, rho, mu):
    """arghmm_sample_posterior of the compiled reference, as [L, 2] rows."""
    P, A, S = _IntMat(ptrees), _IntMat(ages), _IntMat(sprs)
    bl = np.ascontiguousarray(blocklens, np.int32)
    s = _Seqs(seqs)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    out = np.zeros((s.arr.shape[1], 2), np.int32)
    fn = ref().arghmm_sample_posterior
    fn.restype = c_void_p
    fn(P.ptr, A.ptr, S.ptr, _ip(bl), P.arr.shape[0], P.arr.shape[1], _dp(times), len(times),
       _dp(popsizes), c_double(rho), c_double(mu), s.ptr, s.arr.shape[0], s.arr.shape[1],
       out.ctypes.data_as(c_int_p))
    return out


_libc = ctypes.CDLL(None)
_libc.rand.restype = c_int


def srand(seed):
    _libc.srand(int(seed))


def draw_uniforms(n):
    """n draws of rand()/RAND_MAX from the process-global stream."""
    RAND_MAX = 2147483647
    return np.array([_libc.rand() / float(RAND_MAX) for _ in range(n)], np.float64)
