"""ctypes bindings of libargweaver_amd.so.

Mirrors the reference's ``argweaver/argweaverc.py`` (lines 29-343 declare
the exports, 529-744 wrap them) for the threading path: the same C entry
points under the same names, called with the same array conventions
(``int**`` row-pointer matrices, ``char**`` sequences, opaque
``LocalTrees*`` handles).  The reference wraps arglib ARG objects; here the
ARG travels as plain arrays (``ptrees``, ``ages``, ``sprs``, ``blocklens``),
which is exactly what ``arg2ctrees``/``ctrees2arg`` hand to and read from
the C library (argweaverc.py:1285-1353).
"""
import ctypes
from ctypes import (POINTER, c_bool, c_char_p, c_double, c_int, c_longlong,
                    c_ulonglong, c_void_p)

import numpy as np

from . import load_library

c_int_p = POINTER(c_int)
c_int_p_p = POINTER(POINTER(c_int))
c_double_p = POINTER(c_double)
c_double_p_p = POINTER(POINTER(c_double))
c_char_p_p = POINTER(c_char_p)
intstate = c_int * 2

_lib = load_library()


def _export(name, restype, argtypes):
    fn = getattr(_lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


# ARG data structure API (argweaverc.py:300-331)
argweaver_new_trees = _export(
    "arghmm_new_trees", c_void_p,
    [c_int_p_p, c_int_p_p, c_int_p_p, c_int_p, c_int, c_int, c_int])
argweaver_copy_trees = _export("arghmm_copy_trees", c_void_p, [c_void_p])
get_local_trees_ntrees = _export("get_local_trees_ntrees", c_int, [c_void_p])
get_local_trees_nnodes = _export("get_local_trees_nnodes", c_int, [c_void_p])
get_local_trees_ptrees = _export(
    "get_local_trees_ptrees", None,
    [c_void_p, c_int_p_p, c_int_p_p, c_int_p_p, c_int_p])
get_local_trees_blocks = _export(
    "get_local_trees_blocks", None, [c_void_p, c_int_p, c_int_p])
get_treelens = _export(
    "get_treelens", None, [c_void_p, c_double_p, c_int, c_double_p])
delete_local_trees = _export("delete_local_trees", None, [c_void_p])

# thread data structures (argweaverc.py:333-343)
argweaver_get_nstates = _export(
    "arghmm_get_nstates", None, [c_void_p, c_int, c_bool, c_int_p])
get_state_spaces = _export(
    "get_state_spaces", POINTER(POINTER(intstate)), [c_void_p, c_int, c_bool])
delete_state_spaces = _export(
    "delete_state_spaces", None, [POINTER(POINTER(intstate)), c_int])
delete_path = _export("delete_path", None, [c_void_p])
setLogLevel = _export("setLogLevel", None, [c_int])

# forward algorithm and thread sampling (argweaverc.py:165-243)
argweaver_forward_alg = _export(
    "arghmm_forward_alg", c_double_p_p,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double,
     c_char_p_p, c_int, c_int, c_bool, c_double_p, c_bool, c_bool])
delete_forward_matrix = _export(
    "delete_forward_matrix", None, [c_double_p_p, c_int])
delete_double_matrix = _export(
    "delete_double_matrix", None, [c_double_p_p, c_int])
argweaver_sample_posterior = _export(
    "arghmm_sample_posterior", POINTER(intstate),
    [c_int_p_p, c_int_p_p, c_int_p_p, c_int_p, c_int, c_int,
     c_double_p, c_int, c_double_p, c_double, c_double,
     c_char_p_p, c_int, c_int, POINTER(intstate)])
argweaver_sample_thread = _export(
    "arghmm_sample_thread", c_void_p,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double,
     c_char_p_p, c_int, c_int])
argweaver_sample_arg_seq = _export(
    "arghmm_sample_arg_seq", c_void_p,
    [c_double_p, c_int, c_double_p, c_double, c_double,
     c_char_p_p, c_int, c_int])
argweaver_sample_arg_refine = _export(
    "arghmm_sample_arg_refine", c_void_p,
    [c_double_p, c_int, c_double_p, c_double, c_double,
     c_char_p_p, c_int, c_int, c_int, c_int])
argweaver_resample_arg_region = _export(
    "arghmm_resample_arg_region", c_void_p,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double,
     c_char_p_p, c_int, c_int, c_int, c_int, c_int])
argweaver_resample_mcmc_arg = _export(
    "arghmm_resample_mcmc_arg", c_void_p,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double,
     c_char_p_p, c_int, c_int, c_int, c_int, c_int])
argweaver_resample_arg = _export(
    "arghmm_resample_arg", c_void_p,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double,
     c_char_p_p, c_int, c_int, c_int, c_int])

# dense matrices, checks and the generic HMM (argweaverc.py:40-128)
new_transition_probs = _export(
    "new_transition_probs", c_double_p_p,
    [c_int, c_int_p, c_int_p, c_double, POINTER(intstate), c_int, c_int, c_double_p, c_double_p,
     c_int_p, c_int_p, c_int_p, c_double_p, c_double])
new_transition_probs_switch = _export(
    "new_transition_probs_switch", c_double_p_p,
    [c_int_p, c_int_p, c_int, c_int, c_int, c_int, c_int, c_int_p, c_int_p, c_double, c_double,
     POINTER(intstate), c_int, POINTER(intstate), c_int, c_int, c_double_p, c_double_p, c_int_p,
     c_int_p, c_int_p, c_double_p, c_double])
delete_transition_probs = _export("delete_transition_probs", None, [c_double_p_p, c_int])
argweaver_assert_transmat = _export(
    "arghmm_assert_transmat", c_bool,
    [c_int, c_int_p, c_int_p, c_int, c_double_p, c_double_p, c_double])
argweaver_assert_transmat_switch = _export(
    "arghmm_assert_transmat_switch", c_bool,
    [c_int, c_int_p, c_int_p, c_int, c_int, c_int, c_int, c_int, c_double_p, c_double_p, c_double])
argweaver_assert_transmat_internal = _export(
    "arghmm_assert_transmat_internal", c_bool,
    [c_int, c_int_p, c_int_p, c_int, c_double_p, c_double_p, c_double])
argweaver_assert_transmat_switch_internal = _export(
    "arghmm_assert_transmat_switch_internal", c_bool,
    [c_void_p, c_int, c_double_p, c_double_p, c_double])
new_emissions = _export(
    "new_emissions", c_double_p_p,
    [POINTER(intstate), c_int, c_int_p, c_int, c_int_p, c_char_p_p, c_int, c_int, c_double_p,
     c_int, c_double])
delete_emissions = _export("delete_emissions", None, [c_double_p_p, c_int])
argweaver_assert_emit = _export(
    "arghmm_assert_emit", c_bool,
    [c_void_p, c_int, c_double_p, c_double, c_char_p_p, c_int, c_int])
argweaver_assert_emit_internal = _export(
    "arghmm_assert_emit_internal", c_bool,
    [c_void_p, c_int, c_double_p, c_double, c_char_p_p, c_int, c_int])
forward_alg = _export("forward_alg", None, [c_int, c_int, c_double_p_p, c_double_p_p, c_double_p_p])
backward_alg = _export("backward_alg", None, [c_int, c_int, c_double_p_p, c_double_p_p, c_double_p_p])
sample_hmm_posterior = _export(
    "sample_hmm_posterior", None, [c_int, c_int, c_double_p_p, c_double_p_p, c_int_p])
argweaver_est_popsizes_trees = _export(
    "arghmm_est_popsizes_trees", None, [c_void_p, c_double_p, c_int, c_int, c_double_p])

# on-disk formats (argweaverc.py:31-38, 316-324)
read_local_trees = _export("read_local_trees", c_void_p, [c_char_p, c_double_p, c_int])
write_local_trees = _export(
    "write_local_trees", None, [c_char_p, c_void_p, POINTER(c_char_p), c_double_p, c_int])
argweaver_read_sites = _export("arghmm_read_sites", c_void_p, [c_char_p, c_int, c_int])
argweaver_delete_sites = _export("arghmm_delete_sites", None, [c_void_p])
awamd_sites_shape = _export("awamd_sites_shape", c_int, [c_void_p, c_int_p])
awamd_sites_name = _export("awamd_sites_name", c_char_p, [c_void_p, c_int])
awamd_sites_data = _export("awamd_sites_data", c_int, [c_void_p, c_int_p, c_char_p])
awamd_sites_sequences = _export("awamd_sites_sequences", c_int, [c_void_p, ctypes.c_char, c_char_p])
awamd_sites_compress = _export("awamd_sites_compress", c_void_p, [c_void_p, c_int])
awamd_sites_uncompress = _export("awamd_sites_uncompress", c_int, [c_void_p, c_void_p])
awamd_compress_local_trees = _export("awamd_compress_local_trees", c_int, [c_void_p, c_void_p])
awamd_uncompress_local_trees = _export("awamd_uncompress_local_trees", c_int, [c_void_p, c_void_p])
awamd_delete_sites_mapping = _export("awamd_delete_sites_mapping", None, [c_void_p])

# extensions of this library
awamd_remove_thread = _export("awamd_remove_thread", c_void_p, [c_void_p, c_int])
awamd_thread_forward = _export(
    "awamd_thread_forward", c_int,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double, c_char_p_p,
     c_int, c_double_p, c_int, c_double_p, c_int_p, c_double_p, c_double_p])
awamd_thread_forward_noalloc = _export(
    "awamd_thread_forward_noalloc", c_int,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double, c_char_p_p,
     c_int, c_double_p, c_int, c_double_p, c_int_p, c_double_p, c_double_p])
awamd_num_column_states = _export(
    "awamd_num_column_states", c_longlong, [c_void_p, c_int, c_int_p])
awamd_get_mappings = _export("awamd_get_mappings", c_int, [c_void_p, c_int_p])
awamd_get_seqids = _export("awamd_get_seqids", c_int, [c_void_p, c_int_p])
# ARG probability (argweaverc.py:218-241 of the reference binding)
argweaver_likelihood = _export(
    "arghmm_likelihood", c_double, [c_void_p, c_double_p, c_int, c_double, c_char_p_p, c_int, c_int])
awamd_arg_likelihood = _export(
    "awamd_arg_likelihood", c_double, [c_void_p, c_double_p, c_int, c_double, c_char_p_p, c_int, c_int])
awamd_arg_prior = _export(
    "awamd_arg_prior", c_double, [c_void_p, c_double_p, c_int, c_double_p, c_double])
argweaver_prior_prob = _export(
    "arghmm_prior_prob", c_double, [c_void_p, c_double_p, c_int, c_double_p, c_double])
argweaver_tree_prior_prob = _export(
    "arghmm_tree_prior_prob", c_double, [c_void_p, c_double_p, c_int, c_double_p])
argweaver_joint_prob = _export(
    "arghmm_joint_prob", c_double,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double, c_char_p_p, c_int, c_int])
awamd_sample_recombinations = _export(
    "awamd_sample_recombinations", c_int,
    [c_void_p, c_double_p, c_int, c_double_p, c_double, c_double, c_int_p,
     c_int, c_int_p, c_int_p, c_int_p])
awamd_add_thread = _export(
    "awamd_add_thread", c_void_p,
    [c_void_p, c_int, c_int_p, c_int, c_int, c_int_p, c_int_p, c_int_p])
awamd_simulate = _export(
    "awamd_simulate", c_void_p,
    [c_int, c_int, c_double_p, c_int, c_double_p, c_double, c_double,
     c_ulonglong, ctypes.c_char_p])

awamd_device_info = _export("awamd_device_info", c_int, [ctypes.c_char_p, c_int])
awamd_batch_new = _export("awamd_batch_new", c_void_p, [])
awamd_batch_free = _export("awamd_batch_free", None, [c_void_p])
awamd_batch_add = _export(
    "awamd_batch_add", c_int,
    [c_void_p, c_void_p, c_double_p, c_int, c_double_p, c_double, c_double, c_char_p_p,
     c_int, c_double_p])
awamd_batch_upload = _export("awamd_batch_upload", c_int, [c_void_p])
awamd_batch_run = _export("awamd_batch_run", c_int, [c_void_p, c_int, c_int, c_int, c_double_p])
awamd_batch_run_opts = _export(
    "awamd_batch_run_opts", c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double_p])
awamd_batch_run_ex = _export(
    "awamd_batch_run_ex", c_int,
    [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_double, c_double_p])
awamd_batch_set_traceback = _export("awamd_batch_set_traceback", c_int, [c_void_p, c_int, c_int])
awamd_batch_set_tail_tiles = _export("awamd_batch_set_tail_tiles", c_int, [c_void_p, c_int])
awamd_batch_clear = _export("awamd_batch_clear", c_int, [c_void_p])
awamd_batch_tune = _export("awamd_batch_tune", c_int, [c_void_p, c_int, c_int_p, c_int, c_double_p])
awamd_batch_tune_forward = _export(
    "awamd_batch_tune_forward", c_int, [c_void_p, c_int_p, c_int_p, c_int, c_double, c_double_p])
awamd_batch_path = _export("awamd_batch_path", c_int, [c_void_p, c_int, c_int_p])
awamd_batch_alpha_range = _export("awamd_batch_alpha_range", c_int,
                                  [c_void_p, c_int, c_int, c_int, c_double_p])
awamd_batch_alpha = _export("awamd_batch_alpha", c_int, [c_void_p, c_int, c_double_p])
awamd_batch_norms = _export("awamd_batch_norms", c_int, [c_void_p, c_int, c_double_p])
awamd_batch_lnl = _export("awamd_batch_lnl", c_double, [c_void_p, c_int])
awamd_batch_W = _export("awamd_batch_W", c_longlong, [c_void_p, c_int])

awamd_server_new = _export("awamd_server_new", c_void_p, [c_double, c_int, c_int, c_int])
awamd_server_free = _export("awamd_server_free", None, [c_void_p])
awamd_server_submit = _export(
    "awamd_server_submit", c_int,
    [c_void_p, c_void_p, c_double_p, c_int, c_double_p, c_double, c_double, c_char_p_p, c_int,
     c_double_p])
awamd_server_wait = _export("awamd_server_wait", c_int,
                            [c_void_p, c_int, c_double, c_int_p, c_double_p, c_double_p])
awamd_server_now = _export("awamd_server_now", c_double, [c_void_p])
awamd_server_new_ex = _export("awamd_server_new_ex", c_void_p, [c_double, c_int, c_int, c_int, c_int])
awamd_server_wait_ex = _export("awamd_server_wait_ex", c_int,
                               [c_void_p, c_int, c_double, c_int_p, c_double_p, c_double_p])

_libc = ctypes.CDLL(None)


def device_info():
    buf = ctypes.create_string_buffer(512)
    awamd_device_info(buf, 512)
    return buf.value.decode()


def set_random_seed(num):
    """Seed the process-global C rand() stream (argweaverc.py:352-354)."""
    _libc.srand(int(num))


#=============================================================================
# array helpers (ctypes_export.py:29-35 c_matrix conventions)

def _dptr(a):
    return a.ctypes.data_as(c_double_p)


def _iptr(a):
    return a.ctypes.data_as(c_int_p)


class IntMatrix(object):
    """An int** view over a C-contiguous 2-D int32 array (kept alive here)."""

    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr, dtype=np.int32)
        n = self.arr.shape[0]
        # the row addresses, computed in one numpy expression (a ctypes
        # pointer per row costs seconds on large ARGs)
        addr = np.zeros(max(n, 1), np.uint64)
        if n:
            addr[:n] = self.arr.ctypes.data + np.arange(n, dtype=np.uint64) * self.arr.strides[0]
        self.addr = addr
        self.rows = (c_int_p * max(n, 1)).from_buffer(addr)

    @property
    def ptr(self):
        return ctypes.cast(self.rows, c_int_p_p)


class CSeqs(object):
    """A char** over equal-length byte rows (kept alive here)."""

    def __init__(self, seqs):
        if isinstance(seqs, np.ndarray):
            self.arr = np.ascontiguousarray(seqs, dtype=np.uint8)
        else:
            rows = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
            self.arr = np.frombuffer(b"".join(rows), dtype=np.uint8).reshape(len(rows), -1).copy()
        self.nseqs, self.seqlen = self.arr.shape
        ptrs = [self.arr[i].ctypes.data for i in range(self.nseqs)]
        self.rows = (c_char_p * self.nseqs)(*[ctypes.cast(p, c_char_p) for p in ptrs])

    @property
    def ptr(self):
        return ctypes.cast(self.rows, c_char_p_p)


def new_trees(ptrees, ages, sprs, blocklens, start_coord=0):
    """arghmm_new_trees over numpy arrays [ntrees, nnodes] / [ntrees, 4] / [ntrees]."""
    ptrees = np.asarray(ptrees, dtype=np.int32)
    ntrees, nnodes = ptrees.shape
    P, A, S = IntMatrix(ptrees), IntMatrix(ages), IntMatrix(sprs)
    bl = np.ascontiguousarray(blocklens, dtype=np.int32)
    return argweaver_new_trees(P.ptr, A.ptr, S.ptr, _iptr(bl), ntrees, nnodes, start_coord)


def trees_arrays(trees):
    """get_local_trees_ptrees as numpy arrays (ptrees, ages, sprs, blocklens)."""
    ntrees = get_local_trees_ntrees(trees)
    nnodes = get_local_trees_nnodes(trees)
    P = IntMatrix(np.zeros((ntrees, nnodes), np.int32))
    A = IntMatrix(np.zeros((ntrees, nnodes), np.int32))
    S = IntMatrix(np.zeros((ntrees, 4), np.int32))
    bl = np.zeros(ntrees, np.int32)
    get_local_trees_ptrees(trees, P.ptr, A.ptr, S.ptr, _iptr(bl))
    return P.arr, A.arr, S.arr, bl


def trees_internal(trees):
    """Block arrays in internal node numbering plus mappings and seqids."""
    ntrees = get_local_trees_ntrees(trees)
    nnodes = get_local_trees_nnodes(trees)
    maps = np.zeros((ntrees, nnodes), np.int32)
    awamd_get_mappings(trees, _iptr(maps))
    nleaves = (nnodes + 1) // 2
    seqids = np.zeros(max(nleaves, 1), np.int32)
    awamd_get_seqids(trees, _iptr(seqids))
    return maps, seqids[:nleaves]


def seqlen_of(trees):
    """Columns covered by the ARG's blocks (get_local_trees_blocks)."""
    n = get_local_trees_ntrees(trees)
    starts = np.zeros(max(n, 1), np.int32)
    ends = np.zeros(max(n, 1), np.int32)
    get_local_trees_blocks(trees, _iptr(starts), _iptr(ends))
    return int(ends[n - 1] - starts[0]) if n else 0


def num_column_states(trees, ntimes):
    ntrees = get_local_trees_ntrees(trees)
    S = np.zeros(ntrees, np.int32)
    W = awamd_num_column_states(trees, ntimes, _iptr(S))
    return int(W), S


def get_nstates(trees, ntimes, seqlen):
    out = np.zeros(seqlen, np.int32)
    argweaver_get_nstates(trees, ntimes, False, _iptr(out))
    return out


def state_spaces(trees, ntimes):
    """get_state_spaces as a list of [S, 2] arrays (node, time)."""
    ntrees = get_local_trees_ntrees(trees)
    ptr = get_state_spaces(trees, ntimes, False)
    _, S = num_column_states(trees, ntimes)
    out = []
    for i in range(ntrees):
        rows = np.ctypeslib.as_array(ctypes.cast(ptr[i], c_int_p), shape=(int(S[i]) * 2,))
        out.append(rows.reshape(-1, 2).copy())
    delete_state_spaces(ptr, ntrees)
    return out


#=============================================================================
# high-level wrappers (argweaverc.py:529-744 counterparts)

def forward_algorithm(trees, seqs, times, popsizes, rho, mu, prior=None, slow=False):
    """argweaver_forward_algorithm: the forward table as a list of rows."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    pr = None if prior is None else np.ascontiguousarray(prior, np.float64)
    fw = argweaver_forward_alg(
        trees, _dptr(times), len(times), _dptr(popsizes), rho, mu, cs.ptr,
        cs.nseqs, cs.seqlen, pr is not None,
        _dptr(pr) if pr is not None else None, False, slow)
    if not fw:
        raise RuntimeError("arghmm_forward_alg failed")
    nst = get_nstates(trees, len(times), cs.seqlen)
    rows = [np.ctypeslib.as_array(fw[i], shape=(max(int(nst[i]), 1),)).copy()
            for i in range(cs.seqlen)]
    delete_forward_matrix(fw, cs.seqlen)
    return rows


def sample_thread(trees, seqs, times, popsizes, rho, mu):
    """arghmm_sample_thread: thread the last sequence into the ARG (in place)."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return argweaver_sample_thread(trees, _dptr(times), len(times), _dptr(popsizes),
                                   rho, mu, cs.ptr, cs.nseqs, cs.seqlen)


def resample_arg(trees, seqs, times, popsizes, rho, mu, niters=1, nremove=1):
    """arghmm_resample_arg: Gibbs re-threading sweeps (in place)."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return argweaver_resample_arg(trees, _dptr(times), len(times), _dptr(popsizes),
                                  rho, mu, cs.ptr, cs.nseqs, cs.seqlen, niters, nremove)


def resample_arg_region(trees, seqs, times, popsizes, rho, mu, region_start, region_end,
                        niters=1):
    """arghmm_resample_arg_region: conditioned internal re-threadings of the
    columns [region_start, region_end) (MCMC accept / reject)."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    out = argweaver_resample_arg_region(trees, _dptr(times), len(times), _dptr(popsizes), rho, mu,
                                        cs.ptr, cs.nseqs, cs.seqlen, region_start, region_end,
                                        niters)
    if not out:
        raise RuntimeError("arghmm_resample_arg_region failed")
    return out


def resample_mcmc_arg(trees, seqs, times, popsizes, rho, mu, niters=1, niters2=1, window=200):
    """arghmm_resample_mcmc_arg: leaf re-threadings and windowed region
    re-threadings, chosen at random per iteration."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    out = argweaver_resample_mcmc_arg(trees, _dptr(times), len(times), _dptr(popsizes), rho, mu,
                                      cs.ptr, cs.nseqs, cs.seqlen, niters, niters2, window)
    if not out:
        raise RuntimeError("arghmm_resample_mcmc_arg failed")
    return out


def sample_arg_seq(seqs, times, popsizes, rho, mu):
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return argweaver_sample_arg_seq(_dptr(times), len(times), _dptr(popsizes), rho, mu,
                                    cs.ptr, cs.nseqs, cs.seqlen)


def sample_posterior(ptrees, ages, sprs, blocklens, seqs, times, popsizes, rho, mu):
    """arghmm_sample_posterior: forward + traceback of sequence nleaves
    against the ARG given as arrays; returns the path as [L, 2] (node, time)
    rows (the reference's block conversion is kept, so only a one-tree ARG
    gives a fully defined path)."""
    ptrees = np.asarray(ptrees, np.int32)
    ntrees, nnodes = ptrees.shape
    P, A, S = IntMatrix(ptrees), IntMatrix(ages), IntMatrix(sprs)
    bl = np.ascontiguousarray(blocklens, np.int32)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    out = np.zeros((cs.seqlen, 2), np.int32)
    res = argweaver_sample_posterior(
        P.ptr, A.ptr, S.ptr, _iptr(bl), ntrees, nnodes, _dptr(times), len(times),
        _dptr(popsizes), rho, mu, cs.ptr, cs.nseqs, cs.seqlen,
        out.ctypes.data_as(POINTER(intstate)))
    if not res:
        raise RuntimeError("arghmm_sample_posterior failed")
    return out


def sample_arg_refine(seqs, times, popsizes, rho, mu, niters=1, nremove=1):
    """arghmm_sample_arg_refine: sequential sampling, then niters Gibbs sweeps."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return argweaver_sample_arg_refine(_dptr(times), len(times), _dptr(popsizes), rho, mu,
                                       cs.ptr, cs.nseqs, cs.seqlen, niters, nremove)


def thread_forward(trees, seqs, times, popsizes, rho, mu, new_chrom, uniforms=None,
                   precision=0, want_alpha=True, noalloc=False):
    """Forward (+ traceback with explicit uniforms) on the device.

    noalloc: run through awamd_thread_forward_noalloc (the fixed per-GPU
    workspace instead of allocated device buffers).
    Returns dict(alpha=[W] FP64 block-major table or None, path=[L] or None,
    lnl=float, ms=[prep, emit, forward, traceback])."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    L = seqlen_of(trees)
    alpha = np.zeros(num_column_states(trees, len(times))[0], np.float64) if want_alpha else None
    path = np.zeros(L, np.int32) if uniforms is not None else None
    u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float64)
    lnl = c_double(0.0)
    ms = np.zeros(4, np.float64)
    fn = awamd_thread_forward_noalloc if noalloc else awamd_thread_forward
    rc = fn(
        trees, _dptr(times), len(times), _dptr(popsizes), rho, mu, cs.ptr, new_chrom,
        _dptr(u) if u is not None else None, precision,
        _dptr(alpha) if alpha is not None else None,
        _iptr(path) if path is not None else None, ctypes.byref(lnl), _dptr(ms))
    if rc != 0:
        raise RuntimeError("%s failed (%d)" % (fn.__name__, rc))
    return dict(alpha=alpha, path=path, lnl=lnl.value, ms=ms)


def simulate(k, seqlen, times, popsizes, rho, mu, seed):
    """Seeded DSMC workload: returns (trees handle with k leaves, seqs [k, L] uint8)."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    seqs = np.zeros((k, seqlen), np.uint8)
    buf = seqs.ctypes.data_as(ctypes.c_char_p)
    trees = awamd_simulate(k, seqlen, _dptr(times), len(times), _dptr(popsizes), rho, mu,
                           seed, buf)
    if not trees:
        raise RuntimeError("simulation failed")
    return trees, seqs


def remove_thread(trees, seqid):
    return awamd_remove_thread(trees, seqid)


def sample_recombinations(trees, times, popsizes, rho, mu, path, maxrecomb=1 << 20):
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    path = np.ascontiguousarray(path, np.int32)
    cap = min(maxrecomb, len(path))
    pos = np.zeros(cap, np.int32)
    node = np.zeros(cap, np.int32)
    tm = np.zeros(cap, np.int32)
    n = awamd_sample_recombinations(trees, _dptr(times), len(times), _dptr(popsizes), rho,
                                    mu, _iptr(path), cap, _iptr(pos), _iptr(node), _iptr(tm))
    return pos[:n], node[:n], tm[:n]


def add_thread(trees, ntimes, path, seqid, recomb_pos, recomb_node, recomb_time):
    path = np.ascontiguousarray(path, np.int32)
    rp = np.ascontiguousarray(recomb_pos, np.int32)
    rn = np.ascontiguousarray(recomb_node, np.int32)
    rt = np.ascontiguousarray(recomb_time, np.int32)
    return awamd_add_thread(trees, ntimes, _iptr(path), seqid, len(rp), _iptr(rp), _iptr(rn),
                            _iptr(rt))


def arg_likelihood(trees, seqs, times, mu):
    """log P(sequences | ARG) (arghmm_likelihood)."""
    times = np.ascontiguousarray(times, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return argweaver_likelihood(trees, _dptr(times), len(times), mu, cs.ptr, cs.nseqs, cs.seqlen)


def arg_prior_device(trees, times, popsizes, rho):
    """log P(ARG) computed on the device (awamd_arg_prior): the terms of
    arghmm_prior_prob in another summation order."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    return awamd_arg_prior(trees, _dptr(times), len(times), _dptr(popsizes), rho)


def arg_likelihood_device(trees, seqs, times, mu):
    """log P(sequences | ARG) computed on the device (awamd_arg_likelihood):
    the terms of arghmm_likelihood in another summation order."""
    times = np.ascontiguousarray(times, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return awamd_arg_likelihood(trees, _dptr(times), len(times), mu, cs.ptr, cs.nseqs, cs.seqlen)


def arg_prior(trees, times, popsizes, rho):
    """log P(ARG) (arghmm_prior_prob)."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    return argweaver_prior_prob(trees, _dptr(times), len(times), _dptr(popsizes), rho)


def tree_prior(trees, times, popsizes):
    """log prior of the first local tree (arghmm_tree_prior_prob)."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    return argweaver_tree_prior_prob(trees, _dptr(times), len(times), _dptr(popsizes))


def arg_joint_prob(trees, seqs, times, popsizes, mu, rho):
    """log P(sequences, ARG) (arghmm_joint_prob)."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return argweaver_joint_prob(trees, _dptr(times), len(times), _dptr(popsizes), mu, rho,
                                cs.ptr, cs.nseqs, cs.seqlen)


def get_time_points(ntimes=20, maxtime=200e3, delta=.01):
    """Time grid (argweaver/__init__.py:97-105, model.h:21-32)."""
    return [(np.exp(i / float(ntimes - 1) * np.log(1 + delta * maxtime)) - 1) / delta
            for i in range(ntimes)]


#=============================================================================
# dense matrices, checks, generic HMM and file formats

class DoubleMatrix(object):
    """A double** view over a C-contiguous 2-D float64 array (kept alive here)."""

    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr, dtype=np.float64)
        n = self.arr.shape[0]
        addr = np.zeros(max(n, 1), np.uint64)
        if n:
            addr[:n] = self.arr.ctypes.data + np.arange(n, dtype=np.uint64) * self.arr.strides[0]
        self.addr = addr
        self.rows = (c_double_p * max(n, 1)).from_buffer(addr)

    @property
    def ptr(self):
        return ctypes.cast(self.rows, c_double_p_p)


def _rows_to_array(m, nrows, ncols):
    out = np.zeros((nrows, ncols), np.float64)
    for i in range(nrows):
        out[i] = np.ctypeslib.as_array(m[i], shape=(ncols,))
    return out


def _states_arg(states):
    st = np.ascontiguousarray(states, np.int32).reshape(-1, 2)
    return st, st.ctypes.data_as(POINTER(intstate))


def tree_states(ptree, ages, ntimes):
    """States of one tree in node-major order (get_coal_states, states.cpp:53-75)."""
    out = []
    for i, p in enumerate(ptree):
        top = ntimes - 2 if p == -1 else ages[p]
        out.extend((i, a) for a in range(ages[i], top + 1))
    return np.array(out, np.int32).reshape(-1, 2)


def transition_probs(ptree, ages, times, popsizes, rho, states=None):
    """log P(i -> j) within a block (new_transition_probs) as [S, S]."""
    ptree = np.ascontiguousarray(ptree, np.int32)
    ages = np.ascontiguousarray(ages, np.int32)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    st, stp = _states_arg(tree_states(ptree, ages, len(times)) if states is None else states)
    S = len(st)
    z = np.zeros(len(times), np.int32)
    m = new_transition_probs(len(ptree), _iptr(ptree), _iptr(ages), 0.0, stp, S, len(times),
                             _dptr(times), _dptr(times), _iptr(z), _iptr(z), _iptr(z),
                             _dptr(popsizes), rho)
    if not m:
        raise RuntimeError("new_transition_probs failed")
    out = _rows_to_array(m, S, S)
    delete_transition_probs(m, S)
    return out


def transition_probs_switch(ptree, last_ptree, ages, last_ages, spr, times, popsizes, rho):
    """log P(i -> j) across an SPR (new_transition_probs_switch) as [S1, S2];
    spr = (recomb_node, recomb_time, coal_node, coal_time)."""
    ptree = np.ascontiguousarray(ptree, np.int32)
    last_ptree = np.ascontiguousarray(last_ptree, np.int32)
    ages = np.ascontiguousarray(ages, np.int32)
    last_ages = np.ascontiguousarray(last_ages, np.int32)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    s1, s1p = _states_arg(tree_states(last_ptree, last_ages, len(times)))
    s2, s2p = _states_arg(tree_states(ptree, ages, len(times)))
    z = np.zeros(len(times), np.int32)
    m = new_transition_probs_switch(
        _iptr(ptree), _iptr(last_ptree), len(ptree), int(spr[0]), int(spr[1]), int(spr[2]),
        int(spr[3]), _iptr(ages), _iptr(last_ages), 0.0, 0.0, s1p, len(s1), s2p, len(s2),
        len(times), _dptr(times), _dptr(times), _iptr(z), _iptr(z), _iptr(z), _dptr(popsizes), rho)
    if not m:
        raise RuntimeError("new_transition_probs_switch failed")
    out = _rows_to_array(m, len(s1), len(s2))
    delete_transition_probs(m, len(s1))
    return out


def assert_transmat(ptree, ages, times, popsizes, rho):
    ptree = np.ascontiguousarray(ptree, np.int32)
    ages = np.ascontiguousarray(ages, np.int32)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    return bool(argweaver_assert_transmat(len(ptree), _iptr(ptree), _iptr(ages), len(times),
                                          _dptr(times), _dptr(popsizes), rho))


def assert_transmat_switch(ptree, ages, spr, times, popsizes, rho):
    ptree = np.ascontiguousarray(ptree, np.int32)
    ages = np.ascontiguousarray(ages, np.int32)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    return bool(argweaver_assert_transmat_switch(
        len(ptree), _iptr(ptree), _iptr(ages), int(spr[0]), int(spr[1]), int(spr[2]), int(spr[3]),
        len(times), _dptr(times), _dptr(popsizes), rho))


def assert_transmat_internal(ptree, ages, times, popsizes, rho):
    """The internal-branch within-block check (draws one irand for the
    branch lifted above the root)."""
    ptree = np.ascontiguousarray(ptree, np.int32)
    ages = np.ascontiguousarray(ages, np.int32)
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    return bool(argweaver_assert_transmat_internal(len(ptree), _iptr(ptree), _iptr(ages),
                                                   len(times), _dptr(times), _dptr(popsizes), rho))


def assert_transmat_switch_internal(trees, times, popsizes, rho):
    """The internal-branch switch check; leaves trees in subtree/maintree
    form, as the reference does."""
    times = np.ascontiguousarray(times, np.float64)
    popsizes = np.ascontiguousarray(popsizes, np.float64)
    return bool(argweaver_assert_transmat_switch_internal(trees, len(times), _dptr(times),
                                                          _dptr(popsizes), rho))


def emissions(ptree, ages, seqs, times, mu, states=None):
    """emit[column, state] of the new sequence seqs[nleaves] (new_emissions)."""
    ptree = np.ascontiguousarray(ptree, np.int32)
    ages = np.ascontiguousarray(ages, np.int32)
    times = np.ascontiguousarray(times, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    st, stp = _states_arg(tree_states(ptree, ages, len(times)) if states is None else states)
    m = new_emissions(stp, len(st), _iptr(ptree), len(ptree), _iptr(ages), cs.ptr, cs.nseqs,
                      cs.seqlen, _dptr(times), len(times), mu)
    if not m:
        raise RuntimeError("new_emissions failed")
    out = _rows_to_array(m, cs.seqlen, len(st))
    delete_emissions(m, cs.seqlen)
    return out


def assert_emit(trees, seqs, times, mu):
    times = np.ascontiguousarray(times, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return bool(argweaver_assert_emit(trees, len(times), _dptr(times), mu, cs.ptr, cs.nseqs,
                                      cs.seqlen))


def assert_emit_internal(trees, seqs, times, mu):
    times = np.ascontiguousarray(times, np.float64)
    cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
    return bool(argweaver_assert_emit_internal(trees, len(times), _dptr(times), mu, cs.ptr,
                                               cs.nseqs, cs.seqlen))


def hmm_forward(trans, emit, first):
    """forward_alg (hmm.cpp:27-44) in log space; returns [n, S]."""
    emit = np.ascontiguousarray(emit, np.float64)
    n, S = emit.shape
    fw = np.zeros((n, S), np.float64)
    fw[0] = first
    T, E, F = DoubleMatrix(trans), DoubleMatrix(emit), DoubleMatrix(fw)
    forward_alg(n, S, T.ptr, E.ptr, F.ptr)
    return F.arr


def hmm_backward(trans, emit, last=None):
    emit = np.ascontiguousarray(emit, np.float64)
    n, S = emit.shape
    bw = np.zeros((n, S), np.float64)
    if last is not None:
        bw[-1] = last
    T, E, B = DoubleMatrix(trans), DoubleMatrix(emit), DoubleMatrix(bw)
    backward_alg(n, S, T.ptr, E.ptr, B.ptr)
    return B.arr


def hmm_sample_posterior(trans, fw, last_state):
    fw = np.ascontiguousarray(fw, np.float64)
    n = fw.shape[0]
    path = np.zeros(n, np.int32)
    path[-1] = last_state
    T, F = DoubleMatrix(trans), DoubleMatrix(fw)
    sample_hmm_posterior(n, fw.shape[1], T.ptr, F.ptr, _iptr(path))
    return path


def est_popsizes_trees(trees, times, step):
    times = np.ascontiguousarray(times, np.float64)
    out = np.zeros(len(times), np.float64)
    argweaver_est_popsizes_trees(trees, _dptr(times), len(times), step, _dptr(out))
    return out


def read_trees(filename, times):
    """read_local_trees: an ARG handle from a .smc (or .smc.gz) file."""
    times = np.ascontiguousarray(times, np.float64)
    tr = read_local_trees(filename.encode(), _dptr(times), len(times))
    if not tr:
        raise IOError("cannot read local trees from %s" % filename)
    return tr


def write_trees(filename, trees, times, names=None):
    times = np.ascontiguousarray(times, np.float64)
    arr = None
    if names is not None:
        arr = (c_char_p * len(names))(*[n.encode() for n in names])
    write_local_trees(filename.encode(), trees, arr, _dptr(times), len(times))


class Sites(object):
    """A .sites alignment (arghmm_read_sites) with its compression."""

    def __init__(self, filename, start=-1, end=-1):
        self.ptr = argweaver_read_sites(filename.encode(), start, end)
        if not self.ptr:
            raise IOError("cannot read sites from %s" % filename)

    def shape(self):
        sh = np.zeros(4, np.int32)
        awamd_sites_shape(self.ptr, _iptr(sh))
        return tuple(int(x) for x in sh)   # nseqs, nsites, start, end

    @property
    def names(self):
        return [awamd_sites_name(self.ptr, i).decode() for i in range(self.shape()[0])]

    def data(self):
        nseqs, nsites, _, _ = self.shape()
        pos = np.zeros(max(nsites, 1), np.int32)
        cols = np.zeros((max(nsites, 1), nseqs), np.uint8)
        awamd_sites_data(self.ptr, _iptr(pos), cols.ctypes.data_as(c_char_p))
        return pos[:nsites], cols[:nsites]

    def sequences(self, default="A"):
        nseqs, _, start, end = self.shape()
        out = np.zeros((nseqs, end - start), np.uint8)
        awamd_sites_sequences(self.ptr, default.encode(), out.ctypes.data_as(c_char_p))
        return out

    def compress(self, factor):
        """Compress in place; returns the mapping handle (SitesMapping)."""
        m = awamd_sites_compress(self.ptr, factor)
        if not m:
            raise ValueError("sites cannot be compressed by %d" % factor)
        return m

    def uncompress(self, mapping):
        awamd_sites_uncompress(self.ptr, mapping)

    def __del__(self):
        if getattr(self, "ptr", None):
            argweaver_delete_sites(self.ptr)
            self.ptr = None


class Batch(object):
    """Many independent threading instances resident in HBM (awamd_batch_*)."""

    STAT_NAMES = ("ms_prep", "ms_emit", "ms_forward", "ms_traceback", "ms_total", "W",
                  "nvar", "tb_passes", "tb_rerun_cols", "fw_tiles", "fw_reruns", "ms_fw_main",
                  "ms_tb_tail", "ms_tb_side")

    def __init__(self):
        self.ptr = awamd_batch_new()
        self._keep = []
        self.L = []

    def add(self, trees, seqs, times, popsizes, rho, mu, new_chrom, uniforms=None):
        times = np.ascontiguousarray(times, np.float64)
        popsizes = np.ascontiguousarray(popsizes, np.float64)
        cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
        u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float64)
        i = awamd_batch_add(self.ptr, trees, _dptr(times), len(times), _dptr(popsizes), rho,
                            mu, cs.ptr, new_chrom, _dptr(u) if u is not None else None)
        if i < 0:
            raise RuntimeError("awamd_batch_add refused the instance (batch already uploaded)")
        self._keep.append((times, popsizes, cs, u))
        self.L.append(cs.seqlen)
        return i

    def clear(self):
        """Empty the batch for new instances; its device buffers are kept."""
        awamd_batch_clear(self.ptr)
        self._keep = []
        self.L = []

    def upload(self):
        if awamd_batch_upload(self.ptr) != 0:
            raise RuntimeError("awamd_batch_upload failed")
        self._keep = []

    def run(self, precision=0, tile_cols=0, traceback=True, fixup=True, fw_tile=0, fw_warm=0,
            fw_tol=0.0):
        """precision 0: FP64 table; 1: FP32 table with FP64 normalisers and
        traceback CDF.  tile_cols: traceback tile length (0: whole instance).
        fw_tile / fw_warm / fw_tol: FP32 forward tiles (awamd_batch_run_ex)."""
        st = np.zeros(len(self.STAT_NAMES), np.float64)
        if awamd_batch_run_ex(self.ptr, precision, tile_cols, int(traceback), int(fixup),
                              int(fw_tile), int(fw_warm), float(fw_tol), _dptr(st)) != 0:
            raise RuntimeError("awamd_batch_run failed")
        return dict(zip(self.STAT_NAMES, st.tolist()))

    def set_traceback(self, warm=256, options=3):
        """Traceback tiles of later runs: the speculative chain's warm-up
        columns and the options per tile (1..4); the thread does not depend
        on them."""
        if awamd_batch_set_traceback(self.ptr, int(warm), int(options)) != 0:
            raise ValueError("traceback warm-up / options out of range")

    def set_tail_tiles(self, cols):
        """Traceback tile length of the instances whose FP32 forward tiles
        are re-run (0: the run's tile length); the thread does not depend on
        it."""
        if awamd_batch_set_tail_tiles(self.ptr, int(cols)) != 0:
            raise ValueError("tail traceback tile length out of range")

    def tune_forward(self, candidates=((0, 0), (60000, 25000), (80000, 25000), (90000, 35000),
                                       (100000, 40000), (110000, 45000), (120000, 50000),
                                       (140000, 60000)), tol=1e-4):
        """Autotune the FP32 forward tiles on this batch; returns the best
        (tile, warm-up) pair and {pair: forward ms}."""
        tiles = np.ascontiguousarray([c[0] for c in candidates], np.int32)
        warms = np.ascontiguousarray([c[1] for c in candidates], np.int32)
        ms = np.zeros(len(candidates), np.float64)
        best = awamd_batch_tune_forward(self.ptr, _iptr(tiles), _iptr(warms), len(candidates),
                                        float(tol), _dptr(ms))
        if best < 0:
            raise RuntimeError("awamd_batch_tune_forward failed")
        return tuple(candidates[best]), {tuple(c): float(m) for c, m in zip(candidates, ms)}

    def tune(self, precision=1, candidates=(0, 2000, 4000, 8000, 10000, 12000, 14000, 16000)):
        """Autotune the traceback tile length on this batch; returns
        (best, {candidate: ms})."""
        cand = np.ascontiguousarray(candidates, np.int32)
        ms = np.zeros(len(cand), np.float64)
        best = awamd_batch_tune(self.ptr, precision, _iptr(cand), len(cand), _dptr(ms))
        if best < 0:
            raise RuntimeError("awamd_batch_tune failed")
        return int(best), dict(zip(cand.tolist(), ms.tolist()))

    def path(self, i):
        out = np.zeros(self.L[i], np.int32)
        awamd_batch_path(self.ptr, i, _iptr(out))
        return out

    def alpha(self, i):
        out = np.zeros(int(awamd_batch_W(self.ptr, i)), np.float64)
        awamd_batch_alpha(self.ptr, i, _dptr(out))
        return out

    def alpha_range(self, i, c0, c1, nstates):
        """Columns [c0, c1) of instance i's table (nstates: the per-column
        state counts of the instance, as get_nstates gives them)."""
        n = int(np.sum(nstates[c0:c1]))
        out = np.zeros(n, np.float64)
        if awamd_batch_alpha_range(self.ptr, i, int(c0), int(c1), _dptr(out)) != 0:
            raise ValueError("bad column range")
        return out

    def norms(self, i):
        out = np.zeros(self.L[i], np.float64)
        awamd_batch_norms(self.ptr, i, _dptr(out))
        return out

    def lnl(self, i):
        return awamd_batch_lnl(self.ptr, i)

    def W(self, i):
        return int(awamd_batch_W(self.ptr, i))

    def __del__(self):
        if getattr(self, "ptr", None):
            awamd_batch_free(self.ptr)
            self.ptr = None


class Server(object):
    """Continuous-batching server (awamd_server_*): submit() from any thread,
    wait() for a result.  Each of `drivers` driver threads runs a batch of
    its own: the requests queued while one batch is on the GPU are admitted
    together, in arrival order, by a free driver within its share of the
    device-memory budget, and run beside the batch in flight."""

    def __init__(self, hbm_budget_gb=64.0, precision=1, tile_cols=2000, max_batch=64, drivers=2):
        self.ptr = awamd_server_new_ex(hbm_budget_gb, precision, tile_cols, max_batch, drivers)
        if not self.ptr:
            raise RuntimeError("awamd_server_new failed")
        self.L = {}

    def submit(self, trees, seqs, times, popsizes, rho, mu, new_chrom, uniforms):
        times = np.ascontiguousarray(times, np.float64)
        popsizes = np.ascontiguousarray(popsizes, np.float64)
        cs = seqs if isinstance(seqs, CSeqs) else CSeqs(seqs)
        u = np.ascontiguousarray(uniforms, np.float64)
        i = awamd_server_submit(self.ptr, trees, _dptr(times), len(times), _dptr(popsizes), rho, mu,
                                cs.ptr, new_chrom, _dptr(u))
        if i < 0:
            raise RuntimeError("awamd_server_submit failed (unsupported time grid)")
        self.L[i] = cs.seqlen
        return i

    def wait(self, i, timeout=None):
        """Result of request i; raises TimeoutError after `timeout` seconds
        (the request stays pending and can be waited for again)."""
        L = self.L[i]
        path = np.zeros(L, np.int32)
        lnl = c_double(0)
        t = np.zeros(5, np.float64)
        rc = awamd_server_wait_ex(self.ptr, i, float(timeout or 0.0), _iptr(path),
                                  ctypes.byref(lnl), _dptr(t))
        if rc == -2:
            raise TimeoutError("request %d not done after %.1f s" % (i, timeout))
        del self.L[i]
        if rc == 3:
            raise ValueError("request %d refused: over the kernels' limits" % i)
        if rc != 0:
            raise RuntimeError("request %d failed (%d)" % (i, rc))
        return dict(path=path, lnl=lnl.value, t_submit=t[0], t_start=t[1], t_done=t[2],
                    batch_size=int(t[3]), driver=int(t[4]))

    def now(self):
        return awamd_server_now(self.ptr)

    def close(self):
        if getattr(self, "ptr", None):
            awamd_server_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()
