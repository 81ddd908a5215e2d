"""The reference's own transition fixtures (tests/golden/trans_args.json.gz:
the ARGs of test/data/test_trans and test/data/test_trans_switch that
test/test_hmm.py:117-190 checks), run the way that test runs them:

- each ARG's ages rounded to the time grid (discretize_arg,
  argweaver/__init__.py:210-256: the closer of the two neighbouring points,
  the last point left out);
- test_trans: the marginal tree at position 10 (get_marginal_tree,
  arglib.py:578-620) through arghmm_assert_transmat, which checks every
  entry of the within-block matrix against the sum over the SPRs that lead
  from state i to state j; the matrix itself entry by entry against the
  compiled reference's (new_transition_probs);
- test_trans_switch: the tree just before the first recombination and the
  SPR it carries through arghmm_assert_transmat_switch, which checks every
  nonzero entry of the switch matrix against the probability of the SPR
  it implies, in both libraries.
"""
import gzip
import json
import os
from collections import defaultdict
from ctypes import POINTER, c_bool, c_double, c_int

import numpy as np
import pytest

import oracle_bind as O

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
c_int_p = POINTER(c_int)
c_double_p = POINTER(c_double)
c_double_p_p = POINTER(c_double_p)
intstate_p = POINTER(c_int * 2)


def _fixtures():
    with gzip.open(os.path.join(HERE, "golden", "trans_args.json.gz"), "rt") as f:
        return json.load(f)


def _discretize(arg, times):
    grid = list(times[:-1])
    nodes = {n["name"]: dict(n) for n in arg["nodes"]}
    for n in nodes.values():
        a = n["age"]
        if a <= grid[0]:
            n["t"] = 0
        elif a >= grid[-1]:
            n["t"] = len(grid) - 1
        else:
            j = int(np.searchsorted(grid, a))   # grid[j-1] < a <= grid[j]
            if grid[j] == a:
                n["t"] = j
            else:
                i = j - 1
                n["t"] = i if a - grid[i] < grid[j] - a else j
    # recombination positions to increasing integers
    last = 0
    for n in sorted((n for n in nodes.values() if n["event"] == "recomb"), key=lambda x: x["pos"]):
        p = int(n["pos"])
        n["pos"] = p if p > last else last + 1
        last = n["pos"]
    return nodes


def _local_parent(nodes, name, pos):
    n = nodes[name]
    if n["event"] == "recomb":
        return n["parents"][0] if pos < n["pos"] else n["parents"][1]
    return n["parents"][0] if n["parents"] else None


def _marginal(nodes, pos):
    """The marginal tree at pos: its nodes (leaves and nodes with two
    marginal children), each node's tree parent, and the ARG path up."""
    leaves = sorted((k for k, n in nodes.items() if n["event"] == "gene"), key=lambda x: int(x[1:]))
    up = {}
    for leaf in leaves:
        x = leaf
        while True:
            p = _local_parent(nodes, x, pos)
            if p is None or x in up:
                break
            up[x] = p
            x = p
    kids = defaultdict(list)
    for c, p in up.items():
        kids[p].append(c)
    counts = {}

    def nleaves(x):
        if x not in counts:
            counts[x] = 1 if nodes[x]["event"] == "gene" else sum(nleaves(c) for c in kids[x])
        return counts[x]

    tree = [x for x in up.keys() | set(kids) if nodes[x]["event"] == "gene" or len(kids[x]) == 2]
    # the lowest node holding every leaf
    root = min((x for x in tree if nleaves(x) == len(leaves)), key=lambda x: nodes[x]["age"])
    tree = [x for x in tree if x == root or _below(up, x, root)]
    tparent = {}
    for x in tree:
        if x == root:
            tparent[x] = None
            continue
        p = up[x]
        while len(kids[p]) != 2:
            p = up[p]
        tparent[x] = p
    return tree, tparent, root, up, kids


def _below(up, x, root):
    while x in up:
        x = up[x]
        if x == root:
            return True
    return False


def _ptree(nodes, tree, tparent, root):
    """make_ptree (argweaverc.py:1229-1274): nodes in postorder, leaves
    first; parent indices and time indices."""
    children = defaultdict(list)
    for x, p in tparent.items():
        if p is not None:
            children[p].append(x)
    order = []

    def post(x):
        for c in sorted(children[x]):
            post(c)
        order.append(x)

    post(root)
    order = [x for x in order if nodes[x]["event"] == "gene"] + \
            [x for x in order if nodes[x]["event"] != "gene"]
    idx = {x: i for i, x in enumerate(order)}
    ptree = np.array([idx[tparent[x]] if tparent[x] is not None else -1 for x in order], np.int32)
    ages = np.array([nodes[x]["t"] for x in order], np.int32)
    return ptree, ages, idx


def _model(ntimes, aw):
    times = np.array(aw.get_time_points(ntimes, 200000), np.float64)
    return times, np.full(ntimes, 1e4), 1.5e-8 * 20


def _ref_fn(name, restype, argtypes):
    fn = O.ref()[name]
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


def _i(a):
    return a.ctypes.data_as(c_int_p)


def _d(a):
    return a.ctypes.data_as(c_double_p)


@needs_ref
def test_reference_trans_fixtures(aw):
    """test_hmm.py:117-147 on the 40 committed ARGs: the brute-force check
    passes in both libraries, and the matrices agree entry by entry."""
    times, pops, rho = _model(10, aw)
    chk = _ref_fn("arghmm_assert_transmat", c_bool,
                  [c_int, c_int_p, c_int_p, c_int, c_double_p, c_double_p, c_double])
    mat = _ref_fn("new_transition_probs", c_double_p_p,
                  [c_int, c_int_p, c_int_p, c_double, intstate_p, c_int, c_int, c_double_p,
                   c_double_p, c_int_p, c_int_p, c_int_p, c_double_p, c_double])
    z = np.zeros(10, np.int32)
    for i, arg in enumerate(_fixtures()["test_trans"]):
        nodes = _discretize(arg, times)
        tree, tparent, root, _, _ = _marginal(nodes, 10)
        p, a, _ = _ptree(nodes, tree, tparent, root)
        assert len(p) == 15, i
        assert chk(len(p), _i(p), _i(a), 10, _d(times), _d(pops), rho), i
        assert aw.assert_transmat(p, a, times, pops, rho), i
        st = aw.tree_states(p, a, 10)
        mine = aw.transition_probs(p, a, times, pops, rho)
        m = mat(len(p), _i(p), _i(a), 0.0, st.ctypes.data_as(intstate_p), len(st), 10, _d(times),
                _d(times), _i(z), _i(z), _i(z), _d(pops), rho)
        ref = np.array([np.ctypeslib.as_array(m[r], shape=(len(st),)) for r in range(len(st))])
        O.ref().delete_transition_probs(m, len(st))
        fin = np.isfinite(ref)
        assert np.array_equal(fin, np.isfinite(mine)), i
        assert np.allclose(mine[fin], ref[fin], rtol=1e-12, atol=0), i


def _first_spr(nodes):
    """The tree just before the first recombination and its SPR: the branch
    the recombination sits on (walking down from the recombination node to
    the first node of the tree), the recombination time, the branch the
    detached lineage joins (walking down from the coalescence node's other
    child) and the coalescence time.  None for an SPR this walk does not
    resolve (a lineage that re-joins its own branch)."""
    rec = min((n for n in nodes.values() if n["event"] == "recomb"), key=lambda n: n["pos"])
    pos = rec["pos"]
    tree, tparent, root, up, kids = _marginal(nodes, pos - 0.5)
    tset = set(tree)
    # recombination branch
    x = rec["children"][0]
    while x not in tset:
        ch = [c for c in kids[x]]
        if len(ch) != 1:
            return None
        x = ch[0]
    r = x
    # the detached lineage after the recombination, up to where it meets
    # the rest of the tree
    marg = set(up) | set(kids)
    prev, y = rec["name"], rec["parents"][1]
    while y not in marg:
        prev, y = y, _local_parent(nodes, y, pos + 0.5)
        if y is None:
            return None
    others = [c for c in nodes[y]["children"] if c != prev]
    if len(others) != 1:
        return None
    x = others[0]
    while x not in tset:
        ch = kids[x]
        if len(ch) != 1:
            return None
        x = ch[0]
    c = x
    if c == r:
        return None
    ct, rt = nodes[y]["t"], rec["t"]
    cpar = tparent[c]
    if ct < rt or ct < nodes[c]["t"] or (cpar is not None and ct > nodes[cpar]["t"]):
        return None
    return tree, tparent, root, r, rt, c, ct


@needs_ref
def test_reference_trans_switch_fixtures(aw):
    """test_hmm.py:150-190 on the 100 committed ARGs: the switch matrix of
    each ARG's first recombination passes the brute-force check in both
    libraries."""
    times, pops, rho = _model(20, aw)
    chk = _ref_fn("arghmm_assert_transmat_switch", c_bool,
                  [c_int, c_int_p, c_int_p, c_int, c_int, c_int, c_int, c_int, c_double_p,
                   c_double_p, c_double])
    checked = 0
    for i, arg in enumerate(_fixtures()["test_trans_switch"]):
        nodes = _discretize(arg, times)
        got = _first_spr(nodes)
        if got is None:
            continue
        tree, tparent, root, r, rt, c, ct = got
        p, a, idx = _ptree(nodes, tree, tparent, root)
        spr = np.array([idx[r], rt, idx[c], ct], np.int32)
        ref_ok = chk(len(p), _i(p), _i(a), int(spr[0]), int(spr[1]), int(spr[2]), int(spr[3]), 20,
                     _d(times), _d(pops), rho)
        assert ref_ok, i
        assert aw.assert_transmat_switch(p, a, spr, times, pops, rho), i
        checked += 1
    assert checked >= 80, checked
