"""Writes tests/golden/sim1.npz: the reference's own example data
(examples/sim1: sim1.sites and the sampled ARG sim1.sample/out.0.smc.gz, the
8-haplotype, 100 kb, N=1e4, r=1.6e-8, mu=1.8e-8 setting of the README
quick-start) as arrays, plus the outputs of the reference library compiled
from its sources (oracle/_ref) on it:
  fw_cols / fw_rows  forward-table rows (arghmm_forward_alg) at sampled
                     columns, threading leaf 7 back into the 7-leaf ARG
  thread_*           the ARG after arghmm_sample_thread with srand(1)
  resample_*         the ARG after arghmm_resample_arg(niters=1) with srand(2)

Run in a container that has /root/reference and oracle/_ref built:
    python tests/golden/make_sim1.py
The .smc reader follows read_local_trees (src/argweaver/local_tree.cpp:
1430-1545): node names are node indices, ages map to the nearest time point
(find_time, :1019-1034), block lengths are end - start + 1.
"""
import gzip
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
REF = "/root/reference/examples/sim1"
NTIMES, MAXTIME = 20, 200e3
POP, RHO, MU = 1e4, 1.6e-8, 1.8e-8
THREAD_LEAF = 7


def time_points():
    return np.array([(np.exp(i / float(NTIMES - 1) * np.log(1 + .01 * MAXTIME)) - 1) / .01
                     for i in range(NTIMES)])


def find_time(t, times):
    return int(np.argmin(np.abs(times - t)))


_TOKEN = re.compile(r"\(|\)|,|;|:[^,()\[;]*|\[[^\]]*\]|[^,();:\[]+")


def parse_newick(text, times, nnodes):
    """parent and age (time index) per node of one newick tree."""
    parent = np.full(nnodes, -1, np.int32)
    age = np.zeros(nnodes, np.int32)
    stack = [[]]          # children collected per open group
    last = None
    for tok in _TOKEN.findall(text):
        if tok == "(":
            stack.append([])
        elif tok == ",":
            pass
        elif tok == ")":
            last = ("group", stack.pop())
        elif tok == ";":
            break
        elif tok.startswith(":"):
            pass
        elif tok.startswith("["):
            m = re.search(r"age=([-0-9.eE+]+)", tok)
            age[cur] = find_time(float(m.group(1)), times)
        else:
            cur = int(tok)
            if last is not None and last[0] == "group":
                for ch in last[1]:
                    parent[ch] = cur
            last = None
            stack[-1].append(cur)
    return parent, age


def read_smc(path, times):
    names, ptrees, ages, sprs, blocklens = [], [], [], [], []
    spr = [-1, -1, -1, -1]
    for line in gzip.open(path, "rt"):
        f = line.rstrip("\n").split("\t")
        if f[0] == "NAMES":
            names = f[1:]
        elif f[0] == "TREE":
            nnodes = 2 * len(names) - 1
            p, a = parse_newick(f[3], times, nnodes)
            ptrees.append(p)
            ages.append(a)
            sprs.append(list(spr))
            blocklens.append(int(f[2]) - int(f[1]) + 1)
        elif f[0] == "SPR":
            spr = [int(f[2]), find_time(float(f[3]), times), int(f[4]),
                   find_time(float(f[5]), times)]
    return (names, np.array(ptrees, np.int32), np.array(ages, np.int32),
            np.array(sprs, np.int32), np.array(blocklens, np.int32))


def read_sites(path):
    names, pos, cols = [], [], []
    for line in open(path):
        f = line.rstrip("\n").split("\t")
        if f[0] == "NAMES":
            names = f[1:]
        elif f[0] == "REGION":
            length = int(f[3])
        else:
            pos.append(int(f[0]) - 1)
            cols.append(f[1])
    return names, length, np.array(pos, np.int32), cols


def sequences(leaf_names, site_names, length, pos, cols):
    """Rows in leaf order; invariant columns are 'A' (an invariant column's
    emission does not depend on the base)."""
    k = len(leaf_names)
    seqs = np.full((k, length), ord("A"), np.uint8)
    order = [site_names.index(n) for n in leaf_names]
    for p, c in zip(pos, cols):
        for i, j in enumerate(order):
            seqs[i, p] = ord(c[j])
    return seqs


def main():
    import oracle_bind as O   # noqa: E402  (tests/ on the path below)
    times = time_points()
    leaf_names, P, A, S, bl = read_smc(os.path.join(REF, "sim1.sample", "out.0.smc.gz"), times)
    site_names, length, pos, cols = read_sites(os.path.join(REF, "sim1.sites"))
    assert bl.sum() == length
    seqs = sequences(leaf_names, site_names, length, pos, cols)
    pops = np.full(NTIMES, POP)
    out = dict(times=times, ptrees=P, ages=A, sprs=S, blocklens=bl, site_pos=pos,
               site_bases=np.array([[ord(ch) for ch in c] for c in cols], np.uint8),
               site_order=np.array([site_names.index(n) for n in leaf_names], np.int32),
               length=np.int32(length), thread_leaf=np.int32(THREAD_LEAF),
               model=np.array([POP, RHO, MU]))

    # forward rows of the 7-leaf ARG
    rt = O.ref_new_trees(P, A, S, bl)
    O.ref_remove_thread(rt, THREAD_LEAF)
    Pr, Ar, Sr, blr = O.ref_trees_arrays(rt)
    nst = []
    for b in range(len(blr)):
        par, ag = Pr[b], Ar[b]
        n = 0
        for i in range(len(par)):
            top = NTIMES - 2 if par[i] == -1 else ag[par[i]]
            n += top - ag[i] + 1
        nst += [n] * int(blr[b])
    rows = O.ref_forward(rt, seqs, times, pops, RHO, MU, nst)
    # columns 0, every 4999th, the last, and both sides of the first 16 tree switches
    sw = np.cumsum(blr)[:16]
    cols_s = np.unique(np.concatenate([np.arange(0, length, 4999), [length - 1], sw, sw - 1]))
    out["fw_cols"] = cols_s.astype(np.int32)
    out["fw_rows"] = np.concatenate([rows[c] for c in cols_s])
    out["fw_rowlen"] = np.array([len(rows[c]) for c in cols_s], np.int32)

    # sample_thread of leaf 7 into the 7-leaf ARG
    O.srand(1)
    O.ref_sample_thread(rt, seqs, times, pops, RHO, MU)
    for key, arr in zip(("ptrees", "ages", "sprs", "blocklens"), O.ref_trees_arrays(rt)):
        out["thread_" + key] = arr
    O.ref_delete(rt)

    # one resample_arg sweep over the full ARG
    rt = O.ref_new_trees(P, A, S, bl)
    O.srand(2)
    O.ref_resample_arg(rt, seqs, times, pops, RHO, MU, 1)
    for key, arr in zip(("ptrees", "ages", "sprs", "blocklens"), O.ref_trees_arrays(rt)):
        out["resample_" + key] = arr
    O.ref_delete(rt)

    np.savez_compressed(os.path.join(HERE, "sim1.npz"), **out)
    print("wrote sim1.npz: %d trees, %d sites, %d forward rows"
          % (len(bl), len(pos), len(cols_s)))


if __name__ == "__main__":
    sys.path.insert(0, os.path.join(os.path.dirname(HERE)))
    main()
