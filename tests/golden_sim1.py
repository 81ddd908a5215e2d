"""Loader of tests/golden/sim1.npz (written by tests/golden/make_sim1.py
from the reference's examples/sim1 data and its compiled library)."""
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sim1.npz")


def load():
    g = dict(np.load(PATH, allow_pickle=False))
    L = int(g["length"])
    k = g["ptrees"].shape[1] // 2 + 1
    seqs = np.full((k, L), ord("A"), np.uint8)
    for p, row in zip(g["site_pos"], g["site_bases"]):
        seqs[:, p] = row[g["site_order"]]
    g["seqs"] = seqs
    pop, rho, mu = g["model"]
    g["pops"] = np.full(len(g["times"]), pop)
    g["rho"], g["mu"] = float(rho), float(mu)
    return g


def rows(g):
    """{column: golden forward row}"""
    out, o = {}, 0
    for c, n in zip(g["fw_cols"], g["fw_rowlen"]):
        out[int(c)] = g["fw_rows"][o:o + n]
        o += n
    return out
