"""The library's drop-in calls on the GPU against golden vectors that the
compiled reference produced on its own example data (examples/sim1): the
forward table at sampled columns, and the ARGs of sample_thread and
resample_arg under fixed glibc rand() seeds.  Needs no reference build."""
import numpy as np
import pytest

import golden_sim1
import oracle_bind as O

pytestmark = pytest.mark.gpu


def _check(aw, tr, g, prefix):
    for key, arr in zip(("ptrees", "ages", "sprs", "blocklens"), aw.trees_arrays(tr)):
        assert np.array_equal(arr, g[prefix + key]), (prefix, key)


def test_forward_alg_golden(aw):
    print("golden sim1 forward", flush=True)
    g = golden_sim1.load()
    tr = aw.new_trees(g["ptrees"], g["ages"], g["sprs"], g["blocklens"])
    aw.remove_thread(tr, int(g["thread_leaf"]))
    rows = aw.forward_algorithm(tr, g["seqs"], g["times"], g["pops"], g["rho"], g["mu"])
    for c, row in golden_sim1.rows(g).items():
        rel = np.abs(rows[c] - row) / np.maximum(np.abs(row), 1e-300)
        assert rel.max() <= 1e-12, (c, rel.max())


def test_sample_thread_golden(aw):
    print("golden sim1 sample_thread", flush=True)
    g = golden_sim1.load()
    tr = aw.new_trees(g["ptrees"], g["ages"], g["sprs"], g["blocklens"])
    aw.remove_thread(tr, int(g["thread_leaf"]))
    O.srand(1)
    aw.sample_thread(tr, g["seqs"], g["times"], g["pops"], g["rho"], g["mu"])
    _check(aw, tr, g, "thread_")


def test_resample_arg_golden(aw):
    print("golden sim1 resample_arg", flush=True)
    g = golden_sim1.load()
    tr = aw.new_trees(g["ptrees"], g["ages"], g["sprs"], g["blocklens"])
    O.srand(2)
    aw.resample_arg(tr, g["seqs"], g["times"], g["pops"], g["rho"], g["mu"], niters=1)
    _check(aw, tr, g, "resample_")
