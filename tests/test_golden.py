"""The oracle against golden vectors the compiled reference produced on the
reference's own example data (examples/sim1, config-1 parameters): forward
rows at sampled columns, and the ARG that sample_thread builds under a
fixed glibc rand() seed (forward + traceback + recombinations + add-thread)."""
import numpy as np

import golden_sim1
import oracle_bind as O


def _partial(aw, g):
    tr = aw.new_trees(g["ptrees"], g["ages"], g["sprs"], g["blocklens"])
    aw.remove_thread(tr, int(g["thread_leaf"]))
    P, A, S, bl = aw.trees_arrays(tr)
    maps, seqids = aw.trees_internal(tr)
    inst = O.Instance(g["times"], g["pops"], g["rho"], g["mu"], P, A, S, bl, g["seqs"],
                      seqids=seqids, new_chrom=int(g["thread_leaf"]), mappings=maps)
    return tr, inst


def test_oracle_forward_rows(aw):
    g = golden_sim1.load()
    tr, inst = _partial(aw, g)
    fw, _ = inst.forward()
    offs = inst.block_offsets()
    nst = inst.nstates()
    starts = np.concatenate([[0], np.cumsum(inst.blocklens)])
    for c, row in golden_sim1.rows(g).items():
        b = int(np.searchsorted(starts, c, side="right") - 1)
        S = int(nst[b])
        o = int(offs[b]) + (c - int(starts[b])) * S
        mine = fw[o:o + S]
        assert len(row) == S
        assert np.max(np.abs(mine - row) / np.maximum(np.abs(row), 1e-300)) <= 1e-12, c


def test_oracle_thread_chain(aw):
    g = golden_sim1.load()
    tr, inst = _partial(aw, g)
    L = int(g["length"])
    O.srand(1)
    u = O.draw_uniforms(L)
    fw, _ = inst.forward()
    path = inst.traceback(fw, u)
    pos, node, tm = inst.sample_recombinations(path)
    times = g["times"]
    tr2 = aw.add_thread(tr, len(times), path, int(g["thread_leaf"]), pos, node, tm)
    mine = aw.trees_arrays(tr2)
    for key, arr in zip(("ptrees", "ages", "sprs", "blocklens"), mine):
        assert np.array_equal(arr, g["thread_" + key]), key
