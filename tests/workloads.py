"""Seeded threading instances shared by the tests (product simulator + leaf removal)."""
import numpy as np

import oracle_bind as O


def model(ntimes=20, maxtime=200e3, popsize=1e4, c=10, rho=1.6e-8, mu=1.8e-8):
    times = O.time_points(ntimes, maxtime)
    pops = np.full(ntimes, popsize)
    return times, pops, rho * c, mu * c


def instance(A, k, L, seed, ntimes=20, popsizes=None, remove=None, c=10):
    """Simulate k haplotypes, remove one leaf; returns (trees, seqs, Instance, model)."""
    times, pops, rho, mu = model(ntimes, c=c)
    if popsizes is not None:
        pops = np.asarray(popsizes, np.float64)
    tr, seqs = A.simulate(k, L, times, pops, rho, mu, seed)
    rm = k - 1 if remove is None else remove
    A.remove_thread(tr, rm)
    P, Ag, S, bl = A.trees_arrays(tr)
    maps, seqids = A.trees_internal(tr)
    inst = O.Instance(times, pops, rho, mu, P, Ag, S, bl, seqs, seqids=seqids,
                      new_chrom=rm, mappings=maps)
    return tr, seqs, inst, (times, pops, rho, mu)
