"""Threading-HMM throughput on one MI355X.

One step = one pass of the threading path (column classification, per-block
transition/switch/emission preparation, variant-column emissions, forward
and stochastic traceback) over a batch of independent threading instances
whose inputs (ARG blocks, sequences, uniforms) are already resident in HBM.

Workload (default "batched"): 256 instances, k ~ U{8..64} haplotypes,
L log-uniform on [50 kb, 2 Mb] compressed 10 bp -> 5k..200k columns,
ntimes=20, maxtime=200e3, N=1e4, r=1.6e-8, mu=1.8e-8 (rho, mu x 10 per
column); seeded DSMC ARGs with one leaf removed per instance.

value: column-states per second (W = sum over blocks blocklen * nstates)
of the FP32 column-tiled path (normalised FP32 table, FP64 normalisers,
log-likelihood and traceback CDF).  The plain FP64 path (the drop-in
arithmetic) is timed on the same batch and reported beside it.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK = 8.0e12   # B/s, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(A, n, seed, kmin=8, kmax=64, lmin=5000, lmax=200000):
    rng = np.random.RandomState(seed)
    times = np.array(A.get_time_points(20, 200e3), np.float64)
    pops = np.full(20, 1e4)
    rho, mu = 1.6e-8 * 10, 1.8e-8 * 10
    insts = []
    for i in range(n):
        k = int(rng.randint(kmin, kmax + 1))
        L = int(np.exp(rng.uniform(np.log(lmin), np.log(lmax))))
        tr, seqs = A.simulate(k, L, times, pops, rho, mu, seed * 100003 + i)
        full = A.trees_arrays(tr)   # the ARG before removal, for the CPU baseline
        A.remove_thread(tr, k - 1)
        u = rng.uniform(size=L)
        W, _ = A.num_column_states(tr, 20)
        insts.append(dict(k=k, L=L, trees=tr, full=full, seqs=seqs, u=u, W=W))
    return insts, (times, pops, rho, mu)


def _cpu_order(insts, seed):
    """Instances in a W-weighted random order (without replacement): the
    sample follows where the workload's column-states are."""
    W = np.array([float(it["W"]) for it in insts])
    rng = np.random.RandomState(seed + 5)
    return list(rng.choice(len(insts), size=len(insts), replace=False, p=W / W.sum()))


def _port_time(it, model):
    """Oracle forward + stochastic traceback of one instance on one core:
    (column-states, forward seconds, traceback seconds)."""
    import oracle_bind as O
    import argweaver_amd.argweaverc as A
    times, pops, rho, mu = model
    P, Ag, S, bl = A.trees_arrays(it["trees"])
    maps, seqids = A.trees_internal(it["trees"])
    inst = O.Instance(times, pops, rho, mu, P, Ag, S, bl, it["seqs"], seqids=seqids,
                      new_chrom=it["k"] - 1, mappings=maps)
    t0 = time.perf_counter()
    fw, _ = inst.forward()
    t1 = time.perf_counter()
    inst.traceback(fw, it["u"])
    t2 = time.perf_counter()
    return it["W"], t1 - t0, t2 - t1


def _cpu_worker(args):
    """One core's share of the aggregate: instances until `budget` seconds
    of forward + traceback."""
    insts, model, idx, budget = args
    W = 0
    t = 0.0
    for i in idx:
        w, f, b = _port_time(insts[i], model)
        W += w
        t += f + b
        if t >= budget:
            break
    return W, t


def cpu_baseline(A, insts, model, seed, budget_s=10.0, cores=16):
    """The forward and stochastic traceback of the workload on the host
    cores, on a W-weighted random sample: the oracle's C restatement (the
    reference's algorithm, forward + traceback only) on one core, and the
    aggregate of `cores` processes each running its own share in parallel.
    The compiled reference's forward (arghmm_forward_alg, oracle/_ref) is
    timed on the same instances, when built, to show the restatement's
    speed against it.  Runs before the process touches the GPU (the pool
    is forked)."""
    import multiprocessing as mp
    import oracle_bind as O
    order = _cpu_order(insts, seed)
    # one core
    W = 0
    tf = tb = 0.0
    n = 0
    for i in order:
        w, f, b = _port_time(insts[i], model)
        W += w
        tf += f
        tb += b
        n += 1
        if tf + tb >= budget_s:
            break
    single = W / (tf + tb)
    out = dict(value=single, unit="column-states/s", cores=1, kind="port",
               sample="%d instances drawn W-weighted from the workload (W=%.3g column-states, "
                      "%.1f s): oracle forward %.1f s + traceback %.1f s, one core"
               % (n, W, tf + tb, tf, tb))
    # the compiled reference's forward on the same instances
    if O.ref_available():
        times, pops, rho, mu = model
        fn = O.ref()["arghmm_forward_alg"]
        fn.restype = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
        Wr, tr_ = 0, 0.0
        for i in order[:n]:
            it = insts[i]
            rt = O.ref_new_trees(*it["full"])
            O.ref_remove_thread(rt, it["k"] - 1)
            sq = O._Seqs(it["seqs"])
            dp = ctypes.POINTER(ctypes.c_double)
            t0 = time.perf_counter()
            m = fn(ctypes.c_void_p(rt), times.ctypes.data_as(dp), len(times),
                   pops.ctypes.data_as(dp), ctypes.c_double(rho), ctypes.c_double(mu), sq.ptr,
                   sq.arr.shape[0], sq.arr.shape[1], ctypes.c_bool(False), None,
                   ctypes.c_bool(False), ctypes.c_bool(False))
            tr_ += time.perf_counter() - t0
            O.ref()["delete_forward_matrix"](m, sq.arr.shape[1])
            O.ref_delete(rt)
            Wr += it["W"]
        out["reference_forward_only"] = dict(value=Wr / tr_, seconds=tr_,
                                             port_forward_only=W / tf)
    # the host's cores in parallel, each on its own share of the order
    P = max(1, min(cores, os.cpu_count() or 1))
    shares = [order[c::P] for c in range(P)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(P) as pool:
        res = pool.map(_cpu_worker, [(insts, model, sh, budget_s) for sh in shares])
    wall = time.perf_counter() - t0
    out["aggregate"] = dict(value=sum(r[0] for r in res) / wall, cores=P, wall_s=wall,
                            column_states=float(sum(r[0] for r in res)))
    return out


def read_traffic(path):
    """HBM bytes per column-state of the FP32 forward from the committed
    rocprofv3 counter summary (the line 'bytes_per_column_state: X'), or
    None when it is absent."""
    try:
        for line in open(path):
            if line.strip().startswith("bytes_per_column_state:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def plain_fp64(A, insts, model, budget_s, log):
    """The drop-in path: one instance per launch in FP64 (what one
    arghmm_sample_thread call runs on the GPU), inputs resident, timed
    instance after instance on a bounded sample of the workload."""
    times, pops, rho, mu = model
    rng = np.random.RandomState(7)
    idx = rng.permutation(len(insts))
    W = 0
    t_used = 0.0
    n = 0
    for i in idx:
        it = insts[i]
        b = A.Batch()
        b.add(it["trees"], it["seqs"], times, pops, rho, mu, it["k"] - 1, it["u"])
        b.upload()
        # the traceback tiling of a drop-in call (run_instance): 4096-column
        # tiles for instances of 16384 columns or more
        tile = 4096 if it["L"] >= 16384 else 0
        b.run(0, tile)   # warm (allocations)
        t0 = time.perf_counter()
        b.run(0, tile)
        dt = time.perf_counter() - t0
        W += b.W(0)
        t_used += dt
        n += 1
        del b
        if t_used >= budget_s:
            break
    log("plain fp64: %d instances one per launch, %.3g column-states in %.2f s" % (n, W, t_used))
    return dict(value=W / t_used, unit="column-states/s", instances=n, column_states=W,
                seconds=t_used)


def make_serving_pool(A, seed, n=48):
    """Requests of the config-2 size mix: k ~ U{8..50} haplotypes (one
    removed), L log-uniform on 5k..100k columns (50 kb .. 1 Mb at c=10),
    the batched workload's model."""
    return make_workload(A, n, seed + 1000, kmin=8, kmax=50, lmin=5000, lmax=100000)


def serving(A, insts, model, rate, nreq, seed, log, budget_gb=96.0, tile=2000, max_batch=64,
            precision=1):
    """Poisson arrivals at `rate` requests/s of instances drawn from the
    pool, served with continuous batching (two drivers, FP32 or FP64);
    per-request latency = finish - arrival on the server clock."""
    import threading
    rng = np.random.RandomState(seed)
    picks = rng.randint(0, len(insts), size=nreq)
    gaps = rng.exponential(1.0 / rate, size=nreq)
    srv = A.Server(hbm_budget_gb=budget_gb, precision=precision, tile_cols=tile, max_batch=max_batch)
    ids = [None] * nreq

    def producer():
        t_next = srv.now()
        for j, i in enumerate(picks):
            t_next += gaps[j]
            dt = t_next - srv.now()
            if dt > 0:
                time.sleep(dt)
            it = insts[i]
            ids[j] = srv.submit(it["trees"], it["seqs"], *model, it["k"] - 1, it["u"])

    t0 = srv.now()
    th = threading.Thread(target=producer)
    th.start()
    th.join()
    res = []
    for j, i in enumerate(ids):
        res.append(srv.wait(i, timeout=120))
        if (j + 1) % 64 == 0:
            log("serving %.0f req/s (%s): %d/%d done at %.1f s"
                % (rate, "fp32" if precision else "fp64", j + 1, nreq, srv.now() - t0))
    wall = srv.now() - t0
    srv.close()
    lat = np.array([r["t_done"] - r["t_submit"] for r in res]) * 1e3
    nb = len(set((r["driver"], r["t_start"]) for r in res))
    # requests that ran beside a batch of the other driver
    spans = [(r["t_start"], r["t_done"], r["driver"]) for r in res]
    overlapped = sum(1 for a in spans if any(b[2] != a[2] and b[0] < a[1] and a[0] < b[1]
                                             for b in spans))
    W = float(sum(insts[i]["W"] for i in picks))
    out = dict(rate=rate, requests=nreq, p50_ms=float(np.percentile(lat, 50)),
               p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
               throughput=W / wall, unit="column-states/s", wall_s=wall, batches=nb,
               max_batch=max_batch, precision="fp32" if precision else "fp64",
               overlapped_requests=overlapped,
               mean_batch=float(np.mean([r["batch_size"] for r in res])))
    log("serving %.0f req/s (%s): p50 %.1f ms, p99 %.1f ms, %.3g column-states/s, %d batches, "
        "%d of %d requests beside the other driver's batch"
        % (rate, out["precision"], out["p50_ms"], out["p99_ms"], out["throughput"], nb,
           overlapped, nreq))
    return out


def mcmc_model(seed):
    """Config-5 model: 20 time points, N(t) piecewise constant over the 20
    intervals with values log-uniform in 5e3..5e4 (seeded)."""
    rng = np.random.RandomState(seed + 11)
    pops = np.exp(rng.uniform(np.log(5e3), np.log(5e4), size=20))
    return pops, 1.6e-8 * 10, 1.8e-8 * 10


def mcmc_rethread(A, seed, log, k=100, L=500000, nrethreads=2):
    """Config-5 sample: the MCMC step of resample_arg on a k=100, 5 Mb
    (500k columns at c=10) ARG under a piecewise-constant N(t): remove a
    leaf, re-thread it through the drop-in arghmm_sample_thread (forward,
    traceback, recombinations and ARG surgery).  Times nrethreads steps and
    records the number of local trees (recombinations + 1) after each."""
    times = np.array(A.get_time_points(20, 200e3), np.float64)
    pops, rho, mu = mcmc_model(seed)
    t0 = time.time()
    tr, seqs = A.simulate(k, L, times, pops, rho, mu, seed)
    log("mcmc workload: k=%d L=%d simulated in %.1f s" % (k, L, time.time() - t0))
    A.set_random_seed(seed)
    out = []
    for j in range(nrethreads):
        # arghmm_sample_thread threads sequence nseqs-1: remove its leaf and
        # thread it back (one Gibbs step of resample_arg)
        t0 = time.perf_counter()
        A.remove_thread(tr, k - 1)
        W, _ = A.num_column_states(tr, 20)
        A.sample_thread(tr, seqs, times, pops, rho, mu)
        dt = time.perf_counter() - t0
        # the chain's trace statistics (arghmm_likelihood / arghmm_prior_prob,
        # outside the timed step)
        lk, pr = A.arg_likelihood(tr, seqs, times, mu), A.arg_prior(tr, times, pops, rho)
        out.append((dt, W, int(A.get_local_trees_ntrees(tr)), lk, pr))
        log("mcmc re-thread %d: W=%.3g, %.2f s, %d local trees, log P(seqs | ARG) %.8g, "
            "log P(ARG) %.8g" % (j, W, dt, out[-1][2], lk, pr))
    # the first call includes one-time allocations
    dt, W = out[-1][0], out[-1][1]
    return dict(k=k, columns=L, popsizes="piecewise constant, 20 intervals, log-uniform 5e3..5e4",
                rethread_s=dt, column_states=W, value=W / dt, unit="column-states/s",
                local_trees_trace=[o[2] for o in out], log_likelihood_trace=[o[3] for o in out],
                log_prior_trace=[o[4] for o in out])


def mcmc_rethread_fp32(A, seed, log, k=100, L=500000, nsteps=3):
    """Config-5 in FP32: re-thread leaves k-1, k-2, .. in turn (remove, FP32
    forward with forward tiles and traceback, recombinations, add), with the
    log-likelihood and local-tree traces of the chain."""
    times = np.array(A.get_time_points(20, 200e3), np.float64)
    pops, rho, mu = mcmc_model(seed)
    tr, seqs = A.simulate(k, L, times, pops, rho, mu, seed)
    rng = np.random.RandomState(seed + 21)
    A.set_random_seed(seed + 22)
    steps = []
    err = None
    for j in range(nsteps):
        leaf = k - 1 - j
        t0 = time.perf_counter()
        A.remove_thread(tr, leaf)
        t_rm = time.perf_counter() - t0
        uj = rng.uniform(size=L)
        if j == 0:
            # the first step's FP32 forward against FP64 (outside the timing)
            b = A.Batch()
            b.add(tr, seqs, times, pops, rho, mu, leaf, uj)
            b.upload()
            nst = A.get_nstates(tr, 20, L)
            b.run(0, 4096, traceback=False)
            l64, w64 = b.lnl(0), alpha_windows(A, b, 0, L, nst, nwin=8)
            b.run(1, 4096, traceback=False, fw_tile=50000, fw_warm=20000, fw_tol=1e-4)
            l32, w32 = b.lnl(0), alpha_windows(A, b, 0, L, nst, nwin=8)
            ea = er = 0.0
            for (c0, c1, a64), (_, _, a32) in zip(w64, w32):
                x, y = window_errors(a64, a32, nst[c0:c1])
                ea, er = max(ea, x), max(er, y)
            err = dict(rel_dlogL=abs(l32 - l64) / abs(l64), alpha_max_abs_err_normalised=ea,
                       alpha_max_rel_err_large_entries=er,
                       alpha_sample="first step, 8 windows of 256 columns")
            log("mcmc fp32 step 0 vs fp64: |dlogL|/|logL| %.2e, alpha err abs %.2e rel %.2e"
                % (err["rel_dlogL"], ea, er))
            del b
            t0 = time.perf_counter() - t_rm
        t1 = time.perf_counter()
        r = A.thread_forward(tr, seqs, times, pops, rho, mu, leaf, uj,
                             precision=1, want_alpha=False)
        t2 = time.perf_counter()
        pos, node, tm = A.sample_recombinations(tr, times, pops, rho, mu, r["path"])
        t3 = time.perf_counter()
        A.add_thread(tr, len(times), r["path"], leaf, pos, node, tm)
        t4 = time.perf_counter()
        dt = t4 - t0
        steps.append(dict(leaf=leaf, seconds=dt, ms_remove=t_rm * 1e3,
                          ms_thread_forward=(t2 - t1) * 1e3, ms_recombinations=(t3 - t2) * 1e3,
                          ms_add=(t4 - t3) * 1e3, lnl=r["lnl"], recombinations=int(len(pos)),
                          local_trees=int(A.get_local_trees_ntrees(tr)),
                          log_likelihood=A.arg_likelihood(tr, seqs, times, mu),
                          log_prior=A.arg_prior(tr, times, pops, rho)))
        log("mcmc fp32 re-thread leaf %d: %.2f s (remove %.1f ms, thread_forward %.1f ms, "
            "recombinations %.1f ms, add %.1f ms), thread logL %.6g, %d recombinations, %d local "
            "trees, log P(seqs | ARG) %.8g, log P(ARG) %.8g"
            % (leaf, dt, t_rm * 1e3, (t2 - t1) * 1e3, (t3 - t2) * 1e3, (t4 - t3) * 1e3, r["lnl"],
               len(pos), steps[-1]["local_trees"], steps[-1]["log_likelihood"],
               steps[-1]["log_prior"]))
    # the chain statistic log P(seqs | ARG) on the host (the reference's
    # order of terms) and on the device (awamd_arg_likelihood)
    t0 = time.perf_counter()
    lh = A.arg_likelihood(tr, seqs, times, mu)
    t1 = time.perf_counter()
    A.arg_likelihood_device(tr, seqs, times, mu)   # first call: module load
    t2 = time.perf_counter()
    ld = A.arg_likelihood_device(tr, seqs, times, mu)
    t3 = time.perf_counter()
    t4 = time.perf_counter()
    ph = A.arg_prior(tr, times, pops, rho)
    t5 = time.perf_counter()
    pd = A.arg_prior_device(tr, times, pops, rho)
    t6 = time.perf_counter()
    lik = dict(host_s=t1 - t0, device_s=t3 - t2, device_first_call_s=t2 - t1,
               rel_diff=abs(ld - lh) / abs(lh), prior_host_s=t5 - t4, prior_device_s=t6 - t5,
               prior_rel_diff=abs(pd - ph) / abs(ph))
    log("log P(seqs | ARG) of the config-5 ARG: host %.3f s, device %.3f s (first call %.3f s), "
        "relative difference %.1e; log P(ARG): host %.3f s, device %.3f s, relative difference "
        "%.1e" % (lik["host_s"], lik["device_s"], lik["device_first_call_s"], lik["rel_diff"],
                  lik["prior_host_s"], lik["prior_device_s"], lik["prior_rel_diff"]))
    return dict(steps=steps, seconds_per_step=float(np.mean([s["seconds"] for s in steps[1:]])),
                fp32_error=err, arg_likelihood=lik)


def mcmc_batched_chains(A, seed, log, nchains=4, k=100, L=500000, steps=100):
    """Config 5 as batched FP32 chains: nchains independent chains, each a
    k=100, 500k-column ARG simulated from its own seed under the same N(t).
    Every step re-threads one leaf of every chain, the chains' instances in
    one batch (reused across steps): remove the leaf (host), FP32 forward
    with forward tiles and tiled traceback for all chains (GPU), then the
    recombinations and the insertion per chain (host).  The leaves cycle
    k-1, k-2, ..., so steps = k is one sweep re-threading every leaf of
    every chain.  Reports the time per step, the aggregate column-states/s,
    each chain's recombination count (local trees - 1), log P(seqs | ARG) and
    log P(ARG) (computed on the device, outside the step's time) after every
    step, and
    its log P(seqs | ARG) and log P(ARG) on the host before and after."""
    times = np.array(A.get_time_points(20, 200e3), np.float64)
    pops, rho, mu = mcmc_model(seed)
    chains = [A.simulate(k, L, times, pops, rho, mu, seed + 100 + c) for c in range(nchains)]
    rng = np.random.RandomState(seed + 31)
    A.set_random_seed(seed + 32)
    stats0 = [(A.arg_likelihood(tr, sq, times, mu), A.arg_prior(tr, times, pops, rho))
              for tr, sq in chains]
    b = A.Batch()
    secs, Ws, trace, host_ms, gpu_ms, lik_trace, prior_trace = [], [], [], [], [], [], []
    # the chains' host work runs on one thread per chain (the library's
    # calls release the interpreter lock); the recombination draws of all
    # chains come from the one process-global rand() stream, so they run in
    # chain order
    from concurrent.futures import ThreadPoolExecutor
    pool = ThreadPoolExecutor(max_workers=nchains)
    for j in range(steps):
        leaf = k - 1 - (j % k)
        t0 = time.perf_counter()
        list(pool.map(lambda c: A.remove_thread(c[0], leaf), chains))
        b.clear()
        for tr, sq in chains:
            b.add(tr, sq, times, pops, rho, mu, leaf, rng.uniform(size=L))
        b.upload()
        t1 = time.perf_counter()
        st = b.run(1, 4096, fw_tile=50000, fw_warm=20000, fw_tol=1e-4)
        t2 = time.perf_counter()
        paths = [b.path(c) for c in range(nchains)]
        recs = [A.sample_recombinations(tr, times, pops, rho, mu, paths[c])
                for c, (tr, _) in enumerate(chains)]
        list(pool.map(lambda c: A.add_thread(chains[c][0], len(times), paths[c], leaf, *recs[c]),
                      range(nchains)))
        t3 = time.perf_counter()
        secs.append(t3 - t0)
        Ws.append(st["W"])
        host_ms.append(((t1 - t0) + (t3 - t2)) * 1e3)
        gpu_ms.append((t2 - t1) * 1e3)
        trace.append([int(A.get_local_trees_ntrees(tr)) - 1 for tr, _ in chains])
        # log P(seqs | ARG) of every chain after the step, on the device
        # (outside the step's time)
        lik_trace.append([A.arg_likelihood_device(tr, sq, times, mu) for tr, sq in chains])
        prior_trace.append([A.arg_prior_device(tr, times, pops, rho) for tr, _ in chains])
        if j % 20 == 19 or j == steps - 1:
            log("mcmc batched chains: step %d/%d, %.2f s per step (host %.0f ms, batch run %.0f "
                "ms), recombinations %s" % (j + 1, steps, float(np.mean(secs[-20:])),
                                            float(np.mean(host_ms[-20:])),
                                            float(np.mean(gpu_ms[-20:])), trace[-1]))
    pool.shutdown()
    stats1 = [(A.arg_likelihood(tr, sq, times, mu), A.arg_prior(tr, times, pops, rho))
              for tr, sq in chains]
    for tr, _ in chains:
        A.delete_local_trees(tr)
    sps = float(np.mean(secs[1:])) if steps > 1 else secs[0]
    out = dict(chains=nchains, k=k, columns=L, steps=steps,
               leaves_rethreaded_per_chain=steps, seconds_per_step=sps,
               host_ms_per_step=float(np.mean(host_ms[1:] or host_ms)),
               batch_run_ms_per_step=float(np.mean(gpu_ms[1:] or gpu_ms)),
               column_states_per_s=float(np.mean(Ws[1:] or Ws)) / sps,
               recombinations_trace=trace, log_likelihood_trace=lik_trace,
               log_prior_trace=prior_trace,
               log_likelihood_device_vs_host_end=[
                   abs(lik_trace[-1][c] - stats1[c][0]) / abs(stats1[c][0]) for c in range(nchains)],
               log_likelihood_start=[x[0] for x in stats0], log_likelihood_end=[x[0] for x in stats1],
               log_prior_start=[x[1] for x in stats0], log_prior_end=[x[1] for x in stats1])
    log("mcmc batched chains: %d chains x %d steps, %.2f s per step, %.3g column-states/s, "
        "log P(seqs | ARG) %s -> %s" % (nchains, steps, sps, out["column_states_per_s"],
                                        ["%.8g" % x[0] for x in stats0],
                                        ["%.8g" % x[0] for x in stats1]))
    return out


def mcmc_cpu_compare(A, seed, log, k=100, L=500000):
    """The reference chain beside the GPU one (CPU-baseline leg): the same
    config-5 ARG, one re-threading step from the same srand seed in the
    reference library and in the drop-in; reports whether the two ARGs are
    identical, their local-tree counts and the two times."""
    import oracle_bind as O
    if not O.ref_available():
        return None
    times = np.array(A.get_time_points(20, 200e3), np.float64)
    pops, rho, mu = mcmc_model(seed)
    tr, seqs = A.simulate(k, L, times, pops, rho, mu, seed)
    rt = O.ref_new_trees(*A.trees_arrays(tr))
    A.remove_thread(tr, k - 1)
    O.ref_remove_thread(rt, k - 1)
    O.srand(seed + 3)
    t0 = time.perf_counter()
    A.sample_thread(tr, seqs, times, pops, rho, mu)
    t_gpu = time.perf_counter() - t0
    O.srand(seed + 3)
    t0 = time.perf_counter()
    O.ref_sample_thread(rt, seqs, times, pops, rho, mu)
    t_cpu = time.perf_counter() - t0
    mine, ref = A.trees_arrays(tr), O.ref_trees_arrays(rt)
    same = all(np.array_equal(a, b) for a, b in zip(mine, ref))
    out = dict(identical_arg=bool(same), local_trees_gpu=int(len(mine[3])),
               local_trees_cpu=int(len(ref[3])), seconds_gpu=t_gpu, seconds_cpu=t_cpu,
               cpu_cores=1)
    O.ref_delete(rt)
    log("mcmc step vs reference chain: identical ARG %s, local trees %d / %d, "
        "%.2f s GPU, %.2f s CPU" % (same, out["local_trees_gpu"], out["local_trees_cpu"],
                                    t_gpu, t_cpu))
    return out


def window_errors(a64, a32, nst_cols):
    """FP32 against FP64 on normalised columns: the largest absolute error
    and the largest relative error on entries above 1e-3 of the column's
    largest (the bounds stated for the FP32 mode)."""
    ea = er = 0.0
    o = 0
    for n in nst_cols:
        x, y = a64[o:o + n], a32[o:o + n]
        o += n
        if n == 0 or x.sum() <= 0:
            continue
        x = x / x.sum()
        y = y / y.sum()
        ea = max(ea, float(np.abs(x - y).max()))
        big = x > 1e-3 * x.max()
        er = max(er, float((np.abs(x - y)[big] / x[big]).max()))
    return ea, er


def alpha_windows(A, b, i, L, nst, nwin=4, width=256):
    """Columns of nwin windows spread over the instance (table view)."""
    out = []
    for w in range(nwin):
        c0 = int((w + 0.5) * L / nwin) - width // 2
        c0 = max(0, min(c0, L - width))
        c1 = min(L, c0 + width)
        out.append((c0, c1, b.alpha_range(i, c0, c1, nst)))
    return out


def single_large(A, seed, log, k=50, L=100000):
    """Config-2 instance (k=50, 1 Mb at c=10): the FP64 drop-in schedule
    against the FP32 one (forward tiles), with the FP32 table's error
    against FP64 on normalised columns and the log-likelihood error."""
    times = np.array(A.get_time_points(20, 200e3), np.float64)
    pops = np.full(20, 1e4)
    rho, mu = 1.6e-8 * 10, 1.8e-8 * 10
    tr, seqs = A.simulate(k, L, times, pops, rho, mu, seed + 5)
    A.remove_thread(tr, k - 1)
    u = np.random.RandomState(seed + 5).uniform(size=L)
    b = A.Batch()
    b.add(tr, seqs, times, pops, rho, mu, k - 1, u)
    b.upload()
    res = {}
    for prec, kw in ((0, {}), (1, dict(fw_tile=50000, fw_warm=20000, fw_tol=1e-4))):
        b.run(prec, 4096, **kw)
        t0 = time.perf_counter()
        st = b.run(prec, 4096, **kw)
        res[prec] = dict(seconds=time.perf_counter() - t0, lnl=b.lnl(0), alpha=b.alpha(0),
                         W=st["W"])
    W = res[0]["W"]
    # normalised columns: block-major rows of S entries per column
    a64, a32 = res[0]["alpha"], res[1]["alpha"]
    nst = A.get_nstates(tr, 20, L)
    off = np.concatenate([[0], np.cumsum(nst)])
    err_abs = 0.0
    err_rel = 0.0
    for c in range(0, L, 97):   # a sample of columns
        x = a64[off[c]:off[c + 1]]
        y = a32[off[c]:off[c + 1]]
        x = x / x.sum()
        y = y / y.sum()
        err_abs = max(err_abs, float(np.abs(x - y).max()))
        big = x > 1e-3 * x.max()
        err_rel = max(err_rel, float((np.abs(x - y)[big] / x[big]).max()))
    out = dict(k=k, columns=L, column_states=W, fp64_seconds=res[0]["seconds"],
               fp32_seconds=res[1]["seconds"], fp64_value=W / res[0]["seconds"],
               fp32_value=W / res[1]["seconds"],
               rel_dlogL=abs(res[1]["lnl"] - res[0]["lnl"]) / abs(res[0]["lnl"]),
               alpha_max_abs_err_normalised=err_abs, alpha_max_rel_err_large_entries=err_rel,
               sampled_columns=len(range(0, L, 97)))
    log("single large k=%d L=%d: fp64 %.3f s, fp32 %.3f s, |dlogL|/|logL| %.2e, alpha err abs %.2e"
        " rel %.2e" % (k, L, out["fp64_seconds"], out["fp32_seconds"], out["rel_dlogL"], err_abs,
                       err_rel))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tile", type=int, default=-1,
                    help="traceback tile length in columns (-1: autotune, 0: one tile per instance)")
    ap.add_argument("--fw-tile", type=int, default=-1,
                    help="FP32 forward tile length in columns (0: one tile per instance, "
                         "-1: autotune with the warm-up)")
    ap.add_argument("--fw-warm", type=int, default=20000, help="warm-up columns of a forward tile")
    ap.add_argument("--fw-tol", type=float, default=1e-4,
                    help="relative tolerance of a forward tile's start check")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of forward + traceback per core in the CPU baseline")
    ap.add_argument("--cpu-cores", type=int, default=16,
                    help="processes of the CPU baseline's aggregate (capped at the host's cores)")
    ap.add_argument("--plain-budget", type=float, default=10.0)
    ap.add_argument("--serve-rates", default="50,200,2000",
                    help="Poisson request rates (req/s), comma-separated")
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--no-mcmc", action="store_true", help="skip the config-5 re-threading sample")
    ap.add_argument("--mcmc-chains", type=int, default=4,
                    help="config-5 chains re-threaded together in one batch")
    ap.add_argument("--mcmc-steps", type=int, default=100,
                    help="steps of the batched config-5 chains (100: every leaf once; 0: skip)")
    ap.add_argument("--no-large", action="store_true", help="skip the config-2 single instance")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (calibrates HBM byte counters)")
    ap.add_argument("--traffic-profile", default=os.path.join(ROOT, "profiles", "k_fw_all_batched256_traffic.md"),
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE summary giving HBM bytes per column-state")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the threading instances shard across GPUs with no "
              "exchange, so --gpus N>1 is not run here", file=sys.stderr)
        sys.exit(2)

    import argweaver_amd
    import argweaver_amd.argweaverc as A
    t0 = time.time()
    insts, model = make_workload(A, args.instances, args.seed)
    log("workload: %d instances generated in %.1f s" % (len(insts), time.time() - t0))
    # the host-core baseline first: its worker processes are forked before
    # this process touches the GPU
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(A, insts, model, args.seed, args.cpu_budget, args.cpu_cores)
        log("cpu baseline: %.3g column-states/s on one core (%s), %.3g on %d cores"
            % (cpu["value"], cpu["kind"], cpu["aggregate"]["value"], cpu["aggregate"]["cores"]))
    if not argweaver_amd.device_available():
        raise SystemExit("no HIP device visible")
    if args.calib_copy:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read, 1 GiB written")
        del x, y
    log(A.device_info())

    batch = A.Batch()
    for it in insts:
        batch.add(it["trees"], it["seqs"], *model, it["k"] - 1, it["u"])
    batch.upload()
    Wtot = sum(batch.W(i) for i in range(len(insts)))
    for i, it in enumerate(insts):
        assert it["W"] == batch.W(i)
    log("uploaded: W = %.4g column-states, columns = %d" % (Wtot, sum(i["L"] for i in insts)))

    if args.fw_tile < 0:
        (args.fw_tile, args.fw_warm), fms = batch.tune_forward(tol=args.fw_tol)
        log("autotuned forward tiles: %d columns, warm-up %d (%s)"
            % (args.fw_tile, args.fw_warm,
               ", ".join("%d/%d: %.1f ms" % (t, w, v) for (t, w), v in fms.items())))
    if args.tile < 0:
        tile, tms = batch.tune(1)
        log("autotuned traceback tile: %d (%s)" % (tile, ", ".join("%d: %.1f ms" % kv for kv in tms.items())))
        args.tile = tile

    fw = dict(fw_tile=args.fw_tile, fw_warm=args.fw_warm, fw_tol=args.fw_tol)

    def timed(prec):
        kw = fw if prec == 1 else {}
        for _ in range(args.warmup):
            batch.run(prec, args.tile, **kw)
        t0 = time.perf_counter()
        stats = [batch.run(prec, args.tile, **kw) for _ in range(args.steps)]
        wall = (time.perf_counter() - t0) / args.steps
        return wall, stats

    def mean(st, k):
        return float(np.mean([x[k] for x in st]))

    # FP32 error against FP64 on a spread of instances (sorted by W, every
    # 32nd) and windows of their columns
    err_idx = sorted(range(len(insts)), key=lambda i: insts[i]["W"])[::32]
    nsts = {i: A.get_nstates(insts[i]["trees"], 20, insts[i]["L"]) for i in err_idx}

    wall32, st32 = timed(1)
    win32 = {i: alpha_windows(A, batch, i, insts[i]["L"], nsts[i]) for i in err_idx}
    log("batched fp32: %.1f ms/step (prep %.1f, emit %.1f, forward %.1f [main launch %.1f, "
        "%d tiles, %d re-run], traceback %.1f, tb passes %d)"
        % (wall32 * 1e3, mean(st32, "ms_prep"), mean(st32, "ms_emit"), mean(st32, "ms_forward"),
           mean(st32, "ms_fw_main"), st32[-1]["fw_tiles"], st32[-1]["fw_reruns"],
           mean(st32, "ms_traceback"), st32[-1]["tb_passes"]))
    lnl32 = [batch.lnl(i) for i in range(len(insts))]

    extra = {"stages_ms": {k: mean(st32, k) for k in ("ms_prep", "ms_emit", "ms_forward",
                                                      "ms_fw_main", "ms_traceback", "ms_total")},
             "fw_tiles": dict(fw, tiles=int(st32[-1]["fw_tiles"]),
                              reruns=int(st32[-1]["fw_reruns"]))}
    if not args.no_fp64:
        wall64, st64 = timed(0)
        lnl64 = [batch.lnl(i) for i in range(len(insts))]
        rel = max(abs(a - b) / abs(b) for a, b in zip(lnl32, lnl64))
        ea = er = 0.0
        for i in err_idx:
            for (c0, c1, a32), (_, _, a64) in zip(win32[i], alpha_windows(A, batch, i, insts[i]["L"],
                                                                         nsts[i])):
                x, y = window_errors(a64, a32, nsts[i][c0:c1])
                ea, er = max(ea, x), max(er, y)
        extra["fp32_error"] = dict(
            max_rel_dlogL=rel, alpha_max_abs_err_normalised=ea,
            alpha_max_rel_err_large_entries=er,
            alpha_sample="%d instances (every 32nd by W), 4 windows of 256 columns each"
            % len(err_idx))
        log("batched fp32 vs fp64: max |dlogL|/|logL| %.2e, alpha err abs %.2e rel %.2e"
            % (rel, ea, er))
        log("batched fp64: %.1f ms/step; max |dlogL|/|logL| fp32 vs fp64 %.2e"
            % (wall64 * 1e3, rel))
        plain = plain_fp64(A, insts, model, args.plain_budget, log)
        extra["fp64_batched"] = dict(value=Wtot / wall64, ms_per_step=wall64 * 1e3,
                                     ms_forward=mean(st64, "ms_forward"),
                                     ms_traceback=mean(st64, "ms_traceback"))
        extra["fp64_plain"] = plain
        extra["fp32_batched_vs_fp64_plain"] = (Wtot / wall32) / plain["value"]
        extra["max_rel_dlogL_fp32_vs_fp64"] = rel

    if args.serve_rates:
        pool, pmodel = make_serving_pool(A, args.seed)
        extra["serving"] = [serving(A, pool, pmodel, float(r), args.serve_requests,
                                    args.seed + j, log, tile=args.tile, precision=prec)
                            for prec in (1, 0)
                            for j, r in enumerate(args.serve_rates.split(","))]

    if not args.no_large:
        extra["single_large"] = single_large(A, args.seed, log)

    if not args.no_mcmc:
        extra["mcmc_rethread"] = mcmc_rethread(A, args.seed, log)
        extra["mcmc_rethread"]["fp32_chain"] = mcmc_rethread_fp32(A, args.seed, log)
        if args.mcmc_steps > 0:
            extra["mcmc_rethread"]["batched_chains"] = mcmc_batched_chains(
                A, args.seed, log, nchains=args.mcmc_chains, steps=args.mcmc_steps)

    # dominant kernel: the FP32 forward's main launch (k_fw_all<float, 19,
    # false>, one per step; re-run tiles are a separate symbol); algorithmic
    # bytes = one FP32 table entry written per column-state
    fwd_ms = mean(st32, "ms_fw_main")
    achieved = (Wtot * 4) / (fwd_ms * 1e-3)
    traffic = None
    bpcs = read_traffic(args.traffic_profile)
    if bpcs is not None:
        traffic = bpcs * Wtot   # HBM bytes per launch, scaled from the profiled run
    roofline = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                    frac=achieved / HBM_PEAK, traffic=traffic,
                    kernel="k_fw_all<float, 19, false>", algorithmic_bytes_per_launch=Wtot * 4,
                    ms_per_launch=fwd_ms)

    if not args.no_cpu:
        if not args.no_mcmc:
            cmp = mcmc_cpu_compare(A, args.seed, log)
            if cmp is not None:
                extra.setdefault("mcmc_rethread", {})["vs_reference_chain"] = cmp

    out = {
        "metric": "threaded columns*states/s (fwd+traceback) FP32-batched vs FP64; p99 latency",
        "value": Wtot / wall32,
        "unit": "column-states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 table, fp64 normalisers/CDF",
        "data": "synthetic (seeded DSMC ARGs, one leaf removed per instance)",
        "config": {"workload": "batched: %d instances, k~U{8..64}, L log-uniform 50kb-2Mb at c=10, "
                               "ntimes=20" % len(insts),
                   "instances": len(insts), "column_states": Wtot, "tile_cols": args.tile},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
