"""Serving: requests submitted while a batch runs are admitted by a free
driver and run beside it on the device; every result equals the same request
run alone."""
import threading
import time

import numpy as np
import pytest

from workloads import instance

pytestmark = pytest.mark.gpu

CASES = [(5, 900, 71), (40, 2500, 72), (12, 3200, 73), (64, 1500, 74), (8, 600, 75),
         (20, 2000, 76), (30, 1200, 77), (6, 2800, 78), (50, 900, 79), (16, 1800, 80)]


def _items(aw):
    out = []
    for k, L, seed in CASES:
        tr, seqs, _, model = instance(aw, k, L, seed)
        u = np.random.RandomState(seed).uniform(size=L)
        out.append(dict(k=k, trees=tr, seqs=seqs, model=model, u=u))
    return out


def _alone(aw, it, prec, tile):
    b = aw.Batch()
    b.add(it["trees"], it["seqs"], *it["model"], it["k"] - 1, it["u"])
    b.upload()
    b.run(prec, tile)
    return b.path(0), b.lnl(0)


@pytest.mark.parametrize("prec", [0, 1])
def test_served_equals_alone(aw, prec):
    print("serve prec=%d" % prec, flush=True)
    items = _items(aw)
    tile = 700
    srv = aw.Server(hbm_budget_gb=8.0, precision=prec, tile_cols=tile, max_batch=4)
    ids = []

    def producer():
        rng = np.random.RandomState(5)
        for it in items:
            ids.append(srv.submit(it["trees"], it["seqs"], *it["model"], it["k"] - 1, it["u"]))
            time.sleep(rng.exponential(0.003))

    t = threading.Thread(target=producer)
    t.start()
    t.join()
    res = [srv.wait(i, timeout=60) for i in ids]
    srv.close()
    for it, r in zip(items, res):
        path, lnl = _alone(aw, it, prec, tile)
        assert r["lnl"] == lnl
        assert np.array_equal(r["path"], path)
        assert r["t_submit"] <= r["t_start"] <= r["t_done"]


def test_budget_admission(aw):
    """A tiny budget admits one request at a time, still every result."""
    print("serve tiny budget", flush=True)
    items = _items(aw)[:4]
    srv = aw.Server(hbm_budget_gb=1e-6, precision=1, tile_cols=0, max_batch=8)
    ids = [srv.submit(it["trees"], it["seqs"], *it["model"], it["k"] - 1, it["u"]) for it in items]
    res = [srv.wait(i, timeout=60) for i in ids]
    srv.close()
    for it, r in zip(items, res):
        path, lnl = _alone(aw, it, 1, 0)
        assert r["lnl"] == lnl and np.array_equal(r["path"], path)
    # one request per batch, and at most one batch per driver at a time
    assert all(r["batch_size"] == 1 for r in res)
    for r in res:
        t = r["t_start"] + 1e-6
        assert sum(1 for q in res if q["t_start"] <= t < q["t_done"]) <= 2


def test_admitted_mid_flight_equals_alone(aw):
    """A short request submitted while a long one is on the GPU is admitted
    by the other driver at once, runs and finishes while the long one is
    still running, and both results equal the requests run alone."""
    print("serve mid-flight admission", flush=True)
    long_tr, long_seqs, _, long_model = instance(aw, 40, 200000, 91)
    long_u = np.random.RandomState(91).uniform(size=200000)
    short = _items(aw)[0]
    srv = aw.Server(hbm_budget_gb=16.0, precision=0, tile_cols=0, max_batch=8, drivers=2)
    il = srv.submit(long_tr, long_seqs, *long_model, 39, long_u)
    time.sleep(0.05)
    i_s = srv.submit(short["trees"], short["seqs"], *short["model"], short["k"] - 1, short["u"])
    rs = srv.wait(i_s, timeout=120)
    rl = srv.wait(il, timeout=120)
    srv.close()
    print("long %.3f-%.3f driver %d, short %.3f-%.3f driver %d" % (
        rl["t_start"], rl["t_done"], rl["driver"], rs["t_start"], rs["t_done"], rs["driver"]),
        flush=True)
    assert rs["driver"] != rl["driver"]
    assert rl["t_start"] <= rs["t_start"] < rs["t_done"] < rl["t_done"]
    path, lnl = _alone(aw, short, 0, 0)
    assert rs["lnl"] == lnl and np.array_equal(rs["path"], path)
    it = dict(k=40, trees=long_tr, seqs=long_seqs, model=long_model, u=long_u)
    path, lnl = _alone(aw, it, 0, 0)
    assert rl["lnl"] == lnl and np.array_equal(rl["path"], path)
    aw.delete_local_trees(long_tr)


def test_oversized_request_fails_alone(aw):
    """A request over the kernels' limits (257 nodes per tree) is refused at
    submission with its own status; the valid request queued beside it runs."""
    print("serve oversized + valid", flush=True)
    items = _items(aw)[:2]
    big_tr, big_seqs, _, big_model = instance(aw, 130, 300, 81)
    srv = aw.Server(hbm_budget_gb=8.0, precision=1, tile_cols=0, max_batch=8)
    i0 = srv.submit(items[0]["trees"], items[0]["seqs"], *items[0]["model"], items[0]["k"] - 1,
                    items[0]["u"])
    ib = srv.submit(big_tr, big_seqs, *big_model, 129, np.random.RandomState(81).uniform(size=300))
    i1 = srv.submit(items[1]["trees"], items[1]["seqs"], *items[1]["model"], items[1]["k"] - 1,
                    items[1]["u"])
    with pytest.raises(ValueError):
        srv.wait(ib, timeout=60)
    r0, r1 = srv.wait(i0, timeout=60), srv.wait(i1, timeout=60)
    srv.close()
    for it, r in zip(items, (r0, r1)):
        path, lnl = _alone(aw, it, 1, 0)
        assert r["lnl"] == lnl and np.array_equal(r["path"], path)


def test_batch_add_after_upload_refused(aw):
    """upload() releases the staged host rows, so a later add() is refused
    (it would otherwise plan an instance with no rows)."""
    it = _items(aw)[0]
    b = aw.Batch()
    b.add(it["trees"], it["seqs"], *it["model"], it["k"] - 1, it["u"])
    b.upload()
    with pytest.raises(RuntimeError):
        b.add(it["trees"], it["seqs"], *it["model"], it["k"] - 1, it["u"])
    b.run(1, 0)
    path, lnl = _alone(aw, it, 1, 0)
    assert b.lnl(0) == lnl and np.array_equal(b.path(0), path)


def test_served_big_block_request_equals_alone(aw):
    """A request with blocks of more than 1024 states (32 states per lane,
    the kernels' separate symbols) served beside small ones: every result
    equals the request run alone."""
    print("serve with a big-block request", flush=True)
    items = []
    for k, L, seed, pops in [(128, 1200, 91, 3e4), (8, 900, 92, 1e4), (20, 1500, 93, 1e4)]:
        tr, seqs, _, model = instance(aw, k, L, seed, popsizes=np.full(20, pops))
        u = np.random.RandomState(seed).uniform(size=L)
        items.append(dict(k=k, trees=tr, seqs=seqs, model=model, u=u))
    assert aw.num_column_states(items[0]["trees"], 20)[1].max() > 1024
    srv = aw.Server(hbm_budget_gb=8.0, precision=1, tile_cols=300, max_batch=4)
    ids = [srv.submit(it["trees"], it["seqs"], *it["model"], it["k"] - 1, it["u"])
           for it in items]
    res = [srv.wait(i, timeout=120) for i in ids]
    srv.close()
    for it, r in zip(items, res):
        path, lnl = _alone(aw, it, 1, 300)
        assert r["lnl"] == lnl
        assert np.array_equal(r["path"], path)
