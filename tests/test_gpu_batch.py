"""Batched engine on the GPU: mixed-size batches in FP64 (against the oracle)
and FP32 (exact and column-tiled, against FP64 and the oracle), and the
same instance alone versus inside a mix."""
import numpy as np
import pytest

from workloads import instance

pytestmark = pytest.mark.gpu

# k=64 instances need 16 slots per lane in the FP32 schedule, k<=40 need 8:
# the mix exercises per-instance slot tables under one kernel instantiation
MIX = [(4, 700, 21), (64, 2500, 22), (12, 3000, 23), (40, 1800, 24), (8, 1200, 25)]


def _build(aw, cases):
    out = []
    for k, L, seed in cases:
        tr, seqs, inst, model = instance(aw, k, L, seed)
        u = np.random.RandomState(seed).uniform(size=L)
        out.append(dict(k=k, L=L, seed=seed, trees=tr, seqs=seqs, inst=inst, model=model, u=u))
    return out


def _batch(aw, items):
    b = aw.Batch()
    for it in items:
        b.add(it["trees"], it["seqs"], *it["model"], it["k"] - 1, it["u"])
    b.upload()
    return b


@pytest.fixture(scope="module")
def mix(aw):
    items = _build(aw, MIX)
    for it in items:
        fw, lnl = it["inst"].forward()
        it["fw"], it["lnl"] = fw, lnl
        it["path"] = it["inst"].traceback(fw, it["u"])
    return items


def test_batch_fp64_matches_oracle(aw, mix):
    print("batch fp64 mix", [(m["k"], m["L"]) for m in mix], flush=True)
    b = _batch(aw, mix)
    b.run(0, 0)
    for i, it in enumerate(mix):
        a = b.alpha(i)
        rel = np.abs(a - it["fw"]) / np.maximum(np.abs(it["fw"]), 1e-300)
        assert rel.max() <= 1e-12, (i, rel.max())
        assert abs(b.lnl(i) - it["lnl"]) <= 1e-12 * abs(it["lnl"])
        assert np.array_equal(b.path(i), it["path"]), i


@pytest.mark.parametrize("tile", [0, 300, 1000])
def test_batch_fp32_close_to_fp64(aw, mix, tile):
    print("batch fp32 mix tile=%d" % tile, flush=True)
    b = _batch(aw, mix)
    st = b.run(1, tile)
    assert st["W"] == sum(len(m["fw"]) for m in mix)
    for i, it in enumerate(mix):
        lnl = b.lnl(i)
        # FP32 normalised columns, FP64 normalisers: |dlogL| <= 1e-6 relative
        assert abs(lnl - it["lnl"]) <= 1e-6 * abs(it["lnl"]), (i, lnl, it["lnl"])
        # the FP32 table is normalised per column; compare with the oracle's
        # columns normalised the same way
        a32 = b.alpha(i)
        offs = it["inst"].block_offsets()
        nst = it["inst"].nstates()
        fw = it["fw"]
        err = 0.0
        for blk in range(len(offs) - 1):
            o0, o1, S = int(offs[blk]), int(offs[blk + 1]), int(nst[blk])
            ref = fw[o0:o1].reshape(-1, S)
            got = a32[o0:o1].reshape(-1, S)
            refn = ref / ref.sum(axis=1, keepdims=True)
            err = max(err, np.abs(got - refn).max())
        assert err <= 2e-4, (i, err)
        agree = np.mean(b.path(i) == it["path"])
        assert agree >= 0.98, (i, agree)


def _fp32_col_err(b, i, it):
    """max abs difference between b's FP32 columns and the oracle's columns,
    both normalised per column"""
    a32 = b.alpha(i)
    offs = it["inst"].block_offsets()
    nst = it["inst"].nstates()
    fw = it["fw"]
    err = 0.0
    for blk in range(len(offs) - 1):
        o0, o1, S = int(offs[blk]), int(offs[blk + 1]), int(nst[blk])
        ref = fw[o0:o1].reshape(-1, S)
        got = a32[o0:o1].reshape(-1, S)
        err = max(err, np.abs(got - ref / ref.sum(axis=1, keepdims=True)).max())
    return err


@pytest.mark.parametrize("warm", [600, 2])
def test_fp32_forward_tiles(aw, mix, warm):
    """FP32 forward tiles: every tile start is checked against the column the
    tile before computes; with a long warm-up the speculative starts pass,
    with a 2-column warm-up they fail and are re-run from that column.
    Either way the table, the log-likelihood and the thread stay within the
    FP32 bounds of the oracle's FP64 results."""
    print("fp32 forward tiles warm=%d" % warm, flush=True)
    b = _batch(aw, mix)
    st = b.run(1, 300, fw_tile=900, fw_warm=warm, fw_tol=1e-5)
    # tile lengths follow the column cost of each instance's states per
    # lane, so only bounds are fixed: at least the instances longer than the
    # longest tile are split (tiles stay within 900 / 0.94 columns), and no
    # tile is shorter than 1.25 warm-ups (the columns of the first tile) or
    # stores fewer than a quarter of a warm-up (the rest)
    cmax, cmin = int(900 / 0.94), max(warm + warm // 4, 1)
    lo = sum(1 + max(0, -(-(m["L"] - cmax) // (cmax - warm))) for m in mix)
    hi = sum(1 + max(0, -(-(m["L"] - cmin) // max(cmin - warm, 1))) for m in mix)
    assert lo <= st["fw_tiles"] <= hi and st["fw_tiles"] > len(mix), (st["fw_tiles"], lo, hi)
    if warm == 2:
        assert st["fw_reruns"] >= 1
    for i, it in enumerate(mix):
        assert abs(b.lnl(i) - it["lnl"]) <= 1e-6 * abs(it["lnl"]), (i, b.lnl(i), it["lnl"])
        assert _fp32_col_err(b, i, it) <= 2e-4, i
        assert np.mean(b.path(i) == it["path"]) >= 0.98, i
    # the tile plan depends on the instance alone: a second run repeats the
    # first bit for bit
    first = [(b.lnl(i), b.path(i), b.alpha(i)) for i in range(len(mix))]
    st2 = b.run(1, 300, fw_tile=900, fw_warm=warm, fw_tol=1e-5)
    assert st2["fw_tiles"] == st["fw_tiles"]
    for i in range(len(mix)):
        assert b.lnl(i) == first[i][0], i
        assert np.array_equal(b.path(i), first[i][1]), i
        assert np.array_equal(b.alpha(i), first[i][2]), i


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_tiled_traceback_matches_untiled(aw, mix, prec):
    """Tiles with fix-up passes give the same thread as one tile per instance:
    after coalescence the re-sampled path is the stored one."""
    print("batch tiled vs untiled traceback prec=%d" % prec, flush=True)
    b = _batch(aw, mix)
    b.run(prec, 0)
    ref = [(b.lnl(i), b.path(i), b.alpha(i)) for i in range(len(mix))]
    st = b.run(prec, 250)
    for i in range(len(mix)):
        lnl, path, a = ref[i]
        assert b.lnl(i) == lnl, i
        assert np.array_equal(b.alpha(i), a), i
        assert np.array_equal(b.path(i), path), (i, np.mean(b.path(i) == path))
    assert st["tb_passes"] >= 1


@pytest.mark.parametrize("prec,tile,fw", [(0, 0, 0), (1, 0, 0), (1, 400, 0), (1, 400, 900)])
def test_alone_equals_in_mix(aw, mix, prec, tile, fw):
    """An instance's results do not depend on the rest of the batch (FP32
    forward tiles included: their plan depends on the instance alone)."""
    print("alone vs mix prec=%d tile=%d fw_tile=%d" % (prec, tile, fw), flush=True)
    kw = dict(fw_tile=fw, fw_warm=600, fw_tol=1e-5) if fw else {}
    b = _batch(aw, mix)
    b.run(prec, tile, **kw)
    for i in (0, 1, 3):
        one = _batch(aw, [mix[i]])
        one.run(prec, tile, **kw)
        assert one.lnl(0) == b.lnl(i), i
        assert np.array_equal(one.path(0), b.path(i)), i
        assert np.array_equal(one.alpha(0), b.alpha(i)), i


@pytest.mark.parametrize("options,warm,tail", [(1, 256, 0), (2, 0, 60), (3, 256, 100), (4, 50, 4000)])
def test_traceback_options_match_untiled(aw, mix, options, warm, tail):
    """Traceback tiles with options, resolved right to left (alternatives
    adopted, re-runs where none fits), give the thread of one tile per
    instance on the same table, bit for bit.  The FP32 forward runs as tiles
    with re-runs (a 2-column warm-up), so the options of the instances that
    are final early run beside those re-runs; the re-run instances' own
    traceback tiles have the tail length."""
    print("traceback options=%d warm=%d tail=%d" % (options, warm, tail), flush=True)
    b = _batch(aw, mix)
    kw = dict(fw_tile=900, fw_warm=2, fw_tol=1e-5)
    b.run(1, 0, **kw)
    ref = [b.path(i) for i in range(len(mix))]
    b.set_traceback(warm, options)
    b.set_tail_tiles(tail)
    st = b.run(1, 150, **kw)
    assert st["fw_reruns"] >= 1
    assert st["ms_tb_tail"] > 0
    for i in range(len(mix)):
        assert np.array_equal(b.path(i), ref[i]), (i, np.mean(b.path(i) == ref[i]))


def test_tile_autotuner(aw, mix):
    """The tuner times every candidate on the resident tables and returns one
    of them; a run with it reproduces the untiled thread."""
    print("tile autotuner", flush=True)
    b = _batch(aw, mix)
    b.run(1, 0)
    ref = [b.path(i) for i in range(len(mix))]
    best, ms = b.tune(1, (0, 300, 900))
    assert best in (0, 300, 900)
    assert all(v > 0 for v in ms.values())
    b.run(1, best)
    for i in range(len(mix)):
        assert np.array_equal(b.path(i), ref[i]), i


def test_forward_tile_autotuner(aw, mix):
    """The forward-tile tuner times every candidate and returns one of them."""
    print("forward tile autotuner", flush=True)
    b = _batch(aw, mix)
    cands = ((0, 0), (900, 600), (1500, 600))
    best, ms = b.tune_forward(candidates=cands)
    assert best in cands
    assert set(ms) == set(cands) and all(v > 0 for v in ms.values())


def test_fp32_forward_tiles_28_time_points(aw):
    """Forward tiles with re-runs and tiled traceback on a 28-point time grid
    (the kernels built for up to 31 time bins): FP32 bounds against the
    oracle's FP64 results."""
    print("fp32 forward tiles, 28 time points", flush=True)
    items = []
    for k, L, seed in [(6, 1500, 31), (10, 2000, 32), (16, 1200, 33)]:
        tr, seqs, inst, model = instance(aw, k, L, seed, ntimes=28)
        u = np.random.RandomState(seed).uniform(size=L)
        fw, lnl = inst.forward()
        items.append(dict(k=k, L=L, trees=tr, seqs=seqs, inst=inst, model=model, u=u, fw=fw,
                          lnl=lnl, path=inst.traceback(fw, u)))
    b = _batch(aw, items)
    st = b.run(1, 300, fw_tile=700, fw_warm=2, fw_tol=1e-5)
    assert st["fw_tiles"] > len(items) and st["fw_reruns"] >= 1
    for i, it in enumerate(items):
        assert abs(b.lnl(i) - it["lnl"]) <= 1e-6 * abs(it["lnl"]), (i, b.lnl(i), it["lnl"])
        assert _fp32_col_err(b, i, it) <= 2e-4, i
        assert np.mean(b.path(i) == it["path"]) >= 0.98, i


def test_batch_clear_refill(aw, mix):
    """A batch emptied with clear() and filled with other instances (its
    device buffers reused) gives what a fresh batch of them gives, in both
    precisions; the FP64 threads equal the oracle's."""
    print("batch clear and refill", flush=True)
    b = _batch(aw, mix[:2])
    b.run(1, 300, fw_tile=900, fw_warm=200, fw_tol=1e-5)
    b.clear()
    for it in mix[2:]:
        b.add(it["trees"], it["seqs"], *it["model"], it["k"] - 1, it["u"])
    b.upload()
    fresh = _batch(aw, mix[2:])
    for prec in (0, 1):
        b.run(prec, 300, fw_tile=900, fw_warm=200, fw_tol=1e-5)
        fresh.run(prec, 300, fw_tile=900, fw_warm=200, fw_tol=1e-5)
        for i, it in enumerate(mix[2:]):
            assert b.lnl(i) == fresh.lnl(i), (prec, i)
            assert np.array_equal(b.path(i), fresh.path(i)), (prec, i)
            if prec == 0:
                assert np.array_equal(b.path(i), it["path"]), i
