"""Serving loop (bsar_server_*): continuous batching of imaging jobs.

The property under test: a job's image is the same alone and in a mix of
other jobs -- bit-identical, in both precisions -- and equal to the one-image
call (bsar_backproject) over a dataset holding the job's first K pulses, which
the single-image tests pin against the oracle (plain FP64 within 1e-9 of the
peak, mixed within -80 dB)."""
import numpy as np
import pytest

from _scenes import Tiny, bumps_dem, db, peak_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
MIXED_DB = -80.0


def _scene(n=40, big_dem=False, seed=11):
    dem = bumps_dem(n=95, spacing=6.0, seed=5) if big_dem else None    # 95 + 95 + 95*96 doubles > LDS budget
    t = Tiny(n=n, S=512, grid=(7, 5, 4), seed=seed, dem=dem)
    return t


def _dataset(bsar, ctx, t, root, K=None):
    K = t.n if K is None else K
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx[:K], t.rx[:K])
    ds.load_rootdata(root)
    return ds


def _grid(t, shape):
    nx, ny, nz = shape
    gx = np.round(19.6 + 0.05 * np.arange(nx), 6)
    gy = np.round(-0.5 + 0.05 * np.arange(ny), 6)
    gz = np.round(-5.4 + 0.1 * np.arange(nz), 6)
    return gx, gy, gz


@pytest.fixture(scope="module")
def world(bsar, ctx, oracle):
    t = _scene()
    root = t.target_root(oracle, pts=np.array([[19.9, -0.3, -5.1], [20.3, 0.2, -5.0]]))
    tb = _scene(big_dem=True)
    ds = _dataset(bsar, ctx, t, root)
    dsb = _dataset(bsar, ctx, tb, root)
    return dict(t=t, tb=tb, root=root, ds=ds, dsb=dsb)


# (dataset key, grid, K, precision)
JOBS = [
    ("ds", (9, 7, 3), 40, "fp64"),
    ("ds", (20, 15, 3), 17, "mixed"),
    ("dsb", (6, 6, 6), 40, "mixed"),
    ("ds", (31, 3, 2), 5, "fp64"),
    ("dsb", (11, 9, 2), 23, "fp64"),
    ("ds", (4, 4, 4), 1, "mixed"),
    ("ds", (16, 16, 8), 40, "mixed"),
    ("ds", (1, 1, 1), 40, "fp64"),
]


def _prec(bsar, p):
    return bsar.FP64 if p == "fp64" else bsar.MIXED


def _run(bsar, world, jobs, round_seconds, together):
    out = []
    with bsar.Server(round_seconds=round_seconds) as srv:
        if together:
            ids = [srv.submit(world[k], *_grid(world["t"], g), pulses=K, precision=_prec(bsar, p))
                   for k, g, K, p in jobs]
            got = {}
            for _ in ids:                                  # collected in completion order
                jid, img, info = srv.wait_any(timeout=120)
                got[jid] = (img, info)
            assert sorted(got) == sorted(ids)
            out = [got[i] for i in ids]
            with pytest.raises(bsar.BsarError):
                srv.wait_any(timeout=0.1)                  # nothing left to collect
            st = srv.stats()
            assert st["completed"] == len(jobs) and st["failed"] == 0
            assert st["rounds"] >= 1
        else:
            for k, g, K, p in jobs:
                i = srv.submit(world[k], *_grid(world["t"], g), pulses=K, precision=_prec(bsar, p))
                out.append(srv.wait(i, timeout=120))
    return out


def test_job_equals_single_image_call(bsar, ctx, world, oracle):
    """All pulses: the server's image equals bsar_backproject bit for bit, and
    the oracle within the image tolerances."""
    t, ds = world["t"], world["ds"]
    gx, gy, gz = _grid(t, (9, 7, 3))
    with bsar.Server() as srv:
        i64 = srv.submit(ds, gx, gy, gz, precision=bsar.FP64)
        imx = srv.submit(ds, gx, gy, gz, precision=bsar.MIXED)
        a, ia = srv.wait(i64, timeout=120)
        b, ib = srv.wait(imx, timeout=120)
    ref64, _ = ds.backproject(gx, gy, gz, bsar.FP64)
    refmx, rep = ds.backproject(gx, gy, gz, bsar.MIXED)
    assert np.array_equal(a, ref64)
    assert np.array_equal(b, refmx)
    assert ib["r1_exact_frac"] == pytest.approx(rep["r1_exact_frac"])
    want = oracle.backproject(t.bp(), t.dem, t.tx, t.rx, oracle.upsample(world["root"]), gx, gy, gz)
    assert peak_err(a, want) <= FP64_TOL
    assert db(peak_err(b, want)) <= MIXED_DB
    for info in (ia, ib):
        assert info["state"] == "done" and info["status"] == 0
        assert info["t_submit"] <= info["t_admit"] <= info["t_done"]
        assert info["voxel_pulses"] == gx.size * gy.size * gz.size * t.n


@pytest.mark.parametrize("K", [1, 13, 39])
def test_first_k_pulses(bsar, ctx, world, oracle, K):
    """pulses = K images the first K pulses: R1 means over those K positions
    (proc_test_script.cpp:518-548 with Npulses = K)."""
    t = world["t"]
    gx, gy, gz = _grid(t, (8, 6, 3))
    dsk = _dataset(bsar, ctx, t, world["root"], K=K)
    with bsar.Server() as srv:
        ids = [srv.submit(world["ds"], gx, gy, gz, pulses=K, precision=p) for p in (bsar.FP64, bsar.MIXED)]
        (a, _), (b, _) = [srv.wait(i, timeout=120) for i in ids]
    ref64, _ = dsk.backproject(gx, gy, gz, bsar.FP64)
    refmx, _ = dsk.backproject(gx, gy, gz, bsar.MIXED)
    dsk.close()
    assert np.array_equal(a, ref64)
    assert np.array_equal(b, refmx)
    want = oracle.backproject(t.bp(), t.dem, t.tx[:K], t.rx[:K], oracle.upsample(world["root"]), gx, gy, gz)
    assert peak_err(a, want) <= FP64_TOL
    assert db(peak_err(b, want)) <= MIXED_DB


@pytest.mark.parametrize("round_seconds", [0.0005, 0.04])
def test_same_alone_and_in_the_mix(bsar, world, round_seconds):
    """Eight jobs of uneven grids, pulse counts, precisions and DEM sizes (both
    LDS and global-DEM kernel variants) submitted together share rounds; each
    image equals the one the job gets alone, bit for bit."""
    alone = _run(bsar, world, JOBS, round_seconds, together=False)
    mixed = _run(bsar, world, JOBS, round_seconds, together=True)
    for (k, g, K, p), (a, _), (b, ib) in zip(JOBS, alone, mixed):
        assert a.shape == (g[2], g[1], g[0])
        assert np.array_equal(a, b), (k, g, K, p)
        assert np.abs(a).max() > 0
        assert ib["state"] == "done"


def test_large_job_single_group(bsar, world):
    """A grid of 2^19 voxels runs as one pulse group (groups == 1 path:
    accumulators are the image)."""
    t, ds = world["t"], world["ds"]
    gx, gy, gz = _grid(t, (128, 64, 64))
    with bsar.Server(round_seconds=0.01) as srv:
        ids = [srv.submit(ds, gx, gy, gz, pulses=6, precision=p) for p in (bsar.FP64, bsar.MIXED)]
        small = srv.submit(ds, *_grid(t, (5, 5, 5)), pulses=6, precision=bsar.MIXED)
        (a, _), (b, _) = [srv.wait(i, timeout=300) for i in ids]
        c, _ = srv.wait(small, timeout=300)
        assert srv.stats()["rounds"] >= 2
    dsk = _dataset(bsar, ds.ctx, t, world["root"], K=6)
    ref64, _ = dsk.backproject(gx, gy, gz, bsar.FP64)
    refmx, _ = dsk.backproject(gx, gy, gz, bsar.MIXED)
    refc, _ = dsk.backproject(*_grid(t, (5, 5, 5)), bsar.MIXED)
    dsk.close()
    assert np.array_equal(a, ref64)
    assert np.array_equal(b, refmx)
    assert np.array_equal(c, refc)
    assert db(peak_err(b, a)) <= MIXED_DB


def test_admission_by_hbm_cap(bsar, world):
    """A cap that holds one job at a time: jobs queue, run one by one, and the
    peak admitted footprint stays under the cap; a job larger than the cap is
    rejected at submission."""
    t, ds = world["t"], world["ds"]
    grid = _grid(t, (10, 10, 4))
    with bsar.Server() as probe:
        i = probe.submit(ds, *grid, precision=bsar.MIXED)
        ref, info = probe.wait(i, timeout=120)
    fp = info["footprint_bytes"]
    cap = int(fp * 1.5)
    with bsar.Server(hbm_cap_bytes=cap) as srv:
        ids = [srv.submit(ds, *grid, precision=bsar.MIXED) for _ in range(4)]
        with pytest.raises(bsar.BsarError) as e:
            srv.submit(ds, *_grid(t, (100, 100, 100)), precision=bsar.MIXED)
        assert e.value.status == -6
        res = [srv.wait(i, timeout=120) for i in ids]
        st = srv.stats()
    assert st["hbm_peak"] <= cap and st["rejected"] == 1 and st["completed"] == 4
    for img, inf in res:
        assert np.array_equal(img, ref)
    admits = sorted(inf["t_admit"] for _, inf in res)
    dones = sorted(inf["t_done"] for _, inf in res)
    for k in range(1, 4):
        assert admits[k] >= dones[k - 1] - 1e-9     # one at a time


def test_job_errors(bsar, ctx, world):
    t, ds = world["t"], world["ds"]
    g = _grid(t, (3, 3, 3))
    with bsar.Server() as srv:
        for kw in (dict(pulses=t.n + 1), dict(pulses=-1), dict(precision=7)):
            with pytest.raises(bsar.BsarError) as e:
                srv.submit(ds, *g, **kw)
            assert e.value.status == -1
        empty = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)   # no range data
        with pytest.raises(bsar.BsarError) as e:
            srv.submit(empty, *g)
        assert e.value.status == -7
        empty.close()
        i = srv.submit(ds, g[0], g[1], np.empty(0))      # empty grid: done at once
        img, info = srv.wait(i, timeout=10)
        assert img.shape == (0, 3, 3) and info["state"] == "done"
        with pytest.raises(bsar.BsarError):
            srv.wait(i, timeout=1)                       # collected
        srv.drain()
        assert srv.stats()["active"] == 0
