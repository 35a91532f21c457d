"""Back-projection on the GPU against the reference's own output, the CPU
oracle, and size-independent properties.

Tolerances: plain FP64 within 1e-9 of the oracle / reference relative to
the image peak (in practice the difference is sin/cos rounding, ~1e-15);
mixed precision within -80 dB of the FP64 image, peak-normalised."""
import numpy as np
import pytest

from _scenes import Tiny, db, peak_err
from conftest import GOLDEN
from test_oracle_golden import parse_image

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
MIXED_DB = -80.0


@pytest.mark.parametrize("tag", ["random", "target"])
def test_fp64_matches_reference_binary(bsar, ctx, tag):
    """Inputs and output_3D_matrix.bin of the reference's proc_test_script
    (its compiled-in constants: fc = 10 GHz, fs = 60 MHz, x4)."""
    g = np.load(GOLDEN / f"ref_bp_{tag}.npz")
    ref = parse_image(g["out_bin"].tobytes())
    p = bsar.Params(float(g["fc"]), float(g["c"]), float(g["fs"]), float(g["n1"]), float(g["n2"]), int(g["upsample"]))
    ds = ctx.dataset(p, bsar.Dem(g["dem_x"], g["dem_y"], g["dem_z"]), g["tx"], g["rx"])
    ds.load_rootdata(g["root"])
    img, rep = ds.backproject(g["gx"], g["gy"], g["gz"], bsar.FP64)
    assert peak_err(img, ref) <= FP64_TOL
    assert rep["voxels"] == ref.size and rep["pulses"] == len(g["tx"])
    mixed, _ = ds.backproject(g["gx"], g["gy"], g["gz"], bsar.MIXED)
    assert db(peak_err(mixed, ref)) <= MIXED_DB


@pytest.mark.parametrize("kind", ["random", "target"])
def test_fp64_matches_oracle_tiny(bsar, ctx, oracle, kind):
    t = Tiny()
    root = t.random_root() if kind == "random" else t.target_root(oracle)
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    ds.load_rootdata(root)
    img, rep = ds.backproject(t.gx, t.gy, t.gz, bsar.FP64)
    want = oracle.backproject(t.bp(), t.dem, t.tx, t.rx, oracle.upsample(root), t.gx, t.gy, t.gz)
    assert peak_err(img, want) <= FP64_TOL
    assert rep["launches"] >= 1
    mixed, rep = ds.backproject(t.gx, t.gy, t.gz, bsar.MIXED)
    assert db(peak_err(mixed, want)) <= MIXED_DB
    assert rep["r1_exact_frac"] < 0.05


def test_fewer_data_pulses_than_positions(bsar, ctx, oracle):
    """idxAz = min(i, Npulses - 1) (proc_test_script.cpp:607)."""
    t = Tiny(n=20)
    root = t.random_root(ndata=13)
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    ds.load_rootdata(root)
    img, _ = ds.backproject(t.gx, t.gy, t.gz, bsar.FP64)
    want = oracle.backproject(t.bp(), t.dem, t.tx, t.rx, oracle.upsample(root), t.gx, t.gy, t.gz)
    assert peak_err(img, want) <= FP64_TOL


def test_out_of_window_gives_zeros(bsar, ctx):
    t = Tiny(S=8)            # 32 upsampled samples: every delay lands past the window
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    ds.load_rootdata(t.random_root())
    for prec in (bsar.FP64, bsar.MIXED):
        img, _ = ds.backproject(t.gx, t.gy, t.gz, prec)
        assert np.all(img == 0)


@pytest.mark.parametrize("demkind", ["empty", "unsorted", "rough"])
def test_dem_variants(bsar, ctx, oracle, scene_mod, demkind):
    if demkind == "empty":
        e = np.zeros(0)
        dem = (e, e, np.zeros((0, 0)))
    elif demkind == "unsorted":
        dem = (np.array([-300.0, 300.0, 0.0, -150.0, 150.0]), np.array([-300.0, 0.0, 300.0]),
               np.array([[0.1, 0.2, 0.3, 0.4, 0.5], [0.0, -0.3, 0.6, 0.2, 0.1], [0.3, 0.3, -0.2, 0.0, 0.4]]))
    else:
        dem = scene_mod.dem_fbm(np.random.default_rng(3), nx=100, ny=90)
    # antennas 12 m and voxels ~5 m below the terrain height near the grid
    z0 = oracle.dem_interp(dem, [[20.0, 0.0]])[0]
    t = Tiny(n=10, grid=(3, 3, 2), dem=dem, z0=z0)
    root = t.random_root()
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*dem), t.tx, t.rx)
    ds.load_rootdata(root)
    img, _ = ds.backproject(t.gx, t.gy, t.gz, bsar.FP64)
    want = oracle.backproject(t.bp(), dem, t.tx, t.rx, oracle.upsample(root), t.gx, t.gy, t.gz)
    assert np.abs(want).max() > 0
    assert peak_err(img, want) <= FP64_TOL


@pytest.mark.parametrize("demkind", ["fbm_small", "ragged", "dup", "flat", "small_z"])
def test_dem_variants_staged_in_lds(bsar, ctx, oracle, scene_mod, golden, demkind):
    """DEMs small enough for the back-projection kernels' LDS-staged search
    (the 15-wide level in two column parts, (height, x difference) pairs,
    the side stop of the row and column bounds): 40 m RMS terrain, a ragged
    grid with non-uniform axes, duplicated nodes, a flat DEM, and Z smaller
    than its axes.  Any refraction point that differs from the reference's
    moves a path by centimetres, which the 1e-9 bar would show."""
    if demkind == "fbm_small":
        dem = scene_mod.dem_fbm(np.random.default_rng(5), nx=30, ny=28)
    else:
        dem = tuple(np.ascontiguousarray(golden[f"dem_{demkind}_{a}"]) for a in "xyz")
    nx, ny = len(dem[0]), len(dem[1])
    assert nx + ny + 2 * ny * (nx + 1) <= 8192    # 8-byte slots of the LDS-staged DEM
    z0 = oracle.dem_interp(dem, [[20.0, 0.0]])[0]
    t = Tiny(n=12, grid=(4, 3, 2), dem=dem, z0=z0)
    root = t.random_root()
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*dem), t.tx, t.rx)
    ds.load_rootdata(root)
    want = oracle.backproject(t.bp(), dem, t.tx, t.rx, oracle.upsample(root), t.gx, t.gy, t.gz)
    assert np.abs(want).max() > 0
    img, _ = ds.backproject(t.gx, t.gy, t.gz, bsar.FP64)
    assert peak_err(img, want) <= FP64_TOL
    mixed, _ = ds.backproject(t.gx, t.gy, t.gz, bsar.MIXED)
    assert db(peak_err(mixed, want)) <= MIXED_DB


def test_degenerate_grids_and_state_errors(bsar, ctx, oracle):
    t = Tiny(n=8)
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    with pytest.raises(bsar.BsarError) as e:
        ds.backproject(t.gx, t.gy, t.gz)
    assert e.value.status == -7                 # no range data loaded yet
    root = t.random_root()
    ds.load_rootdata(root)
    one, _ = ds.backproject(t.gx[:1], t.gy[:1], t.gz[:1])
    want = oracle.backproject(t.bp(), t.dem, t.tx, t.rx, oracle.upsample(root), t.gx[:1], t.gy[:1], t.gz[:1])
    assert peak_err(one, want) <= FP64_TOL
    empty, rep = ds.backproject(t.gx[:0], t.gy, t.gz)
    assert empty.size == 0 and rep["voxels"] == 0
    with pytest.raises(bsar.BsarError):
        ds.backproject(t.gx, t.gy, t.gz, precision=7)


def test_small_workload_vs_oracle_subset(bsar, ctx, oracle, scene_mod):
    """41x41x21 grid, 256 pulses, 4096 samples: full image on the GPU, a
    strided voxel subset on the CPU oracle (the data comes from the oracle's
    restatement of processRawData)."""
    sc = scene_mod.make_scene("small")
    aux, raw, root = oracle.simulate(sc.sim_params(), sc.dem, sc.tx, sc.rx, sc.scat, sc.rcs)
    root = np.ascontiguousarray(root.T)
    ds = ctx.dataset(bsar.Params(sc.fc, sc.c, sc.fs, sc.n1, sc.n2, 4), bsar.Dem(*sc.dem), sc.tx, sc.rx)
    ds.load_rootdata(root)
    img, rep = ds.backproject(sc.gx, sc.gy, sc.gz, bsar.FP64)
    nx, ny, nz = len(sc.gx), len(sc.gy), len(sc.gz)
    # oracle voxel ids use the reference's z -> x -> y enumeration
    rng = np.random.default_rng(0)
    ids = np.sort(rng.choice(nx * ny * nz, 48, replace=False))
    want = oracle.backproject(sc.bp_params(), sc.dem, sc.tx, sc.rx, oracle.upsample(root), sc.gx, sc.gy, sc.gz,
                              subset=ids)
    yi, xi, zi = ids % ny, (ids // ny) % nx, ids // (nx * ny)
    g, w = img[zi, yi, xi], want[zi, yi, xi]
    assert np.abs(g - w).max() <= FP64_TOL * np.abs(img).max()
    mixed, mrep = ds.backproject(sc.gx, sc.gy, sc.gz, bsar.MIXED)
    assert db(peak_err(mixed, img)) <= MIXED_DB
    assert mrep["r1_exact_frac"] < 0.01


def test_linearity_full_grid(bsar, ctx):
    """Size-independent property on the full 41x41x21 grid: the image is
    linear in the range data, for both precisions."""
    t = Tiny(n=64, S=2048, grid=(41, 41, 21))
    d1, d2 = t.random_root(), t.random_root()
    a, b = 0.5 + 1.5j, -2.0 + 0.25j
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    out = {}
    for k, d in (("1", d1), ("2", d2), ("12", a * d1 + b * d2)):
        ds.load_rootdata(d)
        out[k], _ = ds.backproject(t.gx, t.gy, t.gz, bsar.FP64)
        out[k + "m"], _ = ds.backproject(t.gx, t.gy, t.gz, bsar.MIXED)
    assert peak_err(out["12"], a * out["1"] + b * out["2"]) <= 1e-12
    assert db(peak_err(out["12m"], a * out["1m"] + b * out["2m"])) <= MIXED_DB


@pytest.mark.parametrize("grid", [(1, 19, 11), (3, 9, 17), (2, 8, 8), (5, 17, 1)])
def test_voxel_bands_cover_ragged_grids(bsar, ctx, oracle, grid):
    """The kernels take their voxels in bands of 8 rows x 8 planes; grids with
    several bands and ragged last bands in y and z (and a single x) give every
    voxel of the image once: FP64 against the oracle on all voxels, mixed
    within -80 dB."""
    t = Tiny(n=8, S=256, grid=grid)
    t.gx = t.gx[0] + 0.01 * np.arange(grid[0])
    t.gy = t.gy[0] + 0.01 * np.arange(grid[1])
    t.gz = t.gz[0] + 0.02 * np.arange(grid[2])
    root = t.random_root()
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    ds.load_rootdata(root)
    img, _ = ds.backproject(t.gx, t.gy, t.gz, bsar.FP64)
    mixed, _ = ds.backproject(t.gx, t.gy, t.gz, bsar.MIXED)
    ds.close()
    want = oracle.backproject(t.bp(), t.dem, t.tx, t.rx, oracle.upsample(root), t.gx, t.gy, t.gz)
    assert np.abs(want).min() > 0
    assert peak_err(img, want) <= FP64_TOL
    assert db(peak_err(mixed, want)) <= MIXED_DB
