"""HIP kernels against the oracle and the reference's golden vectors.

Bit-exact for the DEM lookup, the refraction search, the slant-range
refraction points and the spline upsampling (integer/index work and pure
+ - * / sqrt in the reference's order).  Range compression runs a different
(exact-twiddle) FFT: FP64 within 1e-12 of the reference's output relative to
its peak, FP32 within 2e-5."""
import numpy as np
import pytest

from conftest import GOLDEN
from test_oracle_golden import DEMS, bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", DEMS)
def test_dem_interp_bit_exact(bsar, ctx, golden, name):
    g = golden
    dem = bsar.Dem(g[f"dem_{name}_x"], g[f"dem_{name}_y"], g[f"dem_{name}_z"])
    out = ctx.dem_interp(dem, g[f"dem_{name}_xy"])
    assert bits_equal(out, g[f"dem_{name}_out"])


@pytest.mark.parametrize("name", DEMS)
@pytest.mark.parametrize("tag", ["a", "b", "eq"])
def test_refraction_bit_exact(bsar, ctx, golden, name, tag):
    g = golden
    dem = bsar.Dem(g[f"dem_{name}_x"], g[f"dem_{name}_y"], g[f"dem_{name}_z"])
    n1, n2 = g[f"refr_{name}_{tag}_n"]
    out = ctx.refraction_points(dem, g[f"refr_{name}_{tag}_legs"], n1, n2)
    assert bits_equal(out, g[f"refr_{name}_{tag}_out"])


def test_refraction_empty_dem(bsar, ctx, golden):
    e = np.zeros(0)
    out = ctx.refraction_points(bsar.Dem(e, e, np.zeros((0, 0))), golden["refr_empty_legs"], 1.0, 2.0)
    assert bits_equal(out, golden["refr_empty_out"])


def test_refraction_random_legs_vs_oracle(bsar, ctx, oracle, scene_mod):
    rng = np.random.default_rng(42)
    dem = scene_mod.dem_bumps(rng)
    M = 4000
    top = np.stack([rng.uniform(-20, 20, M), rng.uniform(-20, 20, M), rng.uniform(10, 14, M)], 1)
    bot = np.stack([rng.uniform(19, 21, M), rng.uniform(-1, 1, M), rng.uniform(-6, -4, M)], 1)
    legs = np.concatenate([top, bot], 1)
    legs[M // 2:] = np.concatenate([bot, top], 1)[M // 2:]
    gpu = ctx.refraction_points(bsar.Dem(*dem), legs, 1.0, 2.449489742783178)
    ref = oracle.refraction_points(dem, legs, 1.0, 2.449489742783178)
    assert bits_equal(gpu, ref)


def test_rough_dem_refraction_vs_oracle(bsar, ctx, oracle, scene_mod):
    """167 x 100 fBm terrain (read through the cache, not LDS)."""
    rng = np.random.default_rng(1003)
    dem = scene_mod.dem_fbm(rng)
    M = 1500
    top = np.stack([rng.uniform(-15, 15, M), rng.uniform(-15, 15, M), rng.uniform(40, 60, M)], 1)
    bot = np.stack([rng.uniform(-2, 2, M), rng.uniform(-2, 2, M), rng.uniform(-30, -10, M)], 1)
    legs = np.concatenate([top, bot], 1)
    n2 = float(np.sqrt(rng.uniform(4, 9)))
    gpu = ctx.refraction_points(bsar.Dem(*dem), legs, 1.0, n2)
    ref = oracle.refraction_points(dem, legs, 1.0, n2)
    assert bits_equal(gpu, ref)


def test_rough_scene_slant_range_device_vs_oracle(bsar, ctx, oracle, scene_mod):
    """The rough workload's geometry (grid 3-7 m under 167 x 100 fBm terrain,
    antennas 12 m above it) on a sub-grid: the device-buffer entry point, the
    host entry point and the oracle give bit-identical refraction points for
    every (voxel, pulse, leg)."""
    sc = scene_mod.make_scene("rough", n_pulses=40, grid=(11, 9, 7))
    V, N = sc.voxels, sc.n
    nx, ny = len(sc.gx), len(sc.gy)
    v = np.arange(V)
    pts = np.stack([sc.gx[v % nx], sc.gy[(v // nx) % ny], sc.gz[v // (nx * ny)]], 1)
    ds = ctx.dataset(bsar.Params(sc.fc, sc.c, sc.fs, sc.n1, sc.n2), bsar.Dem(*sc.dem), sc.tx, sc.rx)
    bp = bsar.DeviceBuffer.from_array(pts)
    ba, bb = bsar.DeviceBuffer(V * N * 24), bsar.DeviceBuffer(V * N * 24)
    rep = ds.slant_range_fermat_device(bp.ptr, V, ba.ptr, bb.ptr)
    ida, volta = ba.read((V, N, 3)), bb.read((V, N, 3))
    for b in (bp, ba, bb):
        b.close()
    assert rep["voxel_pulses"] == V * N and rep["launches"] == 1 and rep["seconds"] > 0
    hida, hvolta = ds.slant_range_fermat(pts)
    ds.close()
    assert bits_equal(ida, hida) and bits_equal(volta, hvolta)
    P = np.repeat(pts, N, axis=0)
    legs_a = np.concatenate([np.tile(sc.tx, (V, 1)), P], 1)
    legs_b = np.concatenate([P, np.tile(sc.rx, (V, 1))], 1)
    assert bits_equal(ida.reshape(-1, 3), oracle.refraction_batch(sc.dem, legs_a, sc.n1, sc.n2, threads=8))
    assert bits_equal(volta.reshape(-1, 3), oracle.refraction_batch(sc.dem, legs_b, sc.n2, sc.n1, threads=8))


def test_slant_range_bit_exact(bsar, ctx, golden):
    g = golden
    p = bsar.Params(1e9, 3e8, 800e6, 1.0, 2.449489742783178)
    ds = ctx.dataset(p, bsar.Dem(g["dem_bumps_x"], g["dem_bumps_y"], g["dem_bumps_z"]), g["sr_tx"], g["sr_rx"])
    ida, volta = ds.slant_range_fermat(g["sr_P"][None, :])
    assert bits_equal(ida[0], g["sr_ida"])
    assert bits_equal(volta[0], g["sr_volta"])


@pytest.mark.parametrize("nf", [(2, 4), (3, 4), (5, 3), (64, 4), (257, 4), (16, 1), (9, 7)])
def test_spline_bit_exact(ctx, golden, nf):
    n, f = nf
    d = golden[f"spline_{n}_{f}_in"].view(np.complex128)
    out = ctx.upsample(d[None, :], f)
    assert bits_equal(out[0].view(np.float64), golden[f"spline_{n}_{f}_out"])


def test_upsample_batch_vs_oracle(ctx, oracle):
    rng = np.random.default_rng(5)
    root = rng.normal(size=(37, 1000)) + 1j * rng.normal(size=(37, 1000))
    assert bits_equal(ctx.upsample(root, 4).view(np.float64), oracle.upsample(root, 4).view(np.float64))


def test_upsample_more_pulses_than_grid_rows(ctx, oracle):
    # the kernel's grid holds at most 65535 pulses along y and loops over the rest
    rng = np.random.default_rng(6)
    root = rng.normal(size=(70001, 3)) + 1j * rng.normal(size=(70001, 3))
    assert bits_equal(ctx.upsample(root, 3).view(np.float64), oracle.upsample(root, 3).view(np.float64))


def test_upsample_rejects_oversized_columns(bsar, ctx):
    # S * factor >= 2^30 is refused before any work (no pulses: nothing to allocate)
    with pytest.raises(bsar.BsarError):
        ctx.upsample(np.zeros((0, 1 << 15), np.complex128), 1 << 15)


def test_fused_fp32_copy_matches_conversion(bsar, ctx, oracle):
    # load_rootdata writes the complex64 copy in the spline's pass;
    # load_upsampled converts the complex128 samples in a pass of its own:
    # the mixed images (which read only the complex64 copy) are bit-identical
    from _scenes import Tiny
    t = Tiny(n=12, S=256, grid=(5, 4, 3))
    root = t.random_root()
    imgs = []
    for fused in (True, False):
        ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
        if fused:
            ds.load_rootdata(root)
        else:
            ds.load_upsampled(oracle.upsample(root, t.up))
        imgs.append(ds.backproject(t.gx, t.gy, t.gz, bsar.MIXED)[0])
        ds.close()
    assert bits_equal(imgs[0].view(np.float64), imgs[1].view(np.float64))


@pytest.mark.parametrize("S", [2, 8, 64, 1024, 8192, 16384, 32768])   # LDS, split (16384) and global (32768) paths
@pytest.mark.parametrize("conj", [True, False])
def test_range_compress_fp64(ctx, oracle, S, conj):
    rng = np.random.default_rng(S + conj)
    N = 6
    sig = rng.normal(size=(N, S)) + 1j * rng.normal(size=(N, S))
    ref = rng.normal(size=S) + 1j * rng.normal(size=S)
    gpu = ctx.range_compress(sig, ref, conj=conj)
    f = oracle.fft_correlation if conj else oracle.fft_convolution
    want = np.stack([f(s, ref) for s in sig])
    # The reference's twiddles come from repeated multiplication w *= wlen
    # (fft_utils.cpp:40), whose rounding grows linearly with the block length:
    # the oracle, which keeps that recurrence, sits about 3e-17 * S of the peak
    # away from an exactly-twiddled transform (1e-12 at S = 32768).  The GPU
    # computes each twiddle directly, so it is held to max(1e-12, 1e-16 * S)
    # of the oracle and to 1e-13 of numpy's transform.
    assert np.abs(gpu - want).max() <= max(1e-12, 1e-16 * S) * np.abs(want).max()
    H = np.fft.fft(ref)
    exact = np.fft.ifft(np.fft.fft(sig, axis=1) * (np.conj(H) if conj else H), axis=1)
    assert np.abs(gpu - exact).max() <= 1e-13 * np.abs(exact).max()


@pytest.mark.parametrize("S", [16, 4096, 16384, 32768, 65536])   # LDS, split (32768) and global (65536) paths
def test_range_compress_fp32(bsar, ctx, oracle, S):
    rng = np.random.default_rng(S)
    N = 4
    sig = rng.normal(size=(N, S)) + 1j * rng.normal(size=(N, S))
    ref = rng.normal(size=S) + 1j * rng.normal(size=S)
    gpu = ctx.range_compress(sig, ref, conj=True, precision=bsar.MIXED)
    want = np.stack([oracle.fft_correlation(s, ref) for s in sig])
    assert np.abs(gpu - want).max() <= 2e-5 * np.abs(want).max()


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_range_compress_16384_many_columns(bsar, ctx, form):
    """16384-sample columns run in one persistent workgroup per CU that loops
    over the columns (and, complex64, loads the next column while finishing
    the current one): 601 columns, more than the GPU's CUs and not a multiple
    of them, against numpy's transform column by column."""
    rng = np.random.default_rng(601)
    S, N = 16384, 601
    sig = rng.normal(size=(N, S)) + 1j * rng.normal(size=(N, S))
    ref = rng.normal(size=S) + 1j * rng.normal(size=S)
    if form == "fp64":
        b = bsar.DeviceBuffer.from_array(sig)
        ctx.range_compress_device(b, S, N, ref, conj=True, precision=bsar.FP64)
        gpu, base, tol = b.read((N, S), np.complex128), sig, 1e-13
    else:
        s32 = sig.astype(np.complex64)
        b = bsar.DeviceBuffer.from_array(s32)
        ctx.range_compress_device(b, S, N, ref, conj=True, precision=bsar.MIXED)
        gpu, base, tol = b.read((N, S), np.complex64), s32.astype(np.complex128), 2e-6
    b.close()
    exact = np.fft.ifft(np.fft.fft(base, axis=1) * np.conj(np.fft.fft(ref))[None, :], axis=1)
    err = np.abs(gpu - exact).max(axis=1) / np.abs(exact).max()
    assert err.max() <= tol, (int(err.argmax()), float(err.max()))


@pytest.mark.parametrize("n", [1, 7, 12])
def test_range_compress_non_power_of_two(ctx, golden, n):
    """The reference pads to the next power of two and keeps the first n
    (length 1: its transforms are the identity)."""
    a = golden[f"fft_{n}_in"].view(np.complex128)
    b = golden[f"corr_{n}_ref"].view(np.complex128)
    gpu = ctx.range_compress(a[None, :], b, conj=True)[0]
    want = golden[f"corr_{n}_out"].view(np.complex128)
    assert np.abs(gpu - want).max() <= 1e-12 * np.abs(want).max()
    gpu = ctx.range_compress(a[None, :], b, conj=False)[0]
    want = golden[f"conv_{n}_out"].view(np.complex128)
    assert np.abs(gpu - want).max() <= 1e-12 * np.abs(want).max()


def test_pulse_compression_of_reference_simulation(bsar, ctx, golden):
    """processRawData's last step: rootData = corr(rawData, reference)."""
    g = golden
    fc, c, fs_rx, fs_tx, f0, f1, T, n1, n2 = g["sim_args"]
    S = int(g["sim_S"][0])
    raw = (g["sim_raw"][..., 0] + 1j * g["sim_raw"][..., 1]).T        # [pulse][range]
    root = (g["sim_root"][..., 0] + 1j * g["sim_root"][..., 1]).T
    chirp = bsar.chirp(f0, f1, T, fs_tx)
    L = len(chirp)
    ref = np.zeros(S, complex)
    ref[: L - L // 2] = chirp[L // 2:]
    ref[S - L // 2:] = chirp[: L // 2]
    gpu = ctx.range_compress(raw, ref, conj=True)
    assert np.abs(gpu - root).max() <= 1e-12 * np.abs(root).max()


@pytest.mark.parametrize("S", [4096, 16384])
def test_range_compress_device_matches_host(bsar, ctx, S):
    """The device-buffer entry point runs the host entry point's kernels: same
    bits in FP64 (complex128 in place) and FP32 (complex64 in place)."""
    rng = np.random.default_rng(S + 3)
    N = 5
    sig = rng.normal(size=(N, S)) + 1j * rng.normal(size=(N, S))
    ref = rng.normal(size=S) + 1j * rng.normal(size=S)
    b = bsar.DeviceBuffer.from_array(sig)
    rep = ctx.range_compress_device(b, S, N, ref, conj=True, precision=bsar.FP64)
    assert bits_equal(b.read((N, S), np.complex128).view(np.float64),
                      ctx.range_compress(sig, ref, conj=True).view(np.float64))
    assert rep["seconds"] > 0 and rep["pulses"] == N
    b.close()
    s32 = sig.astype(np.complex64)
    b = bsar.DeviceBuffer.from_array(s32)
    ctx.range_compress_device(b, S, N, ref, conj=False, precision=bsar.MIXED)
    host = ctx.range_compress(s32.astype(np.complex128), ref, conj=False, precision=bsar.MIXED)
    assert bits_equal(b.read((N, S), np.complex64).view(np.float32), host.astype(np.complex64).view(np.float32))
    b.close()


def test_load_rootdata_device_matches_host(bsar, ctx):
    from _scenes import Tiny
    t = Tiny(n=12, S=256, grid=(5, 4, 3))
    root = t.random_root()
    a = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    b = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    a.load_rootdata(root)
    buf = bsar.DeviceBuffer.from_array(root)
    b.load_rootdata_device(buf, t.S, len(root), bsar.FP64)
    buf.close()
    for p in (bsar.FP64, bsar.MIXED):
        assert bits_equal(a.backproject(t.gx, t.gy, t.gz, p)[0].view(np.float64),
                          b.backproject(t.gx, t.gy, t.gz, p)[0].view(np.float64))
    r32 = root.astype(np.complex64)
    a.load_rootdata(r32.astype(np.complex128))
    buf = bsar.DeviceBuffer.from_array(r32)
    b.load_rootdata_device(buf, t.S, len(root), bsar.MIXED)
    buf.close()
    # the complex64 samples are widened exactly inside the spline: both forms
    # match the host load of the same samples as complex128
    for p in (bsar.FP64, bsar.MIXED):
        assert bits_equal(a.backproject(t.gx, t.gy, t.gz, p)[0].view(np.float64),
                          b.backproject(t.gx, t.gy, t.gz, p)[0].view(np.float64))
    a.close()
    b.close()


def test_process_raw_matches_reference(bsar, ctx, golden):
    """processRawData of the compiled reference (one target, rcs 1, its radar
    constants, S = 256): same delay bins, responses within the last bit of
    cos/sin, raw and rootData within 1e-12 of their peaks."""
    g = golden
    fc, c, fs_rx, fs_tx, f0, f1, T, n1, n2 = g["sim_args"]
    S = int(g["sim_S"][0])
    ds = ctx.dataset(bsar.Params(fc, c, fs_rx, n1, n2), bsar.Dem(g["dem_bumps_x"], g["dem_bumps_y"], g["dem_bumps_z"]),
                     g["sr_tx"], g["sr_rx"])
    out = ds.process_raw(g["sr_P"][None, :], [1.0], (f0, f1, T, fs_tx), S)
    ds.close()
    cplx = lambda a: (a[..., 0] + 1j * a[..., 1]).T            # [S][n] -> [n][S]
    aux, raw, root = cplx(g["sim_aux"]), cplx(g["sim_raw"]), cplx(g["sim_root"])
    assert np.array_equal(out["aux"] != 0, aux != 0) and np.count_nonzero(aux) == aux.shape[0]
    assert np.abs(out["aux"] - aux).max() <= 4e-16
    for k, want in (("raw", raw), ("root", root)):
        assert np.abs(out[k] - want).max() <= 1e-12 * np.abs(want).max()


def test_process_raw_several_scatterers_vs_oracle(bsar, ctx, oracle):
    from _scenes import Tiny
    t = Tiny(n=20, S=1024)
    rng = np.random.default_rng(17)
    pts = np.stack([rng.uniform(19.8, 20.3, 5), rng.uniform(-0.2, 0.2, 5), rng.uniform(-5.3, -4.8, 5)], 1)
    rcs = rng.normal(size=5) + 1j * rng.normal(size=5)
    sim = dict(fc=t.fc, c=t.c, fs=t.fs, f0=-150e6, f1=150e6, T=1e-7, fs_chirp=t.fs, n1=t.n1, n2=t.n2, S=t.S)
    aux, raw, root = oracle.simulate(sim, t.dem, t.tx, t.rx, pts, rcs)
    ds = ctx.dataset(t.params(bsar), bsar.Dem(*t.dem), t.tx, t.rx)
    out = ds.process_raw(pts, rcs, (-150e6, 150e6, 1e-7, t.fs), t.S)
    ds.close()
    for k, want in (("aux", aux), ("raw", raw), ("root", root)):
        want = np.ascontiguousarray(want.T)
        assert np.abs(out[k] - want).max() <= 1e-12 * np.abs(want).max(), k
    assert np.array_equal(out["aux"] != 0, aux.T != 0)


def test_scene_forward_model_vs_oracle(bsar, ctx, oracle, scene_mod):
    """The workloads' data path (scene.simulate over bsar_process_raw): 64
    scatterers, B = 500 MHz, against the oracle's processRawData restatement."""
    sc = scene_mod.make_scene("wide", n_pulses=16, samples=2048)
    root = scene_mod.simulate(ctx, sc)
    _, raw_o, root_o = oracle.simulate(sc.sim_params(), sc.dem, sc.tx, sc.rx, sc.scat, sc.rcs)
    want = np.ascontiguousarray(root_o.T)
    assert root.shape == want.shape
    assert np.abs(root - want).max() <= 1e-12 * np.abs(want).max()
    raw = scene_mod.simulate(ctx, sc, compress=False)
    assert np.abs(raw - raw_o.T).max() <= 1e-12 * np.abs(raw_o).max()


def test_context_dem_cache_follows_the_dem(bsar, ctx, oracle, golden):
    """The context keeps the last DEM on the device and reuses it only for
    identical arrays: alternating DEMs of the same sizes (and one changed in
    a single height) give each DEM's own bit-exact refraction points and
    heights."""
    g = golden
    a = tuple(np.ascontiguousarray(g[f"dem_bumps_{k}"]) for k in "xyz")
    b = (a[0], a[1], np.ascontiguousarray(a[2] + 0.75))
    c = (a[0], a[1], a[2].copy())
    c[2][3, 4] += 1e-9
    legs = g["refr_bumps_a_legs"]
    n1, n2 = g["refr_bumps_a_n"]
    xy = g["dem_bumps_xy"]
    for dem in (a, b, a, c, a, b):
        assert bits_equal(ctx.refraction_points(bsar.Dem(*dem), legs, n1, n2), oracle.refraction_points(dem, legs, n1, n2))
        assert bits_equal(ctx.dem_interp(bsar.Dem(*dem), xy), oracle.dem_interp(dem, xy))
    e = np.zeros(0)
    empty = (e, e, np.zeros((0, 0)))
    assert bits_equal(ctx.refraction_points(bsar.Dem(*empty), g["refr_empty_legs"], 1.0, 2.0), g["refr_empty_out"])
    assert bits_equal(ctx.refraction_points(bsar.Dem(*a), legs, n1, n2), g["refr_bumps_a_out"])


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_refraction_with_non_finite_heights_vs_oracle(bsar, ctx, oracle, scene_mod, bad):
    """DEM nodes that are NaN or infinite next to the refraction region: the
    reference's strict '<' never selects a NaN path and the screening must not
    drop the candidate that wins; refraction points and images bit-exact /
    within 1e-9 of the oracle."""
    rng = np.random.default_rng(11)
    x, y, z = scene_mod.dem_bumps(rng)
    z = z.copy()
    ix, iy = np.searchsorted(x, 20.0), np.searchsorted(y, 0.0)
    for dj, di in ((0, 0), (0, -1), (1, 1), (-1, 2)):
        z[iy + dj, ix + di] = bad
    dem = (x, y, z)
    M = 1500
    top = np.stack([rng.uniform(-15, 15, M), rng.uniform(-15, 15, M), rng.uniform(10, 14, M)], 1)
    bot = np.stack([rng.uniform(-10, 50, M), rng.uniform(-30, 30, M), rng.uniform(-6, -4, M)], 1)
    legs = np.concatenate([top, bot], 1)
    legs[M // 2:] = np.concatenate([bot, top], 1)[M // 2:]
    n2 = 2.449489742783178
    assert bits_equal(ctx.refraction_points(bsar.Dem(*dem), legs, 1.0, n2), oracle.refraction_points(dem, legs, 1.0, n2))
