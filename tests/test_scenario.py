"""Scenario generators and DEM files (host code, no GPU): bsar_spiral
(funcao_espiral / generate{Transmitter,Receiver}Trajectory), bsar_grid_axis
(crearGrid), bsar_dem_load (loadDEM of the build without GDAL, and GeoTIFF
decoding), bsar_geo_to_utm / bsar_meridian_arc.

Pins: the compiled reference (oracle/_ref, when built here) bit for bit; the
oracle's scalar restatements; the file set the reference's own
gs_test_script writes for its default scenario (tests/golden/
ref_scenario.npz, byte-identical CSV text); the reference's GeoTIFF assets
decoded by an independent TIFF library (PIL/libtiff) and GeoTIFFs that
library encodes; Snyder's transverse Mercator series and numerical
quadrature of the meridian arc for the projection.
"""
import hashlib
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

REF_SPIRAL = (2, 172.5, 147.5, 120.0, 80.0, 120.0, 400.0)          # radar_init.cpp:11-16
SPIRALS = [REF_SPIRAL, (3, 20.0, 30.0, 15.0, 10.0, 25.0, 200.0), (0.5, 100.0, 50.0, 10.0, 20.0, 5.0, 1000.0),
           (1.0, 10.0, 10.0, 5.0, 5.0, 3.0, 100.0), (4, 30.0, 5.0, 50.0, 40.0, 60.0, 1234.5)]
DEM_DIR = GOLDEN / "dem"
ASSETS = sorted(DEM_DIR.glob("*.tif"))


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def ref(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("compiled reference not built here (oracle/_ref)")
    return oracle


@pytest.fixture(scope="module")
def scenario():
    return np.load(GOLDEN / "ref_scenario.npz")


# ---- trajectories ----------------------------------------------------------

@pytest.mark.parametrize("args", SPIRALS)
def test_spiral_matches_compiled_reference(bsar, ref, args):
    want = ref.ref_spiral(*args)
    assert np.array_equal(bits(bsar.spiral(*args)), bits(want))
    assert np.array_equal(bits(ref.spiral(*args)), bits(want))       # the oracle restatement is pinned too


def test_default_trajectories_match_compiled_reference(bsar, ref):
    """generateTransmitterTrajectory(initializeRadarTx()) and the receiver's."""
    assert np.array_equal(bits(bsar.spiral(*REF_SPIRAL)), bits(ref.ref_default_trajectory(False)))
    assert np.array_equal(bits(bsar.spiral(*REF_SPIRAL, mirror_xy=True)), bits(ref.ref_default_trajectory(True)))


@pytest.mark.parametrize("args", SPIRALS)
def test_spiral_matches_oracle(bsar, oracle, args):
    assert np.array_equal(bits(bsar.spiral(*args)), bits(oracle.spiral(*args)))
    assert np.array_equal(bits(bsar.spiral(*args, mirror_xy=True)), bits(oracle.spiral(*args, mirror=True)))


def test_spiral_geometry(bsar):
    """Closed forms of funcao_espiral: starts at (R_major, 0, H_high); the
    radius falls linearly with time, the height too; theta grows as
    (Vt/Vr) log(R_start / R(t)); T = (R_start - R_end) / Vr."""
    N, rmin, rmaj, hh, hl, vt, prf = REF_SPIRAL
    p = bsar.spiral(*REF_SPIRAL)
    vr = vt * (np.log(rmaj) - np.log(rmin)) / (N * 2 * np.pi)
    T = (rmaj - rmin) / vr
    assert abs(len(p) - (np.floor(T * prf) + 1)) <= 1
    assert np.allclose(p[0], [rmaj, 0.0, hh], atol=1e-12)
    t = np.arange(len(p)) / prf
    R = rmaj - vr * t
    assert np.allclose(np.hypot(p[:, 0], p[:, 1]), R, rtol=0, atol=1e-9)
    assert np.allclose(p[:, 2], hh - (hh - hl) / T * t, rtol=0, atol=1e-9)
    theta = np.unwrap(np.arctan2(p[:, 1], p[:, 0]))
    assert np.allclose(theta, (vt / vr) * np.log(rmaj / R), atol=1e-9)
    rx = bsar.spiral(*REF_SPIRAL, mirror_xy=True)
    assert np.array_equal(rx[:, :2], -p[:, :2]) and np.array_equal(rx[:, 2], p[:, 2])


def test_spiral_empty_and_errors(bsar):
    assert bsar.spiral(1.0, 10.0, 10.0, 5.0, 5.0, 3.0, 100.0).shape == (0, 3)   # equal radii: T is NaN
    with pytest.raises(bsar.BsarError, match="does not end"):
        bsar.spiral(2, 172.5, 147.5, 120.0, 80.0, 120.0, 0.0)                   # PRF 0: endless loop in C++
    with pytest.raises(bsar.BsarError, match="does not end"):
        bsar.spiral(2, 172.5, 147.5, 120.0, 80.0, 0.0, 400.0)                   # Vt 0: T = inf


# ---- grid axes ---------------------------------------------------------------

GRIDS = [((20.0, 0.0, -5.0), (0.04, 0.8, 0.08, 0.8)), ((1.3, -2.7, 0.1), (0.1, 1.0, 0.03, 0.31)),
         ((-0.5, 7.25, 3.0), (0.25, 0.0, 0.5, 2.0)), ((0.0, 0.0, 0.0), (0.3, 1.0, 0.7, 0.7))]


@pytest.mark.parametrize("pos,d", GRIDS)
def test_grid_matches_compiled_reference(bsar, ref, pos, d):
    rx, ry, rz = ref.ref_grid(pos, *d)
    assert np.array_equal(bits(bsar.grid_axis(pos[0], d[0], d[1])), bits(rx))
    assert np.array_equal(bits(bsar.grid_axis(pos[1], d[0], d[1])), bits(ry))
    assert np.array_equal(bits(bsar.grid_axis(pos[2], d[2], d[3])), bits(rz))
    assert np.array_equal(bits(ref.grid_axis(pos[2], d[2], d[3])), bits(rz))


@pytest.mark.parametrize("pos,d", GRIDS)
def test_grid_matches_oracle(bsar, oracle, pos, d):
    for c, (step, half) in zip(pos, [(d[0], d[1]), (d[0], d[1]), (d[2], d[3])]):
        got = bsar.grid_axis(c, step, half)
        assert np.array_equal(bits(got), bits(oracle.grid_axis(c, step, half)))
        # MATLAB c-half:step:c+half up to the running sum's rounding
        k = np.arange(len(got))
        assert np.allclose(got, c - half + k * step, atol=1e-12 * (1 + len(got)))
        assert got[-1] <= c + half + step / 2 < got[-1] + step + 1e-12


def test_grid_errors(bsar):
    for step in (0.0, -0.1, float("nan"), float("inf")):
        with pytest.raises(bsar.BsarError):
            bsar.grid_axis(0.0, step, 1.0)


# ---- loadDEM (build without GDAL) -------------------------------------------

ESRI = {
    "plain": "ncols 4\nnrows 3\nxllcorner -30\nyllcorner 10\ncellsize 15\nNODATA_value -9999\n"
             "1 2 3 4\n5 6 7 8\n9 10 11 12\n",
    "short_rows": "ncols 4\nnrows 3\nxllcorner 0\nyllcorner 0\ncellsize 2.5\nnodata_value -1\n"
                  "1 2\n5 6 7 8 99\n\n",
    "bad_tokens": "ncols 3\nnrows 2\ncellsize 1\n1 x 3\n4.5e1 -2 7abc\n",
    "blank_in_header": "ncols 2\nnrows 2\n\n1 2\n3 4\n",
    "extra_rows": "nrows 2\nncols 2\nyllcorner -5\nxllcorner 3.25\ncellsize 0.5\n1 2\n3 4\n5 6\n7 8\n",
    "no_header": "1 2 3\n4 5 6\n",
}


def _load_ref_mode(bsar, path):
    dem, meta = bsar.load_dem(path, bsar.DEM_FILE_REFERENCE)
    return dem, meta


@pytest.mark.parametrize("case", sorted(ESRI))
def test_esri_grid_matches_compiled_reference(bsar, ref, tmp_path, case):
    f = tmp_path / f"{case}.asc"
    f.write_text(ESRI[case])
    dem, meta = _load_ref_mode(bsar, f)
    x, y, z, n = ref.ref_load_dem(f)
    assert meta["source"] == "ascii_grid"
    assert np.array_equal(bits(dem.x), bits(x)) and np.array_equal(bits(dem.y), bits(y))
    assert dem.z.shape == z.shape and np.array_equal(bits(dem.z), bits(z))


def test_esri_grid_values(bsar, tmp_path):
    """load_dem_simple.cpp:136-214 by hand: rows in file order at
    yllcorner + j cellsize, missing or unparseable values read as NODATA."""
    f = tmp_path / "a.asc"
    f.write_text(ESRI["plain"])
    dem, _ = _load_ref_mode(bsar, f)
    assert np.array_equal(dem.x, [-30.0, -15.0, 0.0, 15.0]) and np.array_equal(dem.y, [10.0, 25.0, 40.0])
    assert np.array_equal(dem.z, np.arange(1, 13, dtype=float).reshape(3, 4))
    f.write_text(ESRI["short_rows"])
    dem, _ = _load_ref_mode(bsar, f)
    assert np.array_equal(dem.z, [[1, 2, -1, -1], [5, 6, 7, 8], [-1, -1, -1, -1]])
    f.write_text(ESRI["bad_tokens"])
    dem, _ = _load_ref_mode(bsar, f)
    assert np.array_equal(dem.z, [[1, -9999, -9999], [45.0, -2, 7]])


@pytest.mark.parametrize("name", ["assets/DEM_1x1km_Res30m_Lat-3_6160_Lon-80_4552.tif", "/nonexistent/dem.asc",
                                  "x.TIF"])
def test_default_dem(bsar, oracle, name):
    """A .tif name or a file that cannot be opened gives the flat 32 x 32 DEM
    over [-500, 500] m (load_dem_simple.cpp:16-134)."""
    dem, meta = _load_ref_mode(bsar, name)
    assert meta["source"] == "default"
    assert dem.z.shape == (32, 32) and not dem.z.any()
    assert np.array_equal(dem.x, -500.0 + np.arange(32) * (1000.0 / 31)) and np.array_equal(dem.x, dem.y)
    if oracle.ref_lib() is not None:
        x, y, z, _ = oracle.ref_load_dem(name)
        assert np.array_equal(bits(dem.x), bits(x)) and np.array_equal(bits(dem.y), bits(y))
        assert np.array_equal(bits(dem.z), bits(z))


# ---- the reference's default scenario, file by file -----------------------

def test_reference_scenario_text_files(bsar, scene_mod, scenario, tmp_path):
    """The CSV/text files of bsar_sim --workload reference against those the
    reference's gs_test_script writes: byte-identical (trajectories by
    SHA-256).  Covers the spirals, crearGrid, loadDEM's default DEM, the
    environment defaults and the six-decimal CSV writer."""
    sc = scene_mod.reference_scene()
    assert sc.n == int(scenario["n_pulses"])
    root = np.zeros((sc.n, sc.S), np.complex128)
    scene_mod.write_reference_inputs(tmp_path, sc, root)
    for key in scenario.files:
        if key.startswith("text_"):
            name = key[5:]
            assert (tmp_path / name).read_bytes() == scenario[key].tobytes(), name
        elif key.startswith("sha256_"):
            name = key[7:]
            assert hashlib.sha256((tmp_path / name).read_bytes()).digest() == scenario[key].tobytes(), name
    hdr = (tmp_path / "rootData_complex.bin").read_bytes()[:44]
    assert hdr == scenario["rootData_header"].tobytes()      # same header: dims (S, N)


def test_facade_scenario_program(bsar, tmp_path):
    """gs_test_script's set-up calls (generate*Trajectory, funcao_espiral,
    crearGrid, loadDEM) compiled against bsar_facade.hpp give the same bits as
    the C ABI."""
    exe = tmp_path / "facade_scenario"
    pkg = ROOT / "c-_bistatic_sar_amd"
    r = subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "facade_scenario.cpp"),
                        "-o", str(exe), f"-L{pkg}", "-lbsar", f"-Wl,-rpath,{pkg}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe), str(ASSETS[0]), "geotiff"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = {ln.split()[0]: ln.split()[1:] for ln in out.stdout.splitlines()}
    vec = lambda k: np.array([float(v) for v in rows[k][1:]])
    tx = bsar.spiral(*REF_SPIRAL)
    rx = bsar.spiral(*REF_SPIRAL, mirror_xy=True)
    for k, col in (("txx", tx[:, 0]), ("txy", tx[:, 1]), ("txz", tx[:, 2]), ("rxx", rx[:, 0]), ("rxy", rx[:, 1])):
        assert np.array_equal(bits(vec(k)), bits(col)), k
    assert np.array_equal(vec("spx"), bsar.spiral(3, 20.0, 30.0, 15.0, 10.0, 25.0, 200.0)[:, 0])
    assert np.array_equal(vec("gx"), bsar.grid_axis(20.0, 0.04, 0.8))
    assert np.array_equal(vec("gz"), bsar.grid_axis(-5.0, 0.08, 0.8))
    assert rows["mesh"][:3] == ["41", "41", "21"]
    dem, _ = bsar.load_dem(ASSETS[0], bsar.DEM_FILE_GEOTIFF)
    assert np.array_equal(vec("demx"), dem.x) and np.array_equal(vec("demz"), dem.z.ravel())


# ---- GeoTIFF ------------------------------------------------------------------

def _pil():
    return pytest.importorskip("PIL.Image")


@pytest.mark.parametrize("path", ASSETS, ids=[p.name for p in ASSETS])
def test_geotiff_assets_match_independent_decoder(bsar, path):
    """The reference's DEM assets (1" WGS84, int16, LZW, one 256 x 256 tile)
    decode to the same heights as PIL/libtiff; rows reversed so Y ascends."""
    Image = _pil()
    dem, meta = bsar.load_dem(path, bsar.DEM_FILE_GEOTIFF)
    want = np.array(Image.open(path)).astype(np.float64)[::-1]
    assert meta["source"] == "geotiff" and meta["nodata_count"] == 0
    assert np.array_equal(dem.z, want)
    assert np.all(np.diff(dem.x) > 0) and np.all(np.diff(dem.y) > 0)


def _geotags(scale, tie, model=2, raster=1, nodata=None):
    from PIL import TiffImagePlugin, TiffTags
    ifd = TiffImagePlugin.ImageFileDirectory_v2()
    ifd[33550] = tuple(map(float, scale)) + (0.0,)
    ifd.tagtype[33550] = TiffTags.DOUBLE
    ifd[33922] = (0.0, 0.0, 0.0, float(tie[0]), float(tie[1]), 0.0)
    ifd.tagtype[33922] = TiffTags.DOUBLE
    ifd[34735] = (1, 1, 0, 2, 1024, 0, 1, model, 1025, 0, 1, raster)
    ifd.tagtype[34735] = TiffTags.SHORT
    if nodata is not None:
        ifd[42113] = str(nodata)
        ifd.tagtype[42113] = TiffTags.ASCII
    return ifd


CODECS = ["raw", "tiff_lzw", "tiff_adobe_deflate", "packbits"]


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("pred", [1, 2, 3])
@pytest.mark.parametrize("kind", ["f32", "u8", "u16", "i32"])
def test_geotiff_codecs_round_trip(bsar, tmp_path, codec, pred, kind):
    """Rasters encoded by libtiff (through PIL) with each compression and
    predictor decode to the written samples; projected metres are kept and
    centred."""
    Image = _pil()
    if pred == 3 and kind != "f32":
        pytest.skip("the floating-point predictor is defined for float samples only")
    rng = np.random.default_rng(hash((codec, pred, kind)) % 2 ** 32)
    shape = (23, 37)
    if kind == "f32":
        a = rng.normal(100, 30, shape).astype(np.float32)
        im = Image.fromarray(a)
    elif kind == "u8":
        a = rng.integers(0, 256, shape).astype(np.uint8)
        im = Image.fromarray(a)
    elif kind == "u16":
        a = rng.integers(0, 65536, shape).astype(np.uint16)
        im = Image.fromarray(a)
    else:
        a = rng.integers(-2 ** 31, 2 ** 31, shape).astype(np.int32)
        im = Image.fromarray(a)
    tags = _geotags((30.0, 30.0), (500000.0, 8600000.0), model=1)
    tags[317] = pred
    f = tmp_path / "t.tif"
    try:
        im.save(f, compression=codec, tiffinfo=tags)
    except OSError:
        pytest.skip("libtiff does not write this combination")
    dem, meta = bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)
    assert np.array_equal(dem.z, a.astype(np.float64)[::-1])
    assert np.allclose(dem.x, 30.0 * (np.arange(37) - 18.0)) and np.allclose(dem.y, 30.0 * (np.arange(23) - 11.0))
    assert meta["e0"] == 500000.0 + 15.0 + 18 * 30.0 and meta["utm_zone"] == 0


def _write_tiff(path, a, be=False, tile=None, deflate=False, scale=(1.0, 1.0), tie=(0.0, 0.0), model=1):
    """Minimal TIFF writer (big-endian and tiled layouts PIL does not write)."""
    o = ">" if be else "<"
    H, W = a.shape
    dt = a.dtype.newbyteorder(">" if be else "<")
    fmt = {"f": 3, "i": 2, "u": 1}[a.dtype.kind]
    if tile:
        tw, th = tile
        chunks = []
        for ty in range(0, H, th):
            for tx in range(0, W, tw):
                blk = np.zeros((th, tw), a.dtype)
                sub = a[ty:ty + th, tx:tx + tw]
                blk[:sub.shape[0], :sub.shape[1]] = sub
                chunks.append(blk.astype(dt).tobytes())
    else:
        chunks = [a[r:r + 5].astype(dt).tobytes() for r in range(0, H, 5)]
    if deflate:
        chunks = [zlib.compress(c) for c in chunks]
    data = b"".join(chunks)
    offs, pos = [], 8
    for c in chunks:
        offs.append(pos)
        pos += len(c)
    extra = bytearray()
    entries = []
    base = 8 + len(data)

    def add(tag, typ, vals):
        nonlocal extra
        code = {3: "H", 4: "I", 12: "d"}[typ]
        raw = struct.pack(o + code * len(vals), *vals)
        if len(raw) <= 4:
            entries.append(struct.pack(o + "HHI", tag, typ, len(vals)) + raw.ljust(4, b"\0"))
        else:
            entries.append((tag, typ, len(vals), raw))

    add(256, 4, [W]); add(257, 4, [H]); add(258, 3, [a.dtype.itemsize * 8]); add(259, 3, [8 if deflate else 1])
    add(262, 3, [1]); add(277, 3, [1]); add(339, 3, [fmt])
    if tile:
        add(322, 4, [tile[0]]); add(323, 4, [tile[1]]); add(324, 4, offs); add(325, 4, [len(c) for c in chunks])
    else:
        add(273, 4, offs); add(278, 4, [5]); add(279, 4, [len(c) for c in chunks])
    add(33550, 12, [scale[0], scale[1], 0.0]); add(33922, 12, [0.0, 0.0, 0.0, tie[0], tie[1], 0.0])
    add(34735, 3, [1, 1, 0, 1, 1024, 0, 1, model])
    n = len(entries)
    ifd_off = base
    ext_off = ifd_off + 2 + 12 * n + 4
    out = bytearray()
    for e in entries:
        if isinstance(e, tuple):
            tag, typ, cnt, raw = e
            out += struct.pack(o + "HHII", tag, typ, cnt, ext_off + len(extra))
            extra += raw
        else:
            out += e
    hdr = (b"MM" if be else b"II") + struct.pack(o + "HI", 42, ifd_off)
    blob = hdr + data + struct.pack(o + "H", n) + bytes(out) + struct.pack(o + "I", 0) + bytes(extra)
    Path(path).write_bytes(blob)


@pytest.mark.parametrize("be", [False, True])
@pytest.mark.parametrize("tile", [None, (16, 16)])
@pytest.mark.parametrize("deflate", [False, True])
@pytest.mark.parametrize("dtype", ["<i2", "<f8", "<u4"])
def test_geotiff_layouts(bsar, tmp_path, be, tile, deflate, dtype):
    rng = np.random.default_rng(7)
    a = (rng.normal(0, 1000, (21, 35))).astype(dtype)
    f = tmp_path / "l.tif"
    _write_tiff(f, a, be=be, tile=tile, deflate=deflate, scale=(2.0, 3.0), tie=(100.0, 200.0))
    dem, meta = bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)
    assert np.array_equal(dem.z, a.astype(np.float64)[::-1])
    assert np.allclose(dem.x, 2.0 * (np.arange(35) - 17.0)) and np.allclose(dem.y, 3.0 * (np.arange(21) - 10.0))


def test_geotiff_geographic_axes(bsar, oracle):
    """Geographic rasters: x_i = N(lat_c) cos(lat_c) (lon_i - lon_c),
    y_j = M(lat_j) - M(lat_c) (meridian arc by quadrature), the centre's UTM
    coordinates by Snyder's series (to 5 mm)."""
    dem, meta = bsar.load_dem(ASSETS[-1], bsar.DEM_FILE_GEOTIFF)
    a, f = oracle.WGS84
    e2 = f * (2 - f)
    lat_c, lon_c = meta["lat0"], meta["lon0"]
    d = 1.0 / 3600.0
    n = len(dem.x)
    lon = lon_c + d * (np.arange(n) - (n - 1) / 2)
    N = a / np.sqrt(1 - e2 * np.sin(np.radians(lat_c)) ** 2)
    assert np.allclose(dem.x, N * np.cos(np.radians(lat_c)) * np.radians(lon - lon_c), atol=1e-6)
    m = len(dem.y)
    lat = lat_c + d * (np.arange(m) - (m - 1) / 2)
    Mc = oracle.meridian_arc_quad(lat_c)
    want_y = np.array([oracle.meridian_arc_quad(v) - Mc for v in lat])
    assert np.allclose(dem.y, want_y, atol=1e-6)
    # spacing of a 1" raster at 13 deg S: ~30.1 m east, ~30.7 m north
    assert 29.5 < np.diff(dem.x).mean() < 30.5 and 30.5 < np.diff(dem.y).mean() < 31.0
    zone = abs(meta["utm_zone"])
    e, nn = oracle.utm_snyder(lat_c, lon_c, zone * 6 - 183)
    assert meta["utm_zone"] == -18 and abs(meta["e0"] - e) < 5e-3 and abs(meta["n0"] - nn) < 5e-3


def test_geotiff_nodata_and_point_raster(bsar, tmp_path):
    Image = _pil()
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    a[1, 2] = -32768.0
    f = tmp_path / "n.tif"
    Image.fromarray(a).save(f, compression="tiff_lzw", tiffinfo=_geotags((1.0, 1.0), (0.0, 0.0), model=1, raster=2,
                                                                         nodata=-32768))
    dem, meta = bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)
    valid = np.delete(a.ravel(), 6)
    assert meta["nodata_count"] == 1 and dem.z[1, 2] == np.float64(valid.mean(dtype=np.float64))
    # pixel is point: the tie point is the first sample, no half-pixel shift
    assert meta["e0"] == 1.5 and meta["n0"] == -1.0


def test_geotiff_rejects_bad_files(bsar, tmp_path):
    f = tmp_path / "bad.tif"
    f.write_bytes(b"II*\0" + b"\0" * 4)
    with pytest.raises(bsar.BsarError):
        bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)
    f.write_bytes(b"not a tiff at all")
    with pytest.raises(bsar.BsarError, match="not a TIFF"):
        bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)
    f.write_bytes(b"II+\0" + b"\0" * 12)
    with pytest.raises(bsar.BsarError, match="BigTIFF"):
        bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)
    with pytest.raises(bsar.BsarError, match="cannot open"):
        bsar.load_dem(tmp_path / "missing.tif", bsar.DEM_FILE_GEOTIFF)
    Image = _pil()
    Image.fromarray(np.zeros((4, 4), np.float32)).save(f)           # no georeferencing
    with pytest.raises(bsar.BsarError, match="georeferencing"):
        bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)


def _patch_tag(path, tag, value):
    """Rewrite the inline LONG value of `tag` in a little-endian TIFF's first IFD."""
    b = bytearray(Path(path).read_bytes())
    ifd = struct.unpack_from("<I", b, 4)[0]
    n = struct.unpack_from("<H", b, ifd)[0]
    for k in range(n):
        o = ifd + 2 + 12 * k
        t, typ, cnt = struct.unpack_from("<HHI", b, o)
        if t == tag:
            assert typ == 4 and cnt == 1
            struct.pack_into("<I", b, o + 8, value)
            Path(path).write_bytes(bytes(b))
            return
    raise KeyError(tag)


@pytest.mark.parametrize("tag,value", [(322, 2 ** 31), (323, 2 ** 31), (322, 2 ** 20 + 1), (322, 0), (256, 2 ** 32 - 1),
                                       (257, 0), (278, 0)])
def test_geotiff_rejects_crafted_sizes(bsar, tmp_path, tag, value):
    """Tile / strip / image sizes from the file are bounded before any
    allocation or row offset uses them (a chunk holds at most 2^28 samples):
    crafted tags give a format error, not an overflow, an out-of-bounds read
    or an uncaught allocation failure."""
    a = np.arange(16, dtype="<i2").reshape(4, 4)
    f = tmp_path / "c.tif"
    _write_tiff(f, a, tile=(16, 16) if tag in (322, 323) else None)
    _patch_tag(f, tag, value)
    with pytest.raises(bsar.BsarError) as e:
        bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)
    assert e.value.status == -5


def test_geotiff_large_legal_tile_stays_in_bounds(bsar, tmp_path):
    """A tile of 2^20 x 1 samples over a 4 x 4 raster: the chunk buffer is sized
    for the tile, the file's short chunk zero-filled, every row read inside it."""
    a = np.arange(16, dtype="<i2").reshape(4, 4)
    f = tmp_path / "t.tif"
    _write_tiff(f, a, tile=(16, 16))
    _patch_tag(f, 322, 2 ** 20)
    _patch_tag(f, 323, 1)
    try:
        bsar.load_dem(f, bsar.DEM_FILE_GEOTIFF)
    except bsar.BsarError as e:
        assert e.status == -5


def test_write_reports_close_failure(bsar, tmp_path):
    """A file whose last buffered block cannot be written (here /dev/full)
    is reported as an I/O error, not success."""
    if not Path("/dev/full").exists():
        pytest.skip("no /dev/full")
    img = np.zeros((1, 1, 2), np.complex128)
    with pytest.raises(bsar.BsarError) as e:
        bsar.write_image("/dev/full", img)
    assert e.value.status == -4


# ---- projection ----------------------------------------------------------------

def test_utm_against_snyder_series(bsar, oracle):
    rng = np.random.default_rng(11)
    for _ in range(500):
        lat, lon = rng.uniform(-80, 84), rng.uniform(-180, 180)
        e, n, z = bsar.geo_to_utm(lat, lon)
        e2, n2 = oracle.utm_snyder(lat, lon, abs(z) * 6 - 183)
        assert abs(e - e2) < 5e-3 and abs(n - n2) < 5e-3
        assert (z < 0) == (lat < 0) and 1 <= abs(z) <= 60


def test_utm_closed_forms(bsar, oracle):
    e, n, z = bsar.geo_to_utm(0.0, 3.0)
    assert (e, n, z) == (500000.0, 0.0, 31)                       # equator on the central meridian
    e, n, _ = bsar.geo_to_utm(90.0, 3.0)
    q = oracle.meridian_arc_quad(90.0)
    assert e == 500000.0 and abs(n - 0.9996 * q) < 1e-6           # the pole: k0 x quarter meridian
    assert abs(q - 10001965.729) < 1e-3                           # WGS84 quarter meridian
    for lat in (-47.0, -3.3, 12.5, 66.0):
        assert abs(bsar.meridian_arc(lat) - oracle.meridian_arc_quad(lat)) < 1e-6
        ea, na, _ = bsar.geo_to_utm(lat, 9.0 + 1.7, 32)
        eb, nb, _ = bsar.geo_to_utm(lat, 9.0 - 1.7, 32)
        assert abs((ea - 500000.0) + (eb - 500000.0)) < 1e-6 and abs(na - nb) < 1e-6     # symmetric about the CM
        en, nn_, _ = bsar.geo_to_utm(abs(lat), 9.0 + 1.7, 32)
        es, ns, _ = bsar.geo_to_utm(-abs(lat), 9.0 + 1.7, 32)
        assert abs(en - es) < 1e-6 and abs(nn_ + (ns - 10000000.0)) < 1e-6               # mirror about the equator
