"""Seeded synthetic scenes for the imaging workloads, and the forward model
that turns them into range-compressed data on the GPU.

Geometry and signal follow the reference's simulator (gs_test_script.cpp:
15-137, processRawData raw_data_processing.cpp:292-452): per pulse the
bistatic refraction path of each scatterer gives a delay bin
1 + round(t*fs) and phase -(2*pi/lambda)*path; unit responses are convolved
with the fftshifted chirp and correlated with it (pulse compression).  The
refraction points come from the library's FP64 Fermat kernel and the
convolution/correlation from its batched FFT kernel.

Everything that the reference writes to CSV (antenna positions, DEM, grid
axes) is quantised through "%.6f" text, as data_export.cpp:148,169 does, so
in-memory runs and file-based runs see the same doubles.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def q6(a) -> np.ndarray:
    """Round through the reference's "%.6f" CSV text (std::fixed, setprecision(6))."""
    a = np.asarray(a, dtype=np.float64)
    return np.array([float("%.6f" % v) for v in a.ravel()]).reshape(a.shape)


def spiral(n: int, r0: float = 5.0, r1: float = 15.0, turns: float = 2.0, h: float = 12.0,
           phase: float = 0.0, x0: float = 0.0, y0: float = 0.0) -> np.ndarray:
    """Archimedean spiral, radius r0 -> r1 over `turns` turns at height h: [n][3]."""
    p = np.arange(n, dtype=np.float64)
    s = p / max(n - 1, 1)
    r = r0 + (r1 - r0) * s
    th = 2.0 * np.pi * turns * s + phase
    return q6(np.stack([x0 + r * np.cos(th), y0 + r * np.sin(th), np.full(n, h)], 1))


def dem_bumps(rng: np.random.Generator, n: int = 34, spacing: float = 30.0, bumps: int = 5,
              amp=(0.5, 2.0), sigma=(60.0, 150.0), slope: float = 0.01):
    """Plane plus Gaussian bumps on an n x n grid centred on the origin."""
    ax = -(n - 1) * spacing / 2.0 + spacing * np.arange(n)
    X, Y = np.meshgrid(ax, ax, indexing="xy")
    a, b = rng.uniform(-slope, slope, 2)
    Z = a * X + b * Y
    for _ in range(bumps):
        A = rng.uniform(*amp)
        s = rng.uniform(*sigma)
        cx, cy = rng.uniform(ax[0], ax[-1], 2)
        Z = Z + A * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
    return q6(ax), q6(ax), q6(Z)


def dem_fbm(rng: np.random.Generator, nx: int = 167, ny: int = 100, spacing: float = 30.0, hurst: float = 0.7,
            rms: float = 40.0):
    """Fractional-Brownian relief by spectral synthesis (|k|^-(2H+2) power)."""
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    spec = amp * (rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx)))
    Z = np.real(np.fft.ifft2(spec))
    Z *= rms / np.sqrt(np.mean(Z * Z))
    x = -(nx - 1) * spacing / 2.0 + spacing * np.arange(nx)
    y = -(ny - 1) * spacing / 2.0 + spacing * np.arange(ny)
    return q6(x), q6(y), q6(Z)


def axis(start: float, step: float, count: int) -> np.ndarray:
    return q6(start + step * np.arange(count))


@dataclass
class Scene:
    name: str
    fc: float
    c: float
    fs: float            # ADC rate of the raw data (before the x4 upsampling)
    bw: float
    T: float
    n1: float
    n2: float
    S: int
    tx: np.ndarray
    rx: np.ndarray
    dem: tuple
    gx: np.ndarray
    gy: np.ndarray
    gz: np.ndarray
    scat: np.ndarray
    rcs: np.ndarray
    snr_db: float | None = None
    seed: int = 0
    upsample: int = 4
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.tx)

    @property
    def voxels(self) -> int:
        return len(self.gx) * len(self.gy) * len(self.gz)

    def sim_params(self) -> dict:
        return dict(fc=self.fc, c=self.c, fs=self.fs, f0=-self.bw / 2, f1=self.bw / 2, T=self.T, fs_chirp=self.fs,
                    n1=self.n1, n2=self.n2, S=self.S)

    def bp_params(self) -> dict:
        return dict(c=self.c, fs4=self.fs * self.upsample, lam=self.c / self.fc, n1=self.n1, n2=self.n2)


BOX = dict(x=(19.2, 20.8), y=(-0.8, 0.8), z=(-5.8, -4.2))


def _grid(nx, ny, nz, box=BOX):
    def ax(lo, hi, m):
        return axis(lo, (hi - lo) / (m - 1) if m > 1 else 0.0, m)
    return ax(*box["x"], nx), ax(*box["y"], ny), ax(*box["z"], nz)


def _scatterers(rng, count, box=BOX, margin=0.1):
    lo = np.array([box["x"][0], box["y"][0], box["z"][0]]) + margin
    hi = np.array([box["x"][1], box["y"][1], box["z"][1]]) - margin
    return rng.uniform(lo, hi, size=(count, 3))


def _pkg():
    import importlib.util
    import sys
    from pathlib import Path
    pkg = sys.modules.get("bsar_amd")
    if pkg is None:
        spec = importlib.util.spec_from_file_location("bsar_amd", Path(__file__).parent / "__init__.py")
        pkg = importlib.util.module_from_spec(spec)
        sys.modules["bsar_amd"] = pkg
        spec.loader.exec_module(pkg)
    return pkg


# The reference's default scenario, gs_test_script.cpp:15-137 with the
# constants of radar_init.cpp:5-89: carrier 10 GHz, ADC 60 MHz, chirp -25..25
# MHz over 1.3 us, S = IndiceMaximo = 256 samples; Tx/Rx spirals of 2 turns
# from 147.5 m to 172.5 m radius, 120 m down to 80 m, 120 m/s, PRF 400
# (6689 pulses; Rx mirrored); one target at (20, 0, -5) with rcs 1;
# crearGrid(0.04, 0.8, 0.08, 0.8) around it; loadDEM of the .tif name (the
# flat 32 x 32 DEM of the build without GDAL); n1 = 1, n2 = 4 (the
# environment defaults, raw_data_processing.cpp:17-19).
REFERENCE_DEM = "assets/DEM_1x1km_Res30m_Lat-3_6160_Lon-80_4552.tif"
REFERENCE_SPIRAL = (2, 172.5, 147.5, 120.0, 80.0, 120.0, 400.0)


def reference_scene(n_pulses: int | None = None, dem_path: str | None = None, dem_mode: int = 0) -> Scene:
    """The reference scenario; dem_path/dem_mode replace the DEM by a file
    read with load_dem (mode 1 decodes a GeoTIFF)."""
    pkg = _pkg()
    tx = pkg.spiral(*REFERENCE_SPIRAL)
    rx = pkg.spiral(*REFERENCE_SPIRAL, mirror_xy=True)
    if n_pulses is not None:
        tx, rx = tx[:n_pulses], rx[:n_pulses]
    dem, meta = pkg.load_dem(dem_path or REFERENCE_DEM, dem_mode)
    target = np.array([20.0, 0.0, -5.0])
    gx = pkg.grid_axis(target[0], 0.04, 0.8)
    gy = pkg.grid_axis(target[1], 0.04, 0.8)
    gz = pkg.grid_axis(target[2], 0.08, 0.8)
    return Scene(name="reference", fc=10e9, c=3e8, fs=60e6, bw=50e6, T=1.3e-6, n1=1.0, n2=4.0, S=256, tx=tx, rx=rx,
                 dem=(dem.x, dem.y, dem.z), gx=gx, gy=gy, gz=gz, scat=target[None, :], rcs=np.ones(1, np.complex128),
                 extra={"target": target, "dem_meta": meta})


def make_scene(name: str, seed: int | None = None, n_pulses: int | None = None, grid=None, samples=None,
               clutter: int | None = None) -> Scene:
    """Workloads:
      small   41x41x21 grid, 256 pulses, 4096 samples, 3 scatterers, 20 dB SNR
      large   201x201x101 grid, 2048 pulses, 8192 samples, 16 scatterers + clutter layer
      rough   128^3 grid 4x4x4 m buried 3-7 m under fBm terrain (167x100 DEM), 1024 pulses
      wide    512x512x64 grid, 4096 pulses, 16384 samples, 64 scatterers, B = 500 MHz
      serve   resident 4096-pulse x 8192-sample dataset for the serving stream
              (serve_jobs: grids and pulse counts of the jobs), 16 scatterers
    Every one: fc = 1 GHz, T = 1 us, fs = 800 MHz, c = 3e8, n1 = 1, n2 = sqrt(eps_r).
      reference  the reference's own default scenario (reference_scene)
    """
    if name == "reference":
        if grid is not None or samples is not None:
            raise ValueError("the reference scenario fixes its grid and samples")
        return reference_scene(n_pulses)
    seeds = dict(small=1001, large=1002, rough=1003, wide=1004, serve=1005)
    seed = seeds.get(name, 1000) if seed is None else seed
    rng = np.random.default_rng(seed)
    fc, c, fs, T = 1e9, 3e8, 800e6, 1e-6
    if name == "small":
        N, S, bw, er = 256, 4096, 300e6, 6.0
        gx, gy, gz = _grid(41, 41, 21)
        dem = dem_bumps(rng)
        scat = _scatterers(rng, 3)
        rcs = np.ones(3, np.complex128)
        snr = 20.0
    elif name == "large":
        N, S, bw, er = 2048, 8192, 300e6, 6.0
        gx, gy, gz = _grid(201, 201, 101)
        dem = dem_bumps(rng)
        pts = _scatterers(rng, 16)
        ncl = 256 if clutter is None else clutter
        # clutter layer 0.5 m thick at the top of the box, uniform amplitude, random phase
        cl = np.stack([rng.uniform(*BOX["x"], ncl), rng.uniform(*BOX["y"], ncl),
                       rng.uniform(BOX["z"][1] - 0.5, BOX["z"][1], ncl)], 1)
        scat = np.concatenate([pts, cl])
        rcs = np.concatenate([np.ones(16), 0.1 * np.exp(2j * np.pi * rng.uniform(size=ncl))]).astype(np.complex128)
        snr = None
    elif name == "rough":
        N, S, bw = 1024, 4096, 300e6
        er = float(rng.uniform(4.0, 9.0))
        dem = dem_fbm(rng)
        # grid centre above the DEM's origin; buried 3-7 m below the surface there
        zs = float(np.interp(0.0, dem[0], dem[2][np.searchsorted(dem[1], 0.0)]))
        box = dict(x=(-2.0, 2.0), y=(-2.0, 2.0), z=(zs - 7.0, zs - 3.0))
        gx, gy, gz = _grid(128, 128, 128, box)
        scat = _scatterers(rng, 3, box)
        rcs = np.ones(3, np.complex128)
        snr = None
        tx0 = spiral(N, h=zs + 12.0)
        rx0 = spiral(N, h=zs + 12.0, phase=np.pi / 2)
    elif name == "wide":
        N, S, bw, er = 4096, 16384, 500e6, 6.0
        gx, gy, gz = _grid(512, 512, 64)
        dem = dem_bumps(rng)
        scat = _scatterers(rng, 64)
        rcs = np.ones(64, np.complex128)
        snr = None
    elif name == "serve":
        N, S, bw, er = 4096, 8192, 300e6, 6.0
        gx, gy, gz = _grid(41, 41, 21)
        dem = dem_bumps(rng)
        scat = _scatterers(rng, 16)
        rcs = np.ones(16, np.complex128)
        snr = None
    else:
        raise ValueError(f"unknown workload {name!r}")
    if n_pulses is not None:
        N = n_pulses
    if samples is not None:
        S = samples
    if grid is not None:
        box = BOX if name != "rough" else dict(x=(gx[0], gx[-1]), y=(gy[0], gy[-1]), z=(gz[0], gz[-1]))
        gx, gy, gz = _grid(*grid, box)
    if name == "rough":
        tx = tx0 if n_pulses is None else spiral(N, h=float(tx0[0, 2]))
        rx = rx0 if n_pulses is None else spiral(N, h=float(rx0[0, 2]), phase=np.pi / 2)
    else:
        tx = spiral(N)
        rx = spiral(N, phase=np.pi / 2)
    return Scene(name=name, fc=fc, c=c, fs=fs, bw=bw, T=T, n1=1.0, n2=float(np.sqrt(er)), S=S, tx=tx, rx=rx,
                 dem=dem, gx=gx, gy=gy, gz=gz, scat=scat, rcs=rcs, snr_db=snr, seed=seed)


SERVE_GRIDS = [(16, 16, 16), (32, 32, 32), (64, 64, 64), (96, 96, 48), (128, 128, 64)]
SERVE_PULSES = [128, 512, 1024, 2048]


def serve_jobs(count: int = 256, rate: float = 40.0, seed: int = 1005) -> list[dict]:
    """Serving stream: Poisson arrivals at `rate` jobs/s; each job images a grid
    drawn uniformly from SERVE_GRIDS over the common box with the first K
    pulses, K uniform in SERVE_PULSES."""
    rng = np.random.default_rng(seed + 1)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    gi = rng.integers(0, len(SERVE_GRIDS), count)
    ki = rng.integers(0, len(SERVE_PULSES), count)
    jobs = []
    for k in range(count):
        shape = SERVE_GRIDS[gi[k]]
        gx, gy, gz = _grid(*shape)
        jobs.append(dict(t=float(t[k]), shape=shape, K=SERVE_PULSES[ki[k]], gx=gx, gy=gy, gz=gz,
                         vp=float(np.prod(shape)) * SERVE_PULSES[ki[k]]))
    return jobs


def reference_signal(chirp: np.ndarray, S: int) -> np.ndarray:
    """fftshifted, zero-padded chirp of raw_data_processing.cpp:371-394."""
    L = len(chirp)
    if L > S:
        raise ValueError("chirp longer than the range window")
    half = L // 2
    ref = np.zeros(S, np.complex128)
    ref[: L - half] = chirp[half:]
    ref[S - half:] = chirp[:half]
    return ref


def forward_model(ctx, scene: Scene, precision: int = 0) -> dict:
    """processRawData's three outputs for the scene, [N][S] complex each
    (range fastest, the layout BP reads): aux (placed responses), raw
    (convolved with the chirp) and root (pulse-compressed), from the GPU
    (bsar_process_raw: refraction points, response placement, chirp
    convolution, pulse compression).  With scene.snr_db, complex Gaussian
    noise (power relative to the mean raw-signal power) is added to raw
    before the compression."""
    pkg = _pkg()
    params = pkg.Params(scene.fc, scene.c, scene.fs, scene.n1, scene.n2, scene.upsample)
    ds = ctx.dataset(params, pkg.Dem(*scene.dem), scene.tx, scene.rx)
    try:
        out = ds.process_raw(scene.scat, scene.rcs, (-scene.bw / 2, scene.bw / 2, scene.T, scene.fs), scene.S)
    finally:
        ds.close()
    if scene.snr_db is not None:
        raw = out["raw"]
        rng = np.random.default_rng(scene.seed + 7)
        p = np.mean(np.abs(raw) ** 2)
        sigma = np.sqrt(p / 10 ** (scene.snr_db / 10) / 2)
        out["raw"] = raw + sigma * (rng.normal(size=raw.shape) + 1j * rng.normal(size=raw.shape))
    if scene.snr_db is not None or precision != 0:
        out["root"] = ctx.range_compress(out["raw"], matched_filter(scene), conj=True, precision=precision)
    return out


def simulate(ctx, scene: Scene, precision: int = 0, compress: bool = True) -> np.ndarray:
    """Range-compressed data [N][S] of forward_model (compress=False returns
    the raw data)."""
    out = forward_model(ctx, scene, precision)
    return out["root"] if compress else out["raw"]


def matched_filter(scene: Scene) -> np.ndarray:
    """The pulse-compression reference of processRawData for this scene."""
    pkg = _pkg()
    return reference_signal(pkg.chirp(-scene.bw / 2, scene.bw / 2, scene.T, scene.fs), scene.S)


def write_reference_inputs(proc, sc: Scene, root: np.ndarray, gx=None, gy=None, gz=None, raw=None,
                           aux=None) -> None:
    """The build/output/proc/ file set of exportProcessedData
    (data_export.cpp:16-69), which proc_test_script.cpp reads (:225-292), in
    the reference's formats: rootData_complex.bin (and rawData / auxData when
    given) as data_export.cpp:71-138 writes a 2-D complex matrix, header dims
    (S, N), with the payload range-fastest -- pulse p's samples at r + p*S,
    the order proc_test_script.cpp:330 reads (the reference's writer emits
    r*N + p, which its reader does not undo); CSVs with six decimals
    (data_export.cpp:148,169); target_position.csv is the scene's target (the
    first scatterer); environment.txt as raw_data_processing.cpp:14-46 parses
    it, each value in the shortest "%.17g" text (the reference's default
    stream precision, 6 digits, agrees whenever that is exact)."""
    import struct
    from pathlib import Path
    proc = Path(proc)
    proc.mkdir(parents=True, exist_ok=True)
    gx = sc.gx if gx is None else gx
    gy = sc.gy if gy is None else gy
    gz = sc.gz if gz is None else gz
    N, S = root.shape
    for name, data in (("rootData_complex.bin", root), ("rawData_complex.bin", raw), ("auxData_complex.bin", aux)):
        if data is None:
            continue
        with open(proc / name, "wb") as f:
            f.write(struct.pack("<III", 2, 1, 6) + b"double".ljust(16, b"\0") + struct.pack("<QQ", S, N))
            f.write(np.ascontiguousarray(data, dtype=np.complex128).tobytes())

    def vec(name, v):
        (proc / name).write_text(",".join("%.6f" % x for x in v) + "\n")

    def mat(name, m):
        (proc / name).write_text("".join(",".join("%.6f" % x for x in row) + "\n" for row in m))

    vec("grid_xAxis.csv", gx)
    vec("grid_yAxis.csv", gy)
    vec("grid_zAxis.csv", gz)
    vec("DEM_X_vec.csv", sc.dem[0])
    vec("DEM_Y_vec.csv", sc.dem[1])
    mat("DEM_Z_matrix.csv", sc.dem[2])
    mat("Tx_positions.csv", sc.tx)
    mat("Rx_positions.csv", sc.rx)
    vec("target_position.csv", sc.extra.get("target", sc.scat[0]))
    (proc / "environment.txt").write_text("n1=%s\nn2=%s\n" % (_g17(sc.n1), _g17(sc.n2)))


def _g17(v: float) -> str:
    """Shortest of %.6g / %.17g text that reads back as v."""
    for fmt in ("%.6g", "%.17g"):
        t = fmt % v
        if float(t) == v:
            return t
    return repr(v)
