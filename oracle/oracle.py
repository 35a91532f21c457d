"""ctypes binding of liboracle.so (the CPU FP64 restatement) and of the
compiled reference's shim (_ref/libref.so) when present.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker -- never by the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref.so"
REF_PROC = HERE / "_ref" / "proc_test_script"
REF_GS = HERE / "_ref" / "gs_test_script"
REF_TREE = Path("/root/reference")

_dp = C.POINTER(C.c_double)


class DemO(C.Structure):
    _fields_ = [("x", _dp), ("nx", C.c_int), ("y", _dp), ("ny", C.c_int), ("z", _dp), ("zrows", C.c_int),
                ("zcols", C.c_int)]


class SimParams(C.Structure):
    _fields_ = [("fc", C.c_double), ("c", C.c_double), ("fs", C.c_double), ("f0", C.c_double),
                ("f1", C.c_double), ("T", C.c_double), ("fs_chirp", C.c_double), ("n1", C.c_double),
                ("n2", C.c_double), ("S", C.c_int)]


class BpParams(C.Structure):
    _fields_ = [("c", C.c_double), ("fs4", C.c_double), ("lam", C.c_double), ("n1", C.c_double),
                ("n2", C.c_double)]


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    if ref and REF_TREE.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


_o = None


def lib() -> C.CDLL:
    global _o
    if _o is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.or_dem_interp.restype = C.c_double
        L.or_dem_interp.argtypes = [C.POINTER(DemO), C.c_double, C.c_double]
        L.or_refraction_point.argtypes = [C.c_double] * 6 + [C.POINTER(DemO), C.c_double, C.c_double, _dp]
        L.or_reflection_point.argtypes = [C.c_double] * 6 + [C.POINTER(DemO), _dp]
        L.or_slant_range_fermat.argtypes = [C.POINTER(DemO), _dp, _dp, C.c_int, _dp, C.c_double, C.c_double, _dp,
                                            _dp, _dp]
        L.or_spline.argtypes = [_dp, C.c_size_t, C.c_int, _dp]
        L.or_fft.argtypes = [_dp, C.c_size_t]
        L.or_ifft.argtypes = [_dp, C.c_size_t]
        L.or_fft_correlation.argtypes = [_dp, C.c_size_t, _dp, C.c_size_t, _dp]
        L.or_fft_convolution.argtypes = [_dp, C.c_size_t, _dp, C.c_size_t, _dp]
        L.or_chirp.restype = C.c_int
        L.or_chirp.argtypes = [C.c_double] * 4 + [_dp]
        L.or_simulate.argtypes = [C.POINTER(SimParams), C.POINTER(DemO), _dp, _dp, C.c_int, _dp, _dp, C.c_int,
                                  _dp, _dp, _dp]
        L.or_upsample.argtypes = [_dp, C.c_size_t, C.c_size_t, C.c_int, _dp]
        L.or_refraction_batch.argtypes = [_dp, C.c_size_t, C.POINTER(DemO), C.c_double, C.c_double, _dp, C.c_int]
        L.or_backproject.argtypes = [C.POINTER(BpParams), C.POINTER(DemO), _dp, _dp, C.c_int, _dp, C.c_size_t,
                                     C.c_size_t, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp,
                                     C.POINTER(C.c_int64), C.c_int64, C.c_int, C.c_int]
        L.or_flops_per_voxel_pulse.restype = C.c_double
        L.or_flops_per_voxel_pulse.argtypes = [C.c_int]
        _o = L
    return _o


def _p(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class _DemHold:
    def __init__(self, x, y, z):
        self.x, self.y, self.z = _f(x), _f(y), _f(z)
        zr, zc = self.z.shape if self.z.ndim == 2 else (0, 0)
        self.s = DemO(_p(self.x), len(self.x), _p(self.y), len(self.y), _p(self.z), zr, zc)


def dem_interp(dem, xy):
    d = _DemHold(*dem)
    L = lib()
    xy = _f(xy).reshape(-1, 2)
    return np.array([L.or_dem_interp(C.byref(d.s), a, b) for a, b in xy])


def refraction_points(dem, legs, n1, n2):
    d = _DemHold(*dem)
    L = lib()
    legs = _f(legs).reshape(-1, 6)
    out = np.empty((len(legs), 3))
    q = np.empty(3)
    for k, s in enumerate(legs):
        L.or_refraction_point(*s, C.byref(d.s), n1, n2, _p(q))
        out[k] = q
    return out


def reflection_points(dem, pairs):
    """calculateReflectionPoint (or_reflection_point) for pairs[m] = (tx, ty, tz, rx, ry, rz)."""
    d = _DemHold(*dem)
    L = lib()
    pairs = _f(pairs).reshape(-1, 6)
    out = np.empty((len(pairs), 3))
    q = np.empty(3)
    for k, s in enumerate(pairs):
        L.or_reflection_point(*s, C.byref(d.s), _p(q))
        out[k] = q
    return out


def refraction_batch(dem, legs, n1, n2, threads=0):
    """refraction_points over many legs in C with OpenMP."""
    d = _DemHold(*dem)
    legs = _f(legs).reshape(-1, 6)
    out = np.empty((len(legs), 3))
    lib().or_refraction_batch(_p(legs), len(legs), C.byref(d.s), n1, n2, _p(out), int(threads))
    return out


def ref_refraction(dem, legs, n1, n2, threads=1):
    """The reference's calculateRefractionPoint (oracle/_ref/libref.so) over
    legs, split across `threads` host threads (the function is reentrant);
    None when the reference library has not been built."""
    L = ref_lib()
    if L is None:
        return None
    import concurrent.futures as cf
    x, y, z = _f(dem[0]), _f(dem[1]), _f(dem[2])
    zr, zc = z.shape if z.ndim == 2 else (0, 0)
    legs = _f(legs).reshape(-1, 6)
    out = np.empty((len(legs), 3))
    bounds = np.linspace(0, len(legs), max(int(threads), 1) + 1).astype(int)

    def run(a, b):
        if b > a:
            L.ref_refraction(_p(x), len(x), _p(y), len(y), _p(z), zr, zc, _p(legs[a:b]), b - a, n1, n2,
                             out[a:b].ctypes.data_as(_dp))

    with cf.ThreadPoolExecutor(max_workers=len(bounds) - 1) as ex:
        list(ex.map(lambda ab: run(*ab), zip(bounds[:-1], bounds[1:])))
    return out


def slant_range_fermat(dem, tx, rx, P, n1, n2):
    d = _DemHold(*dem)
    tx, rx, P = _f(tx).reshape(-1, 3), _f(rx).reshape(-1, 3), _f(P)
    n = len(tx)
    ida, volta, r4 = np.empty((n, 3)), np.empty((n, 3)), np.empty((4, n))
    lib().or_slant_range_fermat(C.byref(d.s), _p(tx), _p(rx), n, _p(P), n1, n2, _p(ida), _p(volta), _p(r4))
    return ida, volta, r4


def spline(data, factor=4):
    d = _c(data)
    out = np.empty(len(d) * factor, np.complex128)
    lib().or_spline(_p(d.view(np.float64)), len(d), factor, _p(out.view(np.float64)))
    return out


def upsample(root, factor=4):
    """root [N][S] complex (range fastest) -> [N][S*factor]."""
    r = _c(root)
    N, S = r.shape
    out = np.empty((N, S * factor), np.complex128)
    lib().or_upsample(_p(r.view(np.float64)), S, N, factor, _p(out.view(np.float64)))
    return out


def fft(a):
    x = _c(a).copy()
    lib().or_fft(_p(x.view(np.float64)), len(x))
    return x


def ifft(a):
    x = _c(a).copy()
    lib().or_ifft(_p(x.view(np.float64)), len(x))
    return x


def fft_correlation(sig, ref):
    s, r = _c(sig), _c(ref)
    out = np.empty(max(len(s), len(r)), np.complex128)
    lib().or_fft_correlation(_p(s.view(np.float64)), len(s), _p(r.view(np.float64)), len(r),
                             _p(out.view(np.float64)))
    return out


def fft_convolution(sig, ref):
    s, r = _c(sig), _c(ref)
    out = np.empty(max(len(s), len(r)), np.complex128)
    lib().or_fft_convolution(_p(s.view(np.float64)), len(s), _p(r.view(np.float64)), len(r),
                             _p(out.view(np.float64)))
    return out


def chirp(f0, f1, T, fs):
    L = lib()
    n = L.or_chirp(f0, f1, T, fs, None)
    out = np.empty(n, np.complex128)
    L.or_chirp(f0, f1, T, fs, _p(out.view(np.float64)))
    return out


def simulate(sim: dict, dem, tx, rx, scat, rcs):
    """processRawData over scatterers: returns aux, raw, root as [S][N] complex."""
    d = _DemHold(*dem)
    tx, rx = _f(tx).reshape(-1, 3), _f(rx).reshape(-1, 3)
    scat = _f(scat).reshape(-1, 3)
    rcs = _c(rcs).reshape(-1)
    n, S = len(tx), int(sim["S"])
    sp = SimParams(sim["fc"], sim["c"], sim["fs"], sim["f0"], sim["f1"], sim["T"], sim["fs_chirp"], sim["n1"],
                   sim["n2"], S)
    aux = np.empty((S, n), np.complex128)
    raw = np.empty((S, n), np.complex128)
    root = np.empty((S, n), np.complex128)
    lib().or_simulate(C.byref(sp), C.byref(d.s), _p(tx), _p(rx), n, _p(scat), _p(rcs.view(np.float64)), len(scat),
                      _p(aux.view(np.float64)), _p(raw.view(np.float64)), _p(root.view(np.float64)))
    return aux, raw, root


def backproject(bp: dict, dem, tx, rx, rcomp, gx, gy, gz, subset=None, reflection=False, threads=0):
    """Image [nz][ny][nx] complex (x fastest); voxels outside `subset` stay 0."""
    d = _DemHold(*dem)
    tx, rx = _f(tx).reshape(-1, 3), _f(rx).reshape(-1, 3)
    rc = _c(rcomp)
    ndata, S4 = rc.shape
    gx, gy, gz = _f(gx), _f(gy), _f(gz)
    out = np.zeros((len(gz), len(gy), len(gx)), np.complex128)
    bpp = BpParams(bp["c"], bp["fs4"], bp["lam"], bp["n1"], bp["n2"])
    if subset is None:
        sub_p, nsub = None, 0
    else:
        sub = np.ascontiguousarray(subset, dtype=np.int64)
        sub_p, nsub = sub.ctypes.data_as(C.POINTER(C.c_int64)), len(sub)
    lib().or_backproject(C.byref(bpp), C.byref(d.s), _p(tx), _p(rx), len(tx), _p(rc.view(np.float64)), S4, ndata,
                         _p(gx), len(gx), _p(gy), len(gy), _p(gz), len(gz), _p(out.view(np.float64)), sub_p, nsub,
                         int(reflection), int(threads))
    return out


def flops_per_voxel_pulse(n):
    return lib().or_flops_per_voxel_pulse(n)


# ---- compiled reference (present only where it was built from /root/reference)

_r = None


def ref_lib():
    global _r
    if _r is None:
        if not REF_LIB.exists():
            return None
        L = C.CDLL(str(REF_LIB))
        dem_args = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int]
        L.ref_dem_interp.argtypes = dem_args + [_dp, C.c_int, _dp]
        L.ref_refraction.argtypes = dem_args + [_dp, C.c_int, C.c_double, C.c_double, _dp]
        L.ref_slant_range.argtypes = dem_args + [_dp, _dp, C.c_int, _dp, C.c_double, C.c_double, _dp, _dp, _dp]
        L.ref_reflection.argtypes = dem_args + [_dp, C.c_int, _dp]
        L.ref_slant_reflection.argtypes = dem_args + [_dp, _dp, C.c_int, _dp, C.c_double, C.c_double, _dp]
        L.ref_spline.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.ref_fft.argtypes = [_dp, C.c_int]
        L.ref_ifft.argtypes = [_dp, C.c_int]
        L.ref_fft_correlation.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp]
        L.ref_fft_convolution.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp]
        L.ref_chirp.restype = C.c_int
        L.ref_chirp.argtypes = [C.c_double] * 4 + [_dp]
        L.ref_process_raw.argtypes = dem_args + [_dp, _dp, C.c_int, _dp, C.c_double] + [C.c_double] * 9 + \
            [C.c_int, _dp, _dp, _dp]
        L.ref_spiral.restype = C.c_int64
        L.ref_spiral.argtypes = [C.c_double] * 7 + [_dp, _dp, _dp, C.c_int64]
        L.ref_default_trajectory.restype = C.c_int64
        L.ref_default_trajectory.argtypes = [C.c_int, _dp, C.c_int64]
        L.ref_grid.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, C.c_int,
                               C.POINTER(C.c_int)]
        L.ref_load_dem.argtypes = [C.c_char_p, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.POINTER(C.c_int)]
        _r = L
    return _r


def ref_spiral(*args) -> np.ndarray:
    """funcao_espiral of the compiled reference, [n][3]."""
    L = ref_lib()
    n = L.ref_spiral(*args, None, None, None, 0)
    px, py, pz = (np.empty(max(n, 1)) for _ in range(3))
    L.ref_spiral(*args, _p(px), _p(py), _p(pz), n)
    return np.stack([px[:n], py[:n], pz[:n]], 1)


def ref_default_trajectory(rx: bool) -> np.ndarray:
    """generate{Transmitter,Receiver}Trajectory(initializeRadar{Tx,Rx}()) of the compiled reference."""
    L = ref_lib()
    n = L.ref_default_trajectory(int(rx), None, 0)
    out = np.empty((n, 3))
    L.ref_default_trajectory(int(rx), _p(out), n)
    return out


def ref_grid(pos, dxy, lxy, dz, lz):
    """crearGrid axes of the compiled reference."""
    L = ref_lib()
    n = (C.c_int * 3)()
    pos = np.ascontiguousarray(pos, np.float64)
    buf = np.empty(1 << 20)
    L.ref_grid(_p(pos), dxy, lxy, dz, lz, _p(buf), len(buf), n)
    a, b, c = n[0], n[1], n[2]
    return buf[:a].copy(), buf[a:a + b].copy(), buf[a + b:a + b + c].copy()


def ref_load_dem(path):
    """loadDEM (load_dem_simple.cpp) of the compiled reference: X_vec, Y_vec,
    Z_DEM flattened (rows of Z_DEM[0]'s length) and the four sizes."""
    L = ref_lib()
    n = (C.c_int * 4)()
    cap = 1 << 20
    x, y, z = np.empty(cap), np.empty(cap), np.empty(cap)
    L.ref_load_dem(str(path).encode(), _p(x), cap, _p(y), cap, _p(z), cap, n)
    nx, ny, zr, zc = n[0], n[1], n[2], n[3]
    return x[:nx].copy(), y[:ny].copy(), z[:zr * zc].copy().reshape(zr, zc), (nx, ny, zr, zc)


# ---- scenario generators and map projection (independent restatements) ----

def spiral(N, r_minor, r_major, h_high, h_low, vt, prf, mirror=False) -> np.ndarray:
    """funcao_espiral (funcao_espiral.cpp:5-49) in scalar Python: the same
    operations in the same order, glibc's log and sincos."""
    import math
    m = C.CDLL("libm.so.6")
    m.sincos.argtypes = [C.c_double, _dp, _dp]
    sn, cs = C.c_double(), C.c_double()
    rs, re_ = r_major, r_minor
    vr = vt * (math.log(rs) - math.log(re_)) / (N * 2 * math.pi)
    T = (rs - re_) / vr if vr != 0 else float("nan")
    vd = (h_high - h_low) / T
    dt = 1.0 / prf
    out = []
    t = 0.0
    while t <= T:
        theta = (vt / vr) * (math.log(rs) - math.log(rs - vr * t))
        R = rs - vr * t
        # g++ -O3 (the reference's build) folds cos(theta), sin(theta) into one sincos call
        m.sincos(theta, C.byref(sn), C.byref(cs))
        x, y = R * cs.value, R * sn.value
        out.append((-x if mirror else x, -y if mirror else y, h_high - vd * t))
        t += dt
    return np.array(out).reshape(-1, 3)


def grid_axis(center, step, half) -> np.ndarray:
    """One crearGrid axis (crear_grid.cpp:21-34): a running sum."""
    out = []
    v = center - half
    while v <= center + half + step / 2:
        out.append(v)
        v += step
    return np.array(out)


WGS84 = (6378137.0, 1 / 298.257223563)


def meridian_arc_quad(lat_deg, a=WGS84[0], f=WGS84[1]) -> float:
    """Meridian arc from the equator by numerical quadrature of the meridian
    radius of curvature a(1-e^2)/(1-e^2 sin^2)^1.5."""
    from scipy.integrate import quad
    e2 = f * (2 - f)
    val, _ = quad(lambda p: a * (1 - e2) / (1 - e2 * np.sin(p) ** 2) ** 1.5, 0.0, np.radians(lat_deg),
                  epsabs=1e-9, epsrel=1e-13, limit=200)
    return val


def utm_snyder(lat_deg, lon_deg, lon0_deg, a=WGS84[0], f=WGS84[1]):
    """Transverse Mercator by Snyder's series in A = (lon - lon0) cos(lat)
    (USGS Professional Paper 1395, eqs. 3-21, 4-20, 8-9, 8-10, 8-12..8-15),
    UTM scale 0.9996, false easting 500 km, false northing 10,000 km south."""
    k0 = 0.9996
    e2 = f * (2 - f)
    ep2 = e2 / (1 - e2)
    phi = np.radians(lat_deg)
    lam = np.radians(lon_deg - lon0_deg)
    N = a / np.sqrt(1 - e2 * np.sin(phi) ** 2)
    T = np.tan(phi) ** 2
    Cc = ep2 * np.cos(phi) ** 2
    A = lam * np.cos(phi)
    M = a * ((1 - e2 / 4 - 3 * e2 ** 2 / 64 - 5 * e2 ** 3 / 256) * phi
             - (3 * e2 / 8 + 3 * e2 ** 2 / 32 + 45 * e2 ** 3 / 1024) * np.sin(2 * phi)
             + (15 * e2 ** 2 / 256 + 45 * e2 ** 3 / 1024) * np.sin(4 * phi)
             - (35 * e2 ** 3 / 3072) * np.sin(6 * phi))
    x = k0 * N * (A + (1 - T + Cc) * A ** 3 / 6 + (5 - 18 * T + T ** 2 + 72 * Cc - 58 * ep2) * A ** 5 / 120)
    y = k0 * (M + N * np.tan(phi) * (A ** 2 / 2 + (5 - T + 9 * Cc + 4 * Cc ** 2) * A ** 4 / 24
                                     + (61 - 58 * T + T ** 2 + 600 * Cc - 330 * ep2) * A ** 6 / 720))
    return 500000.0 + x, (10000000.0 if lat_deg < 0 else 0.0) + y
