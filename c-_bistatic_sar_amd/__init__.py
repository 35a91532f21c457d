"""Python binding of libbsar.so, the MI355X-native bistatic-SAR imaging path.

Thin ctypes layer over the C ABI in include/bsar.h: numpy arrays in, numpy
arrays out.  All compute runs in the HIP kernels of libbsar.so; there is no
CPU fallback -- importing raises when the library has not been built.

The directory name is not a Python identifier; load it with
``load_package()`` from ``__graft_entry__`` or ``importlib`` by path.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libbsar.so"

FP64 = 0
MIXED = 1

_STATUS = {
    0: "ok", -1: "invalid argument", -2: "device error", -3: "out of memory", -4: "i/o error",
    -5: "format error", -6: "capacity exceeded", -7: "invalid state",
}


class BsarError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_STATUS.get(status, status)} ({status}): {msg}")
        self.status = status


_dp = C.POINTER(C.c_double)


class DemC(C.Structure):
    _fields_ = [("x", _dp), ("nx", C.c_int), ("y", _dp), ("ny", C.c_int), ("z", _dp),
                ("zrows", C.c_int), ("zcols", C.c_int)]


class DemGridC(C.Structure):
    _fields_ = [("dem", DemC), ("lon0", C.c_double), ("lat0", C.c_double), ("e0", C.c_double), ("n0", C.c_double),
                ("utm_zone", C.c_int), ("nodata_count", C.c_int), ("source", C.c_int), ("reserved", C.c_int)]


class ParamsC(C.Structure):
    _fields_ = [("fc", C.c_double), ("c", C.c_double), ("fs", C.c_double), ("n1", C.c_double),
                ("n2", C.c_double), ("upsample", C.c_int), ("reserved", C.c_int)]


class ReportC(C.Structure):
    _fields_ = [("seconds", C.c_double), ("voxel_pulses", C.c_double), ("voxels", C.c_int64),
                ("pulses", C.c_int32), ("launches", C.c_int32), ("r1_exact_frac", C.c_double)]


class ChirpC(C.Structure):
    _fields_ = [("f0", C.c_double), ("f1", C.c_double), ("T", C.c_double), ("fs", C.c_double)]


class ServerConfigC(C.Structure):
    _fields_ = [("max_jobs", C.c_int), ("reserved0", C.c_int), ("hbm_cap_bytes", C.c_int64),
                ("round_seconds", C.c_double), ("reserved", C.c_double * 4)]


class JobDescC(C.Structure):
    _fields_ = [("ds", C.c_void_p), ("pulses", C.c_int), ("precision", C.c_int),
                ("gx", _dp), ("nx", C.c_int), ("gy", _dp), ("ny", C.c_int), ("gz", _dp), ("nz", C.c_int),
                ("out", _dp)]


class JobInfoC(C.Structure):
    _fields_ = [("id", C.c_int64), ("state", C.c_int), ("status", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("voxel_pulses", C.c_double),
                ("r1_exact_frac", C.c_double), ("footprint_bytes", C.c_int64)]


class ServerStatsC(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("rejected", C.c_int64), ("rounds", C.c_int64), ("items", C.c_int64),
                ("device_seconds", C.c_double), ("voxel_pulses", C.c_double), ("hbm_in_use", C.c_int64),
                ("hbm_peak", C.c_int64), ("active", C.c_int), ("queued", C.c_int)]


# name -> (restype, argtypes); every function declared in include/bsar.h
SIGNATURES = {
    "bsar_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "bsar_ctx_destroy": (None, [C.c_void_p]),
    "bsar_last_error": (C.c_char_p, [C.c_void_p]),
    "bsar_status_string": (C.c_char_p, [C.c_int]),
    "bsar_abi_version": (C.c_int, []),
    "bsar_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "bsar_dem_interp": (C.c_int, [C.c_void_p, C.POINTER(DemC), _dp, C.c_int64, _dp]),
    "bsar_refraction_points": (C.c_int, [C.c_void_p, C.POINTER(DemC), _dp, C.c_int64, C.c_double, C.c_double, _dp]),
    "bsar_refraction_points_wave": (C.c_int, [C.c_void_p, C.POINTER(DemC), _dp, C.c_int64, C.c_double, C.c_double,
                                              _dp, _dp]),
    "bsar_slant_range_fermat": (C.c_int, [C.c_void_p, C.c_void_p, _dp, C.c_int64, _dp, _dp]),
    "bsar_slant_range_fermat_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                                 C.c_void_p, C.POINTER(ReportC)]),
    "bsar_process_raw": (C.c_int, [C.c_void_p, C.c_void_p, _dp, _dp, C.c_int, C.POINTER(ChirpC), C.c_int, _dp, _dp,
                                   _dp]),
    "bsar_upsample": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, _dp]),
    "bsar_range_compress": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, _dp, C.c_int, C.c_int, _dp]),
    "bsar_range_compress_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, C.c_int,
                                             C.POINTER(ReportC)]),
    "bsar_range_compress_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _dp, C.c_int,
                                                     C.c_int, C.POINTER(ReportC)]),
    "bsar_fft": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int]),
    "bsar_chirp": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
    "bsar_dataset_create": (C.c_int, [C.c_void_p, C.POINTER(ParamsC), C.POINTER(DemC), _dp, _dp, C.c_int,
                                      C.POINTER(C.c_void_p)]),
    "bsar_dataset_destroy": (None, [C.c_void_p]),
    "bsar_dataset_load_rootdata": (C.c_int, [C.c_void_p, C.c_void_p, _dp, C.c_int, C.c_int]),
    "bsar_dataset_load_rootdata_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                                    C.c_int]),
    "bsar_dataset_load_upsampled": (C.c_int, [C.c_void_p, C.c_void_p, _dp, C.c_int, C.c_int]),
    "bsar_dataset_device_bytes": (C.c_int64, [C.c_void_p]),
    "bsar_backproject": (C.c_int, [C.c_void_p, C.c_void_p, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp,
                                   C.POINTER(ReportC)]),
    "bsar_backproject_device": (C.c_int, [C.c_void_p, C.c_void_p, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int,
                                          C.c_void_p, C.POINTER(ReportC)]),
    "bsar_read_matrix": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.c_int, C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), _dp, C.POINTER(C.c_uint64)]),
    "bsar_write_image": (C.c_int, [C.c_char_p, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]),
    "bsar_write_matrix2d": (C.c_int, [C.c_char_p, _dp, C.c_uint64, C.c_uint64]),
    "bsar_flops_per_voxel_pulse": (C.c_double, [C.c_int]),
    "bsar_file_error": (C.c_char_p, []),
    "bsar_spiral": (C.c_int, [C.c_double] * 7 + [C.c_int, _dp, _dp, _dp, C.c_int64, C.POINTER(C.c_int64)]),
    "bsar_grid_axis": (C.c_int, [C.c_double, C.c_double, C.c_double, _dp, C.c_int, C.POINTER(C.c_int)]),
    "bsar_dem_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(DemGridC)]),
    "bsar_dem_grid_free": (None, [C.POINTER(DemGridC)]),
    "bsar_geo_to_utm": (C.c_int, [C.c_double, C.c_double, C.c_int, _dp, _dp, C.POINTER(C.c_int)]),
    "bsar_meridian_arc": (C.c_double, [C.c_double]),
    "bsar_server_create": (C.c_int, [C.c_int, C.POINTER(ServerConfigC), C.POINTER(C.c_void_p)]),
    "bsar_server_destroy": (None, [C.c_void_p]),
    "bsar_server_last_error": (C.c_char_p, [C.c_void_p]),
    "bsar_server_clock": (C.c_double, [C.c_void_p]),
    "bsar_server_submit": (C.c_int, [C.c_void_p, C.POINTER(JobDescC), C.POINTER(C.c_int64)]),
    "bsar_server_poll": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(JobInfoC)]),
    "bsar_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.POINTER(JobInfoC)]),
    "bsar_server_wait_any": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(JobInfoC)]),
    "bsar_server_drain": (C.c_int, [C.c_void_p]),
    "bsar_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStatsC)]),
}

_lib = None
_live = weakref.WeakSet()     # open contexts and datasets, closed before interpreter exit


@atexit.register
def _close_all():
    objs = list(_live)
    for kind in (Server, Dataset):       # servers, then datasets, then contexts
        for o in objs:
            if isinstance(o, kind):
                o.close()
    for o in objs:
        o.close()


def lib() -> C.CDLL:
    """Load libbsar.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("BSAR_LIB", LIB_PATH))
        if not path.exists():
            raise ImportError(f"{path} is missing: build it with `python c-_bistatic_sar_amd/build.py`")
        L = C.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


@dataclass
class Dem:
    """ProcessedDEM: x[nx], y[ny] ascending, z[ny][nx] (row = y)."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray

    def c(self):
        self._x, self._y, self._z = _f64(self.x), _f64(self.y), _f64(self.z)
        zr, zc = (self._z.shape if self._z.ndim == 2 else (0, 0))
        s = DemC(_ptr(self._x), len(self._x), _ptr(self._y), len(self._y), _ptr(self._z), zr, zc)
        return s


@dataclass
class Params:
    fc: float
    c: float
    fs: float
    n1: float
    n2: float
    upsample: int = 4

    def c_struct(self):
        return ParamsC(self.fc, self.c, self.fs, self.n1, self.n2, self.upsample, 0)


class Context:
    """One HIP stream on one GPU (bsar_ctx)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        st = self.L.bsar_ctx_create(device, C.byref(h))
        if st != 0:
            raise BsarError(st, f"bsar_ctx_create(device={device})")
        self.h = h
        self._children = weakref.WeakSet()
        _live.add(self)

    def close(self):
        if getattr(self, "h", None):
            for d in list(self._children):
                d.close()
            self.L.bsar_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, st: int, what: str):
        if st != 0:
            raise BsarError(st, f"{what}: {self.L.bsar_last_error(self.h).decode()}")

    # -- stand-alone kernels ------------------------------------------------
    def dem_interp(self, dem: Dem, xy) -> np.ndarray:
        xy = _f64(xy).reshape(-1, 2)
        out = np.empty(len(xy))
        d = dem.c()
        self._chk(self.L.bsar_dem_interp(self.h, C.byref(d), _ptr(xy), len(xy), _ptr(out)), "dem_interp")
        return out

    def refraction_points(self, dem: Dem, legs, n1: float, n2: float) -> np.ndarray:
        legs = _f64(legs).reshape(-1, 6)
        out = np.empty((len(legs), 3))
        d = dem.c()
        self._chk(self.L.bsar_refraction_points(self.h, C.byref(d), _ptr(legs), len(legs), n1, n2, _ptr(out)),
                  "refraction_points")
        return out

    def refraction_points_wave(self, dem: Dem, legs, n1: float, n2: float) -> tuple[np.ndarray, float]:
        """refraction_points by the unscreened wavefront reduction; (points, kernel seconds)."""
        legs = _f64(legs).reshape(-1, 6)
        out = np.empty((len(legs), 3))
        secs = C.c_double(0.0)
        d = dem.c()
        self._chk(self.L.bsar_refraction_points_wave(self.h, C.byref(d), _ptr(legs), len(legs), n1, n2, _ptr(out),
                                                     C.byref(secs)), "refraction_points_wave")
        return out, secs.value

    def upsample(self, data, factor: int = 4) -> np.ndarray:
        """data [N][S] complex -> [N][S*factor] complex."""
        d = _c128(data)
        N, S = d.shape
        out = np.empty((N, S * factor), np.complex128)
        self._chk(self.L.bsar_upsample(self.h, _ptr(d.view(np.float64)), S, N, factor, _ptr(out.view(np.float64))),
                  "upsample")
        return out

    def range_compress(self, sig, ref, conj: bool = True, precision: int = FP64) -> np.ndarray:
        """sig [N][S] complex columns, ref [S] -> [N][S] correlation (conj) / convolution."""
        s = _c128(sig)
        N, S = s.shape
        r = _c128(ref).reshape(-1)
        if len(r) != S:
            raise ValueError("reference length must equal the column length")
        out = np.empty_like(s)
        self._chk(self.L.bsar_range_compress(self.h, _ptr(s.view(np.float64)), S, N, _ptr(r.view(np.float64)),
                                             int(conj), precision, _ptr(out.view(np.float64))), "range_compress")
        return out

    def fft(self, data, inverse: bool = False) -> np.ndarray:
        """FFTProcessor::fft / ::ifft of each row of data ([N][S] or [S]
        complex); returns the transformed copy."""
        d = _c128(data).copy()
        S = d.shape[-1] if d.ndim else 1
        N = d.size // S if S else 0
        ptr = _ptr(d.view(np.float64)) if d.size else None
        self._chk(self.L.bsar_fft(self.h, ptr, S, N, int(inverse)), "fft")
        return d

    def range_compress_device(self, buf: "DeviceBuffer", S: int, N: int, ref, conj: bool = True,
                              precision: int = FP64, noalloc: bool = False) -> dict:
        """In place on a device buffer of [N][S] complex128 (FP64) / complex64
        (MIXED); returns the run report (device time of the filter).  noalloc
        calls bsar_range_compress_device_noalloc (the library's fixed per-GPU
        scratch instead of device allocations)."""
        esz = 16 if precision == FP64 else 8
        if buf.nbytes < S * N * esz:
            raise ValueError("device buffer too small")
        r = _c128(ref).reshape(-1)
        if len(r) != S:
            raise ValueError("reference length must equal the column length")
        rep = ReportC()
        fn = self.L.bsar_range_compress_device_noalloc if noalloc else self.L.bsar_range_compress_device
        self._chk(fn(self.h, C.c_void_p(buf.ptr), S, N, _ptr(r.view(np.float64)), int(conj), precision,
                     C.byref(rep)), "range_compress_device_noalloc" if noalloc else "range_compress_device")
        return {f: getattr(rep, f) for f, _ in ReportC._fields_}

    def dataset(self, params: Params, dem: Dem, tx, rx) -> "Dataset":
        return Dataset(self, params, dem, tx, rx)


class Dataset:
    """bsar_dataset: DEM, antennas and (upsampled) range data resident on the GPU."""

    def __init__(self, ctx: Context, params: Params, dem: Dem, tx, rx):
        self.ctx = ctx
        self.L = ctx.L
        self.tx = _f64(tx).reshape(-1, 3)
        self.rx = _f64(rx).reshape(-1, 3)
        if self.tx.shape != self.rx.shape:
            raise ValueError("Tx and Rx need the same number of positions")
        self.n = len(self.tx)
        self.params = params
        p = params.c_struct()
        d = dem.c()
        h = C.c_void_p()
        ctx._chk(self.L.bsar_dataset_create(ctx.h, C.byref(p), C.byref(d), _ptr(self.tx), _ptr(self.rx), self.n,
                                            C.byref(h)), "dataset_create")
        self.h = h
        ctx._children.add(self)
        _live.add(self)

    def close(self):
        if getattr(self, "h", None):
            self.L.bsar_dataset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_rootdata(self, root):
        """root [N_data][S] complex (range fastest), upsampled on the GPU."""
        r = _c128(root)
        nd, S = r.shape
        self.ctx._chk(self.L.bsar_dataset_load_rootdata(self.ctx.h, self.h, _ptr(r.view(np.float64)), S, nd),
                      "load_rootdata")

    def load_rootdata_device(self, buf: "DeviceBuffer", S: int, n_data: int, precision: int = FP64):
        """rootData from a device buffer ([n_data][S] complex128, or complex64 with MIXED)."""
        esz = 16 if precision == FP64 else 8
        if buf.nbytes < S * n_data * esz:
            raise ValueError("device buffer too small")
        self.ctx._chk(self.L.bsar_dataset_load_rootdata_device(self.ctx.h, self.h, C.c_void_p(buf.ptr), S, n_data,
                                                               precision), "load_rootdata_device")

    def load_upsampled(self, rc):
        r = _c128(rc)
        nd, S4 = r.shape
        self.ctx._chk(self.L.bsar_dataset_load_upsampled(self.ctx.h, self.h, _ptr(r.view(np.float64)), S4, nd),
                      "load_upsampled")

    def device_bytes(self) -> int:
        return int(self.L.bsar_dataset_device_bytes(self.h))

    def slant_range_fermat(self, pts):
        p = _f64(pts).reshape(-1, 3)
        ida = np.empty((len(p), self.n, 3))
        volta = np.empty_like(ida)
        self.ctx._chk(self.L.bsar_slant_range_fermat(self.ctx.h, self.h, _ptr(p), len(p), _ptr(ida), _ptr(volta)),
                      "slant_range_fermat")
        return ida, volta

    def process_raw(self, scat, rcs, chirp: tuple, S: int) -> dict:
        """processRawData on the GPU: scat [ns][3], rcs [ns] complex, chirp =
        (f0, f1, T, fs_tx); returns aux, raw, root as [N][S] complex (pulse-major)."""
        sc = _f64(scat).reshape(-1, 3)
        r = _c128(np.broadcast_to(np.asarray(rcs, np.complex128), (len(sc),)))
        ch = ChirpC(*map(float, chirp))
        out = {k: np.empty((self.n, S), np.complex128) for k in ("aux", "raw", "root")}
        self.ctx._chk(self.L.bsar_process_raw(self.ctx.h, self.h, _ptr(sc), _ptr(r.view(np.float64)), len(sc),
                                              C.byref(ch), S, *(_ptr(out[k].view(np.float64))
                                                                for k in ("aux", "raw", "root"))), "process_raw")
        return out

    def slant_range_fermat_device(self, d_pts: int, V: int, d_ida: int, d_volta: int) -> dict:
        """Device-pointer form (caller-owned buffers: pts [V][3], ida/volta [V][N][3]);
        returns the run report."""
        rep = ReportC()
        self.ctx._chk(self.L.bsar_slant_range_fermat_device(self.ctx.h, self.h, C.c_void_p(d_pts), V,
                                                            C.c_void_p(d_ida), C.c_void_p(d_volta), C.byref(rep)),
                      "slant_range_fermat_device")
        return {f: getattr(rep, f) for f, _ in ReportC._fields_}

    def backproject(self, gx, gy, gz, precision: int = FP64):
        """Image [nz][ny][nx] complex (x fastest) and the run report (dict)."""
        gx, gy, gz = _f64(gx), _f64(gy), _f64(gz)
        out = np.empty((len(gz), len(gy), len(gx)), np.complex128)
        rep = ReportC()
        self.ctx._chk(self.L.bsar_backproject(self.ctx.h, self.h, _ptr(gx), len(gx), _ptr(gy), len(gy), _ptr(gz),
                                              len(gz), precision, _ptr(out.view(np.float64)), C.byref(rep)),
                      "backproject")
        return out, {f: getattr(rep, f) for f, _ in ReportC._fields_}


def chirp(f0: float, f1: float, T: float, fs: float) -> np.ndarray:
    L = lib()
    n = L.bsar_chirp(f0, f1, T, fs, None)
    if n < 0:
        raise BsarError(n, "chirp")
    out = np.empty(n, np.complex128)
    L.bsar_chirp(f0, f1, T, fs, _ptr(out.view(np.float64)))
    return out


def read_matrix(path) -> tuple[np.ndarray, list[int]]:
    """readBinary<double>: flat complex array and header dims."""
    L = lib()
    dims = (C.c_uint64 * 64)()
    nd, cplx, cnt = C.c_int(), C.c_int(), C.c_uint64()
    st = L.bsar_read_matrix(str(path).encode(), dims, 64, C.byref(nd), C.byref(cplx), None, C.byref(cnt))
    if st != 0:
        raise BsarError(st, f"read_matrix({path})")
    out = np.empty(cnt.value, np.complex128)
    st = L.bsar_read_matrix(str(path).encode(), dims, 64, C.byref(nd), C.byref(cplx), _ptr(out.view(np.float64)),
                            C.byref(cnt))
    if st != 0:
        raise BsarError(st, f"read_matrix({path})")
    return out, [dims[i] for i in range(nd.value)]


def write_image(path, img, gx=None, gy=None, gz=None):
    """export3DComplexMatrix: img [nz][ny][nx] complex."""
    L = lib()
    im = _c128(img)
    nz, ny, nx = im.shape
    axes = [None if a is None else _f64(a) for a in (gx, gy, gz)]
    ptrs = [None if a is None else _ptr(a) for a in axes]
    st = L.bsar_write_image(str(path).encode(), _ptr(im.view(np.float64)), nx, ny, nz, *ptrs)
    if st != 0:
        raise BsarError(st, f"write_image({path})")


def write_matrix2d(path, data):
    L = lib()
    d = _c128(data)
    rows, cols = d.shape
    st = L.bsar_write_matrix2d(str(path).encode(), _ptr(d.view(np.float64)), rows, cols)
    if st != 0:
        raise BsarError(st, f"write_matrix2d({path})")


def _file_check(st: int, what: str):
    if st != 0:
        raise BsarError(st, f"{what}: {lib().bsar_file_error().decode()}")


def spiral(turns: float, r_minor: float, r_major: float, h_high: float, h_low: float, vt: float, prf: float,
           mirror_xy: bool = False) -> np.ndarray:
    """funcao_espiral (funcao_espiral.cpp:5-49) as [n][3] positions; mirror_xy
    gives generateReceiverTrajectory's (-x, -y) (trajectory_generator.cpp:17-33)."""
    L = lib()
    n = C.c_int64()
    args = (turns, r_minor, r_major, h_high, h_low, vt, prf, int(bool(mirror_xy)))
    _file_check(L.bsar_spiral(*args, None, None, None, 0, C.byref(n)), "spiral")
    px, py, pz = (np.empty(n.value) for _ in range(3))
    _file_check(L.bsar_spiral(*args, _ptr(px), _ptr(py), _ptr(pz), n.value, C.byref(n)), "spiral")
    return np.stack([px, py, pz], 1)


def grid_axis(center: float, step: float, half: float) -> np.ndarray:
    """One axis of crearGrid (crear_grid.cpp:21-34)."""
    L = lib()
    n = C.c_int()
    _file_check(L.bsar_grid_axis(center, step, half, None, 0, C.byref(n)), "grid_axis")
    out = np.empty(n.value)
    _file_check(L.bsar_grid_axis(center, step, half, _ptr(out), n.value, C.byref(n)), "grid_axis")
    return out


DEM_FILE_REFERENCE, DEM_FILE_GEOTIFF = 0, 1
DEM_SOURCES = {0: "default", 1: "ascii_grid", 2: "geotiff"}


def load_dem(path, mode: int = DEM_FILE_REFERENCE) -> tuple[Dem, dict]:
    """loadDEM (load_dem_simple.cpp:9-224) with mode DEM_FILE_REFERENCE, or
    a GeoTIFF decoded to local metres with DEM_FILE_GEOTIFF.  Returns the DEM
    and its metadata (source, centre, UTM zone, nodata count)."""
    L = lib()
    g = DemGridC()
    _file_check(L.bsar_dem_load(str(path).encode(), mode, C.byref(g)), f"load_dem({path})")
    try:
        d = g.dem
        x = np.ctypeslib.as_array(d.x, (d.nx,)).copy() if d.nx else np.empty(0)
        y = np.ctypeslib.as_array(d.y, (d.ny,)).copy() if d.ny else np.empty(0)
        n = d.zrows * d.zcols
        z = np.ctypeslib.as_array(d.z, (n,)).copy().reshape(d.zrows, d.zcols) if n else np.empty((d.zrows, d.zcols))
        meta = {"source": DEM_SOURCES.get(g.source, g.source), "lon0": g.lon0, "lat0": g.lat0, "e0": g.e0,
                "n0": g.n0, "utm_zone": g.utm_zone, "nodata_count": g.nodata_count}
    finally:
        L.bsar_dem_grid_free(C.byref(g))
    return Dem(x, y, z), meta


def geo_to_utm(lat: float, lon: float, zone: int = 0) -> tuple[float, float, int]:
    """WGS84 (lat, lon) degrees -> UTM (easting, northing, zone; zone < 0 south)."""
    e, n, z = C.c_double(), C.c_double(), C.c_int()
    _file_check(lib().bsar_geo_to_utm(lat, lon, zone, C.byref(e), C.byref(n), C.byref(z)), "geo_to_utm")
    return e.value, n.value, z.value


def meridian_arc(lat: float) -> float:
    return float(lib().bsar_meridian_arc(lat))


def flops_per_voxel_pulse(n: int) -> float:
    return float(lib().bsar_flops_per_voxel_pulse(n))


_hip = None


def _hiprt() -> C.CDLL:
    """The HIP runtime libbsar.so is linked against (device buffers for the
    *_device entry points)."""
    global _hip
    if _hip is None:
        lib()
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                H = C.CDLL(name)
                break
            except OSError:
                H = None
        if H is None:
            raise ImportError("libamdhip64.so not found")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip = H
    return _hip


class DeviceBuffer:
    """A device allocation (hipMalloc) owned by Python."""

    def __init__(self, nbytes: int):
        self.H = _hiprt()
        p = C.c_void_p()
        st = self.H.hipMalloc(C.byref(p), max(int(nbytes), 1))
        if st != 0:
            raise BsarError(-3, f"hipMalloc({nbytes}) failed with HIP error {st}")
        self.ptr, self.nbytes = p.value, int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.write(a)
        return b

    def write(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("write past the end of the buffer")
        st = self.H.hipMemcpy(C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p), a.nbytes, 1)
        if st != 0:
            raise BsarError(-2, f"hipMemcpy H2D failed with HIP error {st}")

    def read(self, shape, dtype=np.float64, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("read past the end of the buffer")
        st = self.H.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), out.nbytes, 2)
        if st != 0:
            raise BsarError(-2, f"hipMemcpy D2H failed with HIP error {st}")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            self.H.hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


JOB_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


def _info(ci: JobInfoC) -> dict:
    d = {f: getattr(ci, f) for f, _ in JobInfoC._fields_}
    d["state"] = JOB_STATES.get(d["state"], d["state"])
    d["latency"] = d["t_done"] - d["t_submit"]
    return d


class Server:
    """bsar_server: continuous batching of imaging jobs on one GPU.

    ``submit`` queues a job (image of a grid over the first K pulses of a
    resident dataset) and returns its id; ``wait`` returns (image [nz][ny][nx]
    complex, info dict) once it is done.  Datasets must outlive their jobs.
    """

    def __init__(self, device: int = 0, max_jobs: int = 0, hbm_cap_bytes: int = 0, round_seconds: float = 0.0):
        self.L = lib()
        cfg = ServerConfigC(max_jobs, 0, int(hbm_cap_bytes), float(round_seconds))
        h = C.c_void_p()
        st = self.L.bsar_server_create(device, C.byref(cfg), C.byref(h))
        if st != 0:
            raise BsarError(st, f"bsar_server_create(device={device})")
        self.h = h
        self._jobs = {}          # id -> (out array, axes, dataset): kept alive until collected
        _live.add(self)

    def close(self):
        if getattr(self, "h", None):
            self.L.bsar_server_destroy(self.h)
            self.h = None
            self._jobs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _err(self) -> str:
        return self.L.bsar_server_last_error(self.h).decode()

    def clock(self) -> float:
        return float(self.L.bsar_server_clock(self.h))

    def submit(self, ds: Dataset, gx, gy, gz, pulses: int = 0, precision: int = FP64) -> int:
        gx, gy, gz = _f64(gx), _f64(gy), _f64(gz)
        out = np.empty((len(gz), len(gy), len(gx)), np.complex128)
        d = JobDescC(ds.h, int(pulses), int(precision), _ptr(gx), len(gx), _ptr(gy), len(gy), _ptr(gz), len(gz),
                     _ptr(out.view(np.float64)))
        jid = C.c_int64()
        st = self.L.bsar_server_submit(self.h, C.byref(d), C.byref(jid))
        if st != 0:
            raise BsarError(st, f"server_submit: {self._err()}")
        self._jobs[jid.value] = (out, ds)
        return jid.value

    def poll(self, jid: int) -> dict:
        ci = JobInfoC()
        st = self.L.bsar_server_poll(self.h, jid, C.byref(ci))
        if st != 0:
            raise BsarError(st, f"server_poll: {self._err()}")
        return _info(ci)

    def wait(self, jid: int, timeout: float = -1.0):
        ci = JobInfoC()
        st = self.L.bsar_server_wait(self.h, jid, float(timeout), C.byref(ci))
        if st == -7 and ci.state not in (2, 3):
            raise TimeoutError(f"job {jid} not done after {timeout} s")
        out, _ = self._jobs.pop(jid, (None, None))
        if st != 0:
            raise BsarError(st, f"job {jid}: {self._err()}")
        return out, _info(ci)

    def wait_any(self, timeout: float = -1.0):
        """Next finished job: (id, image, info)."""
        ci = JobInfoC()
        st = self.L.bsar_server_wait_any(self.h, float(timeout), C.byref(ci))
        if st == -7 and ci.id == 0:
            raise TimeoutError(f"no job done after {timeout} s")
        if ci.id == 0:
            raise BsarError(st, f"server_wait_any: {self._err()}")
        out, _ = self._jobs.pop(ci.id, (None, None))
        if st != 0:
            raise BsarError(st, f"job {ci.id}: {self._err()}")
        return ci.id, out, _info(ci)

    def drain(self):
        st = self.L.bsar_server_drain(self.h)
        if st != 0:
            raise BsarError(st, f"server_drain: {self._err()}")

    def stats(self) -> dict:
        cs = ServerStatsC()
        st = self.L.bsar_server_get_stats(self.h, C.byref(cs))
        if st != 0:
            raise BsarError(st, "server_stats")
        return {f: getattr(cs, f) for f, _ in ServerStatsC._fields_}
