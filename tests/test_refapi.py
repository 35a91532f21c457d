"""The reference's own hot-path declarations defined over the library
(c-_bistatic_sar_amd/refapi/bsar_reference_api.cpp), the facade's
FFTProcessor / reflection point, and the batched bsar_fft entry point,
against the compiled reference's golden vectors (tests/golden/ref_kernels.npz).

oracle/_ref/refapi_demo is built (oracle/Makefile) from tests/refapi_demo.cpp,
whose call sites are written against the reference's own headers, linked
with the adapter and libbsar in place of the reference's
raw_data_processing.cpp / fft_utils.cpp / interpolation.cpp.

Tolerances: refraction points, ranges, reflection points, spline and the
placed responses bit-exact; transforms within 1e-12 of the reference's peak
(its twiddles come from a multiplicative recurrence, the GPU's are computed
directly: ~3e-17 * S apart)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR, ROOT
from test_dropin import demo_exe
from test_oracle_golden import DEMS, bits_equal

REF_INCLUDE = Path("/root/reference/include")
REFAPI = ROOT / "oracle" / "_ref" / "refapi_demo"
SIZES = [1, 2, 8, 64, 1024, 7, 12]
FFT_TOL = 1e-12

hexs = lambda a: " ".join(float(v).hex() for v in np.ravel(a))


def parse_rows(text):
    return [np.array(list(map(float.fromhex, ln.split()))) for ln in text.strip().splitlines()]


def cplx(row):
    return row[0::2] + 1j * row[1::2]


def scene_input(g):
    x, y, z = g["dem_bumps_x"], g["dem_bumps_y"], g["dem_bumps_z"]
    lines = [f"{len(x)} {len(y)}", hexs(x), hexs(y)] + [hexs(r) for r in z]
    lines += [str(len(g["sr_tx"]))] + [hexs(r) for r in g["sr_tx"]] + [hexs(r) for r in g["sr_rx"]]
    lines += [hexs(g["sr_P"]), hexs([1.0, 2.449489742783178])]
    return lines


# ---- CPU ------------------------------------------------------------------

@pytest.mark.parametrize("name", DEMS + ["empty"])
def test_oracle_reflection_point(oracle, golden, name):
    """or_reflection_point against the reference's calculateReflectionPoint
    (raw_data_processing.cpp:575-612)."""
    g = golden
    if name == "empty":
        e = np.zeros(0)
        dem = (e, e, np.zeros((0, 0)))
    else:
        dem = (g[f"dem_{name}_x"], g[f"dem_{name}_y"], g[f"dem_{name}_z"])
    assert bits_equal(oracle.reflection_points(dem, g[f"refl_{name}_pairs"]), g[f"refl_{name}_out"])


@pytest.mark.parametrize("name", DEMS + ["empty"])
def test_facade_reflection_point(tmp_path, golden, name):
    """bsar::facade::calculateReflectionPoint (host code) on the reference's
    vectors, including Z smaller than its axes and an empty DEM."""
    g = golden
    if name == "empty":
        x = y = np.zeros(0)
        z = np.zeros((0, 0))
    else:
        x, y, z = g[f"dem_{name}_x"], g[f"dem_{name}_y"], g[f"dem_{name}_z"]
    pr = g[f"refl_{name}_pairs"]
    lines = [f"{len(x)} {len(y)}", f"{z.shape[0]} {z.shape[1] if z.ndim == 2 else 0}", hexs(x), hexs(y)]
    lines += [hexs(r) for r in z] + [str(len(pr))] + [hexs(r) for r in pr]
    r = subprocess.run([str(demo_exe(tmp_path)), "reflect"], input="\n".join(lines) + "\n", capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert bits_equal(np.array(parse_rows(r.stdout)), g[f"refl_{name}_out"])


@pytest.mark.skipif(not REF_INCLUDE.exists(), reason="the reference's headers are not present")
def test_adapter_compiles_against_the_reference_headers(tmp_path):
    """The adapter defines every function the reference's raw_data_processing.h,
    fft_utils.h and interpolation.h declare: a program written against those
    headers links with it and libbsar alone (no reference implementation of
    them), the reference's own RangeData taking calculaSlantRangeFermat's
    result (proc_test_script.cpp:484-485)."""
    exe = tmp_path / "refapi_demo"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-w", f"-I{REF_INCLUDE}", f"-I{ROOT / 'include'}",
                        str(ROOT / "tests" / "refapi_demo.cpp"), str(PKG_DIR / "refapi" / "bsar_reference_api.cpp"),
                        str(REF_INCLUDE.parent / "src" / "project_paths.cpp"), "-o", str(exe), f"-L{PKG_DIR}",
                        "-lbsar", f"-Wl,-rpath,{PKG_DIR}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    und = subprocess.run(["nm", "-C", "-u", str(exe)], capture_output=True, text=True).stdout
    for sym in ("calculaSlantRangeFermat", "calculateRefractionPoint", "interpolateDEMElevation",
                "splineInterpolation", "FFTProcessor::", "processRawData", "createCHIRP"):
        assert sym not in und, sym
    # malformed input exits 2 before any GPU call
    assert subprocess.run([str(exe), "fft"], input="x\n", capture_output=True, text=True).returncode == 2
    assert subprocess.run([str(exe), "nomode"], input="", capture_output=True, text=True).returncode == 2


# ---- GPU ------------------------------------------------------------------

needs_demo = pytest.mark.skipif(not REFAPI.exists(), reason="oracle/_ref/refapi_demo is not built")


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_bsar_fft_matches_reference(ctx, golden, n):
    """bsar_fft (FFTProcessor::fft / ::ifft, fft_utils.cpp:11-60) on the
    reference's own vectors: forward and inverse, power-of-two lengths and the
    no-op rule for the others."""
    a = golden[f"fft_{n}_in"].view(np.complex128)
    for inv, key in ((False, "fft"), (True, "ifft")):
        want = golden[f"{key}_{n}_out"].view(np.complex128)
        got = ctx.fft(a, inverse=inv)
        assert np.abs(got - want).max() <= FFT_TOL * max(np.abs(want).max(), 1e-300), (key, n)
        if n in (1, 7, 12):            # the reference leaves these untransformed
            assert bits_equal(got.view(np.float64), want.view(np.float64)), (key, n)
    # batched: every row transformed as a column of its own
    rng = np.random.default_rng(n)
    rows = rng.normal(size=(5, n)) + 1j * rng.normal(size=(5, n))
    want = np.stack([ctx.fft(r) for r in rows])
    assert bits_equal(ctx.fft(rows).view(np.float64), want.view(np.float64))


@pytest.mark.gpu
@needs_demo
@pytest.mark.parametrize("n", SIZES)
def test_adapter_fft_family(golden, n):
    g = golden
    a, b = g[f"fft_{n}_in"], g[f"corr_{n}_ref"]
    r = subprocess.run([str(REFAPI), "fft"], input=f"{n}\n{hexs(a)}\n{hexs(b)}\n", capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [cplx(x) for x in parse_rows(r.stdout)] if n else []
    for got, key in zip(rows[:4], (f"fft_{n}_out", f"ifft_{n}_out", f"corr_{n}_out", f"conv_{n}_out")):
        want = g[key].view(np.complex128)
        assert np.abs(got - want).max() <= FFT_TOL * np.abs(want).max(), key
    A, B = a.view(np.complex128), b.view(np.complex128)
    # multiply, :140-150: std::complex's (ac - bd, ad + bc), each product
    # rounded on its own (numpy's complex multiply may fuse them)
    prod = np.empty_like(A)
    prod.real = A.real * B.real - A.imag * B.imag
    prod.imag = A.real * B.imag + A.imag * B.real
    assert bits_equal(rows[4].view(np.float64), prod.view(np.float64))
    P = 1 << int(np.ceil(np.log2(max(n, 1))))
    pad = np.zeros(P, complex)
    pad[:n] = A
    assert bits_equal(rows[5].view(np.float64), pad.view(np.float64))             # padToPowerOfTwo, :154-161
    if n > 1 and n & (n - 1):
        assert "not a power of 2" in r.stderr


@pytest.mark.gpu
@needs_demo
def test_adapter_slant_range_and_reflection_bit_exact(golden):
    g = golden
    r = subprocess.run([str(REFAPI), "slant"], input="\n".join(scene_input(g)) + "\n", capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v = np.array(parse_rows(r.stdout))
    assert np.array_equal(v[:, 0:3], g["sr_ida"]) and np.array_equal(v[:, 3:6], g["sr_volta"])
    assert np.array_equal(v[:, 6:10], g["sr_r4"].T)
    assert np.array_equal(v[:, 10:13], g["sr_refl"])


@pytest.mark.gpu
@needs_demo
@pytest.mark.parametrize("nf", [(2, 4), (5, 3), (257, 4), (9, 7)])
def test_adapter_spline_bit_exact(golden, nf):
    n, f = nf
    d = golden[f"spline_{n}_{f}_in"]
    r = subprocess.run([str(REFAPI), "spline"], input=f"{n} {f}\n{hexs(d)}\n", capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert bits_equal(parse_rows(r.stdout)[0], golden[f"spline_{n}_{f}_out"])


@pytest.mark.gpu
@needs_demo
def test_adapter_process_raw_matches_reference(golden):
    """processRawData (raw_data_processing.cpp:292-452) through the adapter:
    the reference's one-target simulation at its own radar constants."""
    g = golden
    fc, c, fs_rx, fs_tx, f0, f1, T, n1, n2 = g["sim_args"]
    S = int(g["sim_S"][0])
    lines = scene_input(g) + [hexs([fc, c, fs_rx, fs_tx, f0, f1, T, float(S), 1.0])]
    r = subprocess.run([str(REFAPI), "sim"], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [cplx(x) for x in parse_rows(r.stdout)]
    assert len(rows) == 3 * S
    aux, raw, root = (np.array(rows[k * S:(k + 1) * S]) for k in range(3))
    ref = lambda a: a[..., 0] + 1j * a[..., 1]
    assert bits_equal(aux.view(np.float64), ref(g["sim_aux"]).view(np.float64))
    for got, key in ((raw, "sim_raw"), (root, "sim_root")):
        want = ref(g[key])
        assert np.abs(got - want).max() <= FFT_TOL * np.abs(want).max(), key
