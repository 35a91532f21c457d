"""The refraction search of the HIP kernels (csrc/bsar_device.h, fermat_solve)
compiled for the CPU (tests/search_host.cpp through tests/host_shim) against
the oracle's exhaustive 709-candidate search (oracle/bsar_oracle.c,
or_refraction_point = raw_data_processing.cpp:455-572): the argmin must be
bit-identical for every leg, in each template configuration a kernel
instantiates.

This pins the search's pruning -- the one- and two-axis lower bounds with their
side stops, the window's bilinear height range, the FP32 estimates and their
error margins -- over far more geometries than the GPU tests reach: the
workloads' spirals over a bumpy DEM, 40 m RMS rough terrain, random end points
(below the terrain too, random refraction indices, legs leaving the DEM), a
DEM with NaN nodes and a flat DEM (exact ties between candidates).  A planted
mistake (the bounds' FP32 reciprocal root 1 % high, so they overshoot) must
make it fail."""
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_scene

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")

CONFIGS = {0: "mixed BP, DEM in LDS", 1: "plain FP64 BP, DEM in LDS", 2: "refraction kernels, DEM in LDS",
           3: "DEM outside LDS", 4: "plain FP64 BP, DEM outside LDS"}


@pytest.fixture(scope="module")
def search_host(tmp_path_factory):
    exe = tmp_path_factory.mktemp("search") / "search_host"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-w",
                        f"-I{ROOT / 'tests' / 'host_shim'}", f"-I{ROOT / 'c-_bistatic_sar_amd' / 'csrc'}",
                        str(ROOT / "tests" / "search_host.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


def run_host(exe, tmp, dem, legs, cfg, plant=1.0):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in dem)
    src, dst = tmp / f"in{cfg}.bin", tmp / f"out{cfg}.bin"
    with open(src, "wb") as f:
        np.array([len(x), len(y)], np.int32).tofile(f)
        np.array([len(legs)], np.int64).tofile(f)
        x.tofile(f)
        y.tofile(f)
        z.reshape(len(y), len(x)).tofile(f)
        np.ascontiguousarray(legs, np.float64).tofile(f)
    r = subprocess.run([str(exe), str(src), str(dst), str(cfg), repr(plant)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    return np.fromfile(dst, np.float64).reshape(-1, 3)


def oracle_points(oracle, dem, legs):
    """or_refraction_point per leg (n1, n2 may differ between legs)."""
    out = np.empty((len(legs), 3))
    for (n1, n2) in {(a, b) for a, b in legs[:, 6:8]}:
        sel = (legs[:, 6] == n1) & (legs[:, 7] == n2)
        out[sel] = oracle.refraction_batch(dem, legs[sel, :6], n1, n2)
    return out


def spiral_legs(rng, m, gx, gy, gz, h, n2, x0=0.0, y0=0.0):
    """Both legs of m random (voxel, pulse) pairs of the workloads' geometry."""
    p = rng.uniform(0, 1, m)
    r, th = 5 + 10 * p, 4 * np.pi * p
    tx = np.stack([x0 + r * np.cos(th), y0 + r * np.sin(th), np.full(m, h)], 1)
    rx = np.stack([x0 + r * np.cos(th + np.pi / 2), y0 + r * np.sin(th + np.pi / 2), np.full(m, h)], 1)
    vox = np.stack([rng.uniform(*gx, m), rng.uniform(*gy, m), rng.uniform(*gz, m)], 1)
    ones = np.ones((m, 1))
    return np.concatenate([np.concatenate([tx, vox, ones, n2 * ones], 1),
                           np.concatenate([vox, rx, n2 * ones, ones], 1)])


def random_legs(rng, m, span, zlo, zhi):
    """Arbitrary end points, heights above and below the terrain, n2 / n1 one
    of 1.0 .. 3.0 (a few values, so the oracle runs in large batches)."""
    a = np.stack([rng.uniform(-span, span, m), rng.uniform(-span, span, m), rng.uniform(zlo, zhi, m)], 1)
    b = np.stack([rng.uniform(-span, span, m), rng.uniform(-span, span, m), rng.uniform(zlo, zhi, m)], 1)
    n = np.stack([np.ones(m), rng.choice([1.0, 1.5, 2.0, 6 ** 0.5, 3.0], m)], 1)
    flip = rng.uniform(size=m) < 0.5
    n[flip] = n[flip][:, ::-1]
    return np.concatenate([a, b, n], 1)


def scenes():
    scene = load_scene()
    rng = np.random.default_rng(2024)
    bumps = scene.dem_bumps(np.random.default_rng(1002))
    rough = scene.dem_fbm(np.random.default_rng(1003))
    holes = tuple(np.array(a, float) for a in bumps)
    hz = holes[2].copy().reshape(len(holes[1]), len(holes[0]))
    hz[rng.integers(0, hz.shape[0], 40), rng.integers(0, hz.shape[1], 40)] = np.nan
    holes = (holes[0], holes[1], hz)
    flat = (bumps[0], bumps[1], np.zeros((len(bumps[1]), len(bumps[0]))))
    small = (np.linspace(-40, 40, 9), np.linspace(-30, 30, 7), rng.normal(0, 1.5, (7, 9)))
    zr = float(np.asarray(rough[2]).reshape(len(rough[1]), len(rough[0]))[len(rough[1]) // 2, len(rough[0]) // 2])
    return [
        ("workload spirals over bumps", bumps,
         np.concatenate([spiral_legs(rng, 24000, (19.2, 20.8), (-0.8, 0.8), (-5.8, -4.2), 12.0, 6 ** 0.5),
                         spiral_legs(rng, 8000, (19.2, 20.8), (-0.8, 0.8), (-5.8, -4.2), 12.0, 3 ** 0.5)])),
        ("spirals over rough terrain", rough,
         spiral_legs(rng, 24000, (-2, 2), (-2, 2), (zr - 7, zr - 3), zr + 12.0, 2.5)),
        ("random end points, bumps", bumps, random_legs(rng, 24000, 600.0, -20.0, 60.0)),
        ("random end points, rough terrain", rough, random_legs(rng, 24000, 900.0, -150.0, 200.0)),
        ("NaN nodes", holes, random_legs(rng, 16000, 300.0, -10.0, 30.0)),
        ("flat DEM (ties)", flat,
         np.concatenate([spiral_legs(rng, 8000, (-1, 1), (-1, 1), (-3, -1), 10.0, 2.0),
                         random_legs(rng, 8000, 100.0, -5.0, 20.0)])),
        ("small DEM, legs leaving it", small, random_legs(rng, 16000, 200.0, -10.0, 20.0)),
    ]


SCENES = scenes()
_WANT = {}


def want_points(oracle, case):
    if case not in _WANT:
        _, dem, legs = SCENES[case]
        _WANT[case] = oracle_points(oracle, dem, legs)
    return _WANT[case]


@pytest.mark.parametrize("cfg", sorted(CONFIGS))
@pytest.mark.parametrize("case", range(len(SCENES)), ids=[s[0] for s in SCENES])
def test_search_matches_exhaustive_search(search_host, oracle, tmp_path, case, cfg):
    name, dem, legs = SCENES[case]
    got = run_host(search_host, tmp_path, dem, legs, cfg)
    want = want_points(oracle, case)
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    bad = np.flatnonzero(~same.all(1))
    assert bad.size == 0, f"{name} / {CONFIGS[cfg]}: {bad.size} of {len(legs)} legs differ, first {bad[:5]}"


def test_planted_bound_error_is_caught(search_host, oracle, tmp_path):
    name, dem, legs = SCENES[0]
    got = run_host(search_host, tmp_path, dem, legs, 0, plant=1.01)
    want = want_points(oracle, 0)
    assert (got != want).any(1).sum() > 0
