"""Write tests/golden/*.npz from the compiled reference.

Runs only where the reference tree is present (it builds oracle/_ref from
/root/reference with oracle/Makefile).  Inputs are small seeded arrays; the
outputs are whatever the reference's own code returns for them:

  ref_kernels.npz  interpolateDEMElevation, calculateRefractionPoint,
                   calculaSlantRangeFermat, splineInterpolation,
                   FFTProcessor::fft/ifft/fftCorrelation/fftConvolution,
                   createCHIRP, processRawData  (through oracle/ref_shim.cpp)
  ref_bp_*.npz     output_3D_matrix.bin written by the unmodified
                   proc_test_script run on a generated build/output/proc/ set
  ref_scenario.npz the file set the unmodified gs_test_script writes for its
                   default scenario (radar_init.cpp constants, spiral
                   trajectories, crearGrid, loadDEM of a .tif name): the small
                   CSV/text files whole, the trajectory CSVs as SHA-256, the
                   three complex matrices on a subset of pulses

    python tests/golden/make_golden.py [--only scenario]
"""
from __future__ import annotations

import os
import shutil
import struct
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent
C = __import__("ctypes")


def q6(a):
    a = np.asarray(a, dtype=np.float64)
    return np.array([float("%.6f" % v) for v in a.ravel()]).reshape(a.shape)


def p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def dem_args(x, y, z):
    return [p(x), len(x), p(y), len(y), p(z), z.shape[0], z.shape[1]]


def make_dems(rng):
    dems = {}
    ax = q6(-105 + 30 * np.arange(8))
    X, Y = np.meshgrid(ax, ax)
    z = q6(0.004 * X - 0.003 * Y + 1.5 * np.exp(-((X - 10) ** 2 + (Y + 20) ** 2) / (2 * 60.0 ** 2)))
    dems["bumps"] = (ax, ax.copy(), z)
    dems["flat"] = (ax, ax.copy(), np.zeros((8, 8)))
    xr = q6(np.sort(rng.uniform(-120, 120, 9)))
    yr = q6(np.sort(rng.uniform(-100, 100, 7)))
    dems["ragged"] = (xr, yr, q6(rng.normal(0, 3, (7, 9))))
    dems["rough"] = (q6(-150 + 30 * np.arange(11)), q6(-120 + 30 * np.arange(9)), q6(rng.normal(0, 8, (9, 11))))
    dems["unsorted"] = (q6(np.array([-60.0, 30, -30, 0, 60, 90])), q6(np.array([-40.0, 0, 40, 20])),
                        q6(rng.normal(0, 2, (4, 6))))
    dems["one_row"] = (q6(np.array([-30.0, 0.0, 30.0])), q6(np.array([5.0])), q6(np.array([[1.0, 2.0, 3.0]])))
    dems["dup"] = (q6(np.array([-30.0, 0.0, 0.0, 30.0])), q6(np.array([-30.0, 30.0])), q6(rng.normal(0, 1, (2, 4))))
    dems["small_z"] = (ax, ax.copy(), q6(rng.normal(0, 1, (5, 6))))   # Z smaller than X/Y vectors
    return dems


def kernels(rng):
    R = O.ref_lib()
    res = {}
    dems = make_dems(rng)
    for name, (x, y, z) in dems.items():
        x, y, z = map(np.ascontiguousarray, (x, y, z))
        m = 1500
        xy = np.stack([rng.uniform(x.min() - 40, x.max() + 40, m), rng.uniform(y.min() - 40, y.max() + 40, m)], 1)
        xy[:150, 0] = x[rng.integers(0, len(x), 150)]
        xy[100:250, 1] = y[rng.integers(0, len(y), 150)]
        xy[300] = [np.nan, 0.0]
        xy = np.ascontiguousarray(xy)
        zo = np.empty(m)
        R.ref_dem_interp(*dem_args(x, y, z), p(xy), m, p(zo))
        res[f"dem_{name}_x"], res[f"dem_{name}_y"], res[f"dem_{name}_z"] = x, y, z
        res[f"dem_{name}_xy"], res[f"dem_{name}_out"] = xy, zo
        # refraction legs: antenna above, voxel below, both directions
        M = 200
        top = np.stack([rng.uniform(-25, 25, M), rng.uniform(-25, 25, M), rng.uniform(8, 16, M)], 1)
        bot = np.stack([rng.uniform(-10, 30, M), rng.uniform(-10, 10, M), rng.uniform(-8, -2, M)], 1)
        legs = np.ascontiguousarray(np.concatenate([np.concatenate([top, bot], 1)[: M // 2],
                                                    np.concatenate([bot, top], 1)[M // 2:]]))
        for tag, n1, n2 in [("a", 1.0, 2.449489742783178), ("b", 2.449489742783178, 1.0), ("eq", 1.5, 1.5)]:
            ro = np.empty((M, 3))
            R.ref_refraction(*dem_args(x, y, z), p(legs), M, n1, n2, p(ro))
            res[f"refr_{name}_{tag}_legs"], res[f"refr_{name}_{tag}_out"] = legs, ro
            res[f"refr_{name}_{tag}_n"] = np.array([n1, n2])
    # empty DEM
    e = np.zeros(0)
    legs = np.ascontiguousarray(np.array([[3.0, -2.0, 12.0, 20.0, 0.5, -5.0], [20.0, 0.0, -5.0, -7.0, 8.0, 11.0]]))
    ro = np.empty((2, 3))
    R.ref_refraction(p(e), 0, p(e), 0, p(np.zeros(1)), 0, 0, p(legs), 2, 1.0, 2.0, p(ro))
    res["refr_empty_legs"], res["refr_empty_out"] = legs, ro
    # slant range of one target over a short spiral
    x, y, z = dems["bumps"]
    n = 12
    t = np.arange(n) / (n - 1)
    tx = q6(np.stack([(5 + 10 * t) * np.cos(4 * np.pi * t), (5 + 10 * t) * np.sin(4 * np.pi * t), np.full(n, 12.0)], 1))
    rx = q6(np.stack([(5 + 10 * t) * np.cos(4 * np.pi * t + np.pi / 2), (5 + 10 * t) * np.sin(4 * np.pi * t + np.pi / 2),
                      np.full(n, 12.0)], 1))
    P = np.array([20.0, 0.1, -5.0])
    ida, volta, r4 = np.empty((n, 3)), np.empty((n, 3)), np.empty(4 * n)
    R.ref_slant_range(*dem_args(x, y, z), p(tx), p(rx), n, p(P), 1.0, 2.449489742783178, p(ida), p(volta), p(r4))
    res.update(sr_tx=tx, sr_rx=rx, sr_P=P, sr_ida=ida, sr_volta=volta, sr_r4=r4.reshape(4, n))
    # spline
    for nn, f in [(2, 4), (3, 4), (5, 3), (64, 4), (257, 4), (16, 1), (9, 7)]:
        d = np.ascontiguousarray(rng.normal(size=2 * nn))
        o = np.empty(2 * nn * f)
        R.ref_spline(p(d), nn, f, p(o))
        res[f"spline_{nn}_{f}_in"], res[f"spline_{nn}_{f}_out"] = d, o
    # FFT family
    for nn in [1, 2, 8, 64, 1024, 7, 12]:
        a = np.ascontiguousarray(rng.normal(size=2 * nn))
        f1, f2 = a.copy(), a.copy()
        R.ref_fft(p(f1), nn)
        R.ref_ifft(p(f2), nn)
        res[f"fft_{nn}_in"], res[f"fft_{nn}_out"], res[f"ifft_{nn}_out"] = a, f1, f2
        b = np.ascontiguousarray(rng.normal(size=2 * nn))
        N = nn
        oc, ov = np.empty(2 * N), np.empty(2 * N)
        R.ref_fft_correlation(p(a), nn, p(b), nn, p(oc))
        R.ref_fft_convolution(p(a), nn, p(b), nn, p(ov))
        res[f"corr_{nn}_ref"], res[f"corr_{nn}_out"], res[f"conv_{nn}_out"] = b, oc, ov
    # chirps: reference constants and the workload constants
    for tag, (f0, f1, T, fs) in dict(refc=(-25e6, 25e6, 1.3e-6, 60e6), cfg=(-150e6, 150e6, 1e-6, 800e6)).items():
        L = R.ref_chirp(f0, f1, T, fs, None)
        o = np.empty(2 * L)
        R.ref_chirp(f0, f1, T, fs, p(o))
        res[f"chirp_{tag}_args"], res[f"chirp_{tag}_out"] = np.array([f0, f1, T, fs]), o
    # processRawData, one target, reference radar constants, S = 256
    S = 256
    x, y, z = dems["bumps"]
    sim = np.array([10e9, 3e8, 60e6, 60e6, -25e6, 25e6, 1.3e-6, 1.0, 2.449489742783178])
    aux, raw, root = (np.empty(2 * S * n) for _ in range(3))
    R.ref_process_raw(*dem_args(x, y, z), p(tx), p(rx), n, p(P), 1.0, *sim, S, p(aux), p(raw), p(root))
    res.update(sim_args=sim, sim_S=np.array([S]), sim_aux=aux.reshape(S, n, 2), sim_raw=raw.reshape(S, n, 2),
               sim_root=root.reshape(S, n, 2))
    # reflection points (calculateReflectionPoint, and as calculaSlantRangeFermat
    # stores them), from a separate stream so the vectors above stay as they were
    rr = np.random.default_rng(575)
    for name, (x, y, z) in dems.items():
        x, y, z = map(np.ascontiguousarray, (x, y, z))
        m = 300
        pr = np.stack([rr.uniform(x.min() - 60, x.max() + 60, m), rr.uniform(y.min() - 60, y.max() + 60, m),
                       rr.uniform(5, 20, m), rr.uniform(x.min() - 60, x.max() + 60, m),
                       rr.uniform(y.min() - 60, y.max() + 60, m), rr.uniform(5, 20, m)], 1)
        pr[:40, 3:5] = pr[:40, 0:2] * -1.0 + 2.0 * np.stack([x[rr.integers(0, len(x), 40)],
                                                            y[rr.integers(0, len(y), 40)]], 1)   # midpoint on a node
        pr = np.ascontiguousarray(q6(pr))
        ro = np.empty((m, 3))
        R.ref_reflection(*dem_args(x, y, z), p(pr), m, p(ro))
        res[f"refl_{name}_pairs"], res[f"refl_{name}_out"] = pr, ro
    e = np.zeros(0)
    pr = np.ascontiguousarray(np.array([[3.0, -2.0, 12.0, 20.0, 0.5, 11.0]]))
    ro = np.empty((1, 3))
    R.ref_reflection(p(e), 0, p(e), 0, p(np.zeros(1)), 0, 0, p(pr), 1, p(ro))
    res["refl_empty_pairs"], res["refl_empty_out"] = pr, ro
    x, y, z = dems["bumps"]
    rs = np.empty((n, 3))
    R.ref_slant_reflection(*dem_args(x, y, z), p(tx), p(rx), n, p(P), 1.0, 2.449489742783178, p(rs))
    res["sr_refl"] = rs
    np.savez_compressed(OUT / "ref_kernels.npz", **res)
    print("wrote ref_kernels.npz")


def write_bin2d(path, arr_rows_cols):
    a = np.ascontiguousarray(arr_rows_cols, dtype=np.complex128)
    rows, cols = a.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<III", 2, 1, 6) + b"double".ljust(16, b"\0") + struct.pack("<QQ", rows, cols))
        f.write(a.tobytes())


def write_csv_vec(path, v):
    with open(path, "w") as f:
        f.write(",".join("%.6f" % x for x in v) + "\n")


def write_csv_mat(path, m):
    with open(path, "w") as f:
        for row in m:
            f.write(",".join("%.6f" % x for x in row) + "\n")


def bp_case(rng, tag, n, S, grid, data_kind):
    """Runs proc_test_script on one generated input set.  Its constants are
    compile-time: fc = 10 GHz, c = 3e8, fs = 60 MHz x 4 (radar_init.cpp:9,17,44,74)."""
    dems = make_dems(rng)
    dx, dy, dz = dems["bumps"]
    t = np.arange(n) / (n - 1)
    tx = q6(np.stack([(5 + 10 * t) * np.cos(4 * np.pi * t), (5 + 10 * t) * np.sin(4 * np.pi * t), np.full(n, 12.0)], 1))
    rx = q6(np.stack([(5 + 10 * t) * np.cos(4 * np.pi * t + np.pi / 2), (5 + 10 * t) * np.sin(4 * np.pi * t + np.pi / 2),
                      np.full(n, 12.0)], 1))
    gx = q6(19.9 + 0.05 * np.arange(grid[0]))
    gy = q6(-0.1 + 0.05 * np.arange(grid[1]))
    gz = q6(-5.1 + 0.1 * np.arange(grid[2]))
    n2 = 2.449489742783178
    if data_kind == "random":
        root = rng.normal(size=(n, S)) + 1j * rng.normal(size=(n, S))     # [pulse][range]
    else:
        R = O.ref_lib()
        P = np.array([gx[grid[0] // 2], gy[grid[1] // 2], gz[grid[2] // 2]])
        sim = [10e9, 3e8, 60e6, 60e6, -25e6, 25e6, 1.3e-6, 1.0, n2]
        aux, raw, rt = (np.empty(2 * S * n) for _ in range(3))
        R.ref_process_raw(*dem_args(dx, dy, dz), p(tx), p(rx), n, p(P), 1.0, *sim, S, p(aux), p(raw), p(rt))
        rt = rt.reshape(S, n, 2)
        root = (rt[..., 0] + 1j * rt[..., 1]).T                          # [pulse][range]
    work = Path(tempfile.mkdtemp(prefix="bsar_golden_"))
    try:
        proc = work / "build" / "output" / "proc"
        proc.mkdir(parents=True)
        # file dims [S, N]; payload flat index r + p*S (range fastest)
        with open(proc / "rootData_complex.bin", "wb") as f:
            f.write(struct.pack("<III", 2, 1, 6) + b"double".ljust(16, b"\0") + struct.pack("<QQ", S, n))
            f.write(np.ascontiguousarray(root).tobytes())
        write_csv_vec(proc / "grid_xAxis.csv", gx)
        write_csv_vec(proc / "grid_yAxis.csv", gy)
        write_csv_vec(proc / "grid_zAxis.csv", gz)
        write_csv_vec(proc / "DEM_X_vec.csv", dx)
        write_csv_vec(proc / "DEM_Y_vec.csv", dy)
        write_csv_mat(proc / "DEM_Z_matrix.csv", dz)
        write_csv_mat(proc / "Tx_positions.csv", tx)
        write_csv_mat(proc / "Rx_positions.csv", rx)
        with open(proc / "environment.txt", "w") as f:
            f.write("n1=%r\nn2=%r\n" % (1.0, n2))
        env = dict(os.environ, OMP_NUM_THREADS="4")
        r = subprocess.run([str(O.REF_PROC)], cwd=work, capture_output=True, text=True, env=env, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        blob = (proc / "output_3D_matrix.bin").read_bytes()
        gcsv = [(proc / f"output_3D_matrix_grid_{a}.csv").read_text() for a in "xyz"]
    finally:
        shutil.rmtree(work)
    np.savez_compressed(OUT / f"ref_bp_{tag}.npz", root=root, tx=tx, rx=rx, dem_x=dx, dem_y=dy, dem_z=dz,
                        gx=gx, gy=gy, gz=gz, n1=1.0, n2=n2, fc=10e9, c=3e8, fs=60e6, upsample=4,
                        out_bin=np.frombuffer(blob, np.uint8), grid_csv=np.array(gcsv))
    print(f"wrote ref_bp_{tag}.npz ({len(blob)} bytes of image)")


SCENARIO_PULSES = [0, 1, 2, 3, 1000, 2000, 3000, 4000, 5000, 6000, 6685, 6686, 6687, 6688]
SCENARIO_TEXT = ["DEM_X_vec.csv", "DEM_Y_vec.csv", "DEM_Z_matrix.csv", "grid_xAxis.csv", "grid_yAxis.csv",
                 "grid_zAxis.csv", "target_position.csv", "environment.txt"]


def scenario_case():
    """gs_test_script's default run (it writes build/output/ and plot_data/
    under its working directory)."""
    import hashlib
    work = Path(tempfile.mkdtemp(prefix="bsar_golden_gs_"))
    try:
        r = subprocess.run([str(O.REF_GS)], cwd=work, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        proc = work / "build" / "output" / "proc"
        res = {f"text_{n}": np.frombuffer((proc / n).read_bytes(), np.uint8) for n in SCENARIO_TEXT}
        for n in ("Tx_positions.csv", "Rx_positions.csv"):
            b = (proc / n).read_bytes()
            res[f"sha256_{n}"] = np.frombuffer(hashlib.sha256(b).digest(), np.uint8)
            res[f"lines_{n}"] = np.array(b.decode().count("\n"))
        for n in ("rootData", "rawData", "auxData"):
            blob = (proc / f"{n}_complex.bin").read_bytes()
            hdr, ndim = blob[:28], struct.unpack("<I", blob[:4])[0]
            dims = struct.unpack("<%dQ" % ndim, blob[28:28 + 8 * ndim])
            data = np.frombuffer(blob[28 + 8 * ndim:], np.complex128).reshape(dims)   # [S][N], r*N + p
            res[f"{n}_header"] = np.frombuffer(blob[:28 + 8 * ndim], np.uint8)
            res[f"{n}_pulses"] = np.ascontiguousarray(data[:, SCENARIO_PULSES].T)       # [P][S]
            if n == "auxData":
                nz = np.abs(data) > 0
                res["aux_bins"] = np.where(nz.any(0), nz.argmax(0), -1).astype(np.int16)  # per pulse
        res["pulses"] = np.array(SCENARIO_PULSES)
        res["n_pulses"] = np.array(dims[1])
        cxx = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
        res["manifest"] = np.array(f"oracle/_ref/gs_test_script built by oracle/Makefile ({cxx}, -O3 -fopenmp), "
                                   f"run with no arguments in an empty directory")
    finally:
        shutil.rmtree(work)
    np.savez_compressed(OUT / "ref_scenario.npz", **res)
    print("wrote ref_scenario.npz")


def main():
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["scenario"], help="write only this fixture (the others share one seeded "
                    "stream and are written together)")
    a = ap.parse_args()
    O.build(ref=True)
    if O.ref_lib() is None:
        sys.exit("compiled reference missing: oracle/_ref was not built")
    if a.only is None:
        rng = np.random.default_rng(20261203)
        kernels(rng)
        bp_case(rng, "random", n=16, S=128, grid=(5, 4, 3), data_kind="random")
        bp_case(rng, "target", n=24, S=128, grid=(6, 5, 4), data_kind="target")
    scenario_case()


if __name__ == "__main__":
    main()
