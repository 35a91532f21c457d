"""The CPU oracle against golden vectors produced by the compiled reference
(tests/golden/make_golden.py).  Everything here is bit-exact: the oracle
restates the reference's operation order and is built without FMA
contraction, like the reference's x86-64 build."""
import struct

import numpy as np
import pytest

from conftest import GOLDEN


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


DEMS = ["bumps", "flat", "ragged", "rough", "unsorted", "one_row", "dup", "small_z"]


@pytest.mark.parametrize("name", DEMS)
def test_dem_interp(oracle, golden, name):
    g = golden
    dem = (g[f"dem_{name}_x"], g[f"dem_{name}_y"], g[f"dem_{name}_z"])
    out = oracle.dem_interp(dem, g[f"dem_{name}_xy"])
    assert bits_equal(out, g[f"dem_{name}_out"])


@pytest.mark.parametrize("name", DEMS)
@pytest.mark.parametrize("tag", ["a", "b", "eq"])
def test_refraction_point(oracle, golden, name, tag):
    g = golden
    dem = (g[f"dem_{name}_x"], g[f"dem_{name}_y"], g[f"dem_{name}_z"])
    n1, n2 = g[f"refr_{name}_{tag}_n"]
    out = oracle.refraction_points(dem, g[f"refr_{name}_{tag}_legs"], n1, n2)
    assert bits_equal(out, g[f"refr_{name}_{tag}_out"])


def test_refraction_empty_dem(oracle, golden):
    e = np.zeros(0)
    out = oracle.refraction_points((e, e, np.zeros((0, 0))), golden["refr_empty_legs"], 1.0, 2.0)
    assert bits_equal(out, golden["refr_empty_out"])


def test_slant_range(oracle, golden):
    g = golden
    dem = (g["dem_bumps_x"], g["dem_bumps_y"], g["dem_bumps_z"])
    ida, volta, r4 = oracle.slant_range_fermat(dem, g["sr_tx"], g["sr_rx"], g["sr_P"], 1.0, 2.449489742783178)
    assert bits_equal(ida, g["sr_ida"])
    assert bits_equal(volta, g["sr_volta"])
    assert bits_equal(r4, g["sr_r4"])


@pytest.mark.parametrize("nf", [(2, 4), (3, 4), (5, 3), (64, 4), (257, 4), (16, 1), (9, 7)])
def test_spline(oracle, golden, nf):
    n, f = nf
    d = golden[f"spline_{n}_{f}_in"].view(np.complex128)
    out = oracle.spline(d, f)
    assert bits_equal(out.view(np.float64), golden[f"spline_{n}_{f}_out"])


@pytest.mark.parametrize("n", [1, 2, 8, 64, 1024, 7, 12])
def test_fft_family(oracle, golden, n):
    a = golden[f"fft_{n}_in"].view(np.complex128)
    b = golden[f"corr_{n}_ref"].view(np.complex128)
    assert bits_equal(oracle.fft(a).view(np.float64), golden[f"fft_{n}_out"])
    assert bits_equal(oracle.ifft(a).view(np.float64), golden[f"ifft_{n}_out"])
    assert bits_equal(oracle.fft_correlation(a, b).view(np.float64), golden[f"corr_{n}_out"])
    assert bits_equal(oracle.fft_convolution(a, b).view(np.float64), golden[f"conv_{n}_out"])


@pytest.mark.parametrize("tag", ["refc", "cfg"])
def test_chirp(oracle, golden, tag):
    f0, f1, T, fs = golden[f"chirp_{tag}_args"]
    assert bits_equal(oracle.chirp(f0, f1, T, fs).view(np.float64), golden[f"chirp_{tag}_out"])


def test_process_raw(oracle, golden):
    g = golden
    fc, c, fs_rx, fs_tx, f0, f1, T, n1, n2 = g["sim_args"]
    S = int(g["sim_S"][0])
    sim = dict(fc=fc, c=c, fs=fs_rx, f0=f0, f1=f1, T=T, fs_chirp=fs_tx, n1=n1, n2=n2, S=S)
    dem = (g["dem_bumps_x"], g["dem_bumps_y"], g["dem_bumps_z"])
    aux, raw, root = oracle.simulate(sim, dem, g["sr_tx"], g["sr_rx"], g["sr_P"][None, :], np.array([1.0 + 0j]))
    for mine, ref in ((aux, g["sim_aux"]), (raw, g["sim_raw"]), (root, g["sim_root"])):
        r = ref[..., 0] + 1j * ref[..., 1]
        assert np.array_equal(mine, r)


def parse_image(blob):
    nd, cplx, tl = struct.unpack_from("<III", blob, 0)
    assert (nd, cplx, tl) == (3, 1, 6)
    assert blob[12:28] == b"double" + b"\0" * 10
    nx, ny, nz = struct.unpack_from("<QQQ", blob, 28)
    data = np.frombuffer(blob[52:], np.complex128)
    assert len(data) == nx * ny * nz
    return data.reshape(nz, ny, nx)


@pytest.mark.parametrize("tag", ["random", "target"])
def test_backprojection_vs_reference_binary(oracle, tag):
    """proc_test_script's own output_3D_matrix.bin vs the oracle on the
    same inputs (fc = 10 GHz, fs = 60 MHz x 4 as compiled into it)."""
    g = np.load(GOLDEN / f"ref_bp_{tag}.npz")
    ref = parse_image(g["out_bin"].tobytes())
    up = oracle.upsample(g["root"], int(g["upsample"]))
    bp = dict(c=float(g["c"]), fs4=float(g["fs"]) * int(g["upsample"]), lam=float(g["c"]) / float(g["fc"]),
              n1=float(g["n1"]), n2=float(g["n2"]))
    img = oracle.backproject(bp, (g["dem_x"], g["dem_y"], g["dem_z"]), g["tx"], g["rx"], up, g["gx"], g["gy"],
                             g["gz"], threads=2)
    assert np.abs(ref).max() > 0
    err = np.abs(img - ref).max() / np.abs(ref).max()
    assert err <= 1e-14, err


REF_ENV = GOLDEN.parents[1] / "oracle" / "_ref" / "proc_test_script_env"


def _run_ref_env(tmp_path, sc, root, env_extra):
    import os
    import subprocess
    sc_mod = __import__("conftest").load_scene()
    sc_mod.write_reference_inputs(tmp_path / "build" / "output" / "proc", sc, root)
    env = dict(os.environ, OMP_NUM_THREADS="2", **env_extra)
    for k in ("BSAR_REF_FC", "BSAR_REF_FS", "BSAR_REF_C"):
        if k not in env_extra:
            env.pop(k, None)
    r = subprocess.run([str(REF_ENV)], cwd=tmp_path, capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    return parse_image((tmp_path / "build" / "output" / "proc" / "output_3D_matrix.bin").read_bytes())


@pytest.mark.skipif(not REF_ENV.exists(), reason="the compiled reference (oracle/_ref) is not built")
def test_reference_driver_at_workload_constants(tmp_path, oracle, scene_mod):
    """The reference's own proc_test_script, linked with its radar constants
    read from the environment (oracle/ref_radar_env.cpp), images a workload
    file set at the workload's carrier and sampling rate (fc = 1 GHz,
    fs = 800 MHz x 4): its output_3D_matrix.bin matches the oracle's
    back-projection with those constants; with the variables unset it gives
    the compiled-in constants' image (10 GHz, 60 MHz)."""
    sc = scene_mod.make_scene("small", n_pulses=24, grid=(5, 4, 3), samples=1024)
    _, _, root = oracle.simulate(sc.sim_params(), sc.dem, sc.tx, sc.rx, sc.scat, sc.rcs)
    root = np.ascontiguousarray(root.T)
    ref = _run_ref_env(tmp_path / "a", sc, root, {"BSAR_REF_FC": "1e9", "BSAR_REF_FS": "800e6", "BSAR_REF_C": "3e8"})
    up = oracle.upsample(root, 4)
    want = oracle.backproject(sc.bp_params(), sc.dem, sc.tx, sc.rx, up, sc.gx, sc.gy, sc.gz, threads=2)
    assert np.abs(want).max() > 0
    assert np.abs(ref - want).max() <= 1e-14 * np.abs(want).max()
    dflt = _run_ref_env(tmp_path / "b", sc, root, {})
    bp = dict(sc.bp_params(), fs4=60e6 * 4, lam=3e8 / 10e9)
    want = oracle.backproject(bp, sc.dem, sc.tx, sc.rx, up, sc.gx, sc.gy, sc.gz, threads=2)
    assert np.abs(dflt - want).max() <= 1e-14 * np.abs(want).max()
