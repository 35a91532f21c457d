"""The reference's default scenario end to end on the GPU: bsar_sim
--workload reference (spirals, crearGrid, loadDEM, processRawData) against
the file set the reference's own gs_test_script writes
(tests/golden/ref_scenario.npz), then bsar_proc on that file set against
the oracle."""
import hashlib
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR
from test_oracle_golden import parse_image

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenario():
    return np.load(GOLDEN / "ref_scenario.npz")


def _bins(aux):
    nz = np.abs(aux) > 0
    return np.where(nz.any(1), nz.argmax(1), -1)


def test_reference_scenario_forward_model(ctx, scene_mod, scenario):
    """processRawData for gs_test_script's target and trajectories: response
    bins identical for all 6689 pulses; aux, raw and root within 1e-9 of the
    reference's peak on the stored pulses (the reference writes [S][N], r*N + p;
    the fixture holds its columns)."""
    sc = scene_mod.reference_scene()
    out = scene_mod.forward_model(ctx, sc)
    assert out["root"].shape == (int(scenario["n_pulses"]), 256)
    assert np.array_equal(_bins(out["aux"]), scenario["aux_bins"])
    P = scenario["pulses"]
    for name, key in (("aux", "auxData"), ("raw", "rawData"), ("root", "rootData")):
        want = scenario[f"{key}_pulses"]
        assert np.abs(out[name][P] - want).max() <= 1e-9 * np.abs(want).max(), name


def test_sim_reference_workload_then_proc(tmp_path, bsar, oracle, scenario):
    """bsar_sim writes gs_test_script's file set (CSV text byte-identical,
    matrices range-fastest); bsar_proc with the reference's compiled-in
    constants (10 GHz, 60 MHz x 4) images it; a voxel subset equals the
    oracle on the same files' contents."""
    r = subprocess.run(["python", str(PKG_DIR / "sim.py"), "--workload", "reference", "--out", str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    proc = tmp_path / "build" / "output" / "proc"
    for key in scenario.files:
        if key.startswith("text_"):
            assert (proc / key[5:]).read_bytes() == scenario[key].tobytes(), key
        elif key.startswith("sha256_"):
            assert hashlib.sha256((proc / key[7:]).read_bytes()).digest() == scenario[key].tobytes(), key
    P = scenario["pulses"]
    for key in ("rootData", "rawData", "auxData"):
        blob = (proc / f"{key}_complex.bin").read_bytes()
        assert blob[:44] == scenario[f"{key}_header"].tobytes()
        data = np.frombuffer(blob[44:], np.complex128).reshape(int(scenario["n_pulses"]), 256)   # r + p*S
        want = scenario[f"{key}_pulses"]
        assert np.abs(data[P] - want).max() <= 1e-9 * np.abs(want).max(), key
    r = subprocess.run([str(PKG_DIR / "bsar_proc")], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    img = parse_image((proc / "output_3D_matrix.bin").read_bytes())
    assert img.shape == (21, 41, 41)
    rd = lambda n: np.array([[float(x) for x in ln.split(",")] for ln in (proc / n).read_text().split()])
    gx, gy, gz = rd("grid_xAxis.csv")[0], rd("grid_yAxis.csv")[0], rd("grid_zAxis.csv")[0]
    dem = (rd("DEM_X_vec.csv")[0], rd("DEM_Y_vec.csv")[0], rd("DEM_Z_matrix.csv"))
    tx, rx = rd("Tx_positions.csv"), rd("Rx_positions.csv")
    root, dims = bsar.read_matrix(proc / "rootData_complex.bin")
    S, N = dims
    bp = dict(c=3e8, fs4=60e6 * 4, lam=3e8 / 10e9, n1=1.0, n2=4.0)
    ix, iy, iz = [0, 20, 40], [20], [10]
    want = oracle.backproject(bp, dem, tx, rx, oracle.upsample(root.reshape(N, S)), gx[ix], gy[iy], gz[iz],
                              threads=8)
    got = img[np.ix_(iz, iy, ix)]
    assert np.abs(want).max() > 0
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
