"""The drop-in layer: bsar_proc (replaces proc_test_script on the same file
set) and the C++ facade with the reference's function names."""
import subprocess

import numpy as np
import pytest

from _files import write_proc_dir
from conftest import GOLDEN, PKG_DIR, ROOT
from test_oracle_golden import parse_image

TOOL = PKG_DIR / "bsar_proc"


def run_tool(*args, cwd=None):
    return subprocess.run([str(TOOL), *map(str, args)], capture_output=True, text=True, timeout=300, cwd=cwd)


def test_tool_parses_reference_inputs(tmp_path):
    g = np.load(GOLDEN / "ref_bp_random.npz")
    d = tmp_path / "build" / "output" / "proc"
    write_proc_dir(d, g)
    r = run_tool("--check", cwd=tmp_path)          # default directory, as the reference
    assert r.returncode == 0, r.stderr
    N, S = g["root"].shape
    assert f"rootData {S} x {N}" in r.stdout
    assert f"grid {len(g['gx'])} x {len(g['gy'])} x {len(g['gz'])}" in r.stdout
    assert f"{len(g['tx'])} positions" in r.stdout


@pytest.mark.parametrize("missing", ["rootData_complex.bin", "grid_yAxis.csv", "Tx_positions.csv"])
def test_tool_missing_input_exits_1(tmp_path, missing):
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    (tmp_path / missing).unlink()
    r = run_tool("--check", "--dir", tmp_path)
    assert r.returncode == 1 and "[ERROR]" in r.stderr


def test_tool_rejects_non_2d_rootdata(tmp_path):
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    import struct
    (tmp_path / "rootData_complex.bin").write_bytes(
        struct.pack("<III", 1, 1, 6) + b"double".ljust(16, b"\0") + struct.pack("<Q", 4) + np.zeros(8).tobytes())
    r = run_tool("--check", "--dir", tmp_path)
    assert r.returncode == 1


def demo_exe(tmp_path):
    exe = tmp_path / "facade_demo"
    r = subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "facade_demo.cpp"),
                        "-o", str(exe), f"-L{PKG_DIR}", "-lbsar", f"-Wl,-rpath,{PKG_DIR}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_facade_compiles_against_reference_style_types(tmp_path):
    assert demo_exe(tmp_path).exists()


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["random", "target"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tool_output_matches_reference_binary(tmp_path, tag, precision):
    g = np.load(GOLDEN / f"ref_bp_{tag}.npz")
    write_proc_dir(tmp_path, g)
    r = run_tool("--dir", tmp_path, "--precision", precision)
    assert r.returncode == 0, r.stdout + r.stderr
    blob = (tmp_path / "output_3D_matrix.bin").read_bytes()
    ref_blob = g["out_bin"].tobytes()
    assert blob[:52] == ref_blob[:52]                 # identical header
    img, ref = parse_image(blob), parse_image(ref_blob)
    err = np.abs(img - ref).max() / np.abs(ref).max()
    assert err <= (1e-9 if precision == "fp64" else 1e-4), err
    for a, text in zip("xyz", g["grid_csv"]):
        assert (tmp_path / f"output_3D_matrix_grid_{a}.csv").read_text() == str(text)


@pytest.mark.gpu
def test_facade_slant_range_bit_exact(tmp_path, golden):
    g = golden
    exe = demo_exe(tmp_path)
    x, y, z = g["dem_bumps_x"], g["dem_bumps_y"], g["dem_bumps_z"]
    f = lambda a: " ".join(float(v).hex() for v in np.ravel(a))
    lines = [f"{len(x)} {len(y)}", f(x), f(y)] + [f(row) for row in z]
    lines += [str(len(g["sr_tx"]))] + [f(r) for r in g["sr_tx"]] + [f(r) for r in g["sr_rx"]]
    lines += [f(g["sr_P"]), f([1.0, 2.449489742783178])]
    r = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [list(map(float.fromhex, ln.split())) for ln in r.stdout.strip().splitlines()]
    vals = np.array(rows[:-1])
    assert np.array_equal(vals[:, 0:3], g["sr_ida"]) and np.array_equal(vals[:, 3:6], g["sr_volta"])
    assert np.array_equal(vals[:, 6:10], g["sr_r4"].T)
    assert np.array_equal(vals[:, 10:13], g["sr_refl"])         # ReflectionPoints, raw_data_processing.cpp:188-193


def test_facade_demo_rejects_malformed_input(tmp_path):
    exe = demo_exe(tmp_path)
    for bad in ("np.float64(1.0) 2\n", "3 3\n1 2\n", "2 2\n0 1\n0 1\n0 0\n0 0\n999999999\n"):
        r = subprocess.run([str(exe)], input=bad, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2


def test_scene_file_set_parses(tmp_path, scene_mod):
    """scene.write_reference_inputs (behind bsar_sim) writes a file set that
    bsar_proc reads back with the stated sizes."""
    sc = scene_mod.make_scene("small", n_pulses=12, grid=(5, 4, 3), samples=256)
    root = np.random.default_rng(2).normal(size=(12, 256)) + 1j * np.random.default_rng(3).normal(size=(12, 256))
    scene_mod.write_reference_inputs(tmp_path / "build" / "output" / "proc", sc, root)
    r = run_tool("--check", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "rootData 256 x 12" in r.stdout and "grid 5 x 4 x 3" in r.stdout and "12 positions" in r.stdout


@pytest.mark.gpu
def test_sim_then_proc_round_trip(tmp_path, bsar, oracle):
    """bsar_sim writes a workload's file set (GPU forward model); bsar_proc
    images it; the result equals the oracle run on the same files' contents."""
    r = subprocess.run(["python", str(PKG_DIR / "sim.py"), "--workload", "small", "--out", str(tmp_path),
                        "--pulses", "16", "--grid", "6", "5", "4", "--samples", "1024"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = run_tool("--fc", "1e9", "--fs", "800e6", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    proc = tmp_path / "build" / "output" / "proc"
    img = parse_image((proc / "output_3D_matrix.bin").read_bytes())
    rd = lambda n: np.array([[float(x) for x in ln.split(",")] for ln in (proc / n).read_text().split()])
    root, dims = bsar.read_matrix(proc / "rootData_complex.bin")
    S, N = dims
    gx, gy, gz = rd("grid_xAxis.csv")[0], rd("grid_yAxis.csv")[0], rd("grid_zAxis.csv")[0]
    dem = (rd("DEM_X_vec.csv")[0], rd("DEM_Y_vec.csv")[0], rd("DEM_Z_matrix.csv"))
    tx, rx = rd("Tx_positions.csv"), rd("Rx_positions.csv")
    env = dict(ln.split("=") for ln in (proc / "environment.txt").read_text().split())
    bp = dict(c=3e8, fs4=800e6 * 4, lam=3e8 / 1e9, n1=float(env["n1"]), n2=float(env["n2"]))
    want = oracle.backproject(bp, dem, tx, rx, oracle.upsample(root.reshape(N, S)), gx, gy, gz)
    assert np.abs(img - want).max() <= 1e-9 * np.abs(want).max()
    assert np.abs(want).max() > 0


SERVE = PKG_DIR / "bsar_serve"
SERVE_JOBS = ("# tag ix0 ix1 iy0 iy1 iz0 iz1 pulses precision\n"
              "full 0 5 0 4 0 3 0 fp64\n"
              "sub_mixed 1 4 0 3 1 3 10 mixed\n"
              "\n"
              "sub_fp64 0 2 1 4 0 2 12 fp64\n")


def run_serve(*args, stdin=None, cwd=None):
    return subprocess.run([str(SERVE), *map(str, args)], input=stdin, capture_output=True, text=True, timeout=300,
                          cwd=cwd)


def test_serve_parses_inputs_and_jobs(tmp_path):
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    (tmp_path / "jobs.txt").write_text(SERVE_JOBS)
    r = run_serve("--check", "--dir", tmp_path, "--jobs", tmp_path / "jobs.txt")
    assert r.returncode == 0, r.stderr
    assert "jobs: 3" in r.stdout
    assert "job full: 5 x 4 x 3 voxels x 16 pulses, fp64" in r.stdout
    assert "job sub_mixed: 3 x 3 x 2 voxels x 10 pulses, mixed" in r.stdout
    r = run_serve("--check", "--dir", tmp_path, stdin=SERVE_JOBS)      # job list on standard input
    assert r.returncode == 0 and "jobs: 3" in r.stdout


@pytest.mark.parametrize("line", ["a 0 6 0 4 0 3 0 fp64", "a 2 2 0 4 0 3 0 fp64", "a 0 5 0 4 0 3 17 fp64",
                                  "a 0 5 0 4 0 3 0 fp32", "a 0 5 0 4 0 3 fp64", "a/b 0 5 0 4 0 3 0 fp64",
                                  "a 0 5 0 4 0 3 0 fp64\na 0 1 0 1 0 1 0 mixed", "# nothing", "a 0 5 0 4 0 3 0 fp64 x"])
def test_serve_rejects_malformed_jobs(tmp_path, line):
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    r = run_serve("--check", "--dir", tmp_path, stdin=line + "\n")
    assert r.returncode == 2 and "[ERROR]" in r.stderr


@pytest.mark.parametrize("args", [("--fc", "abc"), ("--fs", "1e9x"), ("--factor", "2.5"), ("--factor", "9e99"),
                                  ("--c", ""), ("--precision", "fp32")])
def test_tools_reject_malformed_options(tmp_path, args):
    """A bad option value exits with 2 and names the option, before any file is read."""
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    tools = [run_tool] if args[0] == "--precision" else [run_tool, run_serve]
    for tool in tools:
        r = tool("--check", "--dir", tmp_path, *args) if tool is run_tool else \
            tool("--check", "--dir", tmp_path, *args, stdin=SERVE_JOBS)
        assert r.returncode == 2, (tool.__name__, r.returncode, r.stderr)
        assert args[0] in r.stderr


def test_serve_rejects_malformed_numeric_options(tmp_path):
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    for args in (("--round-ms", "fast"), ("--max-jobs", "1.5"), ("--hbm-cap-gb", "lots")):
        r = run_serve("--check", "--dir", tmp_path, *args, stdin=SERVE_JOBS)
        assert r.returncode == 2 and args[0] in r.stderr


def test_serve_missing_input_exits_1(tmp_path):
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    (tmp_path / "DEM_Z_matrix.csv").unlink()
    r = run_serve("--check", "--dir", tmp_path, stdin=SERVE_JOBS)
    assert r.returncode == 1 and "[ERROR]" in r.stderr


@pytest.mark.gpu
def test_serve_images_match_backproject(tmp_path, bsar, ctx):
    """bsar_serve images each job of the list through one server; every image
    equals bsar_backproject over the same sub-grid with a dataset of the job's
    first K positions, bit for bit, and the full-grid job matches the
    reference's own output_3D_matrix.bin."""
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    r = run_serve("--dir", tmp_path, "--round-ms", "5", stdin=SERVE_JOBS)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "[PERFORMANCE] serving: 3 jobs" in r.stdout
    rd = lambda n: np.array([[float(x) for x in ln.split(",")] for ln in (tmp_path / n).read_text().split()])
    gx, gy, gz = rd("grid_xAxis.csv")[0], rd("grid_yAxis.csv")[0], rd("grid_zAxis.csv")[0]
    dem = bsar.Dem(rd("DEM_X_vec.csv")[0], rd("DEM_Y_vec.csv")[0], rd("DEM_Z_matrix.csv"))
    tx, rx = rd("Tx_positions.csv"), rd("Rx_positions.csv")
    p = bsar.Params(float(g["fc"]), float(g["c"]), float(g["fs"]), float(g["n1"]), float(g["n2"]), int(g["upsample"]))
    for line in SERVE_JOBS.splitlines():
        if not line or line.startswith("#"):
            continue
        tag, ix0, ix1, iy0, iy1, iz0, iz1, K, prec = line.split()
        ix0, ix1, iy0, iy1, iz0, iz1, K = map(int, (ix0, ix1, iy0, iy1, iz0, iz1, K))
        K = K or len(tx)
        ds = ctx.dataset(p, dem, tx[:K], rx[:K])
        ds.load_rootdata(g["root"][:K])
        want, _ = ds.backproject(gx[ix0:ix1], gy[iy0:iy1], gz[iz0:iz1], bsar.MIXED if prec == "mixed" else bsar.FP64)
        ds.close()
        got = parse_image((tmp_path / f"{tag}_3D_matrix.bin").read_bytes())
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), tag
        if tag == "full":
            ref = parse_image(g["out_bin"].tobytes())
            assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.gpu
def test_serve_job_over_capacity_exits_1(tmp_path):
    """A job whose working memory alone exceeds the server's HBM cap is
    rejected at submission (BSAR_ERR_CAPACITY); the tool reports it and
    exits 1 after the other jobs."""
    g = np.load(GOLDEN / "ref_bp_random.npz")
    write_proc_dir(tmp_path, g)
    r = run_serve("--dir", tmp_path, "--hbm-cap-gb", "1e-9", stdin=SERVE_JOBS)
    assert r.returncode == 1
    assert "[ERROR] job full: submit status" in r.stderr
