"""bsar_sim: writes a workload's input file set in the reference's formats.

    python c-_bistatic_sar_amd/sim.py --workload small --out DIR [--pulses N] [--grid NX NY NZ]
                                      [--samples S] [--seed K]

DIR/build/output/proc/ then holds what proc_test_script (or this repository's
bsar_proc) reads: rootData_complex.bin (range-compressed data from the GPU
forward model: refraction points, response placement, chirp convolution,
pulse compression), the grid, DEM and antenna CSVs and environment.txt.  The
radar constants are the workload's (fc 1 GHz, fs 800 MHz, c 3e8, x4); pass
them to bsar_proc with --fc 1e9 --fs 800e6 (the reference's executable has
10 GHz / 60 MHz compiled in).
"""
from __future__ import annotations

import argparse
import importlib.util
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent


def _load(name: str, path: Path, pkg: bool = False):
    if name in sys.modules:
        return sys.modules[name]
    kw = dict(submodule_search_locations=[str(path.parent)]) if pkg else {}
    spec = importlib.util.spec_from_file_location(name, path, **kw)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", default="small", choices=["small", "large", "rough", "wide", "serve"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--pulses", type=int)
    ap.add_argument("--samples", type=int)
    ap.add_argument("--grid", type=int, nargs=3)
    ap.add_argument("--seed", type=int)
    a = ap.parse_args(argv)
    pkg = _load("bsar_amd", PKG / "__init__.py", pkg=True)
    scene = _load("bsar_amd.scene", PKG / "scene.py")
    sc = scene.make_scene(a.workload, seed=a.seed, n_pulses=a.pulses, grid=a.grid, samples=a.samples)
    with pkg.Context(0) as ctx:
        root = scene.simulate(ctx, sc)
    proc = Path(a.out) / "build" / "output" / "proc"
    scene.write_reference_inputs(proc, sc, root)
    print(f"wrote {proc}: {len(sc.gx)}x{len(sc.gy)}x{len(sc.gz)} grid, {sc.n} pulses x {sc.S} samples "
          f"(fc {sc.fc:g} Hz, fs {sc.fs:g} Hz, n2 {sc.n2:.6f})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
