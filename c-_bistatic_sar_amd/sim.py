"""bsar_sim: writes a workload's input file set in the reference's formats.

    python c-_bistatic_sar_amd/sim.py --workload small --out DIR [--pulses N] [--grid NX NY NZ]
                                      [--samples S] [--seed K]

DIR/build/output/proc/ then holds the file set gs_test_script writes and
proc_test_script (or this repository's bsar_proc) reads: rootData_complex.bin
(range-compressed data from the GPU forward model: refraction points,
response placement, chirp convolution, pulse compression), rawData and
auxData, the grid, DEM, antenna and target CSVs and environment.txt.  The
radar constants of the synthetic workloads are fc 1 GHz, fs 800 MHz, c 3e8,
x4: pass them to bsar_proc with --fc 1e9 --fs 800e6 (the reference's
executable has 10 GHz / 60 MHz compiled in).  --workload reference replays
the reference's own default scenario (gs_test_script.cpp:15-137 with
radar_init.cpp:5-89), whose constants are the compiled-in ones; --dem FILE
[--geotiff] swaps its DEM for a file (a GeoTIFF decoded to local metres).
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
    ap.add_argument("--workload", default="small", choices=["small", "large", "rough", "wide", "serve", "reference"])
    ap.add_argument("--dem", help="reference workload: DEM file instead of the reference's .tif name")
    ap.add_argument("--geotiff", action="store_true", help="decode --dem as a GeoTIFF (else loadDEM's rules)")
    ap.add_argument("--out", required=True)
    ap.add_argument("--pulses", type=int)
    ap.add_argument("--samples", type=int)
    ap.add_argument("--grid", type=int, nargs=3)
    ap.add_argument("--seed", type=int)
    a = ap.parse_args(argv)
    pkg = _load("bsar_amd", PKG / "__init__.py", pkg=True)
    scene = _load("bsar_amd.scene", PKG / "scene.py")
    if a.workload == "reference":
        sc = scene.reference_scene(a.pulses, a.dem, pkg.DEM_FILE_GEOTIFF if a.geotiff else pkg.DEM_FILE_REFERENCE)
    else:
        sc = scene.make_scene(a.workload, seed=a.seed, n_pulses=a.pulses, grid=a.grid, samples=a.samples)
    with pkg.Context(0) as ctx:
        out = scene.forward_model(ctx, sc)
    proc = Path(a.out) / "build" / "output" / "proc"
    scene.write_reference_inputs(proc, sc, out["root"], raw=out["raw"], aux=out["aux"])
    print(f"wrote {proc}: {len(sc.gx)}x{len(sc.gy)}x{len(sc.gz)} grid, {sc.n} pulses x {sc.S} samples "
          f"(fc {sc.fc:g} Hz, fs {sc.fs:g} Hz, n2 {sc.n2:.6f})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
