"""Benchmark: voxel-pulse back-projections per second on one MI355X, mixed
precision vs plain FP64, on the 201x201x101-voxel x 2048-pulse image
(8192 fast-time samples, upsampled x4 to 32768), with the error of the mixed
image against the FP64 image in dB and the reference's OpenMP CPU path timed
on the same host.

    python bench.py [--steps K] [--warmup W] [--workload large|small|...]
    python bench.py --workload serve [--jobs 256] [--rate 40] [--no-fp64]
    python bench.py --workload rough [--chunk 65536] [--check-solves 100000]
    python bench.py --workload wide

One step = one full mixed-precision image of the workload with the
range-compressed data already resident in HBM.  After the timed steps the
default run adds: the plain FP64 image (and PSLR/ISLR of the mixed one
against it), parity of both images against the oracle on a strided voxel
subset, the reference's own OpenMP proc_test_script at the workload's radar
constants on a 20^3 strided subset (the CPU baseline, whose image is
compared with the GPU's), and the 256-job serving stream (mixed, plus FP64 on
its first 32 jobs) with p50/p99 job latency.  Prints one JSON line on
stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG_DIR = ROOT / "c-_bistatic_sar_amd"

P64 = 78.6e12     # MI355X FP64 vector peak, FLOP/s
P32 = 157.3e12    # MI355X FP32 vector peak, FLOP/s (reached without packed FP32:
                  # a wave64 FP32 FMA issues every 2 cycles, FP64 every 4;
                  # c-_bistatic_sar_amd/tools/valu_rates.hip)


def mix_peak(e64: float, e32: float) -> float:
    """Vector peak for a kernel executing e64 FP64 and e32 FP32 flop per unit:
    the rate at which both precisions' peaks would finish that mix,
    (e64 + e32) / (e64 / P64 + e32 / P32)."""
    return (e64 + e32) / (e64 / P64 + e32 / P32)

# Algorithmic flop per voxel-pulse (add/sub/mul/div/sqrt = 1, FMA = 2,
# sin = cos = 1; loop-invariant terms hoisted):
#   plain FP64 form: 65,406 + 20 N (two 709-candidate refraction solves at 46
#   flop each + 64 set-up, R1 means 2 (10 N + 1), 48 flop of delay/phase/lerp/
#   accumulate);
#   mixed form: the same two FP64 solves (65,356) + FP64 R2 (18) + two
#   tricubic table lookups (2 x 163: 3 x 10 weights, 64 x 2 + 5 x 2 + ...)
#   + FP64 delay/phase/reduction (14) in FP64; lerp, sin/cos, complex
#   product and two Kahan adds (27) in FP32.
def flops_fp64(n: int) -> tuple[float, float]:
    return 65406.0 + 20.0 * n, 0.0


def flops_mixed(n: int) -> tuple[float, float]:
    return 65356.0 + 18.0 + 2 * 163.0 + 14.0, 27.0


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def load(name: str, path: Path, pkg: bool = False):
    if name in sys.modules:
        return sys.modules[name]
    kw = dict(submodule_search_locations=[str(path.parent)]) if pkg else {}
    spec = importlib.util.spec_from_file_location(name, path, **kw)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


PROFILE_FLOPS = "valu_flops_bp_large.md"
PROFILE_TRAFFIC = "traffic_bp_large.md"
PROFILE_ISSUE = "valu_bp_large.md"


def profile_table(name: str) -> tuple[dict, str | None]:
    """Rows of the first markdown table in profiles/<name> keyed by the first
    cell (each row a dict by header), and the commit the profile states
    ("commit: <sha>" line); ({}, None) when the file is absent."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return {}, None
    rows, head, commit = {}, None, None
    for line in p.read_text().splitlines():
        m = re.match(r"\s*commit:\s*`?([0-9a-f]{7,40})`?", line)
        if m:
            commit = m.group(1)
        if not line.strip().startswith("|"):
            if head is not None and rows:
                break
            continue
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if head is None:
            head = cells
        elif set("".join(cells)) <= set("-: "):
            continue
        else:
            rows[cells[0]] = dict(zip(head, cells))
    return rows, commit


def _num(row: dict, key: str):
    try:
        return float(row[key].replace(",", ""))
    except (KeyError, ValueError, AttributeError):
        return None


def roofline(rep: dict, n: int, form: str, flops_prof: str = PROFILE_FLOPS, traffic_prof: str = PROFILE_TRAFFIC,
             issue_prof: str = PROFILE_ISSUE) -> dict:
    """Roofline of the dominant back-projection kernel of one image.

    achieved = executed FP64 + FP32 flop of one launch / the launch's average
    device time (HIP events on the context's stream, measured in this run).
    The executed flop per voxel-pulse comes from rocprofv3 counters
    (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_{F64,F32}: 64 lanes per wave
    instruction, FMA = 2) over the same image, committed in
    profiles/valu_flops_bp_large.md.  peak = MI355X's published vector peaks
    (78.6 TFLOP/s FP64, 157.3 FP32) combined for the kernel's executed mix
    (mix_peak), so frac is the fraction of the time both peaks would need.
    effective_tflops counts the reference algorithm's flop instead (709 exact
    candidates per refraction solve), which the screened search does not
    execute."""
    f64, f32 = (flops_mixed if form == "mixed" else flops_fp64)(n)
    launches = max(int(rep["launches"]) - 1, 1) if rep["launches"] > 1 else 1
    t_launch = rep["seconds"] / launches
    vp_launch = rep["voxel_pulses"] / launches
    kernel = "k_bp_mixed" if form == "mixed" else "k_bp_fp64"
    flops, fcommit = profile_table(flops_prof)
    traffic, tcommit = profile_table(traffic_prof)
    issue, icommit = profile_table(issue_prof)
    row = flops.get(kernel, {})
    e64, e32 = _num(row, "fp64 flop/vp"), _num(row, "fp32 flop/vp")
    out = {"bound": "valu", "achieved": None, "peak": P64 / 1e12, "unit": "TFLOP/s", "frac": None,
           "traffic": None, "kernel": kernel, "avg_launch_ms": round(t_launch * 1e3, 3), "launches": launches,
           "voxel_pulses_per_launch": vp_launch}
    if e64 is not None and e32 is not None:
        ach = (e64 + e32) * vp_launch / t_launch
        pk = mix_peak(e64, e32)
        out.update(achieved=round(ach / 1e12, 3), peak=round(pk / 1e12, 2), frac=round(ach / pk, 4),
                   executed_flop_per_voxel_pulse={"fp64": e64, "fp32": e32,
                                                  "source": f"profiles/{flops_prof}",
                                                  "commit": fcommit})
    tb = _num(traffic.get(kernel, {}), "bytes per launch")
    if tb is not None:
        out["traffic"] = int(tb)
        out["traffic_source"] = {"source": f"profiles/{traffic_prof}", "commit": tcommit,
                                 "algorithmic_bytes_per_launch": _num(traffic[kernel], "algorithmic bytes")}
    iu = _num(issue.get(kernel, {}), "issue util")
    if iu is not None:
        out["valu_issue_util"] = {"value": iu, "source": f"profiles/{issue_prof}", "commit": icommit}
    eff = (f64 + f32) * vp_launch / t_launch
    out["effective_tflops"] = round(eff / 1e12, 3)
    out["reference_flop_per_voxel_pulse"] = {"fp64": f64, "fp32": f32}
    out["peak_note"] = ("published vector peaks 78.6 (FP64) and 157.3 (FP32) TFLOP/s combined for the executed "
                        "FP64/FP32 mix; non-packed FP32 issues at twice the FP64 rate (c-_bistatic_sar_amd/tools/valu_rates.hip)")
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_port_all_cores(sc, root, side: int = 10) -> dict:
    """The C restatement (oracle/liboracle.so, the reference's per-voxel work
    including its reflection-point scan) over every host core, on a strided
    side^3 voxel subset."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    sub = lambda a, m: a[np.linspace(0, len(a) - 1, m).round().astype(int)]
    gx, gy, gz = sub(sc.gx, side), sub(sc.gy, side), sub(sc.gz, side)
    up = O.upsample(root, sc.upsample)
    t0 = time.time()
    O.backproject(sc.bp_params(), sc.dem, sc.tx, sc.rx, up, gx, gy, gz, reflection=True, threads=cores)
    secs = time.time() - t0
    return {"value": round(side ** 3 * sc.n / secs, 1), "unit": "voxel-pulse/s", "cores": cores, "kind": "port",
            "sample": f"{side}x{side}x{side} strided voxels x {sc.n} pulses in {secs:.1f} s"}


def strided(a, m):
    """m indices spread evenly over a (both ends included)."""
    return np.linspace(0, len(a) - 1, m).round().astype(int)


def cpu_baseline(sc, root, side: int = 20) -> tuple[dict, dict | None]:
    """The reference's own OpenMP back-projection on a strided side^3 voxel
    subset of the same workload: oracle/_ref/proc_test_script_env, the
    reference's proc_test_script built from its sources with the radar
    constants read from the environment, so it runs at this workload's
    carrier, sampling rate and speed of light on the same file set (written
    in the reference's formats).  Its timer covers the voxel loop
    (proc_test_script.cpp:448-801), the reflection-point search included.
    Returns the baseline and the reference's image of the subset (axis indices
    and values) for the parity check; falls back to the C restatement
    (oracle/liboracle.so, image None) where the reference was not built."""
    ref = ROOT / "oracle" / "_ref" / "proc_test_script_env"
    ix, iy, iz = strided(sc.gx, side), strided(sc.gy, side), strided(sc.gz, side)
    gx, gy, gz = sc.gx[ix], sc.gy[iy], sc.gz[iz]
    vox = side ** 3
    cores = os.cpu_count() or 1
    if ref.exists():
        work = Path(tempfile.mkdtemp(prefix="bsar_cpu_"))
        try:
            proc = work / "build" / "output" / "proc"
            load("bsar_amd.scene", PKG_DIR / "scene.py").write_reference_inputs(proc, sc, root, gx, gy, gz)
            threads = min(int(os.environ.get("OMP_NUM_THREADS", cores)), cores)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), BSAR_REF_FC=repr(sc.fc), BSAR_REF_FS=repr(sc.fs),
                       BSAR_REF_C=repr(sc.c))
            t0 = time.time()
            r = subprocess.run([str(ref)], cwd=work, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                               env=env, timeout=900)
            wall = time.time() - t0
            m = re.search(r"Procesamiento paralelo completado en: (\d+) ms", r.stdout) if r.returncode == 0 else None
            if m:
                secs = int(m.group(1)) / 1e3
                used = min(threads, 12)   # proc_test_script.cpp:428 caps the team at 12
                blob = (proc / "output_3D_matrix.bin").read_bytes()
                img = parse_image(blob)
                return ({"value": round(vox * sc.n / secs, 1), "unit": "voxel-pulse/s", "cores": used,
                         "kind": "reference",
                         "sample": f"{vox} voxels ({side}x{side}x{side} strided subset of the grid) x {sc.n} pulses "
                                   f"of the same data; the reference's proc_test_script at fc={sc.fc:g} Hz, "
                                   f"fs={sc.fs:g} Hz x4, c={sc.c:g} m/s; voxel loop {secs:.1f} s by its own timer "
                                   f"(wall {wall:.1f} s with file load + x4 spline); its logging to a pipe"},
                        {"ix": ix, "iy": iy, "iz": iz, "img": img})
            log(f"reference CPU run failed (rc={r.returncode}); using the C restatement")
        finally:
            shutil.rmtree(work, ignore_errors=True)
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    up = O.upsample(root, sc.upsample)
    t0 = time.time()
    O.backproject(sc.bp_params(), sc.dem, sc.tx, sc.rx, up, gx, gy, gz, reflection=True, threads=cores)
    secs = time.time() - t0
    return ({"value": round(vox * sc.n / secs, 1), "unit": "voxel-pulse/s", "cores": cores, "kind": "port",
             "sample": f"{vox} voxels x {sc.n} pulses, C restatement with the reflection search"}, None)


def parse_image(blob: bytes) -> np.ndarray:
    """output_3D_matrix.bin (export3DComplexMatrix, data_export.cpp:180-309):
    52-byte header, dims (nx, ny, nz), complex128 payload x fastest."""
    import struct
    nx, ny, nz = struct.unpack_from("<QQQ", blob, 28)
    return np.frombuffer(blob[52:], np.complex128).reshape(nz, ny, nx)


def oracle_parity(sc, root, img64, imgm) -> dict:
    """The FP64 and mixed images against the oracle's back-projection
    (oracle/liboracle.so, the C restatement of proc_test_script.cpp:451-798)
    on a strided 4x4x4 voxel subset plus the voxel nearest each unit point
    target (where the image peaks), relative to the FP64 image peak."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    nx, ny, nz = len(sc.gx), len(sc.gy), len(sc.gz)
    IX, IY, IZ = (a.ravel() for a in np.meshgrid(strided(sc.gx, 4), strided(sc.gy, 4), strided(sc.gz, 4),
                                                  indexing="ij"))
    pts = sc.scat[np.abs(sc.rcs) >= 0.999]
    near = lambda a, v: int(np.abs(a - v).argmin())
    ix = np.concatenate([IX, [near(sc.gx, p[0]) for p in pts]])
    iy = np.concatenate([IY, [near(sc.gy, p[1]) for p in pts]])
    iz = np.concatenate([IZ, [near(sc.gz, p[2]) for p in pts]])
    ids = np.unique(iz * nx * ny + ix * ny + iy)         # the oracle's z -> x -> y voxel order
    t0 = time.time()
    want = O.backproject(sc.bp_params(), sc.dem, sc.tx, sc.rx, O.upsample(root, sc.upsample), sc.gx, sc.gy, sc.gz,
                         subset=ids)
    secs = time.time() - t0
    yi, xi, zi = ids % ny, (ids // ny) % nx, ids // (nx * ny)
    w = want[zi, yi, xi]
    peak = float(np.abs(img64).max())
    e64 = float(np.abs(img64[zi, yi, xi] - w).max()) / peak
    em = float(np.abs(imgm[zi, yi, xi] - w).max()) / peak
    return {"voxels": int(len(ids)), "subset": "4x4x4 strided + the voxel nearest each of "
                                               f"{len(pts)} point targets", "oracle_s": round(secs, 2),
            "fp64_max_err_rel_peak": e64, "mixed_err_db": round(20 * np.log10(max(em, 1e-300)), 2),
            "subset_peak_rel": round(float(np.abs(w).max()) / peak, 4)}


def serve_stream(pkg, scene, n_jobs: int = 256, rate: float = 40.0, round_ms: float = 40.0,
                 fp64_jobs: int | None = None) -> dict:
    """Serving stream: jobs arrive (Poisson) and run through one bsar_server
    with continuous batching, once in mixed precision and, for the first
    fp64_jobs jobs of the same stream (None: all, 0: none), in plain FP64.
    Latency = submission -> image in the caller's host buffer (server clock);
    throughput = completed voxel-pulses / makespan."""
    import threading
    sc = scene.make_scene("serve")
    jobs = scene.serve_jobs(n_jobs, rate)
    total_vp = sum(j["vp"] for j in jobs)
    log(f"serving stream: {len(jobs)} jobs, Poisson {rate}/s over {jobs[-1]['t']:.2f} s, "
        f"{total_vp:.3e} voxel-pulses; dataset {sc.n} pulses x {sc.S} samples")
    n64 = len(jobs) if fp64_jobs is None else min(fp64_jobs, len(jobs))
    samples = [k for k, j in enumerate(jobs[:max(n64, 1)]) if np.prod(j["shape"]) <= 64 ** 3][:6]
    forms = [("mixed", pkg.MIXED, jobs)] + ([("fp64", pkg.FP64, jobs[:n64])] if n64 > 0 else [])
    res, imgs = {}, {}
    with pkg.Context(0) as ctx:
        t0 = time.time()
        root = scene.simulate(ctx, sc)
        log(f"forward model {time.time() - t0:.1f} s")
        ds = ctx.dataset(pkg.Params(sc.fc, sc.c, sc.fs, sc.n1, sc.n2, sc.upsample), pkg.Dem(*sc.dem), sc.tx, sc.rx)
        ds.load_rootdata(root)
        del root
        log(f"dataset resident: {ds.device_bytes() / 1e9:.2f} GB")
        for name, prec, stream in forms:
            with pkg.Server(round_seconds=round_ms * 1e-3) as srv:
                ids = [None] * len(stream)
                t_start = time.perf_counter()

                def feed():
                    for k, j in enumerate(stream):
                        dt = t_start + j["t"] - time.perf_counter()
                        if dt > 0:
                            time.sleep(dt)
                        ids[k] = srv.submit(ds, j["gx"], j["gy"], j["gz"], pulses=j["K"], precision=prec)

                th = threading.Thread(target=feed)
                th.start()
                infos = []
                while len(infos) < len(stream):
                    try:
                        jid, img, info = srv.wait_any(timeout=1.0)
                    except TimeoutError:
                        continue
                    except pkg.BsarError as e:
                        if e.status == -1 and len(infos) < len(stream):   # nothing submitted yet
                            time.sleep(0.01)
                            continue
                        raise
                    infos.append(info)
                    k = ids.index(jid)
                    if k in samples:
                        imgs[(name, k)] = img
                    if len(infos) % 32 == 0 or len(infos) == len(stream):
                        log(f"{name}: {len(infos)}/{len(stream)} jobs done at {time.perf_counter() - t_start:.1f} s")
                th.join()
                st = srv.stats()
            t_sub = min(i["t_submit"] for i in infos)
            t_end = max(i["t_done"] for i in infos)
            span = t_end - t_sub
            lat = np.array([i["latency"] for i in infos])
            vp = sum(j["vp"] for j in stream)
            res[name] = {"value": round(vp / span, 1), "jobs": len(stream), "voxel_pulses": vp,
                         "makespan_s": round(span, 3),
                         "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 1),
                                        "p99": round(float(np.percentile(lat, 99)) * 1e3, 1),
                                        "mean": round(float(lat.mean()) * 1e3, 1)},
                         "device_busy": round(st["device_seconds"] / span, 3), "rounds": st["rounds"],
                         "items": st["items"], "hbm_peak_gb": round(st["hbm_peak"] / 1e9, 3),
                         "r1_exact_frac_max": max(i["r1_exact_frac"] for i in infos) if name == "mixed" else None}
            log(f"{name}: {res[name]['value']:.3e} voxel-pulse/s, p50 {res[name]['latency_ms']['p50']} ms, "
                f"p99 {res[name]['latency_ms']['p99']} ms, device busy {res[name]['device_busy']}")
        ds.close()
    out = {"workload": f"serve: {len(jobs)} jobs, Poisson {rate}/s, grids {scene.SERVE_GRIDS}, pulses "
                       f"{scene.SERVE_PULSES} (first K of {sc.n} x {sc.S} samples x4)",
           "jobs": len(jobs), "rate": rate, "round_ms": round_ms, "mixed": res["mixed"]}
    if "fp64" in res:
        out["fp64"] = res["fp64"]
        if n64 < len(jobs):
            out["fp64"]["note"] = f"the first {n64} jobs of the stream"
        errs = [float(np.abs(imgs[("mixed", k)] - imgs[("fp64", k)]).max() / np.abs(imgs[("fp64", k)]).max())
                for k in samples]
        out["mixed_vs_fp64"] = {"throughput_ratio": round(res["mixed"]["value"] / res["fp64"]["value"], 3),
                                "peak_err_db_max": round(20 * np.log10(max(max(errs), 1e-300)), 2),
                                "jobs_compared": len(samples)}
    return out


def run_serve(args, pkg, scene) -> dict:
    srv = serve_stream(pkg, scene, args.jobs, args.rate, args.round_ms, 0 if args.no_fp64 else None)
    m = srv["mixed"]
    out = {"metric": "serving stream: aggregate voxel-pulse BP/s (mixed precision) with p50/p99 job latency",
           "value": m["value"], "unit": "voxel-pulse/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": round(m["makespan_s"] * 1e3, 1), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic (seeded scene and job stream)",
           "config": {"workload": srv["workload"], "jobs": srv["jobs"], "rate": srv["rate"],
                      "voxel_pulses": m["voxel_pulses"], "round_ms": srv["round_ms"]},
           "latency_ms": m["latency_ms"]}
    out.update({k: v for k, v in srv.items() if k in ("mixed", "fp64", "mixed_vs_fp64")})
    return out


FLOP_PER_SOLVE = 709 * 46 + 64     # one refraction search: 709 candidates x 46 flop + set-up


def band_order(p: np.ndarray, nx: int, ny: int, nz: int, by: int = 8, bz: int = 8):
    """Grid coordinates of positions p in the back-projection kernels' voxel
    order: bands of `by` rows x `bz` planes, x-major inside a band, the band's
    (y, z) cells innermost -- so the 64 points of a wavefront are one x column
    of an 8 x 8 (y, z) block, which keeps the lanes' refraction searches on the
    same branches."""
    plane = nx * ny
    c = p // (bz * plane)
    bd = np.minimum(bz, nz - bz * c)
    r = p - c * bz * plane
    b = r // (by * bd * nx)
    bh = np.minimum(by, ny - by * b)
    w = r - b * by * bd * nx
    cell = bh * bd
    xi = w // cell
    s = w - xi * cell
    zz = s // bh
    return xi, by * b + (s - zz * bh), bz * c + zz


def run_rough(args, pkg, scene) -> dict:
    """Refraction-dominated workload: every refraction solve of a 128^3 grid
    buried 3-7 m under rough fBm terrain (167x100 DEM) for 1024 pulses, both
    legs (2 V N = 4.29e9 solves) through the device-buffer slant-range entry
    point, in voxel chunks of the back-projection kernels' band order; then the argmin of a seeded sample of >= 1e5
    solves against the reference's own calculateRefractionPoint on the host
    (identical point, delay difference <= 1e-12 s)."""
    sc = scene.make_scene("rough")
    V, N = sc.voxels, sc.n
    nx, ny = len(sc.gx), len(sc.gy)
    log(f"workload rough: {nx}x{ny}x{len(sc.gz)} voxels x {N} pulses, DEM {len(sc.dem[0])}x{len(sc.dem[1])}, "
        f"n2 = {sc.n2:.4f}")
    chunk = min(args.chunk, V)
    rng = np.random.default_rng(1003 + 7)
    per_chunk = max(1, -(-args.check_solves // (2 * N * (-(-V // chunk)))))
    dev_s, legs_a, legs_b, got_a, got_b = 0.0, [], [], [], []
    with pkg.Context(0) as ctx:
        ds = ctx.dataset(pkg.Params(sc.fc, sc.c, sc.fs, sc.n1, sc.n2, sc.upsample), pkg.Dem(*sc.dem), sc.tx, sc.rx)
        bp = pkg.DeviceBuffer(chunk * 3 * 8)
        ba = pkg.DeviceBuffer(chunk * N * 3 * 8)
        bb = pkg.DeviceBuffer(chunk * N * 3 * 8)
        t_start = time.perf_counter()
        for c0 in range(0, V, chunk):
            m = min(chunk, V - c0)
            v = np.arange(c0, c0 + m)
            xi, yi, zi = band_order(v, nx, ny, len(sc.gz))
            pts = np.stack([sc.gx[xi], sc.gy[yi], sc.gz[zi]], 1)
            bp.write(pts)
            rep = ds.slant_range_fermat_device(bp.ptr, m, ba.ptr, bb.ptr)
            dev_s += rep["seconds"]
            for vl in rng.choice(m, size=min(per_chunk, m), replace=False):
                qa = ba.read((N, 3), offset=int(vl) * N * 24)
                qb = bb.read((N, 3), offset=int(vl) * N * 24)
                P = np.broadcast_to(pts[vl], (N, 3))
                legs_a.append(np.concatenate([sc.tx, P], 1))
                legs_b.append(np.concatenate([P, sc.rx], 1))
                got_a.append(qa)
                got_b.append(qb)
            log(f"chunk {c0 // chunk + 1}/{-(-V // chunk)}: {m} voxels x {N} pulses x 2 legs in "
                f"{rep['seconds']:.3f} s ({2 * m * N / rep['seconds']:.3e} solves/s)")
        wall = time.perf_counter() - t_start
        for b in (bp, ba, bb):
            b.close()
        ds.close()
    solves = 2.0 * V * N
    value = solves / dev_s
    la, lb = np.concatenate(legs_a), np.concatenate(legs_b)
    ga, gb = np.concatenate(got_a), np.concatenate(got_b)
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    t0 = time.time()
    ra = O.ref_refraction(sc.dem, la, sc.n1, sc.n2, threads=threads)
    kind = "reference"
    if ra is None:
        kind = "port"
        ra = O.refraction_batch(sc.dem, la, sc.n1, sc.n2, threads=threads)
        rb = O.refraction_batch(sc.dem, lb, sc.n2, sc.n1, threads=threads)
    else:
        rb = O.ref_refraction(sc.dem, lb, sc.n2, sc.n1, threads=threads)
    cpu_s = time.time() - t0
    n_check = len(la) + len(lb)

    def delay(legs, q, na, nb):
        d1 = np.sqrt(((legs[:, 0:3] - q) ** 2).sum(1))
        d2 = np.sqrt(((q - legs[:, 3:6]) ** 2).sum(1))
        return (na * d1 + nb * d2) / sc.c

    # The same sample through the unscreened wavefront reduction (one wave per
    # leg, lanes over all 709 candidates), each leg set tiled to >= 5e5 legs:
    # its rate beside the screened search's, and its points against the
    # screened ones.
    with pkg.Context(0) as ctx:
        dem = pkg.Dem(*sc.dem)
        wave_s, wave_n, wave_same = 0.0, 0, 0
        for legs, got, na, nb in ((la, ga, sc.n1, sc.n2), (lb, gb, sc.n2, sc.n1)):
            reps = max(1, -(-500_000 // len(legs)))
            q, secs = ctx.refraction_points_wave(dem, np.tile(legs, (reps, 1)), na, nb)
            wave_s += secs
            wave_n += reps * len(legs)
            wave_same += int((q[:len(legs)] == got).all(1).sum())
    wave = {"solves": wave_n, "value": round(wave_n / wave_s, 1), "unit": "solve/s", "seconds": round(wave_s, 4),
            "screened_over_wave": round(value / (wave_n / wave_s), 2), "identical_to_screened": wave_same,
            "compared": n_check, "kernel": "k_refraction_wave",
            "note": "one wavefront per leg, lanes over every candidate of each level, argmin with the lowest "
                    "reference index among ties; the argmin-check sample tiled"}
    log(f"wavefront reduction: {wave['value']:.3e} solves/s on {wave_n} solves; screened search "
        f"{wave['screened_over_wave']}x faster; {wave_same}/{n_check} identical to the screened points")
    same = int((ga == ra).all(1).sum() + (gb == rb).all(1).sum())
    dtau = max(float(np.abs(delay(la, ga, sc.n1, sc.n2) - delay(la, ra, sc.n1, sc.n2)).max()),
               float(np.abs(delay(lb, gb, sc.n2, sc.n1) - delay(lb, rb, sc.n2, sc.n1)).max()))
    log(f"argmin check: {same}/{n_check} identical, max |dtau| = {dtau:.3e} s ({kind} CPU, {cpu_s:.1f} s)")
    ach = value * FLOP_PER_SOLVE                 # the reference algorithm's flop
    rows, rcommit = profile_table("valu_flops_slant_rough.md")
    e64, e32 = _num(rows.get("k_slant_range", {}), "fp64 flop/solve"), _num(rows.get("k_slant_range", {}),
                                                                             "fp32 flop/solve")
    executed = None if e64 is None or e32 is None else value * (e64 + e32)
    pk = P64 if executed is None else mix_peak(e64, e32)
    return {"metric": "refraction solves/s (Fermat search over the DEM, FP64), both legs of every voxel-pulse",
            "value": round(value, 1), "unit": "solve/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": round(dev_s * 1e3, 1), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded fBm DEM, spiral antennas)",
            "config": {"workload": f"rough: {nx}x{ny}x{len(sc.gz)} grid 3-7 m under a {len(sc.dem[0])}x"
                                   f"{len(sc.dem[1])} fBm DEM (H 0.7, 40 m RMS), {N} pulses, eps_r "
                                   f"{sc.n2 ** 2:.3f}", "voxels": V, "pulses": N, "solves": solves,
                       "chunk_voxels": chunk},
            "wall_s": round(wall, 2),
            "roofline": {"bound": "valu",
                         "achieved": None if executed is None else round(executed / 1e12, 3),
                         "peak": round(pk / 1e12, 2),
                         "unit": "TFLOP/s", "frac": None if executed is None else round(executed / pk, 4),
                         "traffic": None, "kernel": "k_slant_range",
                         "executed_flop_per_solve": None if executed is None else {
                             "fp64": e64, "fp32": e32, "source": "profiles/valu_flops_slant_rough.md",
                             "commit": rcommit},
                         "effective_tflops": round(ach / 1e12, 3), "reference_flop_per_solve": FLOP_PER_SOLVE},
            "argmin_check": {"solves": n_check, "identical": same, "max_abs_dtau_s": dtau, "cpu": kind},
            "wave_reduction": wave,
            "cpu_baseline": {"value": round(n_check / cpu_s, 1), "unit": "solve/s", "cores": threads, "kind": kind,
                             "sample": f"{n_check} solves (the argmin-check sample: {len(legs_a)} voxels x "
                                       f"{N} pulses x 2 legs), calculateRefractionPoint over {threads} threads"},
            "vs_cpu": round(value / (n_check / cpu_s), 1)}


def run_wide(args, pkg, scene) -> dict:
    """Range compression + back-projection end to end: raw data 4096 pulses x
    16384 samples (64 scatterers, B = 500 MHz) resident on the device as
    complex64, FP32 batched FFT -> x conj(H) -> IFFT in place, x4 spline, then
    the mixed back-projection of the full 512x512x64 grid.  The FP64 chain
    (complex128 compression, plain back-projection) runs on a strided
    64x64x64 voxel subset; the mixed image is compared with it there, in dB of
    the FP64 peak."""
    sc = scene.make_scene("wide")
    V, N, S = sc.voxels, sc.n, sc.S
    log(f"workload wide: {len(sc.gx)}x{len(sc.gy)}x{len(sc.gz)} voxels x {N} pulses, {S} samples x{sc.upsample}")
    out = {}
    with pkg.Context(0) as ctx:
        t0 = time.time()
        raw = scene.simulate(ctx, sc, compress=False)
        ref = scene.matched_filter(sc)
        log(f"forward model (raw data) {time.time() - t0:.1f} s")
        params = pkg.Params(sc.fc, sc.c, sc.fs, sc.n1, sc.n2, sc.upsample)
        # one-pulse warm-up of each precision at this S (builds the cached
        # twiddle tables outside the timed calls)
        for dt, prec in ((np.complex64, pkg.MIXED), (np.complex128, pkg.FP64)):
            bw = pkg.DeviceBuffer.from_array(raw[:1].astype(dt))
            ctx.range_compress_device(bw, S, 1, ref, conj=True, precision=prec)
            bw.close()
        # FP32 chain
        b32 = pkg.DeviceBuffer.from_array(raw.astype(np.complex64))
        rc32 = ctx.range_compress_device(b32, S, N, ref, conj=True, precision=pkg.MIXED)
        comp32 = b32.read((N, S), np.complex64)
        dsm = ctx.dataset(params, pkg.Dem(*sc.dem), sc.tx, sc.rx)
        dsm.load_rootdata_device(b32, S, N, pkg.MIXED)
        b32.close()
        # FP64 chain
        b64 = pkg.DeviceBuffer.from_array(raw)
        rc64 = ctx.range_compress_device(b64, S, N, ref, conj=True, precision=pkg.FP64)
        comp64 = b64.read((N, S), np.complex128)
        ds64 = ctx.dataset(params, pkg.Dem(*sc.dem), sc.tx, sc.rx)
        ds64.load_rootdata_device(b64, S, N, pkg.FP64)
        b64.close()
        del raw
        rc_err = float(np.abs(comp32 - comp64).max() / np.abs(comp64).max())
        bytes32 = 2.0 * N * S * 8
        gbs = bytes32 / rc32["seconds"] / 1e9
        gbs64 = 2.0 * bytes32 / rc64["seconds"] / 1e9
        log(f"range compression FP32: {rc32['seconds'] * 1e3:.2f} ms ({gbs:.0f} GB/s), FP64: "
            f"{rc64['seconds'] * 1e3:.2f} ms; FP32 vs FP64 {20 * np.log10(max(rc_err, 1e-300)):.1f} dB")
        t1 = time.perf_counter()
        img_m, rep = dsm.backproject(sc.gx, sc.gy, sc.gz, pkg.MIXED)
        wall_m = time.perf_counter() - t1
        log(f"mixed image: {rep['seconds']:.2f} s device ({V * N / rep['seconds']:.3e} voxel-pulse/s), "
            f"{rep['launches']} launches, R1 exact fraction {rep['r1_exact_frac']:.2e}")
        sx = slice(None, None, 8)
        gx, gy, gz = sc.gx[sx], sc.gy[sx], sc.gz
        t1 = time.perf_counter()
        img64, rep64 = ds64.backproject(gx, gy, gz, pkg.FP64)
        wall64 = time.perf_counter() - t1
        sub = img_m[:, sx, sx]
        err = float(np.abs(sub - img64).max() / np.abs(img64).max())
        Q = load("bsar_amd.quality", PKG_DIR / "quality.py")
        lob = Q.lobe_deltas(np.ascontiguousarray(sub), img64, gx, gy, gz, sc.scat[np.abs(sc.rcs) >= 0.999])
        v64 = len(gx) * len(gy) * len(gz)
        log(f"fp64 subset image: {rep64['seconds']:.2f} s ({v64 * N / rep64['seconds']:.3e} voxel-pulse/s); "
            f"mixed vs FP64 {20 * np.log10(max(err, 1e-300)):.1f} dB")
        dsm.close()
        ds64.close()
    value = V * N / rep["seconds"]
    out = {"metric": "voxel-pulse BP/s, mixed precision, after FP32 batched range compression (end to end)",
           "value": round(value, 1), "unit": "voxel-pulse/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": round(rep["seconds"] * 1e3, 1), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic (seeded scene, GPU forward model)",
           "config": {"workload": f"wide: {len(sc.gx)}x{len(sc.gy)}x{len(sc.gz)} voxels x {N} pulses, {S} samples "
                                  f"x{sc.upsample}, 64 scatterers, B 500 MHz", "voxels": V, "pulses": N,
                      "samples": S},
           "wall_s": round(wall_m, 2),
           "roofline": roofline(rep, N, "mixed", "valu_flops_bp_wide.md", "none", "none"),
           "range_compression": {"fp32_ms": round(rc32["seconds"] * 1e3, 3), "fp64_ms": round(rc64["seconds"] * 1e3, 3),
                                 "fp32_gb_s": round(gbs, 1), "roofline_frac": round(gbs * 1e9 / 8e12, 4),
                                 "algorithmic_bytes": bytes32, "bound": "hbm",
                                 "fp64_gb_s": round(gbs64, 1), "fp64_roofline_frac": round(gbs64 * 1e9 / 8e12, 4),
                                 "fp64_algorithmic_bytes": 2.0 * bytes32,
                                 "fp32_vs_fp64_db": round(20 * np.log10(max(rc_err, 1e-300)), 2)},
           "fp64_subset": {"voxels": v64, "value": round(v64 * N / rep64["seconds"], 1),
                           "seconds": round(rep64["seconds"], 2)},
           "mixed_vs_fp64": {"peak_err_db": round(20 * np.log10(max(err, 1e-300)), 2),
                             "speedup": round((V * N / rep["seconds"]) / (v64 * N / rep64["seconds"]), 3),
                             "pslr_islr": {"targets": lob["targets"], "grid": "64x64x64 subset",
                                           "max_abs_d_pslr_db": lob["max_abs_d_pslr_db"],
                                           "max_abs_d_islr_db": lob["max_abs_d_islr_db"]}}}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="large")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--jobs", type=int, default=256, help="serve: jobs in the stream")
    ap.add_argument("--rate", type=float, default=40.0, help="serve: Poisson arrival rate, jobs/s")
    ap.add_argument("--round-ms", type=float, default=40.0, help="serve: target device time per round")
    ap.add_argument("--chunk", type=int, default=65536, help="rough: voxels per refraction launch")
    ap.add_argument("--check-solves", type=int, default=100000, help="rough: solves checked against the CPU")
    ap.add_argument("--no-serve", action="store_true", help="large/small: skip the serving stream")
    ap.add_argument("--no-parity", action="store_true", help="large/small: skip the oracle parity check")
    ap.add_argument("--traffic-calibration", action="store_true",
                    help="large/small: also load the upsampled samples through bsar_dataset_load_upsampled, whose "
                         "complex64 conversion (k_to_f32: reads 16 B, writes 8 B per sample) calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    ap.add_argument("--serve-fp64-jobs", type=int, default=32,
                    help="large/small: FP64 serving on this prefix of the stream (0: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "the image shards by voxel slab across GPUs with no exchange, see the design notes.", file=sys.stderr)
        return 2

    pkg = load("bsar_amd", PKG_DIR / "__init__.py", pkg=True)
    scene = load("bsar_amd.scene", PKG_DIR / "scene.py")
    if args.workload == "serve":
        print(json.dumps(run_serve(args, pkg, scene)), flush=True)
        return 0
    if args.workload == "wide":
        print(json.dumps(run_wide(args, pkg, scene)), flush=True)
        return 0
    if args.workload == "rough":
        print(json.dumps(run_rough(args, pkg, scene)), flush=True)
        return 0
    sc = scene.make_scene(args.workload)
    V, N = sc.voxels, sc.n
    log(f"workload {sc.name}: {len(sc.gx)}x{len(sc.gy)}x{len(sc.gz)} voxels x {N} pulses, {sc.S} samples")
    with pkg.Context(0) as ctx:
        t0 = time.time()
        root = scene.simulate(ctx, sc)
        log(f"forward model (GPU Fermat + FFT) {time.time() - t0:.1f} s")
        ds = ctx.dataset(pkg.Params(sc.fc, sc.c, sc.fs, sc.n1, sc.n2, sc.upsample), pkg.Dem(*sc.dem), sc.tx, sc.rx)
        ds.load_rootdata(root)
        if args.traffic_calibration:
            cal = ctx.dataset(pkg.Params(sc.fc, sc.c, sc.fs, sc.n1, sc.n2, sc.upsample), pkg.Dem(*sc.dem), sc.tx,
                              sc.rx)
            up = ctx.upsample(root, sc.upsample)
            cal.load_upsampled(up)
            log(f"traffic calibration: k_to_f32 over {up.size} complex128 samples ({up.size * 16} B read, "
                f"{up.size * 8} B written)")
            del up
            cal.close()
        for w in range(args.warmup):
            _, rep = ds.backproject(sc.gx, sc.gy, sc.gz, pkg.MIXED)
            log(f"warmup {w + 1}/{args.warmup}: {rep['seconds']:.2f} s")
        lat, reps = [], []
        pkg.lib().bsar_ctx_synchronize(ctx.h)
        t_start = time.perf_counter()
        for k in range(args.steps):
            t1 = time.perf_counter()
            img_m, rep = ds.backproject(sc.gx, sc.gy, sc.gz, pkg.MIXED)
            lat.append(time.perf_counter() - t1)
            reps.append(rep)
            log(f"step {k + 1}/{args.steps}: {lat[-1]:.2f} s wall, {rep['seconds']:.2f} s device, "
                f"{rep['launches']} launches, R1 exact fraction {rep['r1_exact_frac']:.2e}")
        pkg.lib().bsar_ctx_synchronize(ctx.h)
        elapsed = time.perf_counter() - t_start
        value = V * N * args.steps / elapsed
        out = {"metric": "voxel-pulse BP/s, mixed precision (FP64 refraction + delay, FP32 interp/Kahan accum)",
               "value": round(value, 1), "unit": "voxel-pulse/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2),
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
               "data": "synthetic (seeded scene, GPU forward model, range-compressed on device)",
               "config": {"workload": f"{sc.name}: {len(sc.gx)}x{len(sc.gy)}x{len(sc.gz)} voxels x {N} pulses, "
                                      f"{sc.S} samples x{sc.upsample}", "voxels": V, "pulses": N,
                          "samples": sc.S, "precision": "mixed"},
               "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 1),
                              "p99": round(float(np.percentile(lat, 99)) * 1e3, 1),
                              "note": "whole-image step time; job latency of the serving stream is under serving"},
               "device_s_per_step": round(float(np.mean([r["seconds"] for r in reps])), 4)}
        out["roofline"] = roofline(reps[-1], N, "mixed")
        img64 = None
        if not args.no_fp64:
            t1 = time.perf_counter()
            img64, rep64 = ds.backproject(sc.gx, sc.gy, sc.gz, pkg.FP64)
            w64 = time.perf_counter() - t1
            log(f"fp64 image: {w64:.2f} s wall, {rep64['seconds']:.2f} s device")
            err = float(np.abs(img_m - img64).max() / np.abs(img64).max())
            out["fp64"] = {"value": round(V * N / w64, 1), "ms": round(w64 * 1e3, 1),
                           "device_s": round(rep64["seconds"], 4), "roofline": roofline(rep64, N, "fp64")}
            out["mixed_vs_fp64"] = {"speedup": round(w64 / (elapsed / args.steps), 3),
                                    "peak_err_db": round(20 * np.log10(max(err, 1e-300)), 2)}
            Q = load("bsar_amd.quality", PKG_DIR / "quality.py")
            pts = sc.scat[np.abs(sc.rcs) >= 0.999]
            lob = Q.lobe_deltas(img_m, img64, sc.gx, sc.gy, sc.gz, pts)
            out["mixed_vs_fp64"]["pslr_islr"] = {
                "targets": lob["targets"], "max_abs_d_pslr_db": lob["max_abs_d_pslr_db"],
                "max_abs_d_islr_db": lob["max_abs_d_islr_db"],
                "fp64_pslr_db": [round(r["pslr_db"], 2) for r in lob["per_target"]],
                "fp64_islr_db": [round(r["islr_db"], 2) for r in lob["per_target"]]}
            log(f"PSLR/ISLR deltas over {lob['targets']} point targets: max |dPSLR| "
                f"{lob['max_abs_d_pslr_db']:.2e} dB, max |dISLR| {lob['max_abs_d_islr_db']:.2e} dB")
        ds.close()
    if img64 is not None and not args.no_parity:
        par = oracle_parity(sc, root, img64, img_m)
        par["tolerance"] = {"fp64": 1e-9, "mixed_db": -80.0}
        out["parity"] = par
        log(f"parity vs oracle on {par['voxels']} voxels: fp64 {par['fp64_max_err_rel_peak']:.2e} of the peak, "
            f"mixed {par['mixed_err_db']} dB")
    if not args.no_cpu:
        log("cpu baseline (reference proc_test_script at the workload's constants) ...")
        cb, refimg = cpu_baseline(sc, root)
        out["cpu_baseline"] = cb
        cb["cpu_model"] = cpu_model()
        out["vs_cpu"] = round(value / cb["value"], 1)
        log(f"cpu baseline: {cb['value']:.3e} voxel-pulse/s on {cb['cores']} threads ({cb['kind']})")
        if refimg is not None and img64 is not None and "parity" in out:
            sub = img64[np.ix_(refimg["iz"], refimg["iy"], refimg["ix"])]
            subm = img_m[np.ix_(refimg["iz"], refimg["iy"], refimg["ix"])]
            peak = float(np.abs(img64).max())
            e = float(np.abs(sub - refimg["img"]).max()) / peak
            em = float(np.abs(subm - refimg["img"]).max()) / peak
            out["parity"]["fp64_vs_reference"] = {
                "voxels": int(refimg["img"].size), "max_err_rel_peak": e,
                "mixed_err_db": round(20 * np.log10(max(em, 1e-300)), 2),
                "subset_peak_rel": round(float(np.abs(refimg["img"]).max()) / peak, 4),
                "binary": "oracle/_ref/proc_test_script_env (the reference's proc_test_script, radar constants "
                          "from the environment)"}
            log(f"fp64 vs the reference's own image on {refimg['img'].size} voxels: {e:.2e} of the peak")
        log("C restatement on every core ...")
        cb["all_cores"] = cpu_port_all_cores(sc, root)
    if not args.no_serve:
        srv = serve_stream(pkg, scene, args.jobs, args.rate, args.round_ms, args.serve_fp64_jobs)
        out["serving"] = srv
        out["latency_ms"]["serving_p50"] = srv["mixed"]["latency_ms"]["p50"]
        out["latency_ms"]["serving_p99"] = srv["mixed"]["latency_ms"]["p99"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
