"""Image quality figures for comparing the mixed and plain FP64 images.

* ``peak_err_db``: 20 log10(max |a - b| / max |b|), peak-normalised (per-voxel
  relative error is undefined at image nulls).
* ``lobe_metrics``: peak-to-sidelobe ratio (PSLR) and integrated sidelobe ratio
  (ISLR) around a point target, from the three axis cuts through the image
  peak nearest the target: the main lobe of a cut runs from the peak to the
  first local minimum on each side; PSLR = largest sidelobe sample / peak,
  ISLR = sidelobe energy / main-lobe energy summed over the three cuts.
"""
from __future__ import annotations

import numpy as np


def peak_err_db(a: np.ndarray, ref: np.ndarray) -> float:
    m = float(np.abs(ref).max()) if ref.size else 0.0
    d = float(np.abs(a - ref).max()) if ref.size else 0.0
    x = d / m if m > 0 else d
    return 20.0 * np.log10(max(x, 1e-300))


def _main_lobe(cut: np.ndarray, k: int) -> tuple[int, int]:
    lo = k
    while lo > 0 and cut[lo - 1] < cut[lo]:
        lo -= 1
    hi = k
    while hi + 1 < len(cut) and cut[hi + 1] < cut[hi]:
        hi += 1
    return lo, hi


def cut_metrics(cut: np.ndarray, k: int) -> tuple[float, float, float]:
    """(largest sidelobe sample, sidelobe energy, main-lobe energy) of a 1-D
    magnitude cut whose peak is at index k."""
    lo, hi = _main_lobe(cut, k)
    main = cut[lo:hi + 1]
    side = np.concatenate([cut[:lo], cut[hi + 1:]])
    return (float(side.max()) if side.size else 0.0), float((side ** 2).sum()), float((main ** 2).sum())


def lobe_metrics(img: np.ndarray, near: tuple[int, int, int], search: int = 3) -> dict:
    """PSLR / ISLR (dB) of the target whose nearest voxel is ``near`` =
    (iz, iy, ix) in an image [nz][ny][nx]; the peak is taken within
    +-``search`` voxels of it."""
    a = np.abs(img)
    sl = tuple(slice(max(c - search, 0), c + search + 1) for c in near)
    sub = a[sl]
    off = np.unravel_index(int(np.argmax(sub)), sub.shape)
    p = tuple(s.start + o for s, o in zip(sl, off))
    peak = float(a[p])
    side_max, e_side, e_main = 0.0, 0.0, 0.0
    for ax in range(3):
        idx = list(p)
        idx[ax] = slice(None)
        sm, es, em = cut_metrics(a[tuple(idx)], p[ax])
        side_max, e_side, e_main = max(side_max, sm), e_side + es, e_main + em
    db = lambda x: 20.0 * np.log10(max(x, 1e-300))
    return {"peak": peak, "at": [int(c) for c in p],
            "pslr_db": db(side_max / peak) if peak > 0 else float("nan"),
            "islr_db": 10.0 * np.log10(max(e_side / e_main, 1e-300)) if e_main > 0 else float("nan")}


def nearest_voxel(gx, gy, gz, pt) -> tuple[int, int, int]:
    return (int(np.argmin(np.abs(np.asarray(gz) - pt[2]))), int(np.argmin(np.abs(np.asarray(gy) - pt[1]))),
            int(np.argmin(np.abs(np.asarray(gx) - pt[0]))))


def lobe_deltas(img: np.ndarray, ref: np.ndarray, gx, gy, gz, targets) -> dict:
    """PSLR / ISLR of each target in ``img`` and ``ref`` and their
    differences (img - ref, dB)."""
    rows = []
    for t in targets:
        near = nearest_voxel(gx, gy, gz, t)
        m, r = lobe_metrics(img, near), lobe_metrics(ref, near)
        rows.append({"target": [float(x) for x in t], "pslr_db": r["pslr_db"], "islr_db": r["islr_db"],
                     "d_pslr_db": m["pslr_db"] - r["pslr_db"], "d_islr_db": m["islr_db"] - r["islr_db"]})
    dp = [abs(x["d_pslr_db"]) for x in rows if np.isfinite(x["d_pslr_db"])]
    di = [abs(x["d_islr_db"]) for x in rows if np.isfinite(x["d_islr_db"])]
    return {"targets": len(rows), "max_abs_d_pslr_db": max(dp) if dp else None,
            "max_abs_d_islr_db": max(di) if di else None, "per_target": rows}
