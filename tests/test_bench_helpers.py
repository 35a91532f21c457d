"""Host-side helpers of bench.py (no GPU): the voxel-band order the rough
benchmark hands to the search is a bijection onto the grid, the mixed
FP64/FP32 vector peak reduces to each precision's peak, and a request for
more than one GPU exits with status 2."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 4), (3, 17, 9), (41, 41, 21), (64, 3, 70), (128, 16, 16)])
@pytest.mark.parametrize("by,bz", [(8, 8), (4, 4), (1, 1)])
def test_band_order_is_a_bijection(shape, by, bz):
    nx, ny, nz = shape
    n = nx * ny * nz
    x, y, z = bench.band_order(np.arange(n), nx, ny, nz, by=by, bz=bz)
    assert x.min() >= 0 and y.min() >= 0 and z.min() >= 0
    assert x.max() < nx and y.max() < ny and z.max() < nz
    assert len(np.unique((z * ny + y) * nx + x)) == n


def test_band_order_groups_of_64_are_compact_blocks():
    # with full bands, 64 consecutive positions are one x column of an 8 x 8 (y, z) block
    x, y, z = bench.band_order(np.arange(64 * 5), 5, 16, 8)
    for g in range(5):
        sl = slice(64 * g, 64 * (g + 1))
        assert len(set(x[sl])) == 1 and np.ptp(y[sl]) == 7 and np.ptp(z[sl]) == 7


def test_band_order_one_by_one_is_x_fastest():
    x, y, z = bench.band_order(np.arange(60), 5, 4, 3, by=1, bz=1)
    v = np.arange(60)
    assert (x == v % 5).all() and (y == (v // 5) % 4).all() and (z == v // 20).all()


def test_mix_peak_limits():
    assert bench.mix_peak(1.0, 0.0) == pytest.approx(bench.P64)
    assert bench.mix_peak(0.0, 1.0) == pytest.approx(bench.P32)
    assert bench.P64 < bench.mix_peak(1.0, 1.0) < bench.P32


def test_more_than_one_gpu_exits_2():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
