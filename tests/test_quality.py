"""Image quality figures (PSLR / ISLR / peak-normalised error) against closed
forms: a separable sinc has its first sidelobe at 0.2172 of the peak
(-13.26 dB) and 9.72 % of its energy outside the main lobe (ISLR -9.68 dB)."""
import importlib.util

import numpy as np
import pytest

from conftest import PKG_DIR


@pytest.fixture(scope="module")
def Q():
    spec = importlib.util.spec_from_file_location("bsar_quality", PKG_DIR / "quality.py")
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def sinc_cube(n=321, per_lobe=8.0, shift=(0.0, 0.0, 0.0)):
    t = (np.arange(n) - n // 2) / per_lobe
    sx, sy, sz = (np.sinc(t - s) for s in shift)
    return (sz[:, None, None] * sy[None, :, None] * sx[None, None, :]).astype(np.complex128)


def test_sinc_cut_closed_form(Q):
    """Finely sampled sinc over +-100 lobes: PSLR and ISLR of the continuous one."""
    per = 64
    t = (np.arange(200 * per + 1) - 100 * per) / per
    side, es, em = Q.cut_metrics(np.abs(np.sinc(t)), 100 * per)
    assert 20 * np.log10(side) == pytest.approx(20 * np.log10(0.217234), abs=0.01)
    assert 10 * np.log10(es / em) == pytest.approx(10 * np.log10(0.0972 / 0.9028), abs=0.15)


def test_cube_is_three_cuts(Q):
    """A separable sinc cube: PSLR / ISLR of the three cuts = those of one cut."""
    img = sinc_cube(n=161)
    c = img.shape[0] // 2
    m = Q.lobe_metrics(img, (c, c, c))
    assert m["at"] == [c, c, c]
    side, es, em = Q.cut_metrics(np.abs(img[c, c, :]), c)
    assert m["pslr_db"] == pytest.approx(20 * np.log10(side), abs=1e-9)
    assert m["islr_db"] == pytest.approx(10 * np.log10(es / em), abs=1e-9)
    assert -13.5 < m["pslr_db"] < -13.2


def test_peak_search_and_deltas(Q):
    img = sinc_cube(n=101, shift=(0.0, 0.0, 0.0))
    gx = gy = gz = np.arange(101) * 0.01
    t = [(0.52, 0.49, 0.51)]                       # nearest voxel off the peak by <= 2: found by the search
    d = Q.lobe_deltas(img, img, gx, gy, gz, t)
    assert d["max_abs_d_pslr_db"] == 0.0 and d["max_abs_d_islr_db"] == 0.0
    noisy = img * (1 + 1e-3 * np.random.default_rng(0).normal(size=img.shape))
    d = Q.lobe_deltas(noisy, img, gx, gy, gz, t)
    assert 0 < d["max_abs_d_pslr_db"] < 0.05
    assert Q.peak_err_db(noisy, img) == pytest.approx(
        20 * np.log10(np.abs(noisy - img).max() / np.abs(img).max()))
    assert Q.peak_err_db(img, img) < -3000
