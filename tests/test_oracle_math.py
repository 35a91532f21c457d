"""The CPU oracle against what the mathematics fixes: brute-force DFT and
circular correlation, spline invariants, DEM closed forms, Snell's law on a
flat interface, and linearity / single-pulse closed forms of the
back-projection.  Each check is chosen so that a dropped term, a wrong sign
or a transposed index in the oracle fails it."""
import numpy as np
import pytest
from scipy.optimize import minimize


def dft(x):
    n = len(x)
    k = np.arange(n)
    return np.exp(-2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [1, 2, 4, 16, 128, 512])
def test_fft_matches_dft(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    X = oracle.fft(x)
    assert np.allclose(X, dft(x), rtol=0, atol=1e-12 * max(1, np.abs(X).max()))
    assert np.allclose(oracle.ifft(X), x, atol=1e-13)
    assert np.isclose(np.sum(np.abs(X) ** 2), n * np.sum(np.abs(x) ** 2), rtol=1e-12)


@pytest.mark.parametrize("S", [1024, 32768])
def test_fft_recurrence_error_grows_with_length(oracle, S):
    """The reference's twiddle recurrence w *= wlen (fft_utils.cpp:40) rounds
    once per butterfly of a block, so the oracle's correlation drifts from an
    exactly-twiddled one by O(eps * S): within 1e-16 * S of the peak, and at
    S = 32768 above the 1e-13 an exact-twiddle transform keeps."""
    rng = np.random.default_rng(S)
    s = rng.normal(size=S) + 1j * rng.normal(size=S)
    r = rng.normal(size=S) + 1j * rng.normal(size=S)
    exact = np.fft.ifft(np.fft.fft(s) * np.conj(np.fft.fft(r)))
    err = np.abs(oracle.fft_correlation(s, r) - exact).max() / np.abs(exact).max()
    assert err <= 1e-16 * S
    if S >= 32768:
        assert err > 1e-13


def test_fft_non_power_of_two_is_noop(oracle):
    x = np.arange(6) + 1j
    assert np.array_equal(oracle.fft(x), x)
    assert np.allclose(oracle.ifft(x), x / 6)


def test_fft_impulse_and_constant(oracle):
    n = 64
    d = np.zeros(n, complex)
    d[0] = 1
    assert np.allclose(oracle.fft(d), np.ones(n), atol=1e-15)
    assert np.allclose(oracle.fft(np.ones(n)), n * d, atol=1e-12)


@pytest.mark.parametrize("n", [8, 64, 100])
def test_correlation_convolution_brute_force(oracle, n):
    rng = np.random.default_rng(7 + n)
    s = rng.normal(size=n) + 1j * rng.normal(size=n)
    r = rng.normal(size=n) + 1j * rng.normal(size=n)
    P = 1 << (n - 1).bit_length()
    sp = np.concatenate([s, np.zeros(P - n)])
    rp = np.concatenate([r, np.zeros(P - n)])
    # circular correlation c[k] = sum_m s[m+k] conj(r[m]); convolution v[k] = sum_m s[m] r[k-m]
    corr = np.array([np.sum(np.roll(sp, -k) * np.conj(rp)) for k in range(P)])[:n]
    conv = np.array([np.sum(sp * np.roll(rp[::-1], k + 1)) for k in range(P)])[:n]
    assert np.allclose(oracle.fft_correlation(s, r), corr, atol=1e-11)
    assert np.allclose(oracle.fft_convolution(s, r), conv, atol=1e-11)


def test_spline_invariants(oracle):
    rng = np.random.default_rng(3)
    n, f = 50, 4
    d = rng.normal(size=n) + 1j * rng.normal(size=n)
    y = oracle.spline(d, f)
    assert len(y) == n * f
    assert np.array_equal(y[: (n - 1) * f: f], d[: n - 1])          # nodes reproduce exactly (h1 = 1)
    assert np.all(y[(n - 1) * f:] == d[n - 1])                       # tail clamp to the last sample
    c = oracle.spline(np.full(n, 2.5 - 1j), f)
    assert np.allclose(c, 2.5 - 1j, atol=1e-15)                       # constants stay constant
    ramp = 3.0 + 0.5 * np.arange(n) + 1j * (1 - 0.25 * np.arange(n))
    yr = oracle.spline(ramp, f)
    i = np.arange(f, (n - 2) * f)                                      # interior segments
    assert np.allclose(yr[i], 3.0 + 0.5 * i / f + 1j * (1 - 0.25 * i / f), atol=1e-13)


def test_dem_closed_forms(oracle):
    x = np.array([-30.0, 0.0, 30.0, 60.0])
    y = np.array([-20.0, 10.0, 40.0])
    X, Y = np.meshgrid(x, y)
    Z = 0.3 * X - 0.7 * Y + 2.0
    dem = (x, y, Z)
    rng = np.random.default_rng(1)
    pts = np.stack([rng.uniform(-30, 60, 500), rng.uniform(-20, 40, 500)], 1)
    z = oracle.dem_interp(dem, pts)
    assert np.allclose(z, 0.3 * pts[:, 0] - 0.7 * pts[:, 1] + 2.0, atol=1e-12)    # plane is exact inside
    nodes = np.array([[xi, yj] for yj in y for xi in x])
    assert np.allclose(oracle.dem_interp(dem, nodes), Z.ravel(), atol=1e-12)
    # x outside the axis: column 0 is used whether x is below or above it
    col0 = np.interp([0.0, 25.0], y, Z[:, 0])
    assert np.allclose(oracle.dem_interp(dem, [[-100.0, 0.0], [-100.0, 25.0]]), col0, atol=1e-12)
    assert np.allclose(oracle.dem_interp(dem, [[500.0, 0.0], [500.0, 25.0]]), col0, atol=1e-12)
    # both outside -> Z[0][0]
    assert oracle.dem_interp(dem, [[-100.0, -100.0]])[0] == Z[0, 0]
    # empty DEM -> 0
    e = np.zeros(0)
    assert oracle.dem_interp((e, e, np.zeros((0, 0))), [[1.0, 2.0]])[0] == 0.0


def snell_point(s, e, n1, n2):
    f = lambda p: n1 * np.linalg.norm([p[0] - s[0], p[1] - s[1], -s[2]]) + n2 * np.linalg.norm(
        [e[0] - p[0], e[1] - p[1], e[2]])
    r = minimize(f, x0=[(s[0] + e[0]) / 2, (s[1] + e[1]) / 2], method="Nelder-Mead",
                 options=dict(xatol=1e-10, fatol=1e-12, maxiter=4000))
    return r.x, r.fun


@pytest.mark.parametrize("n1,n2", [(1.0, 2.449489742783178), (2.449489742783178, 1.0), (1.0, 1.0)])
def test_refraction_flat_interface_snell(oracle, n1, n2):
    ax = -495.0 + 30 * np.arange(34)
    dem = (ax, ax, np.zeros((34, 34)))
    rng = np.random.default_rng(11)
    for _ in range(20):
        s = np.array([rng.uniform(-15, 15), rng.uniform(-15, 15), 12.0])
        e = np.array([rng.uniform(19, 21), rng.uniform(-1, 1), rng.uniform(-6, -4)])
        q = oracle.refraction_points(dem, np.concatenate([s, e])[None], n1, n2)[0]
        p, fmin = snell_point(s, e, n1, n2)
        assert q[2] == 0.0
        # final level spacing is 2*3.7/10 = 0.74 m: the grid argmin is within one cell
        assert np.hypot(q[0] - p[0], q[1] - p[1]) < 0.75
        fq = n1 * np.linalg.norm(q - s) + n2 * np.linalg.norm(e - q)
        # at most half a diagonal cell (0.52 m) from the optimum: the path excess
        # is below 1/2 * (n1/d1 + n2/d2) * 0.52^2 < 0.15 m for these depths
        assert fmin <= fq + 1e-9 and fq - fmin < 0.15
        if n1 == n2:   # straight line: the crossing of z = 0 on segment s-e
            t = s[2] / (s[2] - e[2])
            cross = s[:2] + t * (e[:2] - s[:2])
            assert np.hypot(*(q[:2] - cross)) < 0.75


def _tiny(oracle, n=6, S=64, seed=5):
    rng = np.random.default_rng(seed)
    ax = -105.0 + 30 * np.arange(8)
    X, Y = np.meshgrid(ax, ax)
    dem = (ax, ax, 0.01 * X - 0.02 * Y + 0.5)
    t = np.arange(n) / max(n - 1, 1)
    tx = np.stack([(5 + 10 * t) * np.cos(4 * np.pi * t), (5 + 10 * t) * np.sin(4 * np.pi * t), np.full(n, 12.0)], 1)
    rx = np.stack([(5 + 10 * t) * np.cos(4 * np.pi * t + 1.5), (5 + 10 * t) * np.sin(4 * np.pi * t + 1.5),
                   np.full(n, 12.0)], 1)
    gx, gy, gz = 19.9 + 0.1 * np.arange(3), -0.1 + 0.1 * np.arange(2), -5.0 + 0.2 * np.arange(2)
    bp = dict(c=3e8, fs4=3.2e9, lam=0.3, n1=1.0, n2=2.449489742783178)
    return rng, dem, tx, rx, gx, gy, gz, bp


def test_backprojection_linear_in_data(oracle):
    rng, dem, tx, rx, gx, gy, gz, bp = _tiny(oracle)
    S4 = 1024
    d1 = rng.normal(size=(6, S4)) + 1j * rng.normal(size=(6, S4))
    d2 = rng.normal(size=(6, S4)) + 1j * rng.normal(size=(6, S4))
    a, b = 0.75 - 0.5j, -1.25 + 2j
    i1 = oracle.backproject(bp, dem, tx, rx, d1, gx, gy, gz)
    i2 = oracle.backproject(bp, dem, tx, rx, d2, gx, gy, gz)
    i12 = oracle.backproject(bp, dem, tx, rx, a * d1 + b * d2, gx, gy, gz)
    assert np.allclose(i12, a * i1 + b * i2, atol=1e-12 * np.abs(i12).max())


def test_backprojection_single_pulse_closed_form(oracle):
    """N = 1: R1 is the distance to the one antenna, so the value is the
    interpolated sample times exp(i*2*pi/lambda*path) with path built from
    the two refraction points."""
    rng, dem, tx, rx, gx, gy, gz, bp = _tiny(oracle, n=1)
    tx, rx = np.array([[4.0, -3.0, 12.0]]), np.array([[-2.0, 6.0, 11.0]])
    S4 = 2048
    ramp = (0.3 + 0.01 * np.arange(S4)) * np.exp(1j * 0.2)       # lerp of a ramp is exact-ish
    img = oracle.backproject(bp, dem, tx, rx, ramp[None, :], gx, gy, gz)
    for k, z in enumerate(gz):
        for j, y in enumerate(gy):
            for i, x in enumerate(gx):
                P = np.array([x, y, z])
                ida, volta, _ = oracle.slant_range_fermat(dem, tx, rx, P, bp["n1"], bp["n2"])
                qa, qb = ida[0], volta[0]
                path = (bp["n1"] * np.linalg.norm(tx[0] - qa) + bp["n2"] * np.linalg.norm(P - qa)
                        + bp["n2"] * np.linalg.norm(P - qb) + bp["n1"] * np.linalg.norm(rx[0] - qb))
                r = path / bp["c"] * bp["fs4"]
                want = (0.3 + 0.01 * r) * np.exp(1j * 0.2) * np.exp(1j * 2 * np.pi / bp["lam"] * path)
                assert abs(img[k, j, i] - want) < 1e-9 * abs(want)


def test_backprojection_out_of_window_is_skipped(oracle):
    rng, dem, tx, rx, gx, gy, gz, bp = _tiny(oracle)
    short = np.ones((6, 8), complex)      # every delay lands beyond sample 8
    assert np.all(oracle.backproject(bp, dem, tx, rx, short, gx, gy, gz) == 0)


def test_flops_convention(oracle):
    assert oracle.flops_per_voxel_pulse(256) == 70526.0
    assert oracle.flops_per_voxel_pulse(2048) == 106366.0
