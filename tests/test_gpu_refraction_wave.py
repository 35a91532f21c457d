"""The unscreened wavefront-reduction refraction search (one wave per leg,
lanes over every candidate, argmin with the lowest reference index among
ties) against the reference's golden vectors, the oracle and the screened
search: bit-identical refraction points on every DEM staging path (whole DEM
in LDS, split, read through the caches)."""
import numpy as np
import pytest

from test_oracle_golden import DEMS, bits_equal

pytestmark = pytest.mark.gpu

N2 = 2.449489742783178


@pytest.mark.parametrize("name", DEMS)
@pytest.mark.parametrize("tag", ["a", "b", "eq"])
def test_wave_refraction_golden(bsar, ctx, golden, name, tag):
    g = golden
    dem = bsar.Dem(g[f"dem_{name}_x"], g[f"dem_{name}_y"], g[f"dem_{name}_z"])
    n1, n2 = g[f"refr_{name}_{tag}_n"]
    out, secs = ctx.refraction_points_wave(dem, g[f"refr_{name}_{tag}_legs"], n1, n2)
    assert bits_equal(out, g[f"refr_{name}_{tag}_out"])
    assert secs > 0


def test_wave_refraction_empty_dem(bsar, ctx, golden):
    e = np.zeros(0)
    out, _ = ctx.refraction_points_wave(bsar.Dem(e, e, np.zeros((0, 0))), golden["refr_empty_legs"], 1.0, 2.0)
    assert bits_equal(out, golden["refr_empty_out"])


def _legs(rng, M, z0=0.0):
    top = np.stack([rng.uniform(-20, 20, M), rng.uniform(-20, 20, M), z0 + rng.uniform(10, 14, M)], 1)
    bot = np.stack([rng.uniform(19, 21, M), rng.uniform(-1, 1, M), z0 - rng.uniform(4, 6, M)], 1)
    legs = np.concatenate([top, bot], 1)
    legs[M // 2:] = np.concatenate([bot, top], 1)[M // 2:]
    return legs


def test_wave_refraction_random_legs(bsar, ctx, oracle, scene_mod):
    rng = np.random.default_rng(43)
    dem = scene_mod.dem_bumps(rng)
    legs = _legs(rng, 3000)
    wave, _ = ctx.refraction_points_wave(bsar.Dem(*dem), legs, 1.0, N2)
    assert bits_equal(wave, oracle.refraction_points(dem, legs, 1.0, N2))
    assert bits_equal(wave, ctx.refraction_points(bsar.Dem(*dem), legs, 1.0, N2))


def test_wave_refraction_rough_terrain(bsar, ctx, oracle, scene_mod):
    """167 x 100 fBm terrain (the split staging path)."""
    rng = np.random.default_rng(1004)
    dem = scene_mod.dem_fbm(rng)
    M = 1500
    top = np.stack([rng.uniform(-15, 15, M), rng.uniform(-15, 15, M), rng.uniform(40, 60, M)], 1)
    bot = np.stack([rng.uniform(-2, 2, M), rng.uniform(-2, 2, M), rng.uniform(-30, -10, M)], 1)
    legs = np.concatenate([top, bot], 1)
    n2 = float(np.sqrt(rng.uniform(4, 9)))
    wave, _ = ctx.refraction_points_wave(bsar.Dem(*dem), legs, 1.0, n2)
    assert bits_equal(wave, oracle.refraction_points(dem, legs, 1.0, n2))


def test_wave_refraction_dem_above_lds(bsar, ctx, oracle, scene_mod):
    """200 x 160 fBm terrain, above both LDS staging sizes (cached reads)."""
    dem = scene_mod.dem_fbm(np.random.default_rng(2025), nx=200, ny=160, rms=25.0)
    nx, ny = len(dem[0]), len(dem[1])
    assert nx + ny + 2 * ny * (nx + 1) > 8192 and nx + ny + (ny * (nx + 1) + 1) // 2 > 9728
    z0 = float(oracle.dem_interp(dem, [[20.0, 0.0]])[0])
    legs = _legs(np.random.default_rng(78), 1000, z0)
    wave, _ = ctx.refraction_points_wave(bsar.Dem(*dem), legs, 1.0, N2)
    assert bits_equal(wave, ctx.refraction_points(bsar.Dem(*dem), legs, 1.0, N2))
    assert bits_equal(wave[:200], oracle.refraction_points(dem, legs[:200], 1.0, N2))


def test_wave_refraction_no_legs(bsar, ctx, scene_mod):
    dem = scene_mod.dem_bumps(np.random.default_rng(1))
    out, secs = ctx.refraction_points_wave(bsar.Dem(*dem), np.zeros((0, 6)), 1.0, N2)
    assert out.shape == (0, 3) and secs == 0.0
