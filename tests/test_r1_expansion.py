"""The plain FP64 back-projection expands each squared antenna distance of its
R1 sums about a local origin (bsar_bp.h, r1_mean_fp64) when the point is well
separated from the antennas: a = c - 2 b.(q - o) + |q - o|^2 in one add and
three fmas.  Its error bound (11 eps (bsq + |q - o|^2) <= 352 eps a under the
check bsq + |q - o|^2 <= 32 gap^2) is checked here by running the same
operations on the CPU (tests/r1_expansion_model.c) against the exact squared
distance of the original coordinates, over the workloads' geometry, wide
apertures and UTM-sized coordinates.  Expanding without the check must fail
the same bound at UTM-sized coordinates."""
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")

BOUND_EPS = 352.0


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    exe = tmp_path_factory.mktemp("r1x") / "r1_expansion_model"
    r = subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(ROOT / "tests" / "r1_expansion_model.c"), "-o",
                        str(exe), "-lm"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return exe


def run(exe, family, n, nocheck=0):
    out = subprocess.run([str(exe), str(family), str(n), str(nocheck)], capture_output=True, text=True,
                         timeout=300).stdout
    m = re.match(r"worst_eps (\S+) pairs (\d+) of (\d+)", out)
    assert m, out
    return float(m.group(1)), int(m.group(2)), int(m.group(3))


@pytest.mark.parametrize("family", [0, 1, 2])
def test_expanded_distances_within_bound(model, family):
    worst, used, total = run(model, family, 1_000_000)
    assert used > total // 2          # the check admits most pairs of every geometry
    assert worst <= BOUND_EPS


def test_workload_geometry_always_expands(model):
    # antennas 12 m above points a few metres below ground: the fast form runs
    _, used, total = run(model, 0, 200_000)
    assert used >= 0.99 * total


def test_unchecked_expansion_is_caught(model):
    worst, _, _ = run(model, 2, 200_000, nocheck=1)
    assert worst > BOUND_EPS
