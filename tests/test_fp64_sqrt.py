"""The plain FP64 back-projection sums its R1 means with a square root built
from v_rsq_f64, one Goldschmidt step and one Newton step (bsar_bp.h,
sqrt_faithful).  Its accuracy rests on the seed's documented error, so the
same steps run here on the CPU with the seed modelled as the exact
reciprocal root times (1 + e), |e| <= 2^-21 (4x AMD's 2^-23), against the
exact root: the result must stay within 0.51 ulp, i.e. faithfully rounded.
Dropping the Newton step must fail the same check."""
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    exe = tmp_path_factory.mktemp("sqrt") / "sqrt_faithful_model"
    r = subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(ROOT / "tests" / "sqrt_faithful_model.c"), "-o",
                        str(exe), "-lm"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return exe


def run(exe, n, drop=0):
    out = subprocess.run([str(exe), str(n), str(drop)], capture_output=True, text=True, timeout=300).stdout
    m = re.match(r"worst_ulp (\S+) over (\d+) (\d+)", out)
    assert m, out
    return float(m.group(1)), int(m.group(3))


def test_faithful_within_half_ulp(model):
    worst, bad = run(model, 2_000_000)
    assert bad == 0
    assert worst < 0.51


def test_dropped_newton_step_is_caught(model):
    worst, bad = run(model, 200_000, drop=1)
    assert bad > 0 and worst > 1.0
