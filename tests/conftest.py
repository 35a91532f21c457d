import importlib.util
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "c-_bistatic_sar_amd"
GOLDEN = Path(__file__).resolve().parent / "golden"

sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


def load_pkg():
    if "bsar_amd" in sys.modules:
        return sys.modules["bsar_amd"]
    spec = importlib.util.spec_from_file_location("bsar_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bsar_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_scene():
    if "bsar_amd.scene" in sys.modules:
        return sys.modules["bsar_amd.scene"]
    load_pkg()
    spec = importlib.util.spec_from_file_location("bsar_amd.scene", PKG_DIR / "scene.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bsar_amd.scene"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def bsar():
    return load_pkg()


@pytest.fixture(scope="session")
def scene_mod():
    return load_scene()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN / "ref_kernels.npz")


@pytest.fixture(scope="session")
def ctx(bsar):
    c = bsar.Context(0)
    yield c
    c.close()
