"""Shared test set-up: the `gpu` marker, the oracle import path and a loader for the
bionmf-gpu_amd package (its directory name has a hyphen, so it is loaded by path)."""
from __future__ import annotations

import importlib.util
import os
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")


def load_package():
    name = "bionmf_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, ROOT / "bionmf-gpu_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def nmfk(pkg):
    """The loaded product library on a machine with an MI355X; the HIP path must be the one that runs.

    NMFK_TEST_TUNE="key=value,..." (planner tunables, nmfk_tune_set keys) runs the session
    under other launch shapes, e.g. "4=2,5=2" for 2 x 2 matrix-core workgroups.  Batching
    takes the default shapes only, so such sessions run the single-job test files."""
    if pkg.gpu_count() < 1:
        pytest.skip("no HIP device visible")
    pkg.load()  # raises if libnmfgpu_amd.so is missing: no silent fallback
    for item in filter(None, os.environ.get("NMFK_TEST_TUNE", "").split(",")):
        key, value = (int(x) for x in item.split("="))
        pkg.tune_set(key, value)
    return pkg


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
