"""Shared fixtures: the product package (as ``pacg``), the oracle (checker only)
and the golden fixtures written from the reference's own library."""
from __future__ import annotations

import glob
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "c-permutation-avoiding-convolution_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def load_pacg():
    if "pacg" in sys.modules:
        return sys.modules["pacg"]
    spec = importlib.util.spec_from_file_location("pacg", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pacg"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pacg():
    return load_pacg()


@pytest.fixture(scope="session")
def synth():
    load_pacg()
    import pacg.synth as s  # noqa: E402

    return s


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle  # noqa: E402

    return Oracle()


@pytest.fixture(scope="session")
def gpu(pacg):
    if pacg.device_count() < 1:
        pytest.skip("no GPU visible")
    return pacg


def golden_cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))
