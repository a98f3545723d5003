"""Shared fixtures: library handles, parameter helpers, the oracle (checker).

The oracle (oracle/liboracle.so) is the FP64 CPU restatement of the reference
algorithm; tests use it only as the checker of the HIP library.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "dbec-gp-mpi-cuda_amd")
sys.path.insert(0, os.path.join(PKG, "python"))

import dbecpy  # noqa: E402

ORACLE_LIB = os.path.join(REPO, "oracle", "liboracle.so")
ORACLE_BIN = os.path.join(REPO, "oracle", "_bin")
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long CPU run (golden reference case)")


def _ensure_oracle():
    if not (os.path.exists(ORACLE_LIB) and os.path.exists(os.path.join(ORACLE_BIN, "imag3d"))):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)


@pytest.fixture(scope="session")
def oracle_lib():
    _ensure_oracle()
    return dbecpy.load(ORACLE_LIB)


@pytest.fixture(scope="session")
def oracle_bin():
    _ensure_oracle()
    return ORACLE_BIN


@pytest.fixture(scope="session")
def hip_lib():
    """The product library.  Fails loudly when it is not built."""
    if not os.path.exists(dbecpy.HIP_LIB):
        pytest.fail(f"{dbecpy.HIP_LIB} is not built; run __graft_entry__.build()")
    return dbecpy.load(dbecpy.HIP_LIB)


def params(n=(32, 48, 64), h=(0.25, 0.2, 0.3), dt=0.002, trap=(0.5, 1.0, 1.5), cutoff=6.0, par=2.0):
    return dbecpy.make_params(n[0], n[1], n[2], h[0], h[1], h[2], dt, trap[0], trap[1], trap[2], cutoff, par)


def smooth_field(p, seed=0, complex_=False):
    """A smooth, deterministic, non-symmetric test wave function."""
    rng = np.random.default_rng(seed)
    x, y, z = dbecpy.grid(p)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    a, b, c = rng.uniform(0.4, 0.9, 3)
    sx, sy, sz = rng.uniform(-0.5, 0.5, 3)
    f = np.exp(-(a * (X - sx) ** 2 + b * (Y - sy) ** 2 + c * (Z - sz) ** 2) / 2) * (1 + 0.3 * np.sin(X + 2 * Y - Z))
    if complex_:
        f = f * np.exp(1j * (0.7 * X - 0.4 * Y + 0.25 * Z * Z))
    return f


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
