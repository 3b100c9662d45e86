"""Shared test setup.

Markers: `gpu` (needs a MI355X and the built libaddft_hip.so), `slow`.
The oracle (oracle/) is the checker; the product package is `addft`
(under ad-dft_amd/).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "ad-dft_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

LATTICE = np.eye(3) * 6.0
FRAC_DISTORTED = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])
FRAC_SYMMETRIC = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]])
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def oracle_small_system(psp=None, frac=FRAC_DISTORTED, ecut=2.0, nband=2, factor=4.0, sigma=0.05,
                        kfrac=None, kweights=None):
    """The reference conftest system (tests/conftest.py:26-35) in the oracle."""
    from oracle import psp as OP
    from oracle import solver as OS
    psp = OP.synthetic() if psp is None else psp
    return OS.setup(LATTICE, frac, psp, ecut, nband, kfrac=kfrac, kweights=kweights,
                    sigma=sigma, density_factor=factor)


@pytest.fixture(scope="session")
def small():
    from oracle import solver as OS
    geo, st = oracle_small_system()
    state = OS.solve_scf_dense(geo, st, tol=1e-12)
    return geo, st, state
