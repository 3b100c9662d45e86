"""Krylov solvers of the adjoint equation (addft/response.py) on the CPU:
restarted GMRES and BiCGSTAB against a dense solve, with and without a right
preconditioner, and with a restart length shorter than the iteration count."""

import numpy as np
import pytest
import torch

from addft.response import bicgstab, gmres


def _system(n=120, seed=3):
    rng = np.random.default_rng(seed)
    # nonsymmetric, eigenvalues in the right half-plane, spread like I - K chi0
    A = np.diag(np.linspace(1.0, 30.0, n)) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    return A, b, np.linalg.solve(A, b)


@pytest.mark.parametrize("restart", [200, 15])
@pytest.mark.parametrize("pre", [False, True])
def test_gmres_matches_dense_solve(restart, pre):
    A, b, x_ref = _system()
    At = torch.from_numpy(A)
    dinv = torch.from_numpy(1.0 / np.diag(A))
    x, it = gmres(lambda v: At @ v, torch.from_numpy(b), rtol=1e-12, maxiter=400, restart=restart,
                  precond=(lambda v: dinv * v) if pre else None)
    assert np.abs(x.numpy() - x_ref).max() < 1e-9 * np.abs(x_ref).max()
    assert it < 400


def test_gmres_preconditioner_reduces_products():
    A, b, _ = _system()
    At = torch.from_numpy(A)
    dinv = torch.from_numpy(1.0 / np.diag(A))
    _, it0 = gmres(lambda v: At @ v, torch.from_numpy(b), rtol=1e-10, maxiter=400)
    _, it1 = gmres(lambda v: At @ v, torch.from_numpy(b), rtol=1e-10, maxiter=400, precond=lambda v: dinv * v)
    assert it1 < it0


@pytest.mark.parametrize("pre", [False, True])
def test_bicgstab_matches_dense_solve(pre):
    A, b, x_ref = _system()
    At = torch.from_numpy(A)
    dinv = torch.from_numpy(1.0 / np.diag(A))
    x, _ = bicgstab(lambda v: At @ v, torch.from_numpy(b), rtol=1e-12, maxiter=400,
                    precond=(lambda v: dinv * v) if pre else None)
    assert np.abs(x.numpy() - x_ref).max() < 1e-9 * np.abs(x_ref).max()


def test_gmres_zero_rhs():
    A, _, _ = _system(n=10)
    At = torch.from_numpy(A)
    x, it = gmres(lambda v: At @ v, torch.zeros(10, dtype=torch.float64))
    assert it == 0 and float(x.abs().max()) == 0.0
