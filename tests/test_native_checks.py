"""Host-side shape checks of the tensor wrappers (CPU): a mis-sized buffer
must raise ValueError before any launch, never reach a kernel as an
out-of-bounds pointer.  The plans are built without the device (only the
sizes the checks read are set), and every tensor is a CPU tensor: the size
check comes before the residence check, so no GPU call is made."""

import pytest
import torch

from addft.kernels import native


def _plan(shape=(8, 8, 8), ng=33, cdtype=torch.complex128):
    p = object.__new__(native.SpherePlan)
    p.shape = shape
    p.nfft = shape[0] * shape[1] * shape[2]
    p.ng = ng
    p.cdtype = cdtype
    p.rdtype = torch.float64 if cdtype == torch.complex128 else torch.float32
    p._h = None
    p._work = None
    return p


def _box(shape=(8, 8, 8), ng=33):
    b = object.__new__(native.BoxSphere)
    b.shape = shape
    b.nfft = shape[0] * shape[1] * shape[2]
    b.ng = ng
    return b


def test_sphere_plan_wrappers_reject_wrong_sizes():
    p = _plan()
    with pytest.raises(ValueError, match="v_planar"):
        p.stage(torch.zeros(p.nfft - 1, dtype=torch.float64), 1.0)
    with pytest.raises(ValueError, match="c has"):
        p.to_real(torch.zeros(p.ng + 1, dtype=torch.complex128))
    with pytest.raises(ValueError, match="grid_planar"):
        p.to_sphere(torch.zeros(p.nfft + 8, dtype=torch.float64))
    psi = torch.zeros((2, p.ng), dtype=torch.complex128)
    w = torch.zeros(2, dtype=torch.float64)
    with pytest.raises(ValueError, match="out has"):
        p.density(psi, w, out=torch.zeros(p.nfft // 2, dtype=torch.float64))
    with pytest.raises(TypeError):
        p.density(psi, w, out=torch.zeros(p.nfft, dtype=torch.float32))
    with pytest.raises(ValueError, match="shape mismatch"):
        p.density(psi, torch.zeros(3, dtype=torch.float64))
    tk = torch.zeros(p.ng, dtype=torch.float64)
    vs = torch.zeros(p.nfft, dtype=torch.float64)
    with pytest.raises(ValueError, match="out has"):
        p.hpsi_local(psi, tk, vs, out=torch.zeros((1, p.ng), dtype=torch.complex128))
    with pytest.raises(ValueError, match="shape mismatch"):
        p.hpsi_local(psi, tk[:-1].contiguous(), vs)


def test_box_sphere_and_grid_wrappers_reject_wrong_sizes():
    b = _box()
    with pytest.raises(ValueError, match="c has"):
        b.to_real(torch.zeros(b.ng - 1, dtype=torch.complex128))
    with pytest.raises(ValueError, match="grid_planar"):
        b.to_sphere(torch.zeros(b.nfft - 1, dtype=torch.float64))
    with pytest.raises(ValueError, match="grid has"):
        native.to_planar(torch.zeros(100, dtype=torch.float64), (4, 5, 6))
    with pytest.raises(ValueError, match="grid has"):
        native.from_planar(torch.zeros(121, dtype=torch.float64), (4, 5, 6))
    with pytest.raises(ValueError, match="w has"):
        native.xc_fxc(torch.zeros(10, dtype=torch.float64), torch.zeros(9, dtype=torch.float64))
    with pytest.raises(ValueError, match="pair tables"):
        native.gamma_unpack2(torch.zeros((1, 9), dtype=torch.complex128), 2,
                             torch.zeros(4, dtype=torch.int32), torch.zeros(4, dtype=torch.int32))
    with pytest.raises(ValueError, match="minus has"):
        native.gamma_unpack2(torch.zeros((1, 9), dtype=torch.complex128), 2,
                             torch.zeros(5, dtype=torch.int32), torch.zeros(4, dtype=torch.int32))
