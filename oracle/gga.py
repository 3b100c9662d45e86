"""Oracle: the model GGA and the spectral gradient / divergence.

Test infrastructure only.  Restates in NumPy FP64 reference
physics/xc.py:78-98 (grad_r, div_r: spectral derivatives on the psi-sphere,
n(G) = fftn(n)/N gathered there and v(r) = Re N ifftn) and xc.py:137-158
(the model GGA f(n, sigma) = n eps_xc^LDA(n) (1 + 0.1234 s^2), s^2 =
sigma / (4 k_F^2 n^2), k_F = (3 pi^2 n)^(1/3), sigma = |grad n|^2, and
E_xc = (Omega/N) sum f, whose potential the reference obtains by automatic
differentiation of E_xc times N/Omega, xc.py:159-172).  Here the potential
is written out, v = f_n - div(2 f_sigma grad n), and the strain derivative of
E_xc at fixed psi (rho = rho^/Omega on the fixed grid, gradients of the
strained reciprocal vectors).  Pinned by tests/test_oracle_gga.py: the
adjoint identity <grad f, v> = -<f, div v> (the reference's test I-9,
tests/test_primitives.py:102-111), the gradient of a plane wave, the
potential and the strain derivative against central differences of E_xc,
and mu = 0 reducing to LDA.
"""

from __future__ import annotations

import numpy as np

from . import basis as B
from . import physics as P

MU_MODEL = 0.1234


def grad_r(basis, lattice, f_r):
    """(3, nfft) spectral gradient on the psi-sphere (xc.py:78-85)."""
    fg = B.density_g(basis, f_r)
    g = B.gvectors(basis.miller, lattice)
    return np.stack([B.real_from_g(basis, 1j * g[:, a] * fg) for a in range(3)])


def div_r(basis, lattice, v_r):
    """Spectral divergence of a (3, nfft) field (xc.py:88-93)."""
    g = B.gvectors(basis.miller, lattice)
    total = sum(1j * g[:, a] * B.density_g(basis, v_r[a]) for a in range(3))
    return B.real_from_g(basis, total)


def _s2(nn, sigma):
    kf = np.cbrt(3.0 * np.pi ** 2 * nn)
    return np.maximum(sigma, 0.0) / (4.0 * kf * kf * nn * nn)


def energy_density(n, sigma, mu=MU_MODEL):
    """f(n, sigma) with n floored at 1e-14 (xc.py:137-145)."""
    nn = np.maximum(n, P.RHO_FLOOR)
    return P.lda_energy_density(nn) * (1.0 + mu * _s2(nn, sigma))


def partials(n, sigma, mu=MU_MODEL):
    """(df/dn, df/dsigma) at the floored density."""
    nn = np.maximum(n, P.RHO_FLOOR)
    e = P.lda_energy_density(nn)
    s2 = _s2(nn, sigma)
    kf = np.cbrt(3.0 * np.pi ** 2 * nn)
    fn = P.xc_potential(nn) * (1.0 + mu * s2) - (8.0 / 3.0) * mu * e * s2 / nn
    fs = mu * e / (4.0 * kf * kf * nn * nn)
    return fn, fs


def xc_energy(basis, lattice, rho_r, mu=MU_MODEL):
    """E_xc = (Omega/N) sum f (xc.py:148-152)."""
    g = grad_r(basis, lattice, rho_r)
    sigma = np.sum(g * g, axis=0)
    return B.volume(lattice) / basis.nfft * float(np.sum(energy_density(rho_r, sigma, mu)))


def xc_potential(basis, lattice, rho_r, mu=MU_MODEL):
    """v = dE_xc/dn_j N/Omega = f_n - div(2 f_sigma grad n)."""
    g = grad_r(basis, lattice, rho_r)
    sigma = np.sum(g * g, axis=0)
    fn, fs = partials(rho_r, sigma, mu)
    return fn - div_r(basis, lattice, 2.0 * fs[None, :] * g)


def xc_energy_split(basis, lattice, rho_r, mu=MU_MODEL):
    """(exchange, correlation) parts n eps_x (1 + mu s^2), n eps_c (1 + mu s^2)."""
    g = grad_r(basis, lattice, rho_r)
    sigma = np.sum(g * g, axis=0)
    nn = np.maximum(rho_r, P.RHO_FLOOR)
    fac = 1.0 + mu * _s2(nn, sigma)
    w = B.volume(lattice) / basis.nfft
    return w * float(np.sum(nn * P.eps_x(nn) * fac)), w * float(np.sum(nn * P.eps_c(nn) * fac))


def strain_derivative(basis, lattice, rho_r, mu=MU_MODEL):
    """dE_xc/deps (3, 3) with psi frozen: rho = rho^/Omega on the fixed grid
    and the spectral gradient built on the strained reciprocal vectors,
      dE/deps_ab = delta_ab [E_xc - (Omega/N) sum (n f_n + 2 sigma f_sigma)]
                   - (Omega/N) sum 2 f_sigma d_a n d_b n."""
    g = grad_r(basis, lattice, rho_r)
    sigma = np.sum(g * g, axis=0)
    fn, fs = partials(rho_r, sigma, mu)
    w = B.volume(lattice) / basis.nfft
    e = w * float(np.sum(energy_density(rho_r, sigma, mu)))
    iso = e - w * float(np.sum(rho_r * fn + 2.0 * sigma * fs))
    return np.eye(3) * iso - w * np.einsum("j,aj,bj->ab", 2.0 * fs, g, g)
