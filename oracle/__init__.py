"""CPU FP64 oracle for the Kohn-Sham Hpsi / density / adjoint path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy/SciPy restatement of the
reference algorithm (dyzheng/ad-dft, `src/addft/...`).  It exists to check the
HIP path; it is never the thing measured or shipped.  Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import
it.  The product package (`addft`) must not import anything from here.

Every function cites the reference file:line it restates.  The reference
itself is JAX code and cannot be imported in this image (no jax), so the
oracle is pinned by the reference's own known-answer tests, closed forms,
invariants and finite differences (see tests/test_oracle_*.py).

Parity status per module:
  basis      pinned: published sphere/box counts, smooth-235 KATs, exact
             pack round trips, Hartree plane-wave closed form.
  psp        pinned: Simpson/j_l quadrature vs analytic Gaussian transforms.
  physics    pinned: Ylm KATs, XC pointwise closed form + FD, Ewald
             eta-independence / translation invariance / brute force,
             occupation electron count.
  solver     pinned: matrix-free Hpsi == dense H, Davidson == dense eigh,
             energy forms agree (I-7), SCF residual.
  forces     pinned: central FD with re-SCF, error ~ h^2, sum rule.
  response   pinned: I-15 (response == FD of scf_map), transpose identity,
             dQ adjoint vs FD with re-SCF.
"""
