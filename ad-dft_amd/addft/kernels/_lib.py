"""ctypes binding of libaddft_hip.so (the C ABI declared in include/addft_hip.h).

The library is looked up at $ADDFT_NATIVE_LIB or next to this package
(addft/_lib/libaddft_hip.so), mirroring the reference loader
(src/addft/kernels/pack.py:51-55).  Unlike the reference there is no silent
fallback: if the library is missing every native entry point raises
NativeUnavailable, so a GPU run can never pass on a substitute path.
Tensors are torch device tensors; pointers and the current HIP stream are
passed through as plain integers.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

_LIB = None
_PATH = None

C64, C128, R64 = 0, 1, 2
EIGH_DC, EIGH_JACOBI, EIGH_DCJ, EIGH_TRD, EIGH_LDS = 0, 1, 2, 3, 4
PLANE_LDS, PLANE_LANE = 0, 1
OK, EINVAL, EINTERNAL, EUNSUPPORTED = 0, -1, -2, -3


class NativeUnavailable(RuntimeError):
    pass


class AddftError(RuntimeError):
    pass


def lib_path() -> Path:
    env = os.environ.get("ADDFT_NATIVE_LIB")
    if env:
        return Path(env)
    return Path(__file__).resolve().parents[1] / "_lib" / "libaddft_hip.so"


_VP, _I64, _I32, _D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double

_SIGS = {
    "addft_last_error": ([], ctypes.c_char_p),
    "addft_version": ([], _I32),
    "addft_pack_scatter": ([_I32, _VP, _VP, _I64, _I64, _I64, _VP, _VP], _I32),
    "addft_split_bf16x3": ([_VP, _I64, _I64, _I64, _VP, _VP], _I32),
    "addft_cheb_combine": ([_I32, _I64, _I64, _VP, _D, _D, _D, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_gamma_pair_split": ([_VP, _I64, _I64, _VP, _VP, _I64, _I64, _VP, _VP], _I32),
    "addft_gamma_pair_rows": ([_VP, _I64, _I64, _VP, _VP, _I64, _VP, _VP], _I32),
    "addft_gamma_scatter_add": ([_I32, _VP, _I64, _I64, _VP, _VP, _I64, _VP, _VP], _I32),
    "addft_gamma_scatter_add_real": ([_I32, _VP, _I64, _I64, _VP, _VP, _I64, _VP, _VP], _I32),
    "addft_gamma_pack2": ([_I32, _VP, _I64, _I64, _VP, _VP], _I32),
    "addft_gamma_unpack2": ([_I32, _VP, _I64, _I64, _VP, _VP, _I64, _VP, _VP], _I32),
    "addft_gamma_pack2_half": ([_I32, _VP, _I64, _I64, _I64, _VP, _VP, _I64, _VP, _VP], _I32),
    "addft_gamma_unpack2_half": ([_I32, _VP, _I64, _I64, _VP, _VP, _I64, _VP, _I64, _VP], _I32),
    "addft_pack_gather": ([_I32, _VP, _VP, _I64, _I64, _I64, _VP, _VP], _I32),
    "addft_xc_work_bytes": ([_I64], _I64),
    "addft_xc_lda": ([_VP, _I64, _VP, _VP, _VP, _VP], _I32),
    "addft_xc_lda_fxc": ([_VP, _VP, _I64, _VP, _VP], _I32),
    "addft_cg_shift_dot": ([_I64, _I64, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_cg_shifted_rows": ([_I64, _I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_cg_update": ([_I64, _I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_cg_direction": ([_I64, _I64, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_xc_gga_model": ([_VP, _VP, _I64, _D, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_sphere_plan_create": ([_I32, _I32, _I32, _I64, _VP, _I32, ctypes.POINTER(_VP)], _I32),
    "addft_sphere_plan_destroy": ([_VP], _I32),
    "addft_sphere_plan_info": ([_VP, ctypes.POINTER(_I64), ctypes.POINTER(_I64), _VP], _I32),
    "addft_sphere_work_bytes": ([_VP, _I64], _I64),
    "addft_sphere_plan_set_lane_waves": ([_VP, _I32], _I32),
    "addft_sphere_plan_set_kernel": ([_VP, _I32], _I32),
    "addft_sphere_plan_get_kernel": ([_VP, ctypes.POINTER(_I32), ctypes.POINTER(_I32)], _I32),
    "addft_sphere_permute": ([_VP, _I32, _I32, _VP, _I64, _VP, _VP], _I32),
    "addft_grid_transpose": ([_I32, _I32, _I32, _I32, _VP, _VP, _VP], _I32),
    "addft_grid_stage": ([_VP, _VP, _D, _VP, _VP], _I32),
    "addft_hpsi_local": ([_VP, _VP, _VP, _VP, _I64, _VP, _VP, _VP], _I32),
    "addft_hpsi_local_kbatch": ([_VP, _VP, _VP, _I64, _VP, _I64, _VP, _VP, _VP], _I32),
    "addft_density_accumulate": ([_VP, _VP, _VP, _I64, _VP, _I32, _VP, _VP], _I32),
    "addft_sphere_to_real": ([_VP, _VP, _D, _VP, _VP, _VP], _I32),
    "addft_real_to_sphere": ([_VP, _VP, _D, _VP, _VP, _VP], _I32),
    "addft_gvectors": ([_VP, _VP, _I64, _VP, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_beta_table": ([_I32, _I32, _VP, _VP, _VP, _I64, _VP, _VP, _VP, _I64, _VP, _VP, _VP], _I32),
    "addft_vloc_table": ([_I32, _D, _D, _D, _D, _I64, _VP, _VP, _VP, _I64, _VP, _VP, _VP], _I32),
    "addft_projectors": ([_I32, _I64, _I64, _VP, _VP, _VP, _VP, _VP, _VP, _D, _VP, _VP], _I32),
    "addft_vloc_g": ([_I64, _I32, _VP, _VP, _VP, _D, _VP, _VP], _I32),
    "addft_backward_work_bytes": ([_I64], _I64),
    "addft_reduce_work_bytes": ([], _I64),
    "addft_local_backward": ([_I64, _I32, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_nonlocal_backward": ([_I32, _I64, _I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _D,
                                 _VP, _VP, _VP, _VP], _I32),
    "addft_gsum_outer": ([_I64, _VP, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_ewald_work_bytes": ([_I32, _I32], _I64),
    "addft_ewald": ([_I32, _VP, _VP, _VP, _VP, _D, _D, _I32, _VP, _VP, _VP], _I32),
    "addft_occupations": ([_VP, _VP, _I32, _I32, _D, _D, _D, _VP, _VP, _VP], _I32),
    "addft_hartree_g": ([_I64, _VP, _VP, _VP, _VP, _VP, _VP], _I32),
    "addft_box_fft": ([_I32, _I32, _I32, _VP, _I32, _VP, _VP], _I32),
    "addft_kerker_mix": ([_I32, _I32, _I32, _VP, _VP, _VP, _D, _D, _VP, _VP, _VP], _I32),
    "addft_projector_create": ([_I32, _I64, _I64, _VP, _VP, _VP, _VP, _I64, _VP, ctypes.POINTER(_VP)], _I32),
    "addft_projector_set_tiles": ([_VP, _I32], _I32),
    "addft_projector_destroy": ([_VP], _I32),
    "addft_projector_info": ([_VP, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I32)], _I32),
    "addft_nonlocal_work_bytes": ([_VP, _I64], _I64),
    "addft_nonlocal_becp": ([_VP, _VP, _I64, _I64, _VP, _I64, _I32, _VP, _VP], _I32),
    "addft_nonlocal_apply_add": ([_VP, _VP, _I64, _I64, _VP, _I64, _I32, _VP, _VP], _I32),
    "addft_subspace_gram_work_bytes": ([_I32, _I64, _I64, _I32], _I64),
    "addft_subspace_gram": ([_I32, _I64, _I64, _I64, _VP, _I64, _VP, _I64, _VP, _I64, _I32, _VP, _VP], _I32),
    "addft_subspace_rotate": ([_I32, _I64, _I64, _I64, _VP, _I64, _VP, _I64, _VP, _I64, _VP], _I32),
    "addft_subspace_eigh_work_bytes": ([_I64, _I64], _I64),
    "addft_subspace_eigh": ([_I32, _I32, _I64, _VP, _I64, _I64, _VP, _VP, _VP, _VP], _I32),
    "addft_subspace_chol_inv": ([_I32, _I64, _VP, _I64, _I64, _VP, _VP, _VP], _I32),
    "addft_nonlocal_becp_noalloc": ([_VP, _VP, _I64, _I64, _VP, _I64, _VP], _I32),
    "addft_nonlocal_noalloc_scratch_bytes": ([], _I64),
}


def exported_symbols():
    return tuple(_SIGS)


# The reference's per-primitive switches (kernels/pack.py:68, xc_lda.py:66,
# eigh.py:40): "0" selects its JAX implementation instead of the native one.
# Pack and XC exist here only as HIP kernels, so "0" is refused rather than
# ignored; the native dense eigensolver the EIGH switch governs has no use on
# this path (the Davidson/ChFSI eigensolver replaces the dense Hamiltonian),
# so that switch is accepted either way.
_NO_FALLBACK = ("ADDFT_NATIVE_PACK", "ADDFT_NATIVE_XC")


def check_switches():
    for name in _NO_FALLBACK:
        if os.environ.get(name, "1") == "0":
            raise NativeUnavailable(
                f"{name}=0 selects the reference's JAX implementation, which this build does not have: "
                f"its pack and XC run only as HIP kernels; unset {name}")


def lib():
    """Load (once) and return the library, or raise NativeUnavailable."""
    global _LIB, _PATH
    if _LIB is not None:
        return _LIB
    check_switches()
    path = lib_path()
    if not path.exists():
        raise NativeUnavailable(
            f"libaddft_hip.so not found at {path}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or set ADDFT_NATIVE_LIB")
    handle = ctypes.CDLL(str(path))
    for name, (args, res) in _SIGS.items():
        fn = getattr(handle, name)
        fn.argtypes = args
        fn.restype = res
    _LIB, _PATH = handle, path
    return handle


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().addft_last_error().decode(errors="replace")
        raise AddftError(f"{what} failed (code {rc}): {msg}")


def stream() -> int:
    import torch
    return torch.cuda.current_stream().cuda_stream


def ptr(t) -> int:
    """Device address of a tensor's data.  A lazily conjugated complex view
    (torch's conj bit) is refused: the kernels would read it unconjugated."""
    if t is None:
        return 0
    if t.is_complex() and t.is_conj():
        raise ValueError("a lazily conjugated tensor reached a native call; pass conj_physical() / resolve_conj()")
    return int(t.data_ptr())
