"""addft -- MI355X-native Kohn-Sham Hpsi / density / adjoint path.

Drop-in for the reference `addft` API (src/addft/__init__.py:15-37).  The
numerical work runs in hand-written gfx950 HIP kernels reached through the
C ABI in include/addft_hip.h; PyTorch supplies device memory and streams.
Importing the package needs no GPU; the first numerical call does, and it
raises NativeUnavailable when the HIP library has not been built.
"""

from .geometry.cell import Cell, make_cell, strained, volume  # noqa: F401
from .geometry.kmesh import gamma, monkhorst_pack, time_reversal  # noqa: F401
from .geometry.psp import Pseudo, from_upf, from_upf_text, synthetic  # noqa: F401
from .runtime import units  # noqa: F401


def __getattr__(name):
    # the API module imports torch; load it lazily so geometry-only users
    # (and CPU tests) need not
    import importlib
    if name.startswith("__"):
        raise AttributeError(name)
    api = importlib.import_module(".api", __name__)
    if hasattr(api, name):
        return getattr(api, name)
    raise AttributeError(name)


__all__ = ["System", "setup", "setup_from_case", "scf", "energy", "energy_explicit", "energy_and_forces", "stress",
           "energy_decomposition", "make_energy", "energy_forces_stress", "Cell", "strained", "volume",
           "gamma", "monkhorst_pack", "time_reversal", "Pseudo", "synthetic", "from_upf", "units", "delta_q", "delta_q_grad",
           "make_gaussian_weight", "ablate", "ablate_all", "Ablation"]
