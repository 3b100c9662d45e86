"""addft -- MI355X-native Kohn-Sham Hpsi / density / adjoint path.

Drop-in for the reference `addft` API (src/addft/__init__.py:15-37).  The
numerical work runs in hand-written gfx950 HIP kernels reached through the
C ABI in include/addft_hip.h; PyTorch supplies device memory and streams.
"""

from .geometry.cell import Cell, strained, volume  # noqa: F401
from .geometry.kmesh import gamma, monkhorst_pack  # noqa: F401
from .geometry.psp import Pseudo, synthetic, from_upf  # noqa: F401
from .runtime import units  # noqa: F401
