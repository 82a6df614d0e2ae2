"""snapy_amd: the finite-volume stage of snapy on MI355X (gfx950).

The package mirrors snapy's Python surface for the dycore stage:
``MeshBlockOptions.from_yaml``, ``MeshBlock``, the row constants ``kIDN`` ...
and the block-runtime contracts (StateStore, FluxRegister, PrecisionPolicy,
StepDt).  All cell updates run in hand-written HIP kernels reached through
the C-ABI of include/snapy_amd.h.
"""

from .meshblock import Integrator, MeshBlock
from .options import MeshBlockOptions
from .state import FluxRegister, PrecisionPolicy, StateStore, StepDt

kIDN, kIVX, kIVY, kIVZ, kIPR = 0, 1, 2, 3, 4

__all__ = [
    "MeshBlock",
    "MeshBlockOptions",
    "Integrator",
    "StateStore",
    "FluxRegister",
    "PrecisionPolicy",
    "StepDt",
    "kIDN",
    "kIVX",
    "kIVY",
    "kIVZ",
    "kIPR",
]
