"""snapy_amd: the finite-volume stage of snapy on MI355X (gfx950).

The package mirrors snapy's Python surface for the dycore stage:
``MeshBlockOptions.from_yaml``, ``MeshBlock``, the row constants ``kIDN`` ...
and the block-runtime contracts (StateStore, FluxRegister, PrecisionPolicy,
StepDt).  All cell updates run in hand-written HIP kernels reached through
the C-ABI of include/snapy_amd.h.
"""

from .meshblock import Integrator, MeshBlock, StageExecutor
from .options import MeshBlockOptions, MeshOptions
from .state import FluxRegister, PrecisionPolicy, StateStore, StepDt

kIDN, kIVX, kIVY, kIVZ, kIPR = 0, 1, 2, 3, 4
# the reference's spellings (python/snapy/__init__.pyi)
kIV1, kIV2, kIV3, kICY = kIVX, kIVY, kIVZ, 5
kInnerX1, kOuterX1, kInnerX2, kOuterX2, kInnerX3, kOuterX3 = range(6)

__all__ = [
    "MeshBlock",
    "MeshBlockOptions",
    "MeshOptions",
    "StageExecutor",
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
    "kIV1",
    "kIV2",
    "kIV3",
    "kICY",
    "kInnerX1",
    "kOuterX1",
    "kInnerX2",
    "kOuterX2",
    "kInnerX3",
    "kOuterX3",
]
