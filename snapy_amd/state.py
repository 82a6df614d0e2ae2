"""StateStore, FluxRegister, PrecisionPolicy and StepDt: the block-runtime
contracts of the stage, restated over device tensors.

* StateStore (src/mesh/state_store.hpp:24-125): fixed-address hydro_u and its
  RK backup u0; backup() at stage 0, restore() on a redo; every registered
  buffer bumps ``generation`` so a captured graph knows to re-capture.
* FluxRegister (src/hydro/flux_register.hpp:241-402): the stage face-flux
  totals T_d (the F1..F3 buffers the sweeps write) and the optional FP64 step
  fields S_d <- w1 S_d + w2 dt T_d (end_stage, flux_register.cpp:761-777).
* PrecisionPolicy (src/mesh/precision_policy.hpp:13-71): per-quantity
  {storage, compute} dtypes from dynamics.precision.
* StepDt (src/utils/step_dt.hpp:17-63): a host dt or a 0-d FP64 device dt.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch


class StateStore:
    """Fixed-address conserved state and its RK backup.  In the FP32 state
    form the state is the FP32 pair (u, c), c the compensation of the RK
    accumulation, and the backup is the pair (u0, c0)."""

    def __init__(self):
        self._u = None
        self._u0 = None
        self._c = None
        self._c0 = None
        self.generation = 0
        self._storage = {}

    def adopt(self, u: torch.Tensor, comp: torch.Tensor | None = None) -> None:
        """Take u as hydro_u (its address is then fixed) and size u0; with
        comp, the pair (u, comp) and its backup (u0, c0)."""
        if self._u is None or self._u.data_ptr() != u.data_ptr():
            self._u = u
            self._u0 = torch.empty_like(u)
            self._c = comp
            self._c0 = torch.empty_like(comp) if comp is not None else None
            self.generation += 1

    def comp(self):
        return self._c

    def comp0(self):
        return self._c0

    def register_storage(self, name: str, t: torch.Tensor) -> None:
        self._storage[name] = t
        self.generation += 1

    def u(self) -> torch.Tensor:
        return self._u

    def u0(self) -> torch.Tensor:
        return self._u0

    def backup(self) -> None:
        self._u0.copy_(self._u)
        if self._c is not None:
            self._c0.copy_(self._c)

    def restore(self) -> None:
        self._u.copy_(self._u0)
        if self._c is not None:
            self._c.copy_(self._c0)


class FluxRegister:
    """Stage totals T_d and step fields S_d of the active directions."""

    def __init__(self, nvar, shape, active, dtype, device):
        self.nvar = nvar
        self.active = tuple(active)
        self._total = [None, None, None]
        self._step = [None, None, None]
        for d in range(3):
            if active[d]:
                self._total[d] = torch.zeros((nvar,) + tuple(shape), dtype=dtype, device=device)
        # the x1 face pressure is a reference quantity: FP64 in both forms
        self.face_pressure1 = (
            torch.zeros(tuple(shape), dtype=torch.float64, device=device) if active[0] else None
        )
        self._w1 = 0.0
        self._w2 = 0.0
        self._stage = -1

    def total(self, d: int):
        return self._total[d - 1]

    def step(self, d: int):
        return self._step[d - 1]

    def request_step_field(self, d: int) -> None:
        t = self._total[d - 1]
        if t is not None and self._step[d - 1] is None:
            self._step[d - 1] = torch.zeros(t.shape, dtype=torch.float64, device=t.device)

    def has_step_fields(self) -> bool:
        return any(s is not None for s in self._step)

    def begin_stage(self, stage: int, w1: float, w2: float) -> None:
        self._stage = stage
        self._w1 = w1
        self._w2 = w2


@dataclass
class PrecisionRow:
    storage: torch.dtype
    compute: torch.dtype


class PrecisionPolicy:
    """dtype rows: the state, backup, flux narrow under float32; the
    primitive and the reference stay FP64 (precision_policy.hpp:31-71).  The
    GPU form keeps a compensation pair for the narrowed state and FP64
    implicit and budget rows."""

    QUANTITIES = ("hydro_state", "backup", "flux", "primitive", "implicit", "budget")

    def __init__(self, precision: str = "float64"):
        if precision not in ("float64", "float32"):
            raise ValueError("precision must be float64 or float32")
        self.precision = precision
        narrow = torch.float32 if precision == "float32" else torch.float64
        self.rows = {
            "hydro_state": PrecisionRow(narrow, narrow),
            "backup": PrecisionRow(narrow, narrow),
            "flux": PrecisionRow(narrow, narrow),
            "primitive": PrecisionRow(torch.float64, torch.float64),
            "implicit": PrecisionRow(torch.float64, torch.float64),
            "budget": PrecisionRow(torch.float64, torch.float64),
        }

    def fp32(self) -> bool:
        return self.precision == "float32"

    def row(self, q: str) -> PrecisionRow:
        return self.rows[q]


class StepDt:
    """A host dt, or a device 0-d FP64 dt read by captured kernels."""

    def __init__(self, value=None, tensor: torch.Tensor | None = None):
        self._value = value
        self._tensor = tensor

    def on_device(self) -> bool:
        return self._tensor is not None

    def value(self) -> float:
        if self._tensor is not None:
            raise RuntimeError("StepDt: a device dt has no host value inside a captured stage")
        return float(self._value)

    def tensor(self) -> torch.Tensor:
        return self._tensor
