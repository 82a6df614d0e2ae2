"""The work list: many independent x1 blocks of uneven size, one launch per
phase.

`BlockBatch` takes resident `MeshBlock`s (1D ideal-gas blocks without
sources, FP64: the serving mix's shock tubes) and advances all of them by
one RK3 step with a fixed number of launches, whatever their count and
sizes: one CFL launch (a device dt per block), then per stage the backup
(stage 0), cons2prim, the x1 sweep, the divergence + combine and the
boundaries, each one persistent launch over the work list of every block
(snapy_wl_*).  Each block keeps its own buffers (`vars["hydro_u"]`, its RK
backup, primitives, x1 face totals, outflow background and device dt), so a
block can join or leave a batch between steps; its state is bit-identical to
the same block stepped alone with `device_time_step` and `forward`.

The reference advances the blocks of a Mesh one after another, each with
its own launches (src/mesh/mesh.cpp:74-110, MeshBlockImpl::forward
src/mesh/meshblock_advance.cpp:88-105).
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _capi
from .meshblock import MeshBlock

_BC = {"reflecting": 1, "periodic": 2, "extrapolation": 3, "outflow": 4}


def eligible(block: MeshBlock) -> bool:
    """A block the work list can carry: x1 only, ideal gas (5 rows), FP64,
    no gravity, diffusion, implicit solve or tracers, nghost 3, on a HIP
    device."""
    c = block.options.coord
    h = block.options.hydro
    return (
        c.nx2 == 1
        and c.nx3 == 1
        and c.nx1 >= c.nghost
        and c.nghost == 3
        and block.nvar == 5
        and not block.policy.fp32()
        and h.grav1 == 0.0
        and h.nu_iso == 0.0
        and h.kappa_iso == 0.0
        and block.options.intg.implicit_scheme == 0
        and block.nscalar == 0
        and block.device.type == "cuda"
        and all(b in _BC for b in block.options.bcnames[:2])
    )


def _same_physics(a: _capi.Physics, b: _capi.Physics) -> bool:
    return (
        a.gammad == b.gammad
        and a.riemann == b.riemann
        and a.shock_x1 == b.shock_x1
        and a.density_floor == b.density_floor
        and a.pressure_floor == b.pressure_floor
    )


class BlockBatch:
    """A work list over `members`, a list of (MeshBlock, vars) pairs of
    initialized eligible blocks sharing the deck physics (gamma, Riemann
    solver, shock flag, floors); sizes, cell widths, boundary kinds and CFL
    numbers may differ."""

    def __init__(self, members: list[tuple[MeshBlock, dict]]):
        if not members:
            raise ValueError("BlockBatch: no blocks")
        for blk, vars in members:
            if not eligible(blk):
                raise ValueError("BlockBatch: a block outside the work list's path (x1, ideal gas, FP64, no sources)")
            if "hydro_u" not in vars:
                raise ValueError("BlockBatch: initialize every block first")
            if not _same_physics(blk._phys, members[0][0]._phys):
                raise ValueError("BlockBatch: the blocks do not share the deck physics")
        self.members = list(members)
        self.device = members[0][0].device
        self._phys = members[0][0]._phys
        self._build()

    def _build(self) -> None:
        n = len(self.members)
        descs = (_capi.WlBlock * n)()
        for i, (blk, vars) in enumerate(self.members):
            blk._ensure_buffers()
            g = blk._grid
            d = descs[i]
            d.nx1 = g.nx1
            d.bc_lo = _BC[blk.options.bcnames[0]]
            d.bc_hi = _BC[blk.options.bcnames[1]]
            d.dx1, d.dx2, d.dx3 = g.dx1, g.dx2, g.dx3
            d.scale = math.pow(2.0, -blk.intg.current_redo) * blk.intg.options.cfl
            d.u = _capi.ptr(vars["hydro_u"])
            d.u0 = _capi.ptr(blk._state.u0())
            d.w = _capi.ptr(vars["hydro_w"])
            d.flux = _capi.ptr(blk._freg.total(1))
            d.bref = _capi.ptr(vars.get("boundary_reference_w"))
            d.dt = _capi.ptr(blk._dt_dev)
            if d.bref is None and "outflow" in blk.options.bcnames[:2]:
                raise ValueError("BlockBatch: an outflow block without its boundary reference")
        prefix = (ctypes.c_int32 * (4 * (n + 1)))()
        total = (ctypes.c_int64 * 4)()
        lib = _capi.load()
        _capi.check(lib.snapy_wl_counts(descs, n, 3, prefix, total), "wl_counts")
        raw = bytes(descs)
        self._descs = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.device)
        self._prefix = torch.tensor(list(prefix), dtype=torch.int32, device=self.device)
        self._wl = _capi.WorkList()
        self._wl.blocks = _capi.ptr(self._descs)
        self._wl.prefix = _capi.ptr(self._prefix)
        self._wl.nblocks = n
        self._wl.nghost = 3
        for k in range(4):
            self._wl.total[k] = total[k]
        self.cells = sum(int(blk.options.coord.nx1) for blk, _ in self.members)

    def launches_per_step(self) -> int:
        """Kernel launches of one RK3 step of the whole batch (after its
        first step: the primitives' refresh, the CFL, then per stage
        cons2prim, sweep, update and boundaries, and the stage-0 backup)."""
        return 1 + 1 + 3 * 4 + 1

    def step(self) -> None:
        """One RK3 step of every block on the current stream: the device CFL
        dt of each block, then the three stages."""
        lib = _capi.load()
        wl, p = ctypes.byref(self._wl), ctypes.byref(self._phys)
        s = _capi.stream_of(self.device)
        # the CFL reads prim(u) of each block's current state, as a block
        # alone does (MeshBlock.device_time_step): one launch when every
        # block stepped since, each stale block alone when only some did
        stale = [blk._w_stale for blk, _ in self.members]
        if all(stale):
            _capi.check(lib.snapy_wl_cons2prim(wl, p, s), "wl_cons2prim")
        else:
            for (blk, vars), st in zip(self.members, stale):
                if st:
                    blk._refresh_primitives(vars)
        _capi.check(lib.snapy_wl_cfl(wl, p, s), "wl_cfl")
        stages = self.members[0][0].intg.stages
        for k, st in enumerate(stages):
            _capi.check(lib.snapy_wl_stage(wl, p, k, st.wght0(), st.wght1(), st.wght2(), s), "wl_stage")
        for blk, _ in self.members:
            blk._w_stale = True
