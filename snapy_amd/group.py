"""Many blocks of one deck on different grids, advanced with one launch per
kernel.

The reference steps the blocks of a process on worker threads, one stream
each (src/mesh/mesh.cpp:74-110).  A `BlockGroup` instead records every
block's launches once, per phase of a step (the primitives and the CFL dt,
then each RK stage), through the library's recording mode
(`snapy_record_begin` / `snapy_record_end`), and groups them per launch
(`snapy_group_build`): a step then makes each kernel launch once over the
workgroups of all the blocks (`snapy_group_launch`), every workgroup with
its own block's arguments.  A block's cells see the arithmetic of its own
launches, so a grouped step is bit-for-bit the blocks' own steps.

Cubed-sphere meshes (`CubedSphereMesh`, six panels and their members per
launch already) group the same way, meshes of different panel sizes in one
launch per kernel.

The blocks must make the same launches: the same deck (physics, schemes,
boundaries, state form) on grids of any extents and spacings, as long as
each launch runs the same kernel with the same workgroup shape (the x2/x3
sweeps, whose workgroups cover whole x1 lines, may differ in their
workgroup size: the group takes the largest and their bodies idle the
extra threads).  Each block keeps its own buffers and its own device dt.  The recorded launches carry the
blocks' buffer addresses and the stages' weights, so a group is rebuilt
when a block's buffers change (a new block, a restart).
"""

from __future__ import annotations

import ctypes

import torch

from . import _capi

# stage launches that are torch operations, made per block at each step
# rather than recorded
_TORCH_CALLS = ("vic_status_clear",)


def _is_cubed(block) -> bool:
    from .cubed import CubedSphereMesh

    return isinstance(block, CubedSphereMesh)


def groupable(block) -> bool:
    """A block the group can take: a single block (no member axis) of an
    ideal gas (tracers included), or a cubed-sphere mesh."""
    if _is_cubed(block):
        return True
    return block.members == 1 and block.nvar == 5


def group_key(block):
    """Blocks with equal keys run the same deck: the options but the grid's
    extents, spacings and cell counts (the active axes stay).  The library
    checks the launches themselves when the group is built."""
    if _is_cubed(block):
        b = block.options.block
        return ("cubed-sphere", repr(b.hydro), repr(b.intg), tuple(b.bcnames), b.coord.nghost)
    op = block.options
    c = op.coord
    coord = (c.type, c.nghost, c.nx1 > 1, c.nx2 > 1, c.nx3 > 1)
    return (repr(op.hydro), repr(op.intg), tuple(op.bcnames), coord)


class BlockGroup:
    """Blocks (with their vars) advanced together, one launch per kernel
    for all of them per step.  capture: from the second step on, a step is
    one captured hipGraph replayed (the group's launches and the per-block
    backups), once every block's primitives are stale, as after a step."""

    def __init__(self, members, capture: bool = False):
        self.members = list(members)
        self.capture = capture
        self._graph = None
        if not self.members:
            raise ValueError("BlockGroup: no blocks")
        for blk, _ in self.members:
            if not groupable(blk):
                raise ValueError("BlockGroup: a block of members, or not of an ideal gas, is not grouped")
            ready = blk._u is not None if _is_cubed(blk) else blk._freg is not None
            if blk.device.type != "cuda" or not ready:
                raise ValueError("BlockGroup: initialize every block on the GPU first")
        keys = {group_key(blk) for blk, _ in self.members}
        if len(keys) != 1:
            raise ValueError("BlockGroup: the blocks do not share a deck")
        self.device = self.members[0][0].device
        self._lib = _capi.load()
        self._groups = {}
        try:
            for phase in ("dt", 0, 1, 2):
                self._groups[phase] = self._build(phase)
        except BaseException:
            self.release(sync=False)
            raise

    # ------------------------------------------------------------- building
    def _record(self, blk, vars, phase) -> None:
        if _is_cubed(blk):
            self._record_cubed(blk, phase)
            return
        if phase == "dt":
            # MeshBlock.device_time_step of a block whose primitives are stale
            # (recording launches nothing: the block's own flag stays as it
            # was)
            stale = blk._w_stale
            blk._refresh_primitives(vars)
            blk._w_stale = stale
            g, p = ctypes.byref(blk._grid), ctypes.byref(blk._phys)
            w = vars["hydro_w"]
            scale = 2.0 ** (-blk.intg.current_redo) * blk.intg.options.cfl
            _capi.check(
                self._lib.snapy_cfl(g, p, _capi.dtype_code(w.dtype), _capi.ptr(w), scale, _capi.ptr(blk._dt_dev),
                                    _capi.stream_of(self.device)),
                "cfl",
            )
            return
        # every block's plan with the primitives' conversion at the step's
        # start, whatever its primitives are now (a block just admitted and
        # one mid-run make the same launches, so their plans group)
        fresh = blk._w_interior_fresh
        blk._w_interior_fresh = False
        try:
            calls = blk._stage_calls(vars, phase)
        finally:
            blk._w_interior_fresh = fresh
        for name, fn in calls:
            if name in _TORCH_CALLS:
                continue
            _capi.check(fn(), name)

    def _record_cubed(self, mesh, phase) -> None:
        if phase == "dt":
            # CubedSphereMesh.device_time_step with stale primitives
            stale = mesh._w_stale
            mesh._refresh_primitives()
            mesh._w_stale = stale
            cfl = self._lib.snapy_cs_cfl_f32 if mesh.fp32 else self._lib.snapy_cs_cfl
            scale = 2.0 ** (-mesh.current_redo) * mesh.cfl
            _capi.check(cfl(ctypes.byref(mesh._desc), _capi.ptr(mesh._w), scale, _capi.ptr(mesh._dt_dev),
                            _capi.stream_of(self.device)), "cs_cfl")
            return
        for name, fn in mesh._stage_calls(phase):
            _capi.check(fn(), name)

    def _build(self, phase):
        lib = self._lib
        plans = []
        try:
            for blk, vars in self.members:
                _capi.check(lib.snapy_record_begin(), "record_begin")
                plan = ctypes.c_void_p()
                try:
                    self._record(blk, vars, phase)
                finally:
                    _capi.check(lib.snapy_record_end(ctypes.byref(plan)), "record_end")
                    plans.append(plan)
            arr = (ctypes.c_void_p * len(plans))(*[p.value for p in plans])
            grp = ctypes.c_void_p()
            _capi.check(lib.snapy_group_build(arr, len(plans), ctypes.byref(grp)), f"group_build({phase})")
            return grp
        finally:
            for p in plans:
                lib.snapy_plan_free(p)

    # -------------------------------------------------------------- running
    def launches(self, phase) -> int:
        """The grouped launches of a phase ("dt", 0, 1 or 2)."""
        return int(self._lib.snapy_group_launches(self._groups[phase]))

    def step(self) -> None:
        """One RK3 step of every block on the current stream: each block's
        CFL dt (grouped when every block's primitives are stale, as they are
        after a step; else each block's own), then the three stages."""
        stale = all(blk._w_stale for blk, _ in self.members)
        if self._graph is not None and stale:
            self._graph.replay()
            return
        if self.capture and stale:
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize(self.device)
            with torch.cuda.graph(g):
                self._step()
            self._graph = g
            g.replay()
            return
        self._step()

    def _step(self) -> None:
        s = _capi.stream_of(self.device)
        lib = self._lib
        if all(blk._w_stale for blk, _ in self.members):
            _capi.check(lib.snapy_group_launch(self._groups["dt"], s), "group_launch(dt)")
        else:
            for blk, vars in self.members:
                blk.device_time_step(vars)
        for k in range(3):
            if k == 0:
                # the RK backups (state and tracers) and the implicit
                # solve's status latch, per block (MeshBlock._launch_stage
                # and its first call; CubedSphereMesh._launch_stage)
                for blk, vars in self.members:
                    if _is_cubed(blk):
                        blk._u0.copy_(blk._u)
                        if blk._c is not None:
                            blk._c0.copy_(blk._c)
                        continue
                    blk._state.backup()
                    if blk.nscalar:
                        blk._scal["s0"].copy_(vars["scalar_s"])
                    if blk._vic is not None:
                        blk._vic["status"].zero_()
            _capi.check(lib.snapy_group_launch(self._groups[k], s), f"group_launch({k})")
        for blk, _ in self.members:
            blk._w_stale = True
            if not _is_cubed(blk):
                blk._w_interior_fresh = blk._prims_from_finish()

    def release(self, sync: bool = True) -> None:
        """Free the group's device records (after its launches finished)."""
        if sync and self.device.type == "cuda":
            torch.cuda.synchronize(self.device)
        self._graph = None
        for g in self._groups.values():
            if g:
                self._lib.snapy_group_free(g)
        self._groups = {}

    def __del__(self):
        try:
            if self._groups:
                self.release()
        except Exception:
            pass
