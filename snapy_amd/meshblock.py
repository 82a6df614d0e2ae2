"""MeshBlock and StageExecutor: the block runtime of the stage over the C-ABI.

Mirrors the reference Python surface (python/csrc/pymesh.cpp:159-327,
python/snapy/mesh.pyi:50-113): ``MeshBlock(options)``, ``to(device)``,
``initialize(vars) -> (vars, 0.)``, ``forward(vars, dt, stage)``,
``max_time_step(vars)``, ``check_redo(vars)``, ``part(...)``, ``total_u``,
``execution_report``.  ``vars`` is the same dict of tensors keyed
``hydro_w`` / ``hydro_u`` (+ ``boundary_reference_w`` with outflow faces).

Every cell update runs in the gfx950 kernels behind include/snapy_amd.h; the
host only sequences launches on the current HIP stream.  There is no CPU
fallback: a block on a CPU device, or a missing kernel library, raises.
"""

from __future__ import annotations

import ctypes
import gc
import math
import weakref

import torch

from . import _capi
from .options import MeshBlockOptions
from .state import FluxRegister, PrecisionPolicy, StateStore

IDN, IVX, IVY, IVZ, IPR = 0, 1, 2, 3, 4


class Integrator:
    """harp's SSP-RK3 (u <- w0 u0 + w1 u + w2 du per stage); wght2 =
    {1, 1/4, 2/3} as pinned by tests/test_forcing.cpp:1317-1321."""

    class Stage:
        def __init__(self, w0, w1, w2):
            self._w = (w0, w1, w2)

        def wght0(self):
            return self._w[0]

        def wght1(self):
            return self._w[1]

        def wght2(self):
            return self._w[2]

    def __init__(self, options):
        if options.type != "rk3":
            raise ValueError("integrator: rk3 only")
        self.options = options
        self.stages = [
            Integrator.Stage(1.0, 0.0, 1.0),
            Integrator.Stage(0.75, 0.25, 0.25),
            Integrator.Stage(1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0),
        ]
        self.current_redo = 0
        self.max_redo = 5

    def stop(self, cycle: int, time: float) -> bool:
        o = self.options
        if o.nlim >= 0 and cycle >= o.nlim:
            return True
        return time >= o.tlim


class _Buffers:
    """Named buffers a test or driver can read back (block.module(...))."""

    def __init__(self, **bufs):
        self._b = dict(bufs)

    def buffer(self, name):
        return self._b[name]

    def named_buffers(self):
        return dict(self._b)


class MeshBlock:
    """MeshBlock(options) is one block; MeshBlock(options, members=M) of a
    deck whose x3 is a single cell holds M independent members of that
    block stacked along x3 (tensors (nvar, M, nc2, nc1), no x3 ghosts or
    coupling), advanced together by every launch, each member with its own
    dt: the device dt and the time steps are then M values, and forward
    takes a float (every member) or M values."""

    def __init__(self, options: MeshBlockOptions, members: int = 1):
        self.members = int(members)
        if self.members < 1:
            raise ValueError("members >= 1")
        if self.members > 1:
            if options.coord.nx3 != 1:
                raise ValueError("members: the deck's x3 must be a single cell (members stack along x3)")
            if options.hydro.eos.type != "ideal-gas":
                raise ValueError("members: the member axis is built for the ideal-gas stage")
            if options.scalar.nvar > 0:
                raise ValueError("members: a block of members carries no passive scalars")
        self.options = options
        # passive tracers (ScalarImpl, src/scalar/scalar.cpp): ns FP64 rows
        # scalar_s = rho r with their RK backup, and the stage's fluxes
        self.nscalar = int(options.scalar.nvar)
        self._scal = None
        self._cycle = 0
        self.last_redo_causes = 0
        self.intg = Integrator(options.intg)
        self.pintg = self.intg
        self.policy = PrecisionPolicy(options.hydro.precision)
        self.device = torch.device("cpu")
        self._state = StateStore()
        self._freg = None
        self._ref = None
        self._x1tab = None
        self._vic = None
        self._scratch_w = None
        self._dt_dev = None
        # hydro_w holds prim(hydro_u) (False) or the primitives of the last
        # stage's input (True, after forward until check_redo or a time step
        # refreshes them)
        self._w_stale = False
        # the interior of hydro_w holds prim(hydro_u) bit for bit (written by
        # the last stage's implicit finish); only the ghosts may be stale
        self._w_interior_fresh = False
        self._executor = StageExecutor(self, options.intg.execution)
        c = options.coord
        # ideal gas: (rho, m1, m2, m3, E); shallow water: (phi, phi v1..v3)
        self.nvar = 4 if options.hydro.eos.type == "shallow-water" else 5
        m = self.members
        self._grid = _capi.Grid(
            self.nvar,
            c.nghost,
            c.nx1,
            c.nx2,
            m if m > 1 else c.nx3,
            c.nc1(),
            c.nc2(),
            m if m > 1 else c.nc3(),
            c.dx(1),
            c.dx(2),
            c.dx(3),
            c.x1min,
            c.x2min,
            c.x3min,
            m if m > 1 else 0,
            0,
        )
        h = options.hydro
        bcs = (ctypes.c_int32 * 6)(*[_capi.BC[n] for n in options.bcnames])
        self._phys = _capi.Physics(
            h.eos.gammad,
            h.grav1,
            h.eos.density_floor,
            h.eos.pressure_floor,
            _capi.RIEMANN[h.riemann],
            int(h.recon1.shock),
            int(h.recon23.shock),
            bcs,
            int(options.intg.implicit_scheme),
            options.intg.cfl / options.intg.implicit_advection_cfl,
            h.bot_heat_flux,
            int(h.bot_heat_depth),
            int(h.flux_covariance),
            h.nu_iso,
            h.kappa_iso,
            h.eos.gas_constant_dry(),
            1 if h.eos.type == "shallow-water" else 0,
        )
        self._build_coord()

    # ------------------------------------------------------------------ grid
    def _build_coord(self):
        c = self.options.coord
        bufs = {}
        for d in (1, 2, 3):
            nx = (c.nx1, c.nx2, c.nx3)[d - 1]
            lo = (c.x1min, c.x2min, c.x3min)[d - 1]
            hi = (c.x1max, c.x2max, c.x3max)[d - 1]
            dx = (hi - lo) / nx
            ng = c.nghost if nx > 1 else 0
            f = torch.linspace(lo - ng * dx, hi + ng * dx, nx + 2 * ng + 1, dtype=torch.float64)
            bufs[f"x{d}f"] = f
            bufs[f"x{d}v"] = 0.5 * (f[:-1] + f[1:])
            bufs[f"dx{d}f"] = torch.ones(nx + 2 * ng, dtype=torch.float64) * dx
        self._coord = _Buffers(**bufs)

    def cycle(self) -> int:
        """The accepted-step count (pymesh.cpp:161)."""
        return self._cycle

    def inc_cycle(self) -> int:
        """Advance the cycle after an accepted step; returns the cycle before
        the increment (pymesh.cpp:162-167)."""
        v = self._cycle
        self._cycle += 1
        return v

    def module(self, name: str):
        if name == "coord":
            return self._coord
        if name == "hydro":
            return self._hydro_buffers()
        raise KeyError(name)

    def _hydro_buffers(self):
        f = self._freg
        bufs = {}
        if f is not None:
            for d in (1, 2, 3):
                if f.total(d) is not None:
                    bufs[f"F{d}"] = f.total(d)
            if f.face_pressure1 is not None:
                bufs["P1"] = f.face_pressure1
        return _Buffers(**bufs)

    def shape(self):
        c = self.options.coord
        return (self.nvar, self._nc3(), c.nc2(), c.nc1())

    def _nc3(self) -> int:
        """Cells along x3: the deck's, or the members."""
        return self.members if self.members > 1 else self.options.coord.nc3()

    def part(self, offset=(0, 0, 0), exterior=True, extend_x1=0, extend_x2=0, extend_x3=0):
        """The index slices (all variables, x3, x2, x1) of a part of the
        block, as MeshBlockImpl::part (src/mesh/meshblock.cpp:242-284)
        through its binding (python/csrc/pymesh.cpp:193-215).  offset is
        (o3, o2, o1); along an axis, 0 is the interior widened by extend_xd
        cells on each side, and -1 / +1 the ghost layers beyond the interior
        on that side (exterior=True) or the interior layers next to them
        (exterior=False).  An axis without ghosts (one cell) has empty -1 /
        +1 parts; so has a member x3 axis, whose offset-0 part is every
        member."""
        off3 = tuple(int(x) for x in offset)
        if len(off3) != 3 or any(x not in (-1, 0, 1) for x in off3):
            raise ValueError(f"part: offset {tuple(offset)} must be three of -1, 0, 1")
        c = self.options.coord
        depth = 99  # PartOptions::depth's default; the binding does not set it
        out = [slice(None)]
        for d, off, ext in ((3, off3[0], extend_x3), (2, off3[1], extend_x2), (1, off3[2], extend_x1)):
            if d == 3 and self.members > 1:
                out.append(slice(0, self.members) if off == 0 else slice(0, 0))
                continue
            nc = c.nc(d)
            nx = nc - 2 * c.nghost if nc > 1 else 1
            ng = 0 if nx == 1 else c.nghost
            if off == -1:
                start, n = (0 if exterior else ng), min(ng, depth)
            elif off == 0:
                start, n = ng - int(ext), nx + 2 * int(ext)
            else:
                start, n = nx + (ng if exterior else 0), min(ng, depth)
            out.append(slice(start, start + n))
        return tuple(out)

    def _vic_workspace(self, lib) -> int:
        """FP64 workspace doubles of the deck's implicit scheme."""
        if self.options.intg.implicit_scheme == 1:
            return lib.snapy_vic_partial_workspace(ctypes.byref(self._grid))
        return lib.snapy_vic_workspace(ctypes.byref(self._grid))

    # -------------------------------------------------------------- devices
    def to(self, device, dtype=None):
        device = torch.device(device)
        if device.type != "cuda":
            raise _capi.SnapyError(
                "MeshBlock: the stage runs on a HIP device only (no CPU fallback)"
            )
        _capi.load()
        self.device = device
        return self

    def _dtype(self):
        return self.policy.row("hydro_state").storage

    def hbm_footprint(self) -> int:
        """Device bytes the block holds once initialized: the state and its
        RK backup (with the compensation pair in the FP32 form), the FP64
        primitives, the face totals and face pressure, and the gravity,
        diffusion and implicit buffers its options call for."""
        c = self.options.coord
        h = self.options.hydro
        cells = c.nc1() * c.nc2() * self._nc3()
        ncell = self.nvar * cells
        fp32 = self.policy.fp32()
        total = ncell * ((4 + 4) * 2 if fp32 else 8 * 2) + ncell * 8  # u, u0 (+c, c0); w
        ndim = sum(1 for n in (c.nx1, c.nx2, c.nx3) if n > 1)
        total += ndim * ncell * (4 if fp32 else 8) + (cells * 8 if c.nx1 > 1 else 0)
        total += cells * 8 * 8  # x1 table (nc1 rows, bounded by cells)
        if h.grav1 != 0.0:
            total += 5 * cells * 8
        if h.nu_iso > 0.0 or h.kappa_iso > 0.0:
            total += 12 * cells * 8 + ncell * 8
        if self.options.intg.implicit_scheme in (1, 9):
            lib = _capi.load()
            total += int(self._vic_workspace(lib)) * 8 + ncell * 8
            total += c.nx2 * c.nx3 * self.members * 4
            if self.nscalar:
                total += cells * 8  # the dry-gas transfer
        if self.nscalar:
            # s, r, s0, theta, ds and the face fluxes of the active directions
            total += self.nscalar * cells * 8 * (5 + ndim)
        return int(total)

    def _ensure_buffers(self):
        if self._freg is not None:
            return
        c = self.options.coord
        active = (c.nx1 > 1, c.nx2 > 1, c.nx3 > 1)
        shape = (self._nc3(), c.nc2(), c.nc1())
        self._freg = FluxRegister(self.nvar, shape, active, self.policy.row("flux").storage, self.device)
        if "outflow" in self.options.bcnames:
            self._scratch_w = torch.empty(self.shape(), dtype=torch.float64, device=self.device)
        self._dt_dev = torch.zeros(() if self.members == 1 else (self.members,), dtype=torch.float64,
                                   device=self.device)
        h = self.options.hydro
        self._ref = None
        self._x1tab = None
        from .geometry import x1_gravity_table

        cb = self._coord
        # the x1 table: x1v always (the covariance reads it), the gravity
        # factors when grav1 != 0
        self._x1tab = x1_gravity_table(cb.buffer("x1f"), cb.buffer("x1v"), c.nghost, c.nx1, h.grav1).to(
            self.device
        )
        if h.grav1 != 0.0:
            self._ref = torch.zeros((5, self._nc3(), c.nc2(), c.nc1()), dtype=torch.float64, device=self.device)
        self._diff = None
        if h.nu_iso > 0.0 or h.kappa_iso > 0.0:
            lib = _capi.load()
            nwk = int(lib.snapy_diffusion_workspace(ctypes.byref(self._grid)))
            self._diff = {
                "work": torch.zeros(nwk, dtype=torch.float64, device=self.device),
                "du": torch.zeros(self.shape(), dtype=torch.float64, device=self.device),
            }
        self._vic = None
        if self.options.intg.implicit_scheme in (1, 9):
            from .geometry import vic_table

            cb = self._coord
            a1 = c.dx(3) * c.dx(2)
            vol = c.dx(3) * (c.dx(2) * c.dx(1))
            tab, t0, t1 = vic_table(cb.buffer("x1f"), cb.buffer("x1v"), c.nghost, c.nx1, h.grav1, a1, vol)
            lib = _capi.load()
            nws = int(self._vic_workspace(lib))
            self._vic = {
                "tab": tab.to(self.device),
                "t0": t0,
                "t1": t1,
                "ws": torch.empty(nws, dtype=torch.float64, device=self.device),
                "status": torch.zeros(c.nx2 * c.nx3 * self.members, dtype=torch.int32, device=self.device),
                "du": torch.zeros(self.shape(), dtype=torch.float64, device=self.device),
                "scheme": int(self.options.intg.implicit_scheme),
                "name": "vic_partial" if self.options.intg.implicit_scheme == 1 else "vic_full",
                "moved": torch.zeros(shape, dtype=torch.float64, device=self.device) if self.nscalar else None,
            }
        if self.nscalar:
            sshape = (self.nscalar,) + shape
            z = lambda: torch.zeros(sshape, dtype=torch.float64, device=self.device)  # noqa: E731
            self._scal = {
                "s0": z(),
                "theta": z(),
                "ds": z(),
                "flux": [z() if a else None for a in active],
            }

    def flux_register(self) -> FluxRegister:
        self._ensure_buffers()
        return self._freg

    # ------------------------------------------------------------ lifecycle
    def initialize(self, vars):
        """initialize_local + finalize_initialization
        (src/mesh/meshblock.cpp:306-343): outflow background, primitive
        boundary fill, prim -> cons."""
        if self.device.type != "cuda":
            raise _capi.SnapyError("initialize: call block.to('cuda') first")
        self._ensure_buffers()
        w = vars["hydro_w"]
        if tuple(w.shape) != self.shape():
            raise ValueError(f"initialize: hydro_w has shape {tuple(w.shape)}, expected {self.shape()}")
        # primitives are FP64 in both precision forms
        w = w.to(device=self.device, dtype=torch.float64).contiguous()
        vars["hydro_w"] = w
        if self.nscalar:
            r = vars.get("scalar_r")
            want = (self.nscalar,) + self.shape()[1:]
            if r is None or tuple(r.shape) != want:
                raise ValueError(f"initialize: scalar_r of shape {want} is required (nvar {self.nscalar} scalars)")
            vars["scalar_r"] = r.to(device=self.device, dtype=torch.float64).contiguous()
        if "outflow" in self.options.bcnames:
            vars["boundary_reference_w"] = w.clone()
        lib = _capi.load()
        dt = _capi.dtype_code(w.dtype)
        s = _capi.stream_of(self.device)
        bref = _capi.ptr(vars.get("boundary_reference_w"))
        _capi.check(
            lib.snapy_apply_boundaries_prim(
                ctypes.byref(self._grid), ctypes.byref(self._phys), dt, _capi.ptr(w), bref, s
            ),
            "apply_boundaries(prim)",
        )
        if self.nscalar:
            # seeded with the conserved (dry) density of the filled block on
            # every cell (finalize_initialization, meshblock.cpp:380-383)
            vars["scalar_s"] = w[IDN].unsqueeze(0) * vars["scalar_r"]
        if self.policy.fp32():
            # the FP32 state form: the pair (u, c) of the FP64 conserved state
            u = torch.empty(w.shape, dtype=torch.float32, device=self.device)
            comp = torch.empty_like(u)
            _capi.check(
                lib.snapy_prim2cons_f32(
                    ctypes.byref(self._grid), ctypes.byref(self._phys), _capi.ptr(w), _capi.ptr(u),
                    _capi.ptr(comp), s,
                ),
                "prim2cons_f32",
            )
            vars["hydro_u"] = u
            self._state.adopt(u, comp)
            self._w_stale = False
            self._w_interior_fresh = False
            return vars, 0.0
        u = torch.empty_like(w)
        _capi.check(
            lib.snapy_prim2cons(
                ctypes.byref(self._grid), ctypes.byref(self._phys), dt, _capi.ptr(w), _capi.ptr(u), s
            ),
            "prim2cons",
        )
        vars["hydro_u"] = u
        self._state.adopt(u)
        self._w_stale = False
        self._w_interior_fresh = False
        return vars, 0.0

    def write_restart(self, vars, path: str, time: float, extra: dict | None = None) -> str:
        """Write this block's restart part to `path` (RestartOutput::
        write_output_file, src/output/restart.cpp:42-147); see
        snapy_amd.restart for what it holds.  `extra` adds tensors (an output
        schedule, a budget carry)."""
        from . import restart

        if self.members > 1:
            raise restart.RestartError("restart: a block of members has no restart part (one per member block)")
        torch.cuda.synchronize(self.device)
        return restart.write_restart(self, vars, path, time, extra=extra)

    def initialize_from_restart(self, path_or_data):
        """initialize from a restart part (MeshBlockImpl::_init_from_restart,
        src/mesh/meshblock_restart.cpp:32-60): the metadata and grid checked,
        the stored state taken back with its own bits (the FP32 form's pair
        when the part holds one, else the FP64 totals split into the pair),
        the primitives and the outflow background as saved, the cycle
        restored.  Returns (vars, time)."""
        from . import restart

        if self.device.type != "cuda":
            raise _capi.SnapyError("initialize_from_restart: call block.to('cuda') first")
        if self.members > 1:
            raise restart.RestartError("restart: a block of members has no restart part (one per member block)")
        data = restart.load_restart(path_or_data) if isinstance(path_or_data, str) else path_or_data
        restart.check_restart_metadata(data, restart.grid_faces(self))
        if tuple(data["hydro_u"].shape) != self.shape():
            raise restart.RestartError(f"restart: hydro_u has shape {tuple(data['hydro_u'].shape)}, expected {self.shape()}")
        if "outflow" in self.options.bcnames and "boundary_reference_w" not in data:
            raise restart.RestartError("restart: an outflow block needs its boundary_reference_w")
        self._ensure_buffers()
        dev = self.device
        vars = {"hydro_w": data["hydro_w"].to(device=dev, dtype=torch.float64).contiguous()}
        if "boundary_reference_w" in data:
            vars["boundary_reference_w"] = data["boundary_reference_w"].to(device=dev, dtype=torch.float64).contiguous()
        stored = data["hydro_u_stored"]
        if self.policy.fp32():
            if stored.dtype == torch.float32:
                u = stored.to(dev).contiguous()
                comp = data["hydro_u_comp"].to(dev).contiguous()
            else:
                # an FP64 part resumed in the FP32 form: the pair of the totals
                t = data["hydro_u"].to(dev)
                u = t.float().contiguous()
                comp = (t - u.double()).float().contiguous()
            vars["hydro_u"] = u
            self._state.adopt(u, comp)
        else:
            u = data["hydro_u"].to(device=dev, dtype=torch.float64).contiguous()
            vars["hydro_u"] = u
            self._state.adopt(u)
        if self.nscalar:
            for k in ("scalar_s", "scalar_r"):
                if k not in data:
                    raise restart.RestartError(f"restart: a block with scalars needs its {k}")
                vars[k] = data[k].to(device=dev, dtype=torch.float64).contiguous()
        self._cycle = int(data["last_cycle"][0])
        self._w_stale = False
        self._w_interior_fresh = False
        return vars, float(data["last_time"][0])

    def _adopt(self, vars):
        u = vars["hydro_u"]
        if self._state.u() is None or self._state.u().data_ptr() != u.data_ptr():
            self._state.adopt(u)

    # ---------------------------------------------------------------- stage
    def forward(self, vars, dt, stage: int) -> None:
        """One RK stage of the block (MeshBlockImpl::forward,
        meshblock_advance.cpp:88-105): exchange (a single block has no
        neighbours), advance, combine, boundaries.  dt is a host float or a
        0-d FP64 device tensor (StepDt); either way the kernels read it from
        the block's device dt, so a captured stage replays with a new dt."""
        self._adopt(vars)
        self._ensure_buffers()
        if isinstance(dt, torch.Tensor):
            self._dt_dev.copy_(dt.to(self._dt_dev.device, torch.float64).expand_as(self._dt_dev))
        elif isinstance(dt, (list, tuple)):
            self._dt_dev.copy_(torch.tensor(dt, dtype=torch.float64))
        else:
            self._dt_dev.fill_(float(dt))
        self._executor.run_stage(vars, stage)
        self._w_stale = True
        self._w_interior_fresh = self._prims_from_finish()

    def _increment_in_x1_sweep(self) -> bool:
        """The FP32 pair form with x1 gravity (LMARS faces), a vertical
        implicit solve and no diffusion: the stage increment the solve reads
        is formed by the x1 sweep's launch (snapy_sweep_x1_wb_inc_f32), the
        bits of the separate increment launch, after the x2/x3 sweeps."""
        return bool(
            self.policy.fp32()
            and self._ref is not None
            and self._vic is not None
            and self._diff is None
            and self.nvar == 5
            and self.options.hydro.riemann == "lmars"
        )

    def _prims_from_finish(self) -> bool:
        """The FP32 pair form with the partial implicit solve (and no
        tracers, whose update follows the solve and reads the stage's
        primitives): the solve's finish writes the primitives of the
        combined pair into the interior of hydro_w, the values cons2prim
        would form at the next stage's start, so stages 1 and 2 start with
        the ghost fill alone and a refresh between steps fills the ghosts."""
        return bool(
            self.policy.fp32() and self._vic is not None and self._vic["scheme"] == 1 and not self.nscalar
        )

    def _refresh_primitives(self, vars) -> None:
        """hydro_w <- prim(hydro_u) on every cell, on the current stream
        (the reference's check_redo does it, meshblock_redo.cpp:107-112, so
        that the next time step reads the accepted state's primitives).  The
        FP64 form converts the boundary-filled state; the FP32 pair form
        converts u + c and fills the ghosts on the primitives, as its stages
        do."""
        lib = _capi.load()
        g, p = ctypes.byref(self._grid), ctypes.byref(self._phys)
        s = _capi.stream_of(self.device)
        P = _capi.ptr
        u, w = vars["hydro_u"], vars["hydro_w"]
        if self.policy.fp32():
            if not self._w_interior_fresh:
                _capi.check(lib.snapy_cons2prim_f32(g, p, P(u), P(self._state.comp()), P(w), s), "cons2prim")
            _capi.check(
                lib.snapy_apply_boundaries_prim(g, p, 0, P(w), P(vars.get("boundary_reference_w")), s),
                "apply_boundaries(prim)",
            )
        else:
            _capi.check(lib.snapy_cons2prim(g, p, _capi.dtype_code(u.dtype), P(u), P(w), s), "cons2prim")
        self._w_stale = False

    def _launch_stage(self, vars, stage: int) -> None:
        """The stage's launches on the current stream (the body a hipGraph
        captures)."""
        if stage == 0:
            self._state.backup()
            if self.nscalar:
                self._scal["s0"].copy_(vars["scalar_s"])
        for name, fn in self._stage_calls(vars, stage):
            _capi.check(fn(), name)

    def _stage_calls(self, vars, stage: int):
        """The stage's kernel launches in order, as (name, launch) pairs:
        EOS, the x1 reference and sweep, the x2/x3 sweeps, the step fields,
        the stage increment with its sources and (with an implicit scheme)
        the vertical solve, the RK combine, the boundaries."""
        lib = _capi.load()
        g, p = ctypes.byref(self._grid), ctypes.byref(self._phys)
        u = vars["hydro_u"]
        w = vars["hydro_w"]
        pair = self.policy.fp32()
        # dtc: the compute and flux precision (the primitives are FP64)
        dtc = _capi.dtype_code(self.policy.row("flux").storage)
        s = _capi.stream_of(self.device)
        freg = self._freg
        st = self.intg.stages[stage]
        freg.begin_stage(stage, st.wght1(), st.wght2())
        dtp = _capi.ptr(self._dt_dev)
        grav = self._ref is not None
        P = _capi.ptr
        calls = []
        c, c0 = self._state.comp(), self._state.comp0()
        bref = P(vars.get("boundary_reference_w"))
        if stage == 0 and self._vic is not None:
            # the implicit solve's step-level failure latch starts clear (a
            # fill kernel, so the clear replays inside a captured stage)
            st_words = self._vic["status"]
            calls.append(("vic_status_clear", lambda: (st_words.zero_(), 0)[1]))
        # (1) primitives
        fused = self._prims_from_finish()
        if pair:
            # the FP32 pair form fills the ghosts on the FP64 primitives (a
            # mirror, copy or radiating fill commutes with prim(u + c)); after
            # a stage whose implicit finish wrote the interior primitives (in
            # a step, and at a step's start when nothing changed the state
            # since), the ghost fill alone
            if not (fused and (stage > 0 or self._w_interior_fresh)):
                calls.append(("cons2prim", lambda: lib.snapy_cons2prim_f32(g, p, P(u), P(c), P(w), s)))
            calls.append(("apply_boundaries", lambda: lib.snapy_apply_boundaries_prim(g, p, 0, P(w), bref, s)))
        else:
            calls.append(("cons2prim", lambda: lib.snapy_cons2prim(g, p, dtc, P(u), P(w), s)))
        # (2)-(4) fused reconstruction + Riemann per active direction; with
        # x1 gravity the x1 sweep is the well-balanced one about the
        # hydrostatic reference of every column.  With the stage increment
        # fused into the x1 sweep, x2 and x3 go first (their fluxes are the
        # increment's input)
        inc_x1 = self._increment_in_x1_sweep()
        for d in ((2, 3, 1) if inc_x1 else (1, 2, 3)):
            t = freg.total(d)
            if t is None:
                continue
            fp = freg.face_pressure1 if d == 1 else None
            if d == 1 and inc_x1:
                vic = self._vic
                fl23 = [P(freg.total(2)), P(freg.total(3))]
                calls.append(
                    (
                        "sweep_x1_increment",
                        lambda t=t, fp=fp: lib.snapy_sweep_x1_wb_inc_f32(
                            g, p, P(w), P(self._ref), P(t), P(fp), fl23[0], fl23[1], P(self._x1tab),
                            P(vic["du"]), 0.0, dtp, s,
                        ),
                    )
                )
            elif d == 1 and grav:
                # the reference (kept on chip but for its face pressure) and
                # the well-balanced sweep in one launch
                calls.append(
                    (
                        "sweep_x1_wb",
                        lambda t=t, fp=fp: lib.snapy_sweep_x1_wb_ref(g, p, dtc, P(w), P(self._ref), P(t), P(fp), s),
                    )
                )
            elif d > 1 and self.options.coord.nx1 >= 3 and self.nvar == 5:
                # x2/x3 with the x1 flux covariance (on by default)
                calls.append(
                    (f"sweep_x{d}", lambda d=d, t=t: lib.snapy_sweep_cov(d, g, p, dtc, P(w), P(self._x1tab), P(t), s))
                )
            else:
                calls.append((f"sweep_x{d}", lambda d=d, t=t, fp=fp: lib.snapy_sweep(d, g, p, dtc, P(w), P(t), P(fp), s)))
        if self.nscalar:
            calls += self._scalar_flux_calls(vars, dtc, dtp, s)
        # the diffusion posts: the energy row joins the totals, the momentum
        # rows become the sources' du (before the step fields read the totals)
        dsrc = None
        if self._diff is not None:
            dif = self._diff
            tot = (ctypes.c_void_p * 3)(*[P(freg.total(d)) for d in (1, 2, 3)])
            calls.append(
                (
                    "diffusion",
                    lambda: lib.snapy_diffusion(g, p, dtc, P(w), tot, P(dif["work"]), P(dif["du"]), 0.0, dtp, s),
                )
            )
            dsrc = P(dif["du"])
        # FluxRegister::end_stage: S_d <- w1 S_d + w2 dt T_d
        for d in (1, 2, 3):
            if freg.step(d) is not None:
                calls.append(
                    (
                        f"step_field_x{d}",
                        lambda d=d: lib.snapy_step_field(
                            g, dtc, P(freg.total(d)), P(freg.step(d)), st.wght1(), st.wght2(), 0.0, dtp, s
                        ),
                    )
                )
        # (5)-(7) + combine: divergence, sources, implicit, u <- w0 u0 + w1 u + w2 du
        fl = (ctypes.c_void_p * 3)(*[P(freg.total(d)) for d in (1, 2, 3)])
        u0 = self._state.u0()
        if pair:
            ref = P(self._ref) if grav else None
            if self._vic is not None:
                vic = self._vic
                if not inc_x1:
                    calls.append(
                        (
                            "stage_increment",
                            lambda: lib.snapy_update_gravity_f32(
                                g, p, fl, ref, None, None, None, None, P(vic["du"]), P(w), P(self._x1tab), dsrc,
                                0.0, dtp, 1.0, 0.0, 1.0, s,
                            ),
                        )
                    )
                if fused:
                    vic_fn = lib.snapy_vic_partial_prim_f32
                else:
                    vic_fn = lib.snapy_vic_partial_f32 if vic["scheme"] == 1 else lib.snapy_vic_full_f32
                calls.append(
                    (
                        vic["name"],
                        lambda: vic_fn(
                            g, p, P(w), P(vic["du"]), P(u0), P(c0), P(u), P(c), P(vic["tab"]), P(self._x1tab),
                            vic["t0"], vic["t1"], 0.0, dtp, st.wght2(), st.wght0(), st.wght1(), st.wght2(),
                            P(vic["ws"]), P(vic["status"]), P(vic["moved"]), s,
                        ),
                    )
                )
            elif grav:
                calls.append(
                    (
                        "update_gravity",
                        lambda: lib.snapy_update_gravity_f32(
                            g, p, fl, ref, P(u0), P(c0), P(u), P(c), None, P(w), P(self._x1tab), dsrc, 0.0, dtp,
                            st.wght0(), st.wght1(), st.wght2(), s,
                        ),
                    )
                )
            else:
                calls.append(
                    (
                        "update",
                        lambda: lib.snapy_update_f32(
                            g, fl, P(u0), P(c0), P(u), P(c), dsrc, 0.0, dtp, st.wght0(), st.wght1(), st.wght2(), s
                        ),
                    )
                )
            if self.nscalar:
                calls += self._scalar_update_calls(vars, st, dtp, s)
            return calls
        if self._vic is not None:
            vic = self._vic
            calls.append(
                (
                    "stage_increment",
                    lambda: lib.snapy_update_gravity(
                        g, p, dtc, fl, None, P(vic["du"]), P(w), P(self._x1tab), dsrc, 0.0, dtp, 1.0, 0.0, 1.0, s
                    ),
                )
            )
            # the implicit solve takes the stage weight inside its dt and
            # finishes the RK combine
            vic_fn = lib.snapy_vic_partial if vic["scheme"] == 1 else lib.snapy_vic_full
            calls.append(
                (
                    vic["name"],
                    lambda: vic_fn(
                        g, p, P(w), P(vic["du"]), P(u0), P(u), P(vic["tab"]), P(self._x1tab), vic["t0"], vic["t1"],
                        0.0, dtp, st.wght2(), st.wght0(), st.wght1(), st.wght2(), P(vic["ws"]), P(vic["status"]),
                        P(vic["moved"]), s,
                    ),
                )
            )
        elif grav:
            calls.append(
                (
                    "update_gravity",
                    lambda: lib.snapy_update_gravity(
                        g, p, dtc, fl, P(u0), P(u), P(w), P(self._x1tab), dsrc, 0.0, dtp,
                        st.wght0(), st.wght1(), st.wght2(), s,
                    ),
                )
            )
        else:
            calls.append(
                (
                    "update",
                    lambda: lib.snapy_update(
                        g, dtc, fl, P(u0), P(u), dsrc, 0.0, dtp, st.wght0(), st.wght1(), st.wght2(), s
                    ),
                )
            )
        # boundaries on the combined state
        calls.append(
            (
                "apply_boundaries",
                lambda: lib.snapy_apply_boundaries(
                    g, p, dtc, P(u), P(vars.get("boundary_reference_w")), P(self._scratch_w), s
                ),
            )
        )
        if self.nscalar:
            calls += self._scalar_update_calls(vars, st, dtp, s)
        return calls

    def _scalar_cov(self) -> bool:
        """The scalars' fluxes carry the covariance row, and so the
        positivity limiter runs (scalar.cpp:146-160, 193-205)."""
        c = self.options.coord
        return bool(self.options.hydro.flux_covariance) and c.nc1() >= 3 and (c.nx2 > 1 or c.nx3 > 1)

    def _scalar_flux_calls(self, vars, dtc, dtp, s):
        """The scalars' face fluxes of the stage (ScalarImpl::forward,
        scalar.cpp:104-174) on the hydro's mass fluxes, then with the
        covariance the positivity limiter's theta (scalar.cpp:193-205)."""
        lib = _capi.load()
        g, p, P = ctypes.byref(self._grid), ctypes.byref(self._phys), _capi.ptr
        sc, freg, ns = self._scal, self._freg, self.nscalar
        w, sv = vars["hydro_w"], vars["scalar_s"]
        calls = []
        for d in (1, 2, 3):
            if sc["flux"][d - 1] is None:
                continue
            calls.append(
                (
                    f"scalar_flux_x{d}",
                    lambda d=d: lib.snapy_scalar_flux(
                        d, g, p, ns, P(w), P(sv), dtc, P(freg.total(d)), P(self._x1tab), P(sc["flux"][d - 1]), s
                    ),
                )
            )
        if self._scalar_cov():
            hf = (ctypes.c_void_p * 3)(*[P(freg.total(d)) for d in (1, 2, 3)])
            sf = (ctypes.c_void_p * 3)(*[P(f) for f in sc["flux"]])
            calls.append(
                (
                    "scalar_theta",
                    lambda: lib.snapy_scalar_theta(g, p, ns, P(w), P(sv), dtc, hf, sf, 0.0, dtp, P(sc["theta"]), s),
                )
            )
        return calls

    def _scalar_update_calls(self, vars, st, dtp, s):
        """ds = -dt div(F), the implicit tracer transfer, the RK combine and
        the ghosts of s (meshblock_advance.cpp:212-233, 338-340), then
        r = s / rho of the new state (set_scalar_primitive)."""
        lib = _capi.load()
        g, p, P = ctypes.byref(self._grid), ctypes.byref(self._phys), _capi.ptr
        sc, ns = self._scal, self.nscalar
        sv, rv, u = vars["scalar_s"], vars["scalar_r"], vars["hydro_u"]
        sf = (ctypes.c_void_p * 3)(*[P(f) for f in sc["flux"]])
        theta = P(sc["theta"]) if self._scalar_cov() else None
        moved = P(self._vic["moved"]) if self._vic is not None else None
        comp = self._state.comp()
        return [
            (
                "scalar_update",
                lambda: lib.snapy_scalar_update(
                    g, p, ns, sf, theta, P(rv), moved, P(sc["s0"]), P(sv), P(sc["ds"]), 0.0, dtp,
                    st.wght0(), st.wght1(), st.wght2(), s,
                ),
            ),
            (
                "scalar_prim",
                lambda: lib.snapy_scalar_prim(
                    g, p, ns, _capi.dtype_code(u.dtype), P(u), P(comp), P(sv), P(rv), s
                ),
            ),
        ]

    def device_time_step(self, vars) -> torch.Tensor:
        """max_time_step without the host read: the block's device dt
        (2^-redo * cfl * min dx/(|v|+c)), for a driver that never syncs."""
        self._ensure_buffers()
        if self._w_stale:
            self._refresh_primitives(vars)
        lib = _capi.load()
        w = vars["hydro_w"]
        scale = math.pow(2.0, -self.intg.current_redo) * self.intg.options.cfl
        _capi.check(
            lib.snapy_cfl(
                ctypes.byref(self._grid), ctypes.byref(self._phys), _capi.dtype_code(w.dtype),
                _capi.ptr(w), scale, _capi.ptr(self._dt_dev), _capi.stream_of(self.device),
            ),
            "cfl",
        )
        return self._dt_dev

    def max_time_step(self, vars):
        """global_time_step(local_max_time_step(vars))
        (meshblock_advance.cpp:61-85): 2^-redo * cfl * min dx/(|v|+c) over
        the interior of vars['hydro_w'] (one host read); with members, the
        list of the members' time steps."""
        self._ensure_buffers()
        if self._w_stale:
            self._refresh_primitives(vars)
        lib = _capi.load()
        w = vars["hydro_w"]
        out = torch.empty(self._dt_dev.shape, dtype=torch.float64, device=self.device)
        _capi.check(
            lib.snapy_cfl(
                ctypes.byref(self._grid), ctypes.byref(self._phys), _capi.dtype_code(w.dtype),
                _capi.ptr(w), 1.0, _capi.ptr(out), _capi.stream_of(self.device),
            ),
            "cfl",
        )
        scale = math.pow(2.0, -self.intg.current_redo) * self.intg.options.cfl
        if self.members > 1:
            return [scale * dt for dt in out.cpu().tolist()]
        return scale * float(out.item())

    # retry causes, the reference's bits (meshblock_redo.cpp:44-49)
    REDO_FLOOR, REDO_NAN, REDO_VIC = 1, 8, 32

    def redo_causes(self, vars) -> int:
        """The step's retry causes from one host read
        (MeshBlockImpl::local_redo_flag_tensor, meshblock_redo.cpp:98-160):
        density (and pressure) of hydro_w at or within 0.1% of the floor or
        not a number (1), a non-finite conserved interior (8), a column whose
        implicit solve failed in a stage of this step (32).  Reads hydro_w
        as it stands (check_redo refreshes it first)."""
        inner = self.part(exterior=False)
        w = vars["hydro_w"][inner]
        eos = self.options.hydro.eos
        floor = ~(w[IDN].min() > 1.001 * eos.density_floor)
        if self.nvar > IPR:
            floor = floor | ~(w[IPR].min() > 1.001 * eos.pressure_floor)
        u = vars["hydro_u"][inner]
        nan = ~torch.isfinite(u).all()
        c = self._state.comp()
        if c is not None:
            nan = nan | ~torch.isfinite(c[inner]).all()
        vic = torch.zeros((), dtype=torch.bool, device=self.device)
        if self._vic is not None:
            vic = (self._vic["status"] & 2).any()
        flags = torch.stack([floor, nan, vic]).cpu().tolist()
        return (self.REDO_FLOOR if flags[0] else 0) | (self.REDO_NAN if flags[1] else 0) | (
            self.REDO_VIC if flags[2] else 0
        )

    def check_redo(self, vars) -> int:
        """>0 redo, 0 ok, <0 stop (meshblock.hpp:311; check_redo ->
        decide_retry -> apply_redo, meshblock_redo.cpp:32-89,170-206).
        hydro_w is first refreshed from hydro_u.  On a retry cause the redo
        count goes up; past max_redo the block stops, at once for a floor or
        NaN cause and after restoring the state for a failed implicit
        solve.  A redo restores u (and the pair's compensation) from the
        step's backup and recomputes hydro_w from it, so the next time step
        (2^-redo of the CFL dt) reads the restored state."""
        self._adopt(vars)
        self._ensure_buffers()
        self._refresh_primitives(vars)
        causes = self.redo_causes(vars)
        self.last_redo_causes = causes
        if not causes:
            self.intg.current_redo = 0
            return 0
        self.intg.current_redo += 1
        if self.intg.current_redo > self.intg.max_redo and not (causes & self.REDO_VIC):
            return -1
        self._state.restore()
        self._w_interior_fresh = False
        if self.nscalar:
            # the scalars' backup too (StateStore::restore, state_store.hpp:
            # 62-65), and their primitives of the restored state
            vars["scalar_s"].copy_(self._scal["s0"])
            self._scalar_prim(vars)
        self._refresh_primitives(vars)
        if self.intg.current_redo > self.intg.max_redo:
            return -1
        return 1

    def _scalar_prim(self, vars) -> None:
        """scalar_r <- scalar_s / rho of hydro_u (interior, then its ghosts)."""
        lib = _capi.load()
        u = vars["hydro_u"]
        _capi.check(
            lib.snapy_scalar_prim(
                ctypes.byref(self._grid), ctypes.byref(self._phys), self.nscalar, _capi.dtype_code(u.dtype),
                _capi.ptr(u), _capi.ptr(self._state.comp()), _capi.ptr(vars["scalar_s"]),
                _capi.ptr(vars["scalar_r"]), _capi.stream_of(self.device),
            ),
            "scalar_prim",
        )

    def total_u(self, vars):
        """The conserved state; in the FP32 state form the FP64 value of the
        pair, u + c."""
        u = vars["hydro_u"] if isinstance(vars, dict) else vars
        c = self._state.comp()
        if c is not None and u.data_ptr() == self._state.u().data_ptr():
            return u.double() + c.double()
        return u.clone()

    def budget(self, vars) -> torch.Tensor:
        """FP64 domain totals sum(u * V) per row over the interior."""
        lib = _capi.load()
        u = vars["hydro_u"]
        n = lib.snapy_budget_workspace(ctypes.byref(self._grid))
        ws = torch.empty(int(n), dtype=torch.float64, device=self.device)
        out = torch.empty(self.nvar, dtype=torch.float64, device=self.device)
        parts = [u] + ([self._state.comp()] if self._state.comp() is not None else [])
        total = None
        for q in parts:
            _capi.check(
                lib.snapy_budget(
                    ctypes.byref(self._grid),
                    _capi.dtype_code(q.dtype),
                    _capi.ptr(q),
                    _capi.ptr(ws),
                    _capi.ptr(out),
                    _capi.stream_of(self.device),
                ),
                "budget",
            )
            total = out.clone() if total is None else total + out
        return total

    def execution_report(self) -> dict:
        return self._executor.report()

    def release(self, sync: bool = True) -> None:
        """Free the block's graphs and device buffers (a serving loop calls
        it when a request leaves).  sync=False: the caller has already
        synchronized the stream the block ran on."""
        if sync and self.device.type == "cuda":
            torch.cuda.synchronize(self.device)
        self._executor.release()
        self._freg = None
        self._ref = self._x1tab = self._vic = self._diff = None
        self._scratch_w = self._dt_dev = None
        self._state = StateStore()


class StageExecutor:
    """Runs a block's stages eagerly or as captured hipGraphs
    (StageExecutor::run_stage, src/mesh/stage_executor.cpp:809-1071).

    ``execution``: ``eager``, ``hip-graph`` (``cuda-graph`` is read as its
    HIP spelling) or ``auto`` (= eager).  A graph is captured per stage after
    one eager warm-up and keyed on the state's addresses and the StateStore
    generation (the reference's GraphKey, stage_executor.hpp:54-77); the dt
    is read from the block's device dt, so replays take a new dt without
    re-capture.
    """

    def __init__(self, block: "MeshBlock", execution: str = "auto"):
        if execution == "cuda-graph":
            execution = "hip-graph"
        if execution not in ("auto", "eager", "hip-graph"):
            raise ValueError(f"integration/execution '{execution}'")
        # a weak reference: a block and its executor form no cycle, so a block
        # (and its captured graphs) is freed when its last reference goes,
        # never by the cyclic collector in the middle of another capture
        self._block = weakref.ref(block)
        self.mode = "eager" if execution == "auto" else execution
        self._graphs = {}
        self._warm = set()
        self._counts = {"eager": 0, "captures": 0, "replays": 0}

    def _key(self, vars, stage):
        b = self.block
        return (
            stage,
            # a step's first stage without the primitives' conversion when
            # the last stage left them fresh (MeshBlock._prims_from_finish)
            stage == 0 and b._w_interior_fresh,
            b._state.generation,
            vars["hydro_u"].data_ptr(),
            vars["hydro_w"].data_ptr(),
            b._state.u0().data_ptr(),
            vars["scalar_s"].data_ptr() if "scalar_s" in vars else 0,
            vars["scalar_r"].data_ptr() if "scalar_r" in vars else 0,
        )

    @property
    def block(self) -> "MeshBlock":
        return self._block()

    def release(self) -> None:
        """Drop the captured graphs (after the device has finished them)."""
        self._graphs.clear()
        self._warm.clear()

    def run_stage(self, vars, stage: int) -> None:
        if self.mode == "eager":
            self.block._launch_stage(vars, stage)
            self._counts["eager"] += 1
            return
        key = self._key(vars, stage)
        g = self._graphs.get(key)
        if g is None:
            if key not in self._warm:
                # one eager warm-up before capture (the reference takes 3)
                self.block._launch_stage(vars, stage)
                self._warm.add(key)
                self._counts["eager"] += 1
                return
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize(self.block.device)
            # no collector run inside the capture: a finalizer that destroys a
            # graph or a stream is not permitted while a stream captures
            was = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.graph(g):
                    self.block._launch_stage(vars, stage)
            finally:
                if was:
                    gc.enable()
            self._graphs[key] = g
            self._counts["captures"] += 1
        g.replay()
        self._counts["replays"] += 1

    def report(self) -> dict:
        return dict(self._counts, mode=self.mode)
