"""Output of a block: NetCDF snapshots, the budget series and the cycle line.

Restates, for one Cartesian ideal-gas block:
- the deck's `outputs:` list (OutputOptionsImpl, src/options/output.yaml;
  the file id is `out<position>`, meshblock_options.cpp:78-84) and the
  schedule (MeshBlockImpl::make_outputs, src/mesh/meshblock_output.cpp:
  33-46: write when the time reaches the next output time, then advance it
  by dt; every output writes at the end of a run);
- the gridded writer's field selection (loadHydroOutputData,
  src/output/load_hydro_output_data.cpp; loadScalarOutputData,
  load_scalar_output_data.cpp) and its NetCDF layout (NetcdfSink::write,
  src/output/netcdf.cpp:66-235): `<basename>.block<rank>.<file_id>.<NNNNN>.nc`
  with the dimensions time (unlimited), x1, x2, x3 and the face dimensions
  x1f, x2f, x3f of the axes with cells; the coordinate variables with their
  axis, units, long_name and domain_decomposition attributes; the global
  NumFilesInSet; each field as (time, x1, x3, x2), NC_FLOAT unless
  double_precision, with the units and long_name of the reference's
  metadata table (output_utils.cpp:28-45);
- the budget consumer (BudgetConsumer, src/diagnostics/budget.cpp:58-254):
  per record the FP64 domain totals of every conserved row, their change
  over the interval, and the outward transfer through the domain's
  boundary faces (wall, outflow and extrapolation faces; never periodic
  ones) from the FP64 step fields S_d of the flux register; and its series
  file `<basename>.budget.nc` (BudgetSeries, src/output/budget_series.cpp:
  46-180: time, dt, interval, cycle, row_name, total, change, boundary);
- the cycle line (print_cycle_line, budget.cpp:283-394) without its
  run-to-date meters (this stage has no hydro positivity limiter to meter).

The files are NetCDF classic with 64-bit offsets, written with
scipy.io.netcdf_file: libnetcdf (the reference's NC_NETCDF4 container) is
not in this image.  The dimensions, variables, attributes and names are the
reference's, so a NetCDF reader sees the same dataset; classic NetCDF has no
64-bit integer, so the series' `cycle` is a double.  Not written: slices,
sums over an axis, ghost zones, statistics (prim_stat), user output
variables, the budget's counters and minima (a restart's budget carry holds
them empty).
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np
import torch

IDN, IVX, IVY, IVZ, IPR = 0, 1, 2, 3, 4

# short name -> (long name, units) of the reference's metadata table
METADATA = {
    "x1": ("height at cell center", "m"),
    "x1f": ("height at cell boundary", "m"),
    "x2": ("distance at cell center", "m"),
    "x2f": ("distance at cell boundary", "m"),
    "x3": ("distance at cell center", "m"),
    "x3f": ("distance at cell boundary", "m"),
    "rho": ("density", "kg/m^3"),
    "press": ("pressure", "Pa"),
    "vel1": ("vertical velocity", "m/s"),
    "vel2": ("horizontal velocity", "m/s"),
    "vel3": ("horizontal velocity", "m/s"),
}

SERIES_NAME_LENGTH = 48
SERIES_VERSION = 1


@dataclass
class OutputOptions:
    """One entry of the deck's `outputs:` list (src/options/output.yaml)."""

    fid: int = 0
    file_type: str = ""
    dt: float = 0.0
    variables: list = field(default_factory=list)
    double_precision: bool = False
    combine: bool = True

    def file_id(self) -> str:
        return f"out{self.fid}"


OUTPUT_KEYS = {"type", "dt", "variables", "double_precision", "combine", "output_sumx1", "output_sumx2",
               "output_sumx3", "include_ghost_zones", "x1_slice", "x2_slice", "x3_slice"}


def read_outputs(nodes) -> list:
    """The `outputs:` list, each entry's file id its position.  What the
    writers here do not do is refused by name."""
    out = []
    for fid, node in enumerate(nodes or []):
        node = node or {}
        unknown = set(node) - OUTPUT_KEYS
        if unknown:
            raise ValueError(f"outputs/{fid}: unknown keys {sorted(unknown)}")
        if "type" not in node:
            raise ValueError("OutputOptions::from_yaml: output file type must be specified")
        for k in ("output_sumx1", "output_sumx2", "output_sumx3", "include_ghost_zones"):
            if bool(node.get(k, False)):
                raise ValueError(f"outputs/{fid}/{k}: not implemented")
        for k in ("x1_slice", "x2_slice", "x3_slice"):
            if node.get(k):
                raise ValueError(f"outputs/{fid}/{k}: not implemented")
        o = OutputOptions(fid=fid, file_type=str(node["type"]), dt=float(node.get("dt", 0.0)),
                          variables=[str(v) for v in node.get("variables") or []],
                          double_precision=bool(node.get("double_precision", False)),
                          combine=bool(node.get("combine", True)))
        if o.file_type not in ("netcdf", "pnetcdf", "restart", "budget"):
            raise ValueError(f"Output type '{o.file_type}' is not implemented")
        out.append(o)
    return out


# ---------------------------------------------------------------------------
# fields
# ---------------------------------------------------------------------------
def _names(name: str, n: int) -> list:
    """Field::names (writer_core.cpp:55-67): one name for n components takes
    the component number (at '?' or the end)."""
    parts = name.split(";")
    if len(parts) >= n:
        return parts[:n]
    pos = name.find("?")
    head, tail = (name, "") if pos < 0 else (name[:pos], name[pos + 1 :])
    return [head + ("" if n == 1 else str(k + 1)) + tail for k in range(n)]


def gridded_fields(block, vars, variables) -> list:
    """(name, FP64 host array (nc3, nc2, nc1)) of the requested variables, in
    the reference's order (loadHydroOutputData, loadScalarOutputData)."""
    want = set(variables)
    prim = lambda *k: "prim" in want or any(x in want for x in k)  # noqa: E731
    cons = lambda *k: "cons" in want or any(x in want for x in k)  # noqa: E731
    if block._w_stale:
        block._refresh_primitives(vars)
    w = vars["hydro_w"].detach().to("cpu", torch.float64).numpy()
    u = block.total_u(vars).detach().to("cpu", torch.float64).numpy()
    out = []

    def add(name, t, start, count):
        for k, nm in enumerate(_names(name, count)):
            out.append((nm, t[start + k]))

    if cons("D"):
        add("dens", u, IDN, 1)
    if prim("d"):
        add("rho", w, IDN, 1)
    if block.nvar > 4:
        if cons("E"):
            add("Etot", u, IPR, 1)
        if prim("p"):
            add("press", w, IPR, 1)
    if cons("m"):
        add("mom", u, IVX, 3)
    for k, key in enumerate(("m1", "m2", "m3")):
        if key in want:
            add(f"mom{k + 1}", u, IVX + k, 1)
    if prim("v"):
        add("vel", w, IVX, 3)
    for k, keys in enumerate((("vx", "v1"), ("vy", "v2"), ("vz", "v3"))):
        if any(x in want for x in keys):
            add(f"vel{k + 1}", w, IVX + k, 1)
    ns = getattr(block, "nscalar", 0)
    if ns and ("scalar_r" in vars or "scalar_s" in vars):
        r = vars["scalar_r"].detach().to("cpu", torch.float64).numpy()
        s = vars["scalar_s"].detach().to("cpu", torch.float64).numpy()
        every = "scalar" in want
        names = block.options.scalar.names
        for n in range(ns):
            tag = f"_{names[n]}" if n < len(names) and names[n] else str(n)
            nc, nr = "s" + tag, "r" + tag
            if every or "scalar_cons" in want or nc in want:
                out.append((nc, s[n]))
            if every or "scalar_prim" in want or nr in want:
                out.append((nr, r[n]))
    return out


def sanitize(name: str) -> str:
    """sanitize_netcdf_name (writer_core.cpp:239-247)."""
    s = "".join(c if (c.isalnum() or c == "_") else "_" for c in name) or "_"
    if not (s[0].isalpha() or s[0] == "_"):
        s = "_" + s
    return s


def write_netcdf(block, vars, opt: OutputOptions, file_number: int, time: float, basename: str,
                 output_dir: str = ".", rank: int = 0) -> str:
    """One snapshot file (NetcdfSink::write); returns its path."""
    from scipy.io import netcdf_file

    c = block.options.coord
    ng = c.nghost
    nx = (c.nx1, c.nx2, c.nx3)
    lo = [ng if n > 1 else 0 for n in nx]
    fname = os.path.join(output_dir, f"{basename}.block{rank}.{opt.file_id()}.{file_number:05d}.nc")
    os.makedirs(output_dir, exist_ok=True)
    dt_code = "d" if opt.double_precision else "f"
    npdt = np.float64 if opt.double_precision else np.float32
    hunits = "radian" if c.type in ("gnomonic-equiangle", "spherical-polar") else "m"
    fields = gridded_fields(block, vars, opt.variables)
    cb = block.module("coord")
    with netcdf_file(fname, "w", version=2) as f:
        f.createDimension("time", None)
        for d in range(3):
            f.createDimension(f"x{d + 1}", nx[d])
            if nx[d] > 1:
                f.createDimension(f"x{d + 1}f", nx[d] + 1)
        tv = f.createVariable("time", dt_code, ("time",))
        tv.axis, tv.units, tv.long_name = "T", "s", "time"
        for d in range(3):
            axis = ("Z", "X", "Y")[d]
            name = f"x{d + 1}"
            v = f.createVariable(name, dt_code, (name,))
            v.axis = axis
            v.units = hunits if d else "m"
            v.long_name = f"{axis}-coordinate at cell center"
            v.domain_decomposition = np.array([1, nx[d], 1, nx[d]], dtype=np.int32)
            v[:] = cb.buffer(f"x{d + 1}v").numpy()[lo[d] : lo[d] + nx[d]].astype(npdt)
            if nx[d] > 1:
                vf = f.createVariable(name + "f", dt_code, (name + "f",))
                vf.units = hunits if d else "m"
                vf.long_name = f"{axis}-coordinate at cell face"
                vf.domain_decomposition = np.array([0, nx[d], 0, nx[d]], dtype=np.int32)
                vf[:] = cb.buffer(f"x{d + 1}f").numpy()[lo[d] : lo[d] + nx[d] + 1].astype(npdt)
        f.NumFilesInSet = np.int32(1)
        tv[0] = npdt(time)
        for name, a in fields:
            v = f.createVariable(sanitize(name), dt_code, ("time", "x1", "x3", "x2"))
            meta = METADATA.get(name)
            if meta:
                v.long_name, v.units = meta
            # cells: i outermost, then k, then j (Frame::pack)
            cells = a[lo[2] : lo[2] + nx[2], lo[1] : lo[1] + nx[1], lo[0] : lo[0] + nx[0]]
            v[0] = np.transpose(cells, (2, 0, 1)).astype(npdt)
    return fname


# ---------------------------------------------------------------------------
# budget
# ---------------------------------------------------------------------------
def row_names(nvar: int) -> list:
    """names_of (budget.cpp:29-45) for the conserved layout."""
    rows = []
    for n in range(nvar):
        if n == IDN:
            rows.append("mass")
        elif n in (IVX, IVY, IVZ):
            rows.append(f"momentum{n - IVX + 1}")
        elif n == IPR:
            rows.append("energy")
        else:
            rows.append(f"species/{n - 5 + 1}")
    return rows


@dataclass
class BudgetRecord:
    rows: list
    total: np.ndarray
    change: np.ndarray
    boundary: np.ndarray
    time: float = 0.0
    interval: float = 0.0
    dt: float = 0.0
    cycle: int = 0


class BudgetConsumer:
    """The budget consumer of one block: observe after every accepted step,
    flush at a record.  The block's flux register keeps the FP64 step fields
    S_d of its active directions (requested here), whose boundary faces give
    the outward transfer."""

    def __init__(self, block):
        self.block = block
        block._ensure_buffers()
        reg = block.flux_register()
        for d in (1, 2, 3):
            if reg.total(d) is not None and reg.step(d) is None:
                reg.request_step_field(d)
                # a new buffer in the stage: a captured stage is taken again
                block._state.register_storage(f"step_field_x{d}", reg.step(d))
        self._started = False
        self._weights = None

    def _boundary_faces(self, d):
        """The outward sign of the two end planes of direction d's faces:
        -1 below, +1 above, on the domain's boundary faces that a boundary
        function acts on without wrapping (budget.cpp:47-56)."""
        names = self.block.options.bcnames
        lo = names[2 * (d - 1)] not in ("periodic",)
        hi = names[2 * (d - 1) + 1] not in ("periodic",)
        return lo, hi

    def boundary(self) -> torch.Tensor:
        """The outward transfer of the step's fluxes through the boundary
        faces, per row (budget.cpp:75-158)."""
        b = self.block
        c = b.options.coord
        reg = b.flux_register()
        ng = c.nghost
        nx = (c.nx1, c.nx2, c.nx3)
        lo = [ng if n > 1 else 0 for n in nx]
        area = {1: c.dx(3) * c.dx(2), 2: c.dx(3) * c.dx(1), 3: c.dx(2) * c.dx(1)}
        out = torch.zeros(b.nvar, dtype=torch.float64, device=b.device)
        for d in (1, 2, 3):
            S = reg.step(d)
            if S is None:
                continue
            ax = 4 - d  # tensor axis of (nvar, nc3, nc2, nc1)
            idx = [slice(None)] + [slice(lo[2], lo[2] + nx[2]), slice(lo[1], lo[1] + nx[1]),
                                   slice(lo[0], lo[0] + nx[0])]
            flo, fhi = self._boundary_faces(d)
            if flo:
                i = list(idx)
                i[ax] = slice(lo[d - 1], lo[d - 1] + 1)
                out = out - S[tuple(i)].sum(dim=(1, 2, 3)) * area[d]
            if fhi:
                i = list(idx)
                i[ax] = slice(lo[d - 1] + nx[d - 1], lo[d - 1] + nx[d - 1] + 1)
                out = out + S[tuple(i)].sum(dim=(1, 2, 3)) * area[d]
        return out

    def _start(self, vars):
        self._total_start = self.block.budget(vars).clone()
        self._pending = torch.zeros_like(self._total_start)
        self._interval = 0.0
        self._started = True

    def observe(self, vars, dt: float) -> None:
        """After an accepted step of length dt (BudgetConsumer::observe): the
        first observe starts the interval after its own step."""
        self._dt = float(dt)
        if not self._started:
            self._start(vars)
            return
        self._pending = self._pending + self.boundary()
        self._interval += float(dt)

    def carry(self, vars, fires_now: bool):
        """BudgetConsumer::carry (budget.cpp:214-234): the open interval's
        tensors for a restart, or None before the first observe; a record
        firing at this time closes the interval first."""
        from .restart import carry_tensors

        if not self._started:
            return None
        rows = row_names(self.block.nvar)
        if fires_now:
            total = self.block.budget(vars)
            return carry_tensors(rows, total, torch.zeros_like(self._pending), 0.0)
        return carry_tensors(rows, self._total_start, self._pending, self._interval)

    def resume(self, data: dict) -> None:
        """BudgetConsumer::resume (budget.cpp:236-245): an interval carried by
        a restart with the same rows continues; otherwise a new one starts at
        the first observe."""
        from .restart import read_carry

        c = read_carry(data, row_names(self.block.nvar))
        if c is None:
            return
        dev = self.block.device
        self._total_start = c[0].to(dev)
        self._pending = c[1].to(dev)
        self._interval = c[2]
        self._started = True

    def flush(self, vars, time: float) -> BudgetRecord:
        """Close the interval (BudgetConsumer::flush)."""
        if not self._started:
            self._start(vars)
        total = self.block.budget(vars)
        rec = BudgetRecord(
            rows=row_names(self.block.nvar),
            total=total.cpu().numpy().copy(),
            change=(total - self._total_start).cpu().numpy().copy(),
            boundary=self._pending.cpu().numpy().copy(),
            time=float(time),
            interval=self._interval,
            dt=getattr(self, "_dt", 0.0),
            cycle=int(self.block.cycle()),
        )
        self._total_start = total.clone()
        self._pending.zero_()
        self._interval = 0.0
        return rec


def append_budget(path: str, rec: BudgetRecord, new: bool) -> None:
    """BudgetSeries::append (budget_series.cpp:101-160): the file is created
    with the record's names at the run's first record, then each record is
    one more time slice."""
    from scipy.io import netcdf_file

    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    if new or not os.path.exists(path):
        with netcdf_file(path, "w", version=2) as f:
            f.createDimension("time", None)
            f.createDimension("row", len(rec.rows))
            f.createDimension("nchar", SERIES_NAME_LENGTH)
            for name in ("time", "dt", "interval", "cycle"):
                f.createVariable(name, "d", ("time",))
            rn = f.createVariable("row_name", "c", ("row", "nchar"))
            names = np.zeros((len(rec.rows), SERIES_NAME_LENGTH), dtype="S1")
            for n, s in enumerate(rec.rows):
                b = s.encode()[:SERIES_NAME_LENGTH]
                names[n, : len(b)] = np.frombuffer(b, dtype="S1")
            rn[:] = names
            for name in ("total", "change", "boundary"):
                f.createVariable(name, "d", ("time", "row"))
            f.snapy_budget_version = np.int32(SERIES_VERSION)
    with netcdf_file(path, "a", version=2) as f:
        n = f.variables["time"].shape[0]
        f.variables["time"][n] = rec.time
        f.variables["dt"][n] = rec.dt
        f.variables["interval"][n] = rec.interval
        f.variables["cycle"][n] = float(rec.cycle)
        f.variables["total"][n, :] = rec.total
        f.variables["change"][n, :] = rec.change
        f.variables["boundary"][n, :] = rec.boundary


# ---------------------------------------------------------------------------
# the schedule
# ---------------------------------------------------------------------------
class BlockOutputs:
    """The deck's outputs of one block (MeshBlockImpl::make_outputs): call
    `observe(vars, dt)` after every accepted step and `make_outputs(vars,
    time)` after it; `finalize(vars, time)` writes every output once more.
    Restart outputs write restart files (snapy_amd.restart); `resume(data,
    time)` takes the schedule and the budget intervals back from one."""

    def __init__(self, block, outputs=None, basename=None, output_dir="."):
        self.block = block
        self.outputs = list(block.options.outputs if outputs is None else outputs)
        if basename is None:
            fn = os.path.basename(block.options.filename or "snapy")
            basename = fn.rsplit(".", 1)[0] if "." in fn else fn
        self.basename = basename
        self.output_dir = output_dir
        self.next_time = [0.0] * len(self.outputs)
        self.file_number = [0] * len(self.outputs)
        self.budget = {k: BudgetConsumer(block) for k, o in enumerate(self.outputs) if o.file_type == "budget"}
        self._series_started = set()
        self.written = []

    def observe(self, vars, dt: float) -> None:
        for b in self.budget.values():
            b.observe(vars, dt)

    def _write_restart(self, k, vars, time, final):
        """RestartOutput::write_output_file (restart.cpp:42-147): the part
        with the advanced schedule of this output and of the outputs after it
        that fire in the same pass, and each budget output's open interval;
        then the parts combined into the restart file."""
        from . import restart

        file_number, next_time = [], []
        firing = False
        extra = {}
        for j, oj in enumerate(self.outputs):
            if j == k:
                firing = not final
            advance = firing and time >= self.next_time[j]
            file_number.append(self.file_number[j] + (1 if advance else 0))
            next_time.append(self.next_time[j] + (oj.dt if advance else 0.0))
            if j in self.budget:
                c = self.budget[j].carry(vars, advance)
                if c is not None:
                    extra.update(c)
        extra.update(restart.schedule_tensors(self.outputs, file_number, next_time))
        number = None if final else self.file_number[k]
        self.block.write_restart(vars, os.path.join(self.output_dir, restart.part_name(self.basename, 0, number)),
                                 time, extra=extra)
        return restart.combine_parts(self.output_dir, self.basename, number)

    def resume(self, data: dict, time: float) -> None:
        """The schedule and the budget intervals of a restart
        (meshblock_restart.cpp:61-125)."""
        from . import restart

        self.file_number, self.next_time = restart.restore_schedule(data, self.outputs, time)
        for b in self.budget.values():
            b.resume(data)

    def _write(self, k, vars, time, final=False):
        o = self.outputs[k]
        if o.file_type in ("netcdf", "pnetcdf"):
            p = write_netcdf(self.block, vars, o, self.file_number[k], time, self.basename, self.output_dir)
        elif o.file_type == "restart":
            p = self._write_restart(k, vars, time, final)
        else:
            rec = self.budget[k].flush(vars, time)
            p = os.path.join(self.output_dir, f"{self.basename}.budget.nc")
            append_budget(p, rec, p not in self._series_started)
            self._series_started.add(p)
        self.written.append(p)
        return p

    def make_outputs(self, vars, time: float, final: bool = False) -> None:
        for k, o in enumerate(self.outputs):
            if final:
                self._write(k, vars, time, final=True)
            elif time >= self.next_time[k]:
                self._write(k, vars, time)
                self.next_time[k] += o.dt
                self.file_number[k] += 1

    def finalize(self, vars, time: float) -> None:
        self.make_outputs(vars, time, final=True)


def cycle_line(block, vars, time: float, dt: float, precision: int = 13, energy_label: str = " ie=") -> str:
    """The cycle line (print_cycle_line, budget.cpp:283-394) of one block:
    cycle, redo, time, dt, the conserved mass and energy totals, the kinetic
    energy and, with gravity, sum rho (-g1) x1 V.  The corrected-PE part of
    pe and the run-to-date meters are not formed."""
    fmt = lambda v: f"{v:.{precision}e}"  # noqa: E731
    s = f"cycle={block.cycle()} redo={block.intg.current_redo} time={fmt(time)} dt={fmt(dt)}"
    c = block.options.coord
    tot = block.budget(vars).cpu().numpy()
    u = block.total_u(vars).to(torch.float64)
    inner = block.part(exterior=False)
    ui = u[inner]
    vol = c.dx(3) * (c.dx(2) * c.dx(1))
    rho = ui[IDN]
    ke = float((0.5 * (ui[IVX : IVX + 3] * ui[IVX : IVX + 3]).sum(0) / rho * vol).sum())
    s += f" mass0={fmt(tot[IDN])} ke={fmt(ke)}"
    if block.nvar > IPR:
        s += f"{energy_label}{fmt(tot[IPR])}"
    g1 = block.options.hydro.grav1
    if g1 != 0.0:
        x1v = block.module("coord").buffer("x1v").to(u.device)[inner[3]]
        pe = float((rho * (-g1 * x1v) * vol).sum())
        s += f" pe={fmt(pe)}"
    return s
