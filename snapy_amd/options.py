"""Deck readers: MeshBlockOptions.from_yaml and the option structs it fills.

Restates the YAML surface the reference reads for this path
(MeshBlockOptionsImpl::from_yaml, src/mesh/meshblock_options.cpp:18-77, and
the per-module readers it calls).  Keys the stage does not implement raise a
ValueError naming the key, as the reference's strict key checks do
(src/input/check_keys.cpp).  The `scalar` section (passive tracers,
ScalarOptionsImpl::from_yaml, src/scalar/scalar_options.cpp:11-57) is read;
sections that would add state rows or physics the stage does not carry are
refused by name: `species` beyond the single dry gas, `reactions`, and a
`distribute` that splits the domain over blocks.  Sections read by the driver rather than the
stage (`problem`) and the dry gas's inert `reference-state` are accepted; the
`outputs` list is read into `MeshBlockOptions.outputs` (snapy_amd.output).
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import yaml

# the gas constant kintera uses (kintera::constants::Rgas), J/(mol K)
RGAS = 8.31446261815324


def _check_keys(node, path, allowed):
    if node is None:
        return
    if not isinstance(node, dict):
        raise ValueError(f"{path} must be a map")
    bad = [k for k in node if k not in allowed]
    if bad:
        raise ValueError(f"{path}: unsupported key(s) {bad}; allowed {sorted(allowed)}")


def _pi(v):
    """Numbers may be written with a _pi alias (options/coord.yaml)."""
    if isinstance(v, str):
        s = v.strip()
        if s.endswith("_pi"):
            return float(s[:-3] or 1.0) * math.pi
        return float(s)
    return float(v)


@dataclass
class CoordinateOptions:
    """geometry: (src/coord/coordinate.cpp, src/options/coord.yaml)."""

    type: str = "cartesian"
    x1min: float = 0.0
    x1max: float = 1.0
    x2min: float = 0.0
    x2max: float = 1.0
    x3min: float = 0.0
    x3max: float = 1.0
    nx1: int = 1
    nx2: int = 1
    nx3: int = 1
    nghost: int = 3
    interp_order: int = 2  # cells/interp_order: panel-edge ghost interpolation

    def nc(self, d: int) -> int:
        nx = (self.nx1, self.nx2, self.nx3)[d - 1]
        return nx + 2 * self.nghost if nx > 1 else 1

    def nc1(self):
        return self.nc(1)

    def nc2(self):
        return self.nc(2)

    def nc3(self):
        return self.nc(3)

    def dx(self, d: int) -> float:
        lo, hi, n = [
            (self.x1min, self.x1max, self.nx1),
            (self.x2min, self.x2max, self.nx2),
            (self.x3min, self.x3max, self.nx3),
        ][d - 1]
        return (hi - lo) / n


@dataclass
class EquationOfStateOptions:
    type: str = "ideal-gas"
    gammad: float = 1.4
    weight: float = 29.0e-3
    density_floor: float = 1.0e-10
    pressure_floor: float = 1.0e-10
    temperature_floor: float = 20.0
    limiter: bool = False

    def gas_constant_dry(self) -> float:
        # EquationOfStateImpl::gas_constant_dry (equation_of_state.cpp:213-215)
        return RGAS / self.weight


@dataclass
class ReconstructOptions:
    type: str = "weno5"
    scale: bool = False
    shock: bool = True


@dataclass
class HydroOptions:
    eos: EquationOfStateOptions = field(default_factory=EquationOfStateOptions)
    recon1: ReconstructOptions = field(default_factory=ReconstructOptions)
    recon23: ReconstructOptions = field(default_factory=ReconstructOptions)
    riemann: str = "hllc"
    riemann_dir: str = "xy"  # riemann-solver/dir of shallow-roe
    grav1: float = 0.0
    bot_heat_flux: float = 0.0
    bot_heat_depth: int = 1
    precision: str = "float64"
    flux_covariance: bool = True
    # forcing/diffusion, kinematic isotropic coefficients (m^2/s)
    nu_iso: float = 0.0
    kappa_iso: float = 0.0


@dataclass
class IntegratorOptions:
    type: str = "rk3"
    cfl: float = 0.9
    tlim: float = 1.0
    nlim: int = -1
    implicit_scheme: int = 0
    implicit_advection_cfl: float = 1.0
    execution: str = "auto"


@dataclass
class ScalarOptions:
    """ScalarOptionsImpl (src/scalar/scalar.hpp:24-58): nvar passive tracers
    named `names`, reconstructed with the horizontal reconstruction's type
    scale-invariant (meshblock_options.cpp:37-40, scalar.cpp:31-37) and
    transported by the upwind solver on the hydro mass flux."""

    nvar: int = 0
    names: list = field(default_factory=list)
    upper_bound: float = -1.0
    recon_type: str = "weno5"
    riemann: str = "upwind"


@dataclass
class MeshBlockOptions:
    coord: CoordinateOptions = field(default_factory=CoordinateOptions)
    hydro: HydroOptions = field(default_factory=HydroOptions)
    intg: IntegratorOptions = field(default_factory=IntegratorOptions)
    scalar: ScalarOptions = field(default_factory=ScalarOptions)
    # the deck's outputs (snapy_amd.output.OutputOptions), in deck order
    outputs: list = field(default_factory=list)
    bcnames: list = field(default_factory=lambda: ["reflecting"] * 6)
    verbose: bool = False
    filename: str = ""

    @staticmethod
    def from_yaml(filename: str, verbose: bool = False) -> "MeshBlockOptions":
        """snapy.MeshBlockOptions.from_yaml(filename, verbose=False)
        (python/csrc/pymesh.cpp:28; meshblock_options.cpp:18-77)."""
        with open(filename) as f:
            cfg = yaml.safe_load(f) or {}
        op = MeshBlockOptions(verbose=verbose, filename=os.fspath(filename))
        _check_keys(
            cfg,
            "deck",
            {
                "geometry",
                "dynamics",
                "boundary-condition",
                "integration",
                "forcing",
                "outputs",
                "species",
                "reference-state",
                "distribute",
                "problem",
                "scalar",
                "reactions",
            },
        )
        _check_sections(cfg)
        dist = cfg.get("distribute") or {}
        _check_keys(dist, "distribute", _DIST_KEYS)
        for k in ("nb1", "nb2", "nb3"):
            if int(dist.get(k, 1)) != 1:
                raise ValueError(
                    f"distribute/{k} = {dist[k]}: a MeshBlock here covers the whole domain "
                    "(nb1 = nb2 = nb3 = 1; the cubed sphere goes through MeshOptions)"
                )
        op.coord = _read_geometry(cfg.get("geometry") or {})
        op.hydro = _read_dynamics(cfg.get("dynamics") or {}, cfg.get("forcing") or {})
        op.intg = _read_integration(cfg.get("integration") or {})
        op.bcnames = _read_bcs(cfg.get("boundary-condition") or {})
        op.scalar = _read_scalar(cfg.get("scalar"), op.hydro.recon23)
        from .output import read_outputs

        op.outputs = read_outputs(cfg.get("outputs"))
        _check_supported(op)
        return op


_DIST_KEYS = {"layout", "nb1", "nb2", "nb3", "blocks_per_process", "verbose"}


def _read_scalar(node, recon23) -> ScalarOptions:
    """ScalarOptionsImpl::from_yaml (scalar_options.cpp:11-57): nvar, names
    (their count sets nvar), upper-bound, reconstruct {type, shock} (default:
    the horizontal reconstruction), riemann-solver (upwind)."""
    sc = ScalarOptions()
    sc.recon_type = recon23.type
    if not node:
        return sc
    _check_keys(node, "scalar", {"verbose", "nvar", "upper-bound", "names", "reconstruct", "riemann-solver"})
    sc.nvar = int(node.get("nvar", 0))
    sc.upper_bound = float(node.get("upper-bound", sc.upper_bound))
    if node.get("names"):
        sc.names = [str(x) for x in node["names"]]
        sc.nvar = len(sc.names)
    rc = node.get("reconstruct")
    if rc is not None:
        _check_keys(rc, "scalar/reconstruct", {"shock", "type"})
        sc.recon_type = rc.get("type", sc.recon_type)
    rs = node.get("riemann-solver")
    if rs is not None:
        _check_keys(rs, "scalar/riemann-solver", {"type", "dir"})
        sc.riemann = rs.get("type", "upwind")
    if not sc.names:
        sc.names = [f"scalar{n}" for n in range(sc.nvar)]
    return sc


def _check_scalar(op) -> None:
    """What the tracer stage carries (ScalarImpl::reset, scalar.cpp:21-48,
    plus this build's scope)."""
    sc = op.scalar
    if sc.nvar <= 0:
        return
    if sc.riemann != "upwind":
        raise ValueError("[Scalar] Passive scalars currently require the upwind solver")
    if sc.upper_bound == 0.0:
        raise ValueError("[Scalar] upper-bound = 0 pins the tracer to zero and freezes transport; "
                         "use a negative value to disable the bound")
    if sc.upper_bound > 0.0:
        raise ValueError("[Scalar] upper-bound needs the eos limiter, which is not implemented")
    if sc.recon_type != "weno5":
        raise ValueError(f"scalar/reconstruct/type '{sc.recon_type}': weno5 only")
    if op.hydro.eos.type != "ideal-gas" or op.coord.type != "cartesian":
        raise ValueError("scalar: tracers ride the Cartesian ideal-gas stage")
    if "outflow" in op.bcnames:
        raise ValueError("scalar: a radiating (outflow) boundary for the tracers is not implemented")


def _check_sections(cfg) -> None:
    """The deck sections that would add state rows or physics this stage
    does not carry are refused by name: species beyond the single dry gas,
    reactions.  `outputs` (I/O) and `problem` (driver parameters) are read
    by the caller, not the stage; `reference-state` is the thermodynamic
    reference of moist species, inert for the dry ideal gas."""
    sp = cfg.get("species")
    if sp is not None:
        if not isinstance(sp, list):
            raise ValueError("species: a list of species")
        if len(sp) > 1:
            names = [str((x or {}).get("name", "?")) for x in sp]
            raise ValueError(f"species: {names}: only the single dry species is implemented")
    if cfg.get("reactions"):
        raise ValueError("reactions: chemistry is not implemented")


def _read_geometry(g) -> CoordinateOptions:
    _check_keys(g, "geometry", {"type", "bounds", "cells"})
    c = CoordinateOptions()
    c.type = g.get("type", "cartesian")
    b = g.get("bounds") or {}
    names = ("x1min", "x1max", "x2min", "x2max", "x3min", "x3max")
    _check_keys(b, "geometry/bounds", set(names) | {k + "_pi" for k in names})
    for k in names:
        if k in b:
            setattr(c, k, _pi(b[k]))
        # the _pi alias, scaled by pi; given both, the alias wins
        if k + "_pi" in b:
            setattr(c, k, float(b[k + "_pi"]) * math.pi)
    cells = g.get("cells") or {}
    _check_keys(cells, "geometry/cells", {"nx1", "nx2", "nx3", "nghost", "interp_order"})
    for k in ("nx1", "nx2", "nx3", "nghost", "interp_order"):
        if k in cells:
            setattr(c, k, int(cells[k]))
    return c


def _read_recon(node, path) -> ReconstructOptions:
    r = ReconstructOptions()
    if node is None:
        return r
    _check_keys(node, path, {"type", "scale", "shock"})
    r.type = node.get("type", r.type)
    r.scale = bool(node.get("scale", r.scale))
    r.shock = bool(node.get("shock", r.shock))
    return r


def _read_dynamics(d, forcing) -> HydroOptions:
    _check_keys(
        d,
        "dynamics",
        {
            "equation-of-state",
            "reconstruct",
            "riemann-solver",
            "precision",
            "flux-covariance",
            "verbose",
        },
    )
    h = HydroOptions()
    e = d.get("equation-of-state") or {}
    _check_keys(
        e,
        "dynamics/equation-of-state",
        {"type", "gammad", "weight", "density-floor", "pressure-floor", "temperature-floor", "limiter"},
    )
    h.eos.type = e.get("type", h.eos.type)
    h.eos.gammad = float(e.get("gammad", h.eos.gammad))
    h.eos.weight = float(e.get("weight", h.eos.weight))
    h.eos.density_floor = float(e.get("density-floor", h.eos.density_floor))
    h.eos.pressure_floor = float(e.get("pressure-floor", h.eos.pressure_floor))
    h.eos.limiter = bool(e.get("limiter", h.eos.limiter))
    # read for the record: the floor acts only inside the conserved limiter
    # (EquationOfStateImpl::apply_conserved_limiter_, equation_of_state.cpp:
    # 305-360), and limiter: true is refused here
    h.eos.temperature_floor = float(e.get("temperature-floor", h.eos.temperature_floor))
    rc = d.get("reconstruct") or {}
    _check_keys(rc, "dynamics/reconstruct", {"vertical", "horizontal"})
    h.recon1 = _read_recon(rc.get("vertical"), "dynamics/reconstruct/vertical")
    h.recon23 = _read_recon(rc.get("horizontal"), "dynamics/reconstruct/horizontal")
    rs = d.get("riemann-solver") or {}
    _check_keys(rs, "dynamics/riemann-solver", {"type", "dir"})
    h.riemann = rs.get("type", h.riemann)
    h.riemann_dir = str(rs.get("dir", h.riemann_dir))
    h.precision = d.get("precision", h.precision)
    h.flux_covariance = bool(d.get("flux-covariance", h.flux_covariance))
    _check_keys(forcing, "forcing", {"const-gravity", "bot-heat", "diffusion", "coriolis"})
    if "coriolis" in forcing:
        co = forcing["coriolis"] or {}
        _check_keys(co, "forcing/coriolis", {"omega1", "omega2", "omega3", "omegax", "omegay", "omegaz"})
        if any(float(v) != 0.0 for v in co.values()):
            raise ValueError("forcing/coriolis with a nonzero rotation rate is not implemented")
    if "diffusion" in forcing:
        # DiffusionOptionsImpl::from_yaml (src/forcing/diffusion.cpp:280-366):
        # the kinematic isotropic form with stress-free walls is built here
        df = forcing["diffusion"] or {}
        _check_keys(df, "forcing/diffusion", {"nu_iso", "kappa_iso", "dynamic", "wall_velocity"})
        if str(df.get("dynamic", "false")).lower() not in ("false",):
            raise ValueError("forcing/diffusion dynamic: true is not implemented")
        if str(df.get("wall_velocity", "stress-free")) != "stress-free":
            raise ValueError("forcing/diffusion wall_velocity must be stress-free on a Cartesian grid")
        for key in ("nu_iso", "kappa_iso"):
            v = float(df.get(key, 0.0))
            if not (v >= 0.0) or v != v or v == float("inf"):
                raise ValueError(f"DiffusionOptions: {key} must be a finite number >= 0")
            setattr(h, key, v)
    if "bot-heat" in forcing:
        bh = forcing["bot-heat"] or {}
        _check_keys(bh, "forcing/bot-heat", {"flux", "depth", "thickness"})
        if float(bh.get("thickness", 0.0)) != 0.0:
            raise ValueError("forcing/bot-heat thickness > 0 is not implemented (use depth)")
        h.bot_heat_flux = float(bh.get("flux", 0.0))
        h.bot_heat_depth = int(bh.get("depth", 1))
    if "const-gravity" in forcing:
        cg = forcing["const-gravity"] or {}
        _check_keys(cg, "forcing/const-gravity", {"grav1"})
        h.grav1 = float(cg.get("grav1", 0.0))
    return h


def _read_integration(i) -> IntegratorOptions:
    _check_keys(
        i,
        "integration",
        {
            "type",
            "cfl",
            "tlim",
            "nlim",
            "ncycle_out",
            "implicit-scheme",
            "implicit-advection-cfl",
            "execution",
            "verbose",
        },
    )
    o = IntegratorOptions()
    o.type = i.get("type", o.type)
    o.cfl = float(i.get("cfl", o.cfl))
    o.tlim = float(i.get("tlim", o.tlim))
    o.nlim = int(i.get("nlim", o.nlim))
    o.implicit_scheme = int(i.get("implicit-scheme", 0))
    o.implicit_advection_cfl = float(i.get("implicit-advection-cfl", 1.0))
    o.execution = i.get("execution", o.execution)
    return o


_FACES = ["x1-inner", "x1-outer", "x2-inner", "x2-outer", "x3-inner", "x3-outer"]


def _read_bcs(bc) -> list:
    _check_keys(bc, "boundary-condition", {"internal", "external"})
    ext = bc.get("external") or {}
    _check_keys(ext, "boundary-condition/external", set(_FACES))
    return [ext.get(f, "reflecting") for f in _FACES]


def _check_supported(op: MeshBlockOptions, layout: str = "") -> None:
    """Refuse deck features outside the implemented stage, loudly."""
    _check_scalar(op)
    if op.coord.type == "gnomonic-equiangle":
        _check_cubed(op, layout)
        return
    if op.coord.type != "cartesian":
        raise ValueError(f"geometry/type '{op.coord.type}': cartesian or gnomonic-equiangle")
    if op.hydro.eos.type not in ("ideal-gas", "shallow-water"):
        raise ValueError(f"equation-of-state type '{op.hydro.eos.type}' is not implemented")
    if op.hydro.eos.type == "shallow-water":
        # the Cartesian shallow-water stage (rows phi, phi v1, phi v2, phi v3)
        if op.hydro.riemann != "shallow-roe" or op.hydro.riemann_dir != "xy":
            raise ValueError("shallow-water pairs with riemann-solver shallow-roe, dir: xy, on a Cartesian grid")
        if op.coord.nx3 > 1:
            raise ValueError("shallow-water dir xy: x3 must be a single cell")
        h = op.hydro
        if h.grav1 != 0.0 or h.nu_iso > 0.0 or h.kappa_iso > 0.0 or h.bot_heat_flux != 0.0:
            raise ValueError("shallow-water: forcing is not implemented for the shallow-water stage")
        if op.intg.implicit_scheme != 0:
            raise ValueError("shallow-water: no vertical implicit correction")
        if h.precision != "float64":
            raise ValueError("shallow-water: the FP32 state form is built for the ideal-gas stage")
        if "outflow" in op.bcnames:
            raise ValueError("shallow-water: outflow (radiating) boundaries are ideal-gas only")
    elif op.hydro.riemann == "shallow-roe":
        raise ValueError("riemann-solver shallow-roe needs equation-of-state shallow-water")
    if op.hydro.eos.limiter:
        raise ValueError("equation-of-state/limiter: true is not implemented")
    for r, name in ((op.hydro.recon1, "vertical"), (op.hydro.recon23, "horizontal")):
        if r.type != "weno5":
            raise ValueError(f"reconstruct/{name}/type '{r.type}': weno5 only")
        if r.scale:
            raise ValueError(f"reconstruct/{name}/scale: true is not implemented")
    if op.hydro.riemann not in ("hllc", "lmars", "shallow-roe"):
        raise ValueError(f"riemann-solver '{op.hydro.riemann}' is not implemented here")
    if op.hydro.precision not in ("float64", "float32"):
        raise ValueError("dynamics/precision must be float64 or float32")
    c = op.coord
    if op.hydro.grav1 != 0.0:
        if c.nx1 < 6:
            raise ValueError("const-gravity grav1 needs a resolved x1 (nx1 >= 6)")
        if "periodic" in op.bcnames[0:2]:
            raise ValueError("const-gravity grav1 with a periodic x1 column is not implemented")
    if (op.hydro.nu_iso > 0.0 or op.hydro.kappa_iso > 0.0) and op.coord.nghost < 2:
        raise ValueError("[Diffusion] Isotropic diffusion requires nghost >= 2.")
    if op.hydro.bot_heat_flux != 0.0 and op.hydro.grav1 == 0.0:
        raise ValueError("forcing/bot-heat is built on the gravity path here (needs grav1)")
    if op.intg.type != "rk3":
        raise ValueError(f"integration/type '{op.intg.type}': rk3 only")
    if op.intg.implicit_scheme not in (0, 1, 9):
        raise ValueError(
            f"integration/implicit-scheme {op.intg.implicit_scheme}: schemes 1 (vic-partial) and 9 (vic-full) "
            "are implemented"
        )
    if op.intg.implicit_scheme in (1, 9):
        if op.hydro.grav1 == 0.0:
            raise ValueError("implicit-scheme 9 needs forcing/const-gravity grav1 (ImplicitHydro requires it)")
        if c.nx1 < 5:
            raise ValueError("implicit-scheme 9 needs nx1 >= 5")
    for f, n in zip(_FACES, op.bcnames):
        if n not in ("reflecting", "periodic", "extrapolation", "outflow", "custom", "none"):
            raise ValueError(f"boundary {f} '{n}' is not implemented")


def _check_cubed(op: MeshBlockOptions, layout: str) -> None:
    """The cubed-sphere shallow-water stage: six panels, one block each."""
    c, h = op.coord, op.hydro
    if layout != "cubed-sphere":
        raise ValueError("geometry/type gnomonic-equiangle needs distribute/layout: cubed-sphere (use MeshOptions)")
    if h.eos.type != "shallow-water" or h.riemann != "shallow-roe" or h.riemann_dir != "yz":
        raise ValueError("the cubed sphere is built for shallow-water with riemann-solver shallow-roe, dir: yz")
    if c.nx1 != 1 or c.nx2 != c.nx3 or c.nx2 < 4:
        raise ValueError("cubed sphere: nx1 = 1 and nx2 = nx3 >= 4")
    if c.nghost != 3:
        raise ValueError("cubed sphere: nghost 3")
    if c.interp_order not in (2, 4):
        raise ValueError("GnomonicEquiangle: interp_order must be 2 (linear panel-edge ghosts) or 4 (cubic)")
    quarter = 0.25 * math.pi
    for v, want in ((c.x2min, -quarter), (c.x2max, quarter), (c.x3min, -quarter), (c.x3max, quarter)):
        if abs(v - want) > 1e-12:
            raise ValueError("cubed sphere: x2 and x3 span [-pi/4, pi/4] (one block per panel)")
    if h.precision not in ("float64", "float32"):
        raise ValueError("cubed sphere: dynamics/precision float64 or float32")
    if h.grav1 != 0.0 or h.nu_iso > 0.0 or h.kappa_iso > 0.0 or h.bot_heat_flux != 0.0:
        raise ValueError("cubed sphere: no forcing is implemented")
    if h.eos.limiter:
        raise ValueError("equation-of-state/limiter: true is not implemented")
    if op.intg.type != "rk3" or op.intg.implicit_scheme != 0:
        raise ValueError("cubed sphere: rk3, explicit")
    for f, n in zip(_FACES[2:], op.bcnames[2:]):
        if n not in ("custom", "none"):
            raise ValueError(f"cubed sphere: boundary {f} is a panel seam (custom), got '{n}'")


@dataclass
class MeshOptions:
    """MeshOptionsImpl::from_yaml (src/mesh/mesh.hpp): the block options and
    the layout that places the blocks (distribute/)."""

    block: MeshBlockOptions = field(default_factory=MeshBlockOptions)
    layout: str = "slab"
    nb2: int = 1
    nb3: int = 1
    blocks_per_process: int = 1

    @staticmethod
    def from_yaml(filename: str, verbose: bool = False) -> "MeshOptions":
        with open(filename) as f:
            cfg = yaml.safe_load(f) or {}
        _check_sections(cfg)
        dist = cfg.get("distribute") or {}
        _check_keys(dist, "distribute", _DIST_KEYS)
        if int(dist.get("nb1", 1)) != 1:
            raise ValueError("distribute/nb1: the vertical implicit solve needs one block along x1")
        mo = MeshOptions()
        mo.layout = str(dist.get("layout", "slab"))
        mo.nb2 = int(dist.get("nb2", 1))
        mo.nb3 = int(dist.get("nb3", 1))
        mo.blocks_per_process = int(dist.get("blocks_per_process", 1))
        if mo.layout == "cubed-sphere" and (mo.nb2 != 1 or mo.nb3 != 1):
            raise ValueError("cubed sphere: one block per panel (nb2 = nb3 = 1)")
        op = MeshBlockOptions(verbose=verbose, filename=os.fspath(filename))
        op.coord = _read_geometry(cfg.get("geometry") or {})
        op.hydro = _read_dynamics(cfg.get("dynamics") or {}, cfg.get("forcing") or {})
        op.intg = _read_integration(cfg.get("integration") or {})
        op.bcnames = _read_bcs(cfg.get("boundary-condition") or {})
        op.scalar = _read_scalar(cfg.get("scalar"), op.hydro.recon23)
        _check_supported(op, mo.layout)
        mo.block = op
        return mo
