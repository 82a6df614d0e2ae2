"""A request loop over resident blocks: continuous batching of simulation
requests on one GPU.

A request is (kind, parameters, seed, steps).  The server admits a waiting
request once it has arrived, a slot is free and its HBM footprint
(MeshBlock.hbm_footprint: state, RK backups, primitives, face totals,
implicit workspace) fits in the free part of the budget.  Admitted requests
keep their state on the device between steps; every round advances each
active request by one RK3 step, each request on its own HIP stream (all at
the default priority), so the stages of different requests overlap on the
device while each request's own stages stay ordered.  The shock tubes of a
round are batched: one work list (`worklist.BlockBatch`) advances all of
them with one launch per phase on a stream of their own, rebuilt when a
tube joins or leaves; a tube's result is bit-identical batched or alone.
The density-current requests of one state form are grouped too, whatever
their grids, and so are the shallow-water meshes of any panel size
(`group.BlockGroup`):
every kernel launch of their step is one launch over all of them, each
block with its own buffers and dt, the group rebuilt when one joins or
leaves; a result is bit-identical grouped or alone.  A request leaves when
its step count is reached, freeing its slot and memory for the next
arrival; its latency is completion time minus arrival time.  With
`preempt`, a request that does not fit may instead park the active block
request with the most steps left in a host restart part; the parked request
resumes from it when memory frees, bit for bit.

Request kinds (the serving mix of the benchmark):
  * ``sod``              -- a 1D Sod-type tube with seeded left/right states;
  * ``density_current``  -- the 2D density current (gravity, diffusion, the
    vertical implicit solve) at a seeded grid and bubble amplitude, in the
    FP64 or the FP32 state form;
  * ``swe``              -- cubed-sphere shallow water (six gnomonic panels
    of N x N) with a seeded Gaussian hill of h at a seeded point of the
    sphere, at rest.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import problems
from .cubed import CubedSphereMesh
from .meshblock import MeshBlock
from ._capi import SnapyError
from .group import BlockGroup, group_key, groupable
from .worklist import BlockBatch, eligible
from .options import CoordinateOptions, HydroOptions, IntegratorOptions, MeshBlockOptions, MeshOptions

RD, CP, GRAV = 287.0, 1004.0, 9.81
RADIUS = 6.371e6
GAMMA_DRY = CP / (CP - RD)
WEIGHT_DRY = 8.31446261815324 / RD


@dataclass
class Request:
    rid: int
    kind: str
    params: dict
    seed: int
    steps: int
    arrival: float = 0.0  # seconds after the stream's start


@dataclass
class Result:
    rid: int
    kind: str
    steps: int
    cells: int
    arrival: float
    admitted: float
    done: float
    state: np.ndarray  # the final conserved state (FP64 value), interior
    finite: bool

    @property
    def latency(self) -> float:
        return self.done - self.arrival


# ----------------------------------------------------------------------------
# building blocks for requests
# ----------------------------------------------------------------------------
def _options(kind: str, p: dict) -> MeshBlockOptions:
    op = MeshBlockOptions()
    op.intg = IntegratorOptions(cfl=0.9, execution=p.get("execution", "hip-graph"))
    if kind == "sod":
        n = int(p["n"])
        op.coord = CoordinateOptions("cartesian", 0.0, 1.0, -0.5, 0.5, -0.5, 0.5, n, 1, 1, 3)
        op.hydro = HydroOptions(riemann="hllc")
        op.hydro.eos.gammad = 1.4
        op.bcnames = ["outflow"] * 6
    elif kind == "density_current":
        nx, nz = int(p["nx"]), int(p["nz"])
        op.coord = CoordinateOptions("cartesian", 0.0, 6400.0, -25600.0, 25600.0, -50.0, 50.0, nz, nx, 1, 3)
        op.hydro = HydroOptions(riemann="lmars", grav1=-GRAV, nu_iso=75.0,
                                precision=p.get("precision", "float64"))
        op.hydro.eos.gammad = GAMMA_DRY
        op.hydro.eos.weight = WEIGHT_DRY
        op.hydro.recon1.shock = False
        op.hydro.recon23.shock = False
        op.intg.implicit_scheme = 9
        op.bcnames = ["reflecting"] * 6
    elif kind == "swe":
        n = int(p["n"])
        q = 0.25 * math.pi
        op.coord = CoordinateOptions("gnomonic-equiangle", RADIUS, RADIUS + 1.0, -q, q, -q, q, 1, n, n, 3)
        op.hydro = HydroOptions(riemann="shallow-roe", riemann_dir="yz")
        op.hydro.eos.type = "shallow-water"
        op.hydro.recon1.shock = False
        op.hydro.recon23.shock = False
        op.bcnames = ["reflecting"] * 2 + ["custom"] * 4
        return MeshOptions(block=op, layout="cubed-sphere")
    else:
        raise ValueError(f"unknown request kind '{kind}'")
    return op


def _initial_state(kind: str, block, p: dict, seed: int):
    if kind == "sod":
        rng = np.random.default_rng(seed)
        left = (rng.uniform(0.1, 1.0), 0.0, rng.uniform(0.1, 1.0))
        right = (rng.uniform(0.1, 1.0), 0.0, rng.uniform(0.1, 1.0))
        return problems.sod(block, left=left, right=right)
    if kind == "density_current":
        return problems.density_current(block, dT=float(p["dT"]), grav=GRAV)
    # swe: a seeded Gaussian hill of h on the sphere, at rest; one w per panel
    return problems.swe_hill(block, seed, radius=RADIUS, grav=GRAV)


def make_block(req: Request, device="cuda"):
    """A MeshBlock, or for shallow water a six-panel CubedSphereMesh."""
    op = _options(req.kind, req.params)
    if isinstance(op, MeshOptions):
        op.block.intg = IntegratorOptions(cfl=0.9, execution=req.params.get("execution", "hip-graph"))
        return CubedSphereMesh(op).to(device)
    block = MeshBlock(op)
    block.to(device)
    return block


def _initialize(block, w):
    if isinstance(block, CubedSphereMesh):
        vars, _ = block.initialize([{"hydro_w": x} for x in w])
        return vars
    vars, _ = block.initialize({"hydro_w": w.to("cuda")})
    return vars


def _final_state(block, vars) -> np.ndarray:
    if isinstance(block, CubedSphereMesh):
        return block.interior_state().cpu().numpy()
    sub = tuple(block.part(exterior=False))
    return block.total_u(vars)[sub].double().cpu().numpy()


def request_cells(req: Request) -> int:
    p = req.params
    if req.kind == "sod":
        return int(p["n"])
    if req.kind == "density_current":
        return int(p["nx"]) * int(p["nz"])
    return 6 * int(p["n"]) ** 2


# ----------------------------------------------------------------------------
# the mix and the arrival process
# ----------------------------------------------------------------------------
def sample_requests(n: int, seed: int = 7, rate: float = 20.0, kmin: int = 50, kmax: int = 500,
                    sizes: str = "full") -> list[Request]:
    """Poisson arrivals at `rate` per second; 40% Sod (N ~ logU[128, 8192]),
    40% density current (nx in {256, 512, 1024}, nz in {32, 64, 128},
    dT ~ U(-20, -5), FP64 or FP32 state form), 20% shallow water (N in
    {32, 64, 96, 128, 192}); K ~ U[kmin, kmax] steps.  One generator for the
    arrivals and the mix, one seed per request for its content.  sizes
    'small' scales the grids down for a quick stream."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        u = rng.random()
        k = int(rng.integers(kmin, kmax + 1))
        s = int(rng.integers(0, 2**31 - 1))
        small = sizes == "small"
        if u < 0.4:
            n1 = int(round(math.exp(rng.uniform(math.log(128), math.log(8192)))))
            if small:
                n1 = max(32, n1 // 16)
            out.append(Request(rid, "sod", {"n": n1}, s, k, t))
        elif u < 0.8:
            nx = int(rng.choice([256, 512, 1024]))
            nz = int(rng.choice([32, 64, 128]))
            if small:
                nx, nz = nx // 8, max(8, nz // 4)
            prec = "float32" if rng.random() < 0.5 else "float64"
            out.append(Request(rid, "density_current", {"nx": nx, "nz": nz, "dT": float(rng.uniform(-20.0, -5.0)),
                                                        "precision": prec}, s, k, t))
        else:
            n2 = int(rng.choice([32, 64, 96, 128, 192]))
            if small:
                n2 = max(16, n2 // 4)
            out.append(Request(rid, "swe", {"n": n2}, s, k, t))
    return out


# ----------------------------------------------------------------------------
# the server
# ----------------------------------------------------------------------------
@dataclass
class _Slot:
    req: Request
    block: object
    vars: object
    stream: torch.cuda.Stream
    bytes: int
    admitted: float
    done_steps: int = 0
    si: int = 0  # the index of the stream the request was admitted on
    batched: bool = False  # stepped by the round's work list
    key: object = None  # the block group it joins (group_key), or None


@dataclass
class Server:
    budget_bytes: int = 200 * 10**9
    max_slots: int = 16
    realtime: bool = True  # honour arrival times against the wall clock
    batch_tubes: bool = True  # advance the shock tubes as one work list
    # advance the density-current requests of one state form, and the
    # shallow-water meshes, as groups (one launch per kernel), whatever their
    # grids
    batch_groups: bool = True
    # a request that does not fit may preempt the active block request with
    # the most steps left (more than the newcomer's): that request is written
    # to a host restart part, its memory freed, and it resumes later from
    # the part, bit for bit (snapy_amd.restart)
    preempt: bool = False
    log: list = field(default_factory=list)

    def run(self, requests: list[Request]) -> list[Result]:
        """Serve every request; return their results in completion order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        streams = [torch.cuda.Stream() for _ in range(self.max_slots)]
        tube_stream = torch.cuda.Stream()
        batch, batch_ids = None, ()
        # block groups and their streams, by group_key
        self._groups = {}
        group_streams = {}
        free_streams = list(range(self.max_slots))
        active: list[_Slot] = []
        used = 0
        results = []
        parked = {}  # rid -> (restart part, steps done, first admission)
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        while pending or active:
            # admit the arrived requests that fit (first come, first served;
            # a request too big for the free memory waits for departures)
            while pending and free_streams and (not self.realtime or pending[0].arrival <= now()):
                req = pending[0]
                block = make_block(req)
                need = block.hbm_footprint()
                if need > self.budget_bytes:
                    raise RuntimeError(f"request {req.rid} needs {need} bytes, above the budget")
                if used + need > self.budget_bytes:
                    victim = self._victim(active, req, used + need - self.budget_bytes) if self.preempt else None
                    if victim is None:
                        break
                    if victim.key in self._groups:
                        self._drop_group(victim.key, group_streams[victim.key])
                    used -= self._park(victim, parked, now())
                    active.remove(victim)
                    free_streams.append(victim.si)
                    pending.insert(1, victim.req)
                    self.log.append(("preempt", victim.req.rid, now(), used))
                pending.pop(0)
                si = free_streams.pop(0)
                done0, admitted = 0, now()
                with torch.cuda.stream(streams[si]):
                    if req.rid in parked:
                        part, done0, admitted = parked.pop(req.rid)
                        vars, _ = block.initialize_from_restart(part)
                    else:
                        w = _initial_state(req.kind, block, req.params, req.seed)
                        vars = _initialize(block, w)
                used += need
                sl = _Slot(req, block, vars, streams[si], need, admitted, done_steps=done0, si=si)
                if self.batch_tubes and isinstance(block, MeshBlock) and eligible(block):
                    # the tube's steps run on the work list's stream, after
                    # its initialization
                    tube_stream.wait_stream(streams[si])
                    sl.stream = tube_stream
                    sl.batched = True
                elif self.batch_groups and req.kind in ("density_current", "swe") and groupable(block):
                    # its steps run on its group's stream, alone or in the
                    # group
                    sl.key = group_key(block)
                    gs = group_streams.setdefault(sl.key, torch.cuda.Stream())
                    gs.wait_stream(streams[si])
                    sl.stream = gs
                active.append(sl)
                self.log.append(("admit", req.rid, now(), used))
            if not active:
                # nothing to run until the next arrival
                if pending and self.realtime:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
                continue
            # one RK3 step of every active request: the tubes as one work
            # list, every other request on its own stream
            tubes = [sl for sl in active if sl.batched]
            ids = tuple(sl.req.rid for sl in tubes)
            if ids != batch_ids:
                batch = BlockBatch([(sl.block, sl.vars) for sl in tubes]) if tubes else None
                batch_ids = ids
            if batch is not None:
                with torch.cuda.stream(tube_stream):
                    batch.step()
                for sl in tubes:
                    sl.done_steps += 1
            # the block groups: two or more requests of one key as one group,
            # rebuilt when the key's requests changed
            groups = {}
            for sl in active:
                if sl.key is not None:
                    groups.setdefault(sl.key, []).append(sl)
            for key in list(self._groups):
                ids = tuple(x.req.rid for x in groups.get(key, []))
                if self._groups[key][0] != ids:
                    self._drop_group(key, group_streams[key])
            stepped = set()
            for key, sls in groups.items():
                if len(sls) < 2:
                    continue
                with torch.cuda.stream(group_streams[key]):
                    if key not in self._groups:
                        ids = tuple(x.req.rid for x in sls)
                        try:
                            grp = BlockGroup([(x.block, x.vars) for x in sls])
                            self.log.append(("group", ids, now(), used))
                        except SnapyError:
                            # launches the library cannot group (another
                            # kernel choice on one of the grids): these
                            # requests step alone until the group changes
                            grp = None
                            self.log.append(("ungrouped", ids, now(), used))
                        self._groups[key] = (ids, grp)
                    grp = self._groups[key][1]
                    if grp is None:
                        continue
                    grp.step()
                for sl in sls:
                    sl.done_steps += 1
                    stepped.add(sl.req.rid)
            for sl in active:
                if sl.batched or sl.req.rid in stepped:
                    continue
                with torch.cuda.stream(sl.stream):
                    dt = sl.block.device_time_step(sl.vars)
                    for s in range(3):
                        sl.block.forward(sl.vars, dt, s)
                sl.done_steps += 1
            # departures (a departing request's group is freed first and
            # rebuilt without it next round)
            keep = []
            for sl in active:
                if sl.done_steps < sl.req.steps:
                    keep.append(sl)
                    continue
                if sl.key in self._groups:
                    self._drop_group(sl.key, group_streams[sl.key])
                sl.stream.synchronize()
                b = sl.block
                u = _final_state(b, sl.vars)
                # without real-time arrivals every request is there at the start
                arrival = sl.req.arrival if self.realtime else 0.0
                results.append(Result(sl.req.rid, sl.req.kind, sl.req.steps, request_cells(sl.req),
                                      arrival, sl.admitted, now(), u, bool(np.isfinite(u).all())))
                b.release(sync=False)  # its stream was synchronized above
                sl.vars.clear()
                used -= sl.bytes
                free_streams.append(sl.si)
                self.log.append(("done", sl.req.rid, now(), used))
            active = keep
        torch.cuda.synchronize()
        return results

    def _drop_group(self, key, stream) -> None:
        """Free the block group of `key` once its launches have finished (its
        blocks keep their states)."""
        _, grp = self._groups.pop(key)
        if grp is not None:
            stream.synchronize()
            grp.release(sync=False)

    @staticmethod
    def _victim(active, req, short):
        """The active block request with the most steps left, if it has more
        left than `req` needs and freeing it covers the `short` bytes."""
        best = None
        for sl in active:
            if not isinstance(sl.block, MeshBlock) or sl.bytes < short:
                continue
            left = sl.req.steps - sl.done_steps
            if left > req.steps and (best is None or left > best.req.steps - best.done_steps):
                best = sl
        return best

    @staticmethod
    def _park(sl, parked, t):
        """Write the slot's request to a host restart part and free its
        block; returns the bytes freed."""
        from .restart import restart_variables

        sl.stream.synchronize()
        part = restart_variables(sl.block, sl.vars, t)
        parked[sl.req.rid] = (part, sl.done_steps, sl.admitted)
        sl.block.release(sync=False)
        sl.vars.clear()
        return sl.bytes


def summarize(results: list[Result], wall: float) -> dict:
    lat = np.array([r.latency for r in results])
    cell_updates = sum(r.cells * r.steps * 3 for r in results)
    return {
        "requests": len(results),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "cell_updates_per_s": cell_updates / wall,
        "wall_s": wall,
        "all_finite": all(r.finite for r in results),
        "mix": {k: sum(1 for r in results if r.kind == k) for k in ("sod", "density_current", "swe")},
    }
