#!/usr/bin/env python3
"""Benchmark of the finite-volume stage on one MI355X.

Prints ONE JSON line: cell-updates/s per RK3 stage of the named workload
(one interior cell advanced through one RK3 stage), with the dominant
kernel's roofline (HIP-event timed, algorithmic bytes) and the CPU oracle
timed on this host as the baseline.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A "step" is one RK3 step (3 stages) of the workload, its CFL reduction
included, run as captured hipGraphs with a device dt (no host read inside the
timed region).  Inputs are resident in HBM before timing starts.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
FP64_PEAK_TFLOPS = 78.6  # MI355X published FP64 vector peak

WORKLOADS = ("convection3d", "sod")


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="convection3d", choices=WORKLOADS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel timing")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-stream leg")
    ap.add_argument("--serving-requests", type=int, default=24)
    ap.add_argument("--no-cubed", action="store_true", help="skip the batched cubed-sphere leg")
    ap.add_argument("--implicit-scheme", type=int, default=None, choices=(1, 9),
                    help="override the convection deck's implicit scheme")
    ap.add_argument("--cubed-steps", type=int, default=50)
    ap.add_argument("--accuracy-steps", type=int, default=200,
                    help="steps of the FP32-vs-FP64 accuracy run (the deck's nlim)")
    ap.add_argument("--ensemble-members", type=int, default=32,
                    help="members of the density-current ensemble line (one launch per kernel)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default workgroup sizes instead of tuning them for the workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain elementwise copy of 1 GiB (the HBM counter calibration)")
    ap.add_argument("--no-other-scheme", action="store_true",
                    help="skip the FP32 leg with the other vertical implicit scheme")
    return ap.parse_args()


# --------------------------------------------------------------------------
SCHEME = {"value": None}  # --implicit-scheme


def make_block(deck, execution="hip-graph", precision="float64", members=1, **coord):
    from snapy_amd import MeshBlock, MeshBlockOptions

    op = MeshBlockOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", deck))
    for k, v in coord.items():
        setattr(op.coord, k, v)
    op.intg.execution = execution
    op.hydro.precision = precision
    if SCHEME["value"] is not None and op.intg.implicit_scheme:
        op.intg.implicit_scheme = SCHEME["value"]
    block = MeshBlock(op, members=members)
    block.to("cuda")
    return block


def workload(name, precision="float64", execution="hip-graph"):
    """(block, vars, cells, config, dtype, B_alg bytes per cell-update)."""
    from snapy_amd import problems

    if name == "sod":
        block = make_block("sod.yaml")
        vars, _ = block.initialize({"hydro_w": problems.sod(block).cuda()})
        c = block.options.coord
        cfg = {
            "workload": "sod_1d_256",
            "cells": [c.nx1, c.nx2, c.nx3],
            "method": "weno5(shock)+hllc+ssp-rk3, outflow, cfl 0.9",
            "form": "plain FP64",
        }
        return block, vars, c.nx1, cfg, "f64", 160.0
    block = make_block("convection3d.yaml", execution=execution, precision=precision)
    vars, _ = block.initialize({"hydro_w": problems.convection3d(block).cuda()})
    c = block.options.coord
    scheme = block.options.intg.implicit_scheme
    solve = "9 (full 5x5" if scheme == 9 else "1 (partial 3x3 with the tangential fold"
    cfg = {
        "workload": "convection3d_256x256x128",
        "cells": [c.nx1, c.nx2, c.nx3],
        "method": "weno5+lmars+ssp-rk3, well-balanced x1 gravity, flux covariance, bot-heat, "
        f"vertical implicit scheme {solve}, Woodbury), periodic x2/x3, cfl 0.8",
        "implicit_scheme": scheme,
    }
    if precision == "float32":
        cfg["form"] = ("FP32 state form: FP32 sweeps and face fluxes, compensated FP32 state pair, "
                       "FP64 primitives, x1 reference, vertical implicit solve and budgets; one block")
        # B_alg: FP32 state read u, u0, write u; Kahan compensation read and
        # write; 3 FP32 face totals; the FP64 vertical-solve column I/O
        return block, vars, c.nx1 * c.nx2 * c.nx3, cfg, "f32", 5 * (3 * 4 + 3 * 4) + 2 * 5 * 4 + 3 * 5 * 8
    cfg["form"] = "plain FP64 (FP64 storage and compute, one block)"
    # B_alg (cell-update bytes): FP64 state read u, u0, write u; 3 face
    # totals; the FP64 vertical-solve column I/O (3 rows x 5 x 8)
    return block, vars, c.nx1 * c.nx2 * c.nx3, cfg, "f64", 5 * (3 * 8 + 3 * 8) + 3 * 5 * 8


# --------------------------------------------------------------------------
def _oracle_leg(name, budget_s):
    """The CPU FP64 oracle (numpy) on a bounded sample of the workload, run
    for about `budget_s` seconds in this process with its BLAS pools held to
    one thread; returns (cell-updates, seconds, sample description)."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import fv

    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        stages = runs = 0
        if name == "sod":
            while time.perf_counter() - t0 < budget_s:
                d = fv.sod_deck(256)
                b = fv.Block(d)
                b.initialize(fv.sod_initial(d.grid))
                _, n = fv.run(b, 0.25)
                stages += 3 * n
                runs += 1
            el = time.perf_counter() - t0
            return 256 * stages, el, f"{runs} full Sod runs (256 cells, t_end 0.25) per worker"
        # convection3d: a 24 x 24 column sub-domain of the same 128-layer
        # column physics (periodic x2/x3), whole RK3 steps
        from snapy_amd import MeshBlock, MeshBlockOptions, problems

        op = MeshBlockOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", "convection3d.yaml"))
        op.coord.nx2 = op.coord.nx3 = 24
        op.coord.x2max = op.coord.x3max = 2400.0
        blk = MeshBlock(op)
        w = problems.convection3d(blk).numpy()
        c = op.coord
        g = fv.Grid(c.nx1, c.nx2, c.nx3, c.x1min, c.x1max, c.x2min, c.x2max, c.x3min, c.x3max)
        h = op.hydro
        d = fv.Deck(grid=g, gammad=h.eos.gammad, riemann="lmars", shock_x1=False, shock_x23=False,
                    bcs=tuple(op.bcnames), cfl=op.intg.cfl, grav1=h.grav1,
                    implicit_scheme=SCHEME["value"] or op.intg.implicit_scheme,
                    flux_covariance=True, bot_heat_flux=h.bot_heat_flux, bot_heat_depth=h.bot_heat_depth)
        b = fv.Block(d)
        b.initialize(w)
        steps = 0
        while time.perf_counter() - t0 < budget_s or steps == 0:
            dt = b.max_time_step()
            for s in range(3):
                b.forward(dt, s)
            steps += 1
        el = time.perf_counter() - t0
        assert np.isfinite(b.u).all()
        n = c.nx1 * c.nx2 * c.nx3
        return n * 3 * steps, el, f"{steps} RK3 steps of a 128x24x24 sub-domain of the same deck per worker"


def _oracle_worker(job):
    name, budget_s, scheme = job
    SCHEME["value"] = scheme
    return _oracle_leg(name, budget_s)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def cpu_baseline(name, budget_s=10.0, max_workers=16):
    """The CPU FP64 oracle timed on this host: one worker alone, then one
    worker per core of the process's affinity mask (at most `max_workers`,
    the host share of one GPU) together, each on its own copy of the sample
    with one thread.  Run before the GPU is touched: the workers are forked
    from this process.  `value` is the all-cores aggregate."""
    import multiprocessing as mp

    import torch

    ncores = len(os.sched_getaffinity(0))
    n = max(1, min(ncores, max_workers))
    ctx = mp.get_context("fork")
    job = (name, budget_s, SCHEME["value"])
    with ctx.Pool(1) as pool:
        cu1, el1, sample = pool.map(_oracle_worker, [job])[0]
    with ctx.Pool(n) as pool:
        res = pool.map(_oracle_worker, [job] * n)
    agg = sum(cu / el for cu, el, _ in res)
    return {
        "value": agg,
        "unit": "cell-updates/s",
        "cores": n,
        "kind": "port",
        "value_1_core": cu1 / el1,
        "cores_in_affinity_mask": ncores,
        "cpu_model": _cpu_model(),
        "torch_num_threads": torch.get_num_threads(),
        "sample": f"numpy FP64 oracle, {sample}, about {budget_s:.0f} s; {n} single-threaded workers "
                  f"together (each its own copy), and one alone",
    }


# --------------------------------------------------------------------------
def kernel_times(block, vars, reps=20):
    """Average launch duration (ms) of each stage kernel, timed alone with
    HIP events on the stream the kernels are launched on (stage-0 weights,
    so repeated launches are idempotent)."""
    import torch

    calls = block._stage_calls(vars, 0)
    stream = torch.cuda.current_stream()
    out = {}
    for name, fn in calls:
        for _ in range(2):
            fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
    return out


def kernel_bytes(block, name, sb=8, pair=False):
    """Algorithmic HBM bytes of one launch: each input value read once and
    each output written once (stencil re-reads and solver workspace
    excluded)."""
    c = block.options.coord
    ncell = c.nc1() * c.nc2() * c.nc3()
    nint = c.nx1 * c.nx2 * c.nx3
    nd = sum(1 for n in (c.nx1, c.nx2, c.nx3) if n > 1)
    table = {
        "cons2prim": 10 * ncell * sb,
        "hydro_ref_x1": 7 * ncell * 8,
        # the reference and the well-balanced sweep in one launch: w read,
        # psf_lo written; five flux rows and the face pressure written
        "sweep_x1_wb": (5 * 8 + 8) * ncell + (5 * sb + 8) * nint,
        "stage_increment": (5 * nd + 2 + 5) * nint * sb,
        # the x1 sweep with the stage increment: w read, psf_lo written; the x1
        # flux rows and face pressure written, the x2/x3 fluxes read, du
        # written (FP64)
        "sweep_x1_increment": (5 * 8 + 8) * ncell + (5 * sb + 8 + 5 * (nd - 1) * sb + 5 * 8) * nint,
        # du, w (FP64) read; u0, u read and u written: FP64, or FP32 pairs;
        # with the FP32 pair and scheme 1 the interior primitives written too
        "vic_full": (10 * 8 + 15 * (2 * 4 if pair else 8)) * nint,
        "vic_partial": (10 * 8 + 15 * (2 * 4 if pair else 8) + (5 * 8 if block._prims_from_finish() else 0))
        * nint,
        "update": (3 + nd) * 5 * nint * sb,
        # outflow: the primitive view of the block (u read, w written), then
        # the ghost layers; the gathers read and write the ghost layers only
        "apply_boundaries": (10 * ncell * sb if "outflow" in block.options.bcnames else 0)
        + 2 * 2 * 3 * 5 * (ncell // max(c.nc1(), 1)) * sb * (nd > 0),
        "update_gravity": ((3 + nd) * 5 + 2) * nint * sb,
    }
    if name.startswith("sweep_x"):
        return table.get(name, 5 * ncell * sb + 5 * nint * sb)
    return table.get(name, 0)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "convection3d_hbm_traffic.md")
# stage launch -> the kernels it runs (FP64 form, FP32 pair form)
STAGE_KERNELS = {
    "vic_partial": (["vicp_forward_kernel", "vicp_finish_kernel<false>"],
                    ["vicp_forward_kernel", "vicp_finish_kernel<true>"]),
    "sweep_x1_wb": (["sweep_x1_wb_ref_kernel<double, double>"], ["sweep_x1_wb_ref_kernel<float, float>"]),
    "stage_increment": (["update_gravity_kernel<double, true, false>"], ["update_gravity_kernel<float, true, true>"]),
    "sweep_x1_increment": ([], ["sweep_x1_wb_inc_kernel"]),
}


def hbm_traffic(name, pair):
    """HBM bytes per launch of a stage launch, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (calibrated against a plain copy kernel);
    None when the launch was not profiled."""
    try:
        text = open(TRAFFIC_PROFILE).read()
        table = json.loads(text.split("```json", 1)[1].split("```", 1)[0])["kernels"]
        kernels = STAGE_KERNELS[name][1 if pair else 0]
        return sum(table[k]["read_bytes"] + table[k]["written_bytes"] for k in kernels)
    except (OSError, IndexError, KeyError, ValueError):
        return None


def timed_run(block, vars, steps, warmup):
    """Warm up, then time exactly `steps` RK3 steps (device dt, captured
    stages) between two synchronizations; returns seconds."""
    import torch

    for _ in range(warmup):
        dt = block.device_time_step(vars)
        for s in range(3):
            block.forward(vars, dt, s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = block.device_time_step(vars)
        for s in range(3):
            block.forward(vars, dt, s)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _rel_l2_rows(b64, v64, b32, v32):
    """Relative L2 of the FP32 form against the FP64 form per row group."""
    sl = tuple(b64.part(exterior=False))
    u64 = b64.total_u(v64)[sl].double()
    u32 = b32.total_u(v32)[sl].double()
    out = {}
    for key, rows in (("density", [0]), ("momentum", [1, 2, 3]), ("energy", [4])):
        d = (u32[rows] - u64[rows]).square().sum().sqrt()
        n = u64[rows].square().sum().sqrt()
        out[key] = float((d / n).item())
    return out


def _advance_pair(b64, v64, b32, v32, nsteps=None, tlim=None):
    """Advance both forms from the same state with the FP64 form's host dt
    (the last step cut to reach tlim); returns (steps, time)."""
    t, n = 0.0, 0
    while (nsteps is None or n < nsteps) and (tlim is None or t < tlim * (1.0 - 1e-12)):
        dt = b64.max_time_step(v64)
        if tlim is not None:
            dt = min(dt, tlim - t)
        for s in range(3):
            b64.forward(v64, dt, s)
            b32.forward(v32, dt, s)
        t += dt
        n += 1
    return n, t


def fp32_accuracy(name, nsteps=200):
    """Relative L2 of the FP32 form against the FP64 GPU form (itself within
    1e-12 of the FP64 oracle) after `nsteps` steps (the deck's nlim) from the
    same state with the same host dt, per row group; and the FP32 form's
    relative mass drift."""
    b64, v64, _, _, _, _ = workload(name, "float64", execution="eager")
    b32, v32, _, _, _, _ = workload(name, "float32", execution="eager")
    m0 = float(b32.budget(v32)[0].item())
    f0 = float(b64.budget(v64)[0].item())
    n, t = _advance_pair(b64, v64, b32, v32, nsteps=nsteps)
    m1 = float(b32.budget(v32)[0].item())
    f1 = float(b64.budget(v64)[0].item())
    out = _rel_l2_rows(b64, v64, b32, v32)
    out["mass_drift"] = abs(m1 - m0) / abs(m0)
    out["mass_drift_fp64_form"] = abs(f1 - f0) / abs(f0)
    out["steps"] = n
    out["time_s"] = t
    b64.release()
    b32.release()
    return out


def straka_diagnostics(block, vars, theta0=300.0, p0=1.0e5):
    """The density current's published diagnostics: the minimum potential
    temperature perturbation and the front position, the largest x2 where
    theta' reaches -1 K in the lowest layer (linear between cells)."""
    import numpy as np

    sl = tuple(block.part(exterior=False))
    w = vars["hydro_w"][sl].double().cpu().numpy()  # (5, x3, x2, x1)
    op = block.options.hydro.eos
    rd = 8.31446261815324 / op.weight
    cp = op.gammad / (op.gammad - 1.0) * rd
    t = w[4] / (rd * w[0])
    th = t * (p0 / w[4]) ** (rd / cp) - theta0
    x2 = block.module("coord").buffer("x2v").numpy()
    c = block.options.coord
    x2 = x2[c.nghost : c.nghost + c.nx2]
    row = th[:, :, 0].min(axis=0)  # the lowest layer along x2
    front = None
    for j in range(len(x2) - 1, 0, -1):
        if x2[j - 1] >= 0.0 and row[j - 1] <= -1.0 < row[j]:
            a = (-1.0 - row[j - 1]) / (row[j] - row[j - 1])
            front = float(x2[j - 1] + a * (x2[j] - x2[j - 1]))
            break
    return {"theta_prime_min_K": float(th.min()), "front_km": None if front is None else front / 1e3}


def density_current_accuracy(tlim=900.0):
    """The density current deck (512 x 64, scheme 9, viscosity) run to its
    tlim in both forms from the same state with the FP64 form's host dt:
    the FP32 form's relative L2 against the FP64 GPU form, its mass drift,
    and each form's minimum theta' and front position beside the published
    ones (Straka et al. 1993, 25 m reference solution: -9.77 K, 15.54 km;
    that solution also diffuses potential temperature, which this deck does
    not, so its minimum is deeper)."""
    from snapy_amd import problems

    blocks = {}
    for prec in ("float64", "float32"):
        b = make_block("density_current.yaml", execution="eager", precision=prec)
        v, _ = b.initialize({"hydro_w": problems.density_current(b).cuda()})
        blocks[prec] = (b, v)
    (b64, v64), (b32, v32) = blocks["float64"], blocks["float32"]
    m0 = float(b32.budget(v32)[0].item())
    f0 = float(b64.budget(v64)[0].item())
    n, t = _advance_pair(b64, v64, b32, v32, tlim=tlim)
    m1 = float(b32.budget(v32)[0].item())
    f1 = float(b64.budget(v64)[0].item())
    out = _rel_l2_rows(b64, v64, b32, v32)
    out["mass_drift"] = abs(m1 - m0) / abs(m0)
    # the FP64 form's own drift over the same steps, for scale
    out["mass_drift_fp64_form"] = abs(f1 - f0) / abs(f0)
    out["steps"] = n
    out["time_s"] = t
    out["fp64_form"] = straka_diagnostics(b64, v64)
    out["fp32_form"] = straka_diagnostics(b32, v32)
    out["published"] = {"theta_prime_min_K": -9.77, "front_km": 15.54,
                        "source": "Straka et al. 1993, Int. J. Numer. Meth. Fluids 17, 25 m reference, "
                                  "K = 75 m^2/s on momentum and potential temperature"}
    b64.release()
    b32.release()
    return out


def density_current_leg(steps=50, warmup=3, members=1, seed=2025):
    """The density current deck (512 x 64, dx = dz = 100 m, LMARS, isotropic
    viscosity, the vertical implicit scheme the deck states) in the FP32
    state form, captured stages: one block, or `members` ensemble members
    (cold bubbles dT ~ U(-20, -5) K from `seed`) on the member x3 axis, each
    with its own dt, advanced by one launch per kernel.  The throughput is
    against the B_alg (FP32 + compensation + FP64 implicit solve, 2 active
    directions: 260 B per cell-update)."""
    import numpy as np
    import torch

    from snapy_amd import problems

    block = make_block("density_current.yaml", precision="float32", members=members)
    if members == 1:
        w = problems.density_current(block)
    else:
        dTs = np.random.default_rng(seed).uniform(-20.0, -5.0, members)
        w = torch.cat([problems.density_current(block, dT=float(d))[:, i : i + 1] for i, d in enumerate(dTs)], dim=1)
    vars, _ = block.initialize({"hydro_w": w.cuda()})
    for _ in range(warmup):
        dt = block.device_time_step(vars)
        for s in range(3):
            block.forward(vars, dt, s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = block.device_time_step(vars)
        for s in range(3):
            block.forward(vars, dt, s)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    c = block.options.coord
    cells = c.nx1 * c.nx2 * c.nx3 * members
    value = cells * 3 * steps / el
    b_alg = 260
    finite = bool(torch.isfinite(block.total_u(vars)).all())
    out = {
        "workload": f"density_current_{c.nx2}x{c.nx1}" + (f"_x{members}_members" if members > 1 else ""),
        "value": value,
        "unit": "cell-updates/s",
        "steps": steps,
        "ms_per_step": el / steps * 1e3,
        "dtype": "f32",
        "implicit_scheme": block.options.intg.implicit_scheme,
        "form": "FP32 state form, viscosity, FP64 vertical implicit solve, captured stages",
        "stage_b_alg_per_cell": b_alg,
        "stage_frac": value * b_alg / (HBM_PEAK_GBS * 1e9),
        "state_finite": finite,
    }
    block.release()
    return out


def grouped_leg(steps=20, warmup=2, precision="float32", seed=2025):
    """The serving mix's density currents as one group: nine blocks, every
    (nx, nz) of nx in {256, 512, 1024} and nz in {32, 64, 128} (dT from the
    seed), stepped with one launch per kernel over all nine
    (snapy_amd.group.BlockGroup, its step captured), against the same blocks
    stepped one after another on their own captured stages.  Bit-identical
    either way (tests/test_gpu_group.py)."""
    import numpy as np
    import torch

    from snapy_amd import problems, serving
    from snapy_amd.group import BlockGroup
    from snapy_amd.meshblock import MeshBlock

    rng = np.random.default_rng(seed)
    shapes = [(nx, nz) for nz in (32, 64, 128) for nx in (256, 512, 1024)]

    def blocks(execution):
        out = []
        for nx, nz in shapes:
            op = serving._options("density_current", {"nx": nx, "nz": nz, "dT": -15.0, "precision": precision,
                                                     "execution": execution})
            blk = MeshBlock(op)
            blk.to("cuda")
            w = problems.density_current(blk, dT=float(rng.uniform(-20.0, -5.0)), grav=serving.GRAV)
            vars, _ = blk.initialize({"hydro_w": w.cuda()})
            out.append((blk, vars))
        return out

    cells = sum(nx * nz for nx, nz in shapes)
    solo = blocks("hip-graph")

    def solo_steps(n):
        for _ in range(n):
            for blk, vars in solo:
                dt = blk.device_time_step(vars)
                for k in range(3):
                    blk.forward(vars, dt, k)

    solo_steps(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solo_steps(steps)
    torch.cuda.synchronize()
    el_solo = time.perf_counter() - t0
    for blk, _ in solo:
        blk.release()
    members = blocks("eager")
    grp = BlockGroup(members, capture=True)
    for _ in range(warmup):
        grp.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        grp.step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    finite = all(bool(torch.isfinite(v["hydro_u"]).all()) for _, v in members)
    launches = {k: grp.launches(k) for k in ("dt", 0, 1, 2)}
    grp.release()
    for blk, _ in members:
        blk.release()
    return {
        "workload": "density_current_group_9_grids",
        "grids": [f"{nx}x{nz}" for nx, nz in shapes],
        "value": cells * 3 * steps / el,
        "unit": "cell-updates/s",
        "steps": steps,
        "ms_per_step": el / steps * 1e3,
        "one_after_another": {"value": cells * 3 * steps / el_solo, "ms_per_step": el_solo / steps * 1e3,
                              "form": "each block on its own captured stages, in turn"},
        "speedup": el_solo / el,
        "dtype": "f32" if precision == "float32" else "f64",
        "launches_per_phase": launches,
        "state_finite": finite,
    }


def cubed_leg(steps=50, warmup=3, members=32, seed0=1000, precision="float32"):
    """The batched cubed-sphere shallow-water workload: `members` six-panel
    meshes of 192 x 192 cells per panel, each seeded with its own Gaussian
    hill (seed0 + member), advanced together: every stage kernel is one
    launch over all members and panels, the stages are captured hipGraphs,
    each member has its own device CFL dt.  precision float32: the FP32
    form (FP32 pair state, primitives, fluxes and sweeps; FP64 EOS,
    gathers, seams, update and budgets); float64: FP64 throughout.
    Returns the throughput, the per-kernel HIP-event times and the
    roofline."""
    import torch

    from snapy_amd import problems
    from snapy_amd.cubed import CubedSphereMesh
    from snapy_amd.options import MeshOptions

    mo = MeshOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", "swe_cubed.yaml"))
    mo.block.intg.execution = "hip-graph"
    mo.block.hydro.precision = precision
    mesh = CubedSphereMesh(mo, members=members).to("cuda")
    fp32 = precision == "float32"
    vars = [{"hydro_w": w} for m in range(members) for w in problems.swe_hill(mesh, seed0 + m)]
    mesh.initialize(vars)
    m0 = mesh.mass().cpu()
    for _ in range(warmup):
        dt = mesh.device_time_step(vars)
        for s in range(3):
            mesh.forward(vars, dt, s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = mesh.device_time_step(vars)
        for s in range(3):
            mesh.forward(vars, dt, s)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    m1 = mesh.mass().cpu()
    cells = members * 6 * mesh.n * mesh.n
    value = cells * 3 * steps / el
    # per-kernel HIP-event times on the launch stream (stage-0 weights)
    stream = torch.cuda.current_stream()
    kt = {}
    for name, fn in mesh._stage_calls(0):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(20):
            fn()
        e1.record(stream)
        e1.synchronize()
        kt[name] = e0.elapsed_time(e1) / 20
    nc = mesh.nc
    blk = members * 6
    # algorithmic bytes per launch: each input read once, each output written
    # once (4 rows; FP64, or in the FP32 form 4-byte words with the pair's
    # two words per state value)
    sw = 4 if fp32 else 8  # primitive / flux word
    su = 8  # one state value: FP64, or the FP32 pair
    kbytes = {
        "cs_cons2prim": 4 * blk * nc * nc * (su + sw),
        "cs_sweep_x2": (4 * blk * nc * nc + 4 * blk * mesh.n * (mesh.n + 1)) * sw,
        "cs_sweep_x3": (4 * blk * nc * nc + 4 * blk * mesh.n * (mesh.n + 1)) * sw,
        # u0, u read and u written; w and the two face totals read
        "cs_update": (3 * su + 3 * sw) * 4 * cells,
    }
    dom = max((k for k in kt if k in kbytes), key=kt.get)
    achieved = kbytes[dom] / (kt[dom] * 1e-3) / 1e9
    # B_alg: nv (3 s_u + n_dim s_F) + k 2 nv s_c, nv = 4 stored rows
    b_alg = 4 * (3 * 4 + 2 * 4) + 2 * 4 * 4 if fp32 else 4 * (3 * 8 + 2 * 8)
    out = {
        "workload": "swe_cubed_sphere_6x192x192_x32",
        "value": value,
        "unit": "cell-updates/s",
        "steps": steps,
        "ms_per_step": el / steps * 1e3,
        "members": members,
        "cells_per_member": 6 * mesh.n * mesh.n,
        "dtype": "f32" if fp32 else "f64",
        "form": ("FP32 form: FP32 pair state, primitives, fluxes and sweeps; FP64 EOS, ghost gather, "
                 "seams, update and budgets" if fp32 else "FP64")
        + "; every member and panel in each launch, captured stages, per-member device dt",
        "kernel_ms": kt,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS},
        "stage_b_alg_per_cell": b_alg,
        "stage_frac": value * b_alg / (HBM_PEAK_GBS * 1e9),
        "mass_drift_max": float(((m1 - m0).abs() / m0.abs()).max()),
        "state_finite": bool(torch.isfinite(mesh.interior_state()).all().item()),
    }
    mesh.release()
    return out


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain elementwise copy of `nbytes` FP64 (y = -x, one streaming
    kernel per repetition that reads x once and writes y once): the
    known-size reference that the HBM counters FETCH_SIZE / WRITE_SIZE are
    scaled by, and the achievable streaming rate."""
    import torch

    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.neg(x, out=y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.neg(x, out=y)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": reps + 1,
            "ms_per_copy": ms, "copy_GBps": 2 * nbytes / (ms * 1e-3) / 1e9}


def other_scheme_leg(steps, warmup):
    """The FP32 form of the same workload with the other vertical implicit
    scheme (9, the full 5x5 blocks, when the deck runs 1; and the reverse):
    throughput and the implicit kernel's HIP-event time."""
    import torch

    from snapy_amd import MeshBlockOptions

    keep = SCHEME["value"]
    deck = MeshBlockOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", "convection3d.yaml"))
    scheme = 9 if (keep or deck.intg.implicit_scheme) == 1 else 1
    SCHEME["value"] = scheme
    try:
        block, vars, cells, cfg, _, _ = workload("convection3d", "float32")
        el = timed_run(block, vars, steps, warmup)
        kt = kernel_times(block, vars)
        ok = bool(torch.isfinite(vars["hydro_u"]).all().item()) and block.check_redo(vars) == 0
    finally:
        SCHEME["value"] = keep
    vic = "vic_full" if scheme == 9 else "vic_partial"
    out = {"implicit_scheme": scheme, "value": cells * 3 * steps / el, "unit": "cell-updates/s",
           "ms_per_step": el / steps * 1e3, "form": "FP32 state form, as the headline",
           "vic_ms": kt.get(vic), "state_finite": ok}
    del block, vars
    torch.cuda.empty_cache()
    return out


def serving_leg(nreq, seed=2025):
    """A bounded serving stream: Poisson arrivals (20/s, seeded) of the
    Sod / density-current / shallow-water mix, continuous batching with
    admission by HBM footprint (200 GB budget; the shock tubes as one work
    list, the density currents of one deck, form and height as one group);
    p50/p99 request latency."""
    from snapy_amd import serving

    reqs = serving.sample_requests(nreq, seed=seed, rate=20.0)
    srv = serving.Server(budget_bytes=200 * 10**9, max_slots=16, realtime=True)
    t0 = time.perf_counter()
    res = srv.run(reqs)
    wall = time.perf_counter() - t0
    out = serving.summarize(res, wall)
    out["arrival_rate_per_s"] = 20.0
    out["seed"] = seed
    # density-current requests of one deck, state form and height stepped as
    # one group (built when a group's requests changed)
    out["groups_built"] = sum(1 for e in srv.log if e[0] == "group")
    return out


def main():
    args = parse()
    SCHEME["value"] = args.implicit_scheme
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; run with "
                          "--gpus 1 (the path shards by whole members or x2/x3 column groups; not measured here)"}), flush=True)
        sys.exit(2)
    # the CPU baseline first, from worker processes forked before this
    # process touches the GPU
    cpu = None
    if not (args.no_cpu or args.copy_calibration):
        cpu = cpu_baseline(args.workload)
        print(f"# cpu baseline {cpu}", file=sys.stderr, flush=True)
    import torch

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    if args.copy_calibration:
        print(json.dumps(copy_calibration()), flush=True)
        return
    defaults = {"sod": (200, 10), "convection3d": (20, 3)}[args.workload]
    steps = args.steps or defaults[0]
    warmup = args.warmup if args.warmup is not None else defaults[1]

    # the FP64-plain form first (the comparison point), then the headline
    # FP32 form of the same workload
    plain = None
    precision = "float32" if args.workload == "convection3d" else "float64"
    if precision == "float32":
        b64, v64, cells64, cfg64, _, balg64 = workload(args.workload, "float64")
        print(f"# workload {cfg64['workload']} FP64-plain ready", file=sys.stderr, flush=True)
        el64 = timed_run(b64, v64, steps, warmup)
        plain = {
            "value": cells64 * 3 * steps / el64,
            "ms_per_step": el64 / steps * 1e3,
            "stage_b_alg_per_cell": balg64,
            "stage_frac": cells64 * 3 * steps / el64 * balg64 / (HBM_PEAK_GBS * 1e9),
            "form": cfg64["form"],
        }
        print(f"# FP64-plain {steps} steps in {el64:.3f} s", file=sys.stderr, flush=True)
        del b64, v64
        torch.cuda.empty_cache()
    block, vars, cells, cfg, dtype, b_alg = workload(args.workload, precision)
    print(f"# workload {cfg['workload']} {precision} ready", file=sys.stderr, flush=True)
    if not args.no_autotune:
        from snapy_amd import autotune

        tuned = autotune.tune_block(block, vars, reps=5)
        cfg["launch_shapes"] = {f: r["choice"] for f, r in tuned.items()}
        print(f"# launch shapes {cfg['launch_shapes']}", file=sys.stderr, flush=True)
    el = timed_run(block, vars, steps, warmup)
    ms_per_step = el / steps * 1e3
    value = cells * 3 * steps / el
    ok = bool(torch.isfinite(vars["hydro_u"]).all().item()) and block.check_redo(vars) == 0
    print(f"# timed {steps} steps in {el:.3f} s, state finite: {ok}", file=sys.stderr, flush=True)
    execution = block.execution_report()

    roofline = None
    if not args.no_kernels:
        kt = kernel_times(block, vars)
        dom = max(kt, key=kt.get)
        kb = kernel_bytes(block, dom, sb=4 if dtype == "f32" else 8, pair=dtype == "f32")
        achieved = kb / (kt[dom] * 1e-3) / 1e9
        roofline = {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": hbm_traffic(dom, dtype == "f32"),
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/convection3d_hbm_traffic.md)",
            "kernel_ms": kt,
            "stage_b_alg_per_cell": b_alg,
            "stage_frac": value * b_alg / (HBM_PEAK_GBS * 1e9),
        }
        # the measured streaming rate of this box (a plain 1 GiB copy), beside
        # the nominal peak
        cp = copy_calibration()
        roofline["copy_GBps"] = cp["copy_GBps"]
        roofline["frac_of_copy"] = achieved / cp["copy_GBps"]
        torch.cuda.empty_cache()
    accuracy = None
    if precision == "float32":
        del block, vars
        torch.cuda.empty_cache()
        accuracy = fp32_accuracy(args.workload, nsteps=args.accuracy_steps)
        print(f"# FP32 vs FP64 {accuracy}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        if not args.no_other_scheme:
            accuracy["density_current"] = density_current_accuracy()
            print(f"# density current to 900 s {accuracy['density_current']}", file=sys.stderr, flush=True)
            torch.cuda.empty_cache()
    # the metric as the baseline names it: the FP32 form is `value`, the
    # FP64-plain form sits beside it
    serving_summary = None
    if not args.no_serving and args.workload == "convection3d":
        print("# serving stream", file=sys.stderr, flush=True)
        serving_summary = serving_leg(args.serving_requests)
        print(f"# serving {serving_summary}", file=sys.stderr, flush=True)
    other = None
    if not args.no_other_scheme and args.workload == "convection3d":
        other = other_scheme_leg(steps, warmup)
        print(f"# other scheme {other}", file=sys.stderr, flush=True)
    c2 = None
    if not args.no_other_scheme and args.workload == "convection3d":
        print("# density current", file=sys.stderr, flush=True)
        c2 = density_current_leg()
        print(f"# density current {c2['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        ens = density_current_leg(members=args.ensemble_members)
        c2["ensemble"] = ens
        c2["ensemble_speedup_per_member"] = ens["value"] / c2["value"]
        print(f"# density current ensemble x{args.ensemble_members} {ens['value']:.3e} cell-updates/s",
              file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    grouped = None
    if not args.no_other_scheme and args.workload == "convection3d":
        print("# grouped density currents", file=sys.stderr, flush=True)
        grouped = grouped_leg()
        print(f"# grouped {grouped['value']:.3e} cell-updates/s ({grouped['speedup']:.2f}x one after another)",
              file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    cubed = None
    if not args.no_cubed and args.workload == "convection3d":
        print("# batched cubed-sphere shallow water", file=sys.stderr, flush=True)
        cubed = cubed_leg(steps=args.cubed_steps)
        print(f"# cubed {cubed['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        c64 = cubed_leg(steps=args.cubed_steps, precision="float64")
        cubed["fp64_form"] = {k: c64[k] for k in ("value", "ms_per_step", "dtype", "stage_frac", "roofline",
                                                  "mass_drift_max")}
        cubed["speedup_vs_fp64"] = cubed["value"] / c64["value"]
        print(f"# cubed FP64 {c64['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    metric = "cell-updates/s per RK3 stage, FP32-batched vs FP64-plain; p99 request latency"
    line = {
        "metric": metric,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": cfg,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "state_finite": ok,
        "execution": execution,
    }
    if serving_summary is not None:
        line["serving"] = serving_summary
    if cubed is not None:
        line["cubed_sphere"] = cubed
    if other is not None:
        line["other_implicit_scheme"] = other
    if c2 is not None:
        line["density_current"] = c2
    if grouped is not None:
        line["grouped"] = grouped
    if plain is not None:
        line["fp64_plain"] = plain
        line["speedup_vs_fp64_plain"] = value / plain["value"]
        line["fp32_vs_fp64_rel_l2"] = accuracy
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
