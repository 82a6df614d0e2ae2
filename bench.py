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
    return ap.parse_args()


# --------------------------------------------------------------------------
def make_block(deck, execution="hip-graph", **coord):
    from snapy_amd import MeshBlock, MeshBlockOptions

    op = MeshBlockOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", deck))
    for k, v in coord.items():
        setattr(op.coord, k, v)
    op.intg.execution = execution
    block = MeshBlock(op)
    block.to("cuda")
    return block


def workload(name):
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
    block = make_block("convection3d.yaml")
    vars, _ = block.initialize({"hydro_w": problems.convection3d(block).cuda()})
    c = block.options.coord
    cfg = {
        "workload": "convection3d_256x256x128",
        "cells": [c.nx1, c.nx2, c.nx3],
        "method": "weno5+lmars+ssp-rk3, well-balanced x1 gravity, flux covariance, bot-heat, "
        "vertical implicit scheme 9 (full 5x5, Woodbury), periodic x2/x3, cfl 0.8",
        "form": "plain FP64 (FP64 storage and compute, one block)",
    }
    # B_alg (cell-update bytes): FP64 state read u, u0, write u; 3 face
    # totals; the FP64 vertical-solve column I/O (3 rows x 5 x 8)
    return block, vars, c.nx1 * c.nx2 * c.nx3, cfg, "f64", 5 * (3 * 8 + 3 * 8) + 3 * 5 * 8


# --------------------------------------------------------------------------
def cpu_baseline(name, budget_s=10.0):
    """The CPU FP64 oracle (numpy, one thread) on a bounded sample of the
    same workload."""
    import numpy as np

    from oracle import fv

    t0 = time.perf_counter()
    stages = cells = runs = 0
    if name == "sod":
        while time.perf_counter() - t0 < budget_s:
            d = fv.sod_deck(256)
            b = fv.Block(d)
            b.initialize(fv.sod_initial(d.grid))
            _, n = fv.run(b, 0.25)
            stages += 3 * n
            runs += 1
        el = time.perf_counter() - t0
        return {
            "value": 256 * stages / el,
            "unit": "cell-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{runs} full Sod runs (256 cells, t_end 0.25), numpy FP64 oracle, {el:.1f} s",
        }
    # convection3d: a 24 x 24 column sub-domain of the same 128-layer column
    # physics (periodic x2/x3), whole RK3 steps
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
                bcs=tuple(op.bcnames), cfl=op.intg.cfl, grav1=h.grav1, implicit_scheme=9,
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
    n = c.nx1 * c.nx2 * c.nx3
    assert np.isfinite(b.u).all()
    return {
        "value": n * 3 * steps / el,
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} RK3 steps of a 128x24x24 sub-domain of the same deck, numpy FP64 oracle, {el:.1f} s",
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


def kernel_bytes(block, name, sb=8):
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
        "sweep_x1_wb": (5 * sb + 4 * 8) * ncell + 6 * nint * sb,
        "stage_increment": (5 * nd + 2 + 5) * nint * sb,
        "vic_full": 25 * nint * 8,
        "update": (3 + nd) * 5 * nint * sb,
        "update_gravity": ((3 + nd) * 5 + 2) * nint * sb,
    }
    if name.startswith("sweep_x"):
        return table.get(name, 5 * ncell * sb + 5 * nint * sb)
    return table.get(name, 0)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every GPU box here has one; run with "
                          "--gpus 1 (how the path shards over more GPUs is described in DESIGN.md)"}), flush=True)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    defaults = {"sod": (200, 10), "convection3d": (20, 3)}[args.workload]
    steps = args.steps or defaults[0]
    warmup = args.warmup if args.warmup is not None else defaults[1]

    block, vars, cells, cfg, dtype, b_alg = workload(args.workload)
    print(f"# workload {cfg['workload']} ready", file=sys.stderr, flush=True)
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
    ms_per_step = el / steps * 1e3
    value = cells * 3 * steps / el
    ok = bool(torch.isfinite(vars["hydro_u"]).all().item()) and block.check_redo(vars) == 0
    print(f"# timed {steps} steps in {el:.3f} s, state finite: {ok}", file=sys.stderr, flush=True)

    roofline = None
    if not args.no_kernels:
        kt = kernel_times(block, vars)
        dom = max(kt, key=kt.get)
        kb = kernel_bytes(block, dom)
        achieved = kb / (kt[dom] * 1e-3) / 1e9
        roofline = {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel_ms": kt,
            "stage_b_alg_per_cell": b_alg,
            "stage_frac": value * b_alg / (HBM_PEAK_GBS * 1e9),
        }
    cpu = None if args.no_cpu else cpu_baseline(args.workload)
    line = {
        "metric": "cell-updates/s per RK3 stage",
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
        "execution": block.execution_report(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
