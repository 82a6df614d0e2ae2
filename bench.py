#!/usr/bin/env python3
"""Benchmark of the finite-volume stage on one MI355X.

Prints ONE JSON line: cell-updates/s per RK3 stage of the named workload,
with the dominant kernel's roofline (HIP-event timed, algorithmic bytes) and
the CPU oracle timed on this host as the baseline.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A "step" is one RK3 step (3 stages) of the workload, the CFL reduction
included, run as captured hipGraphs with a device dt (no host read inside the
timed region).  Inputs are resident in HBM before timing starts.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="sod", choices=["sod"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    return ap.parse_args()


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
def sod_workload():
    import torch

    from snapy_amd import MeshBlock, MeshBlockOptions, kIDN, kIPR

    deck = os.path.join(ROOT, "snapy_amd", "decks", "sod.yaml")
    op = MeshBlockOptions.from_yaml(deck)
    op.intg.execution = "hip-graph"
    block = MeshBlock(op)
    block._executor.__init__(block, "hip-graph")
    block.to("cuda")
    x1v = block.module("coord").buffer("x1v").view(1, 1, -1)
    w = torch.zeros(block.shape(), dtype=torch.float64)
    w[kIDN] = torch.where(x1v < 0.5, torch.tensor(1.0, dtype=torch.float64), torch.tensor(0.125, dtype=torch.float64))
    w[kIPR] = torch.where(x1v < 0.5, torch.tensor(1.0, dtype=torch.float64), torch.tensor(0.1, dtype=torch.float64))
    vars, _ = block.initialize({"hydro_w": w.cuda()})
    cells = op.coord.nx1 * op.coord.nx2 * op.coord.nx3
    cfg = {
        "workload": "sod_1d_c1",
        "cells": [op.coord.nx1, op.coord.nx2, op.coord.nx3],
        "method": "weno5(shock)+hllc+ssp-rk3, outflow, cfl 0.9",
        "form": "plain FP64",
        "members": 1,
    }
    return block, vars, cells, cfg, "f64", 160.0


def sod_cpu_baseline(budget_s=10.0):
    """The CPU FP64 oracle (numpy) on the same Sod case, repeated whole runs
    until about budget_s seconds of CPU work."""
    from oracle import fv

    t0 = time.perf_counter()
    stages = 0
    runs = 0
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
        "sample": f"{runs} full Sod runs (256 cells, t_end 0.25) of the numpy FP64 oracle, {el:.1f} s",
    }


# --------------------------------------------------------------------------
def kernel_times(block, vars, reps=200):
    """Average launch duration (ms) of each stage kernel, timed alone with
    HIP events on the stream the kernels are launched on."""
    import torch

    from snapy_amd import _capi

    lib = _capi.load()
    g, p = ctypes.byref(block._grid), ctypes.byref(block._phys)
    u, w = vars["hydro_u"], vars["hydro_w"]
    dtc = _capi.dtype_code(u.dtype)
    s = _capi.stream_of(block.device)
    freg = block._freg
    calls = {}
    calls["cons2prim"] = lambda: lib.snapy_cons2prim(g, p, dtc, u.data_ptr(), w.data_ptr(), s)
    for d in (1, 2, 3):
        t = freg.total(d)
        if t is None:
            continue
        fp = freg.face_pressure1 if d == 1 else None
        calls[f"sweep_x{d}"] = (
            lambda d=d, t=t, fp=fp: lib.snapy_sweep(d, g, p, dtc, w.data_ptr(), t.data_ptr(), _capi.ptr(fp), s)
        )
    ucopy = u.clone()
    fl = (ctypes.c_void_p * 3)(*[_capi.ptr(freg.total(d)) for d in (1, 2, 3)])
    calls["update"] = lambda: lib.snapy_update(
        g, dtc, fl, block._state.u0().data_ptr(), ucopy.data_ptr(), None, 1e-9, None, 1.0, 0.0, 1.0, s
    )
    out = {}
    stream = torch.cuda.current_stream()
    for name, fn in calls.items():
        for _ in range(5):
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


def kernel_bytes(block, name, sbytes):
    """Algorithmic HBM bytes of one launch: every cell value read once, every
    output written once (no stencil re-reads)."""
    c = block.options.coord
    ncell = c.nc1() * c.nc2() * c.nc3()
    nint = c.nx1 * c.nx2 * c.nx3
    nv = 5
    if name == "cons2prim":
        return 2 * nv * ncell * sbytes
    if name.startswith("sweep"):
        return nv * ncell * sbytes + nv * nint * sbytes
    if name == "update":
        nd = sum(1 for n in (c.nx1, c.nx2, c.nx3) if n > 1)
        return nv * nint * sbytes * (3 + nd)
    return 0


def main():
    args = parse()
    if args.gpus != 1:
        print(
            json.dumps({"error": "this build measures one GPU: every box here has one; "
                        "run with --gpus 1 (multi-GPU sharding is described in DESIGN.md)"}),
            flush=True,
        )
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    steps = args.steps or 200
    warmup = args.warmup if args.warmup is not None else 10

    block, vars, cells, cfg, dtype, b_alg = sod_workload()
    # warm-up: eager, capture, replays
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

    kt = kernel_times(block, vars)
    dom = max(kt, key=kt.get)
    sb = 8 if dtype == "f64" else 4
    kb = kernel_bytes(block, dom, sb)
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
    cpu = None if args.no_cpu else sod_cpu_baseline()
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
        "execution": block.execution_report(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
