"""The launch-shape autotuner.

For a block's shape and dtype it chooses the workgroup size of three kernel
families -- the EOS (cons2prim), the face sweeps (x1, and the plain x2/x3
sweeps) and the divergence + combine updates -- and the columns per
workgroup of the fused x1 reference sweep -- by timing every candidate on
the device
with HIP events, and it chooses the persistent grid of the work list
(`worklist.BlockBatch`).  A choice goes into the library's tuning
table (snapy_tune_set), which every later launch of that family for that
shape and dtype reads, and into a JSON cache, so a later process applies it
without timing again (`apply_cache`).

A workgroup size or a column count changes only how the cells (columns)
are grouped into workgroups, never the arithmetic of a cell, so tuned and
untuned runs give the same bits (tests/test_gpu_autotune.py).
"""

from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path

import torch

from . import _capi

TUNE_EOS, TUNE_SWEEP, TUNE_UPDATE, TUNE_WL_GRID, TUNE_VIC_COLS, TUNE_REF_COLS = 0, 1, 2, 3, 4, 5
WORKGROUPS = (64, 128, 256)
WL_GRIDS = (256, 512, 1024, 2048, 4096)

# family -> (tunable, the stage launches it covers, candidates, default)
FAMILIES = {
    "eos": (TUNE_EOS, ("cons2prim",), WORKGROUPS, 256),
    "sweep": (TUNE_SWEEP, ("sweep_x1", "sweep_x2", "sweep_x3"), WORKGROUPS, 256),
    "update": (TUNE_UPDATE, ("update", "update_gravity", "stage_increment"), WORKGROUPS, 256),
    "ref_cols": (TUNE_REF_COLS, ("sweep_x1_wb", "sweep_x1_increment"), (1, 2, 4), 4),
}
# set by hand only (set_choice): timed inside the whole vertical solve, whose
# forward kernel dominates, the finish's column count is not resolved
# reliably (the tuner has picked 2, 4 and 8 on one shape; 8 is fastest there)
MANUAL = {
    "vic_cols": (TUNE_VIC_COLS, ("vic_partial",), (1, 2, 4, 8), 8),
}


def default_cache() -> Path:
    """$SNAPY_TUNE_CACHE, else ~/.cache/snapy_amd/tune.json."""
    env = os.environ.get("SNAPY_TUNE_CACHE")
    return Path(env) if env else Path.home() / ".cache" / "snapy_amd" / "tune.json"


def shape_key(block) -> str:
    """The cache key of a block: its interior shape and compute dtype."""
    c = block.options.coord
    dt = "f32" if block.policy.fp32() else "f64"
    return f"{c.nx1}x{c.nx2}x{c.nx3}:{dt}"


def _dtype_code(block) -> int:
    return 1 if block.policy.fp32() else 0


def _tunable(family: str) -> int:
    return (FAMILIES.get(family) or MANUAL[family])[0]


def set_choice(block, family: str, value: int) -> None:
    """Put one family's workgroup size (column count) for this block's shape
    in the table (0 restores the default)."""
    lib = _capi.load()
    t = _tunable(family)
    _capi.check(lib.snapy_tune_set(t, ctypes.byref(block._grid), _dtype_code(block), int(value)), "tune_set")


def get_choice(block, family: str) -> int:
    lib = _capi.load()
    return int(lib.snapy_tune_get(_tunable(family), ctypes.byref(block._grid), _dtype_code(block)))


def _time(fns, reps: int) -> float:
    stream = torch.cuda.current_stream()
    for _ in range(2):
        for fn in fns:
            _capi.check(fn(), "autotune")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        for fn in fns:
            _capi.check(fn(), "autotune")
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_block(block, vars, reps: int = 10, cache: Path | str | None = None,
               rounds: int = 3, tol: float = 0.02) -> dict:
    """Time every candidate of every family on this block (its stage-0
    launches; the block's state is restored afterwards), keep the fastest in
    the library's table and, when `cache` is given (or True: the default
    path), in the JSON cache.  The candidates are timed in `rounds` passes
    in turn and each keeps its best pass; the default stays unless another
    candidate beats it by more than `tol`, so timing noise does not move a
    shape.  Returns {family: {"choice": value, "ms": {value: ms}}}."""
    out = {}
    # the stage-0 launches advance the state: keep it and restore it after
    keep = [vars["hydro_u"], vars["hydro_w"], block._state.u0()]
    keep += [t for t in (block._state.comp(), block._state.comp0()) if t is not None]
    saved = [t.clone() for t in keep]
    calls = dict(block._stage_calls(vars, 0))
    for fam, (_, names, cands, dflt) in FAMILIES.items():
        fns = [calls[n] for n in names if n in calls]
        if not fns:
            continue
        ms = {v: float("inf") for v in cands}
        for _ in range(max(1, rounds)):
            for v in cands:
                set_choice(block, fam, v)
                ms[v] = min(ms[v], _time(fns, reps))
        best = min(cands, key=lambda v: ms[v])
        if ms[dflt] <= ms[best] * (1.0 + tol):
            best = dflt
        set_choice(block, fam, best)
        out[fam] = {"choice": best, "ms": ms}
    for t, c in zip(keep, saved):
        t.copy_(c)
    if cache:
        path = default_cache() if cache is True else Path(cache)
        save(path, shape_key(block), {f: r["choice"] for f, r in out.items()})
    return out


def tune_work_list(batch, reps: int = 10) -> dict:
    """Time the work list's idempotent phases (cons2prim and the x1 sweep)
    at every persistent grid cap and keep the fastest."""
    lib = _capi.load()
    wl, p = ctypes.byref(batch._wl), ctypes.byref(batch._phys)
    s = _capi.stream_of(batch.device)
    fns = [lambda: lib.snapy_wl_cons2prim(wl, p, s), lambda: lib.snapy_wl_sweep_x1(wl, p, s)]
    ms = {}
    for v in WL_GRIDS:
        _capi.check(lib.snapy_tune_set(TUNE_WL_GRID, None, 0, v), "tune_set")
        ms[v] = _time(fns, reps)
    best = min(WL_GRIDS, key=lambda v: (ms[v], v != 2048))
    _capi.check(lib.snapy_tune_set(TUNE_WL_GRID, None, 0, best), "tune_set")
    return {"choice": best, "ms": ms}


def load(path: Path | str | None = None) -> dict:
    p = Path(path) if path else default_cache()
    if not p.exists():
        return {}
    with open(p) as f:
        return json.load(f)


def save(path: Path | str, key: str, choices: dict) -> None:
    p = Path(path)
    table = load(p)
    table[key] = {f: int(v) for f, v in choices.items()}
    p.parent.mkdir(parents=True, exist_ok=True)
    tmp = p.with_suffix(".tmp")
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    tmp.replace(p)


def apply_cache(block, path: Path | str | None = None) -> dict:
    """Apply the cached choices for this block's shape and dtype, if any;
    returns them ({} when the shape was never tuned)."""
    choices = load(path).get(shape_key(block), {})
    for fam, v in choices.items():
        if fam in FAMILIES or fam in MANUAL:
            set_choice(block, fam, v)
    return choices
