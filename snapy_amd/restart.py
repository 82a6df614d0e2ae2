"""Restart files of a block: the state a resume needs to continue bit for bit.

One part file per block, `<basename>.block<rank>.<number>.part` (or
`.final.part`), as the reference names them (RestartOutput::
write_output_file, src/output/restart.cpp:42-147).  The reference writes
them with kintera's tensor archive; here the archive is safetensors, which
loads without executing anything from the file.

What a part holds (the reference's keys where it has them):
- every tensor of the block's variables (`hydro_w`, `boundary_reference_w`,
  and with passive scalars `scalar_s`, `scalar_r`);
- `hydro_u`: the conserved totals, FP64;
- `hydro_u_stored`: the stored state's own bits, and for the FP32 state form
  `hydro_u_comp`, its compensation word (this build's pair), so a resume of
  the same form reuses the same bits (restart.cpp:60-70);
- `hydro_restart_metadata` = [version, stored bits] and the grid faces
  `hydro_ref_x1f/x2f/x3f`, checked on load (check_restart_metadata,
  restart.cpp:323-358);
- `last_time`, `last_cycle`.

The reference's output schedules, budget carry and part bundles belong to
its output layer and are not written.
"""

from __future__ import annotations

import os

import torch

VERSION = 1
METADATA = "hydro_restart_metadata"
REF_GRID = ("hydro_ref_x1f", "hydro_ref_x2f", "hydro_ref_x3f")
REQUIRED = ("hydro_u", "hydro_w", "last_time", "last_cycle")


class RestartError(RuntimeError):
    pass


def part_name(basename: str, rank: int = 0, number: int | None = None) -> str:
    """`<basename>.block<rank>.<number:05d>.part`, `.final.part` for None."""
    tag = "final" if number is None else f"{number:05d}"
    return f"{basename}.block{rank}.{tag}.part"


def grid_faces(block) -> tuple:
    return tuple(block._coord.buffer(k).detach().to("cpu", torch.float64).contiguous() for k in ("x1f", "x2f", "x3f"))


def restart_variables(block, vars, time: float) -> dict:
    """The part's tensors (on the CPU) for `block` at `time`.  The
    primitives saved are those of the saved state (refreshed first when the
    block has stepped since, as the reference's check_redo leaves them)."""
    if getattr(block, "_w_stale", False):
        block._refresh_primitives(vars)
    out = {}
    for k in ("hydro_w", "boundary_reference_w", "scalar_s", "scalar_r"):
        if vars.get(k) is not None:
            out[k] = vars[k].detach().to("cpu").contiguous()
    u = vars["hydro_u"]
    out["hydro_u"] = block.total_u(vars).detach().to("cpu", torch.float64).contiguous()
    out["hydro_u_stored"] = u.detach().to("cpu").contiguous()
    comp = block._state.comp()
    if comp is not None:
        out["hydro_u_comp"] = comp.detach().to("cpu").contiguous()
    bits = 32 if u.dtype == torch.float32 else 64
    out[METADATA] = torch.tensor([VERSION, bits], dtype=torch.int64)
    for key, f in zip(REF_GRID, grid_faces(block)):
        out[key] = f
    out["last_time"] = torch.tensor([float(time)], dtype=torch.float64)
    out["last_cycle"] = torch.tensor([int(block.cycle())], dtype=torch.int64)
    return out


def write_restart(block, vars, path: str, time: float) -> str:
    """Write the part file of `block` to `path` (its directory is made) and
    return the path; the file appears whole (written aside, then renamed)."""
    from safetensors.torch import save_file

    data = restart_variables(block, vars, time)
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    save_file(data, tmp, metadata={"format": "snapy_amd restart", "version": str(VERSION)})
    os.replace(tmp, path)
    return path


def load_restart(path: str) -> dict:
    """The part's tensors, on the CPU; raises RestartError when the file is
    missing or not a restart part."""
    from safetensors import safe_open
    from safetensors.torch import load_file

    if not os.path.exists(path):
        raise RestartError(f"{path}: failed to open restart file")
    try:
        with safe_open(path, framework="pt") as f:
            meta = f.metadata() or {}
        data = load_file(path)
    except Exception as e:  # a truncated or foreign file
        raise RestartError(f"{path}: not a readable restart part ({e})") from e
    if meta.get("format") != "snapy_amd restart":
        raise RestartError(f"{path}: not a restart part")
    return data


def check_restart_metadata(data: dict, faces: tuple) -> None:
    """check_restart_metadata (restart.cpp:323-358): the required keys, the
    metadata version and stored bits against the stored tensor, and the grid
    faces against this block's (a same-shaped but different grid is refused)."""
    for k in REQUIRED:
        if k not in data:
            raise RestartError(f"restart file is missing required variable: {k}")
    meta = data.get(METADATA)
    if meta is None or meta.dtype != torch.int64 or meta.dim() != 1 or meta.numel() != 2:
        raise RestartError("restart: invalid hydro restart metadata")
    if int(meta[0]) != VERSION:
        raise RestartError("restart: unsupported hydro restart metadata version")
    stored = data.get("hydro_u_stored")
    if stored is None:
        raise RestartError("restart: metadata requires hydro_u_stored")
    bits = {torch.float32: 32, torch.float64: 64}.get(stored.dtype, 0)
    if not bits or int(meta[1]) != bits or stored.shape != data["hydro_u"].shape:
        raise RestartError("restart: stored hydro dtype/shape disagrees with metadata")
    if bits == 32 and "hydro_u_comp" not in data:
        raise RestartError("restart: an FP32 stored state needs its compensation hydro_u_comp")
    for key, f in zip(REF_GRID, faces):
        if key not in data or not torch.equal(data[key], f):
            raise RestartError(f"restart: reference grid mismatch at {key}")
