"""Restart files of a block: the state a resume needs to continue bit for bit.

One part file per block, `<basename>.block<rank>.<number>.part` (or
`.final.part`), as the reference names them (RestartOutput::
write_output_file, src/output/restart.cpp:42-147), in the reference's
archive: a TorchScript module whose buffers are the part's tensors, the
format kintera::save_tensors writes and kintera::load_tensors reads
(tests/test_fp32_restart.py:22-29 writes the same archive from Python).  A
dotted key (`budget_carry.rows`) is a buffer of a child module, so the
archive's recursive buffer names are the keys.

A restart output combines the parts of one number into
`<basename>.<number>.restart` (`.final.restart`): a single part is renamed,
several become one bundle (make_restart_bundle, restart.cpp:150-187): the
line `SNAPY_RESTART_BUNDLE_V1`, the part count, one `<part name>\\t<bytes>`
line per part, an empty line, then the parts in index order.  load_restart
reads either form (load_restart, restart.cpp:270-321), and a part written by
an earlier build of this package as safetensors.

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
- `last_time`, `last_cycle`;
- the output schedule: `file_number`, `next_time` and `output_key` (each
  output's legacy key, then its exact key; restart.cpp:91-121), restored
  by key on resume (meshblock_restart.cpp:61-123);
- with a budget output, the open budget interval as `budget_carry.*`
  (write_budget_carry, restart.cpp:382-407).
"""

from __future__ import annotations

import io
import math
import os
import struct
import warnings

import torch

VERSION = 1
METADATA = "hydro_restart_metadata"
REF_GRID = ("hydro_ref_x1f", "hydro_ref_x2f", "hydro_ref_x3f")
REQUIRED = ("hydro_u", "hydro_w", "last_time", "last_cycle")
BUNDLE_MAGIC = "SNAPY_RESTART_BUNDLE_V1"
CARRY_PREFIX = "budget_carry."


class RestartError(RuntimeError):
    pass


def part_name(basename: str, rank: int = 0, number: int | None = None) -> str:
    """`<basename>.block<rank>.<number:05d>.part`, `.final.part` for None."""
    tag = "final" if number is None else f"{number:05d}"
    return f"{basename}.block{rank}.{tag}.part"


def restart_name(basename: str, number: int | None = None) -> str:
    """`<basename>.<number:05d>.restart`, `.final.restart` for None."""
    tag = "final" if number is None else f"{number:05d}"
    return f"{basename}.{tag}.restart"


def part_rank(name: str) -> int:
    """The rank in a part name `<basename>.block<rank>.<number>.part`, or -1
    (part_rank, restart.cpp:245-255)."""
    if not name.endswith(".part"):
        return -1
    stem = name[: -len(".part")]
    dot2 = stem.rfind(".")
    dot1 = stem.rfind(".", 0, dot2) if dot2 > 0 else -1
    if dot1 < 0 or not stem.startswith("block", dot1 + 1):
        return -1
    try:
        return int(stem[dot1 + 6 : dot2])
    except ValueError:
        return -1


# ---------------------------------------------------------------------------
# the archive
# ---------------------------------------------------------------------------
class _Archive(torch.nn.Module):
    """Buffers only; a dotted key is a buffer of a child module."""

    def __init__(self, tensors: dict):
        super().__init__()
        children: dict = {}
        for name, t in tensors.items():
            head, dot, rest = name.partition(".")
            if dot:
                children.setdefault(head, {})[rest] = t
            else:
                self.register_buffer(name, t)
        for head, sub in children.items():
            self.add_module(head, _Archive(sub))

    def forward(self, x):
        return x


def save_tensors(data: dict, path: str) -> None:
    """Write `data` (CPU tensors) to `path` as a TorchScript tensor archive;
    the file appears whole (written aside, then renamed)."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with warnings.catch_warnings():
        # torch marks TorchScript deprecated; the archive is the reference's
        warnings.simplefilter("ignore", DeprecationWarning)
        mod = torch.jit.trace(_Archive({k: v.detach().to("cpu").contiguous() for k, v in data.items()}),
                              torch.zeros(1), check_trace=False)
        torch.jit.save(mod, tmp)
    os.replace(tmp, path)


def load_tensors(src) -> dict:
    """The buffers of a TorchScript tensor archive (a path or bytes), on the
    CPU, by their recursive names."""
    f = io.BytesIO(src) if isinstance(src, (bytes, bytearray)) else src
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", DeprecationWarning)
        mod = torch.jit.load(f, map_location="cpu")
    return {k: v for k, v in mod.named_buffers()}


def _load_safetensors_part(path: str) -> dict:
    from safetensors import safe_open
    from safetensors.torch import load_file

    with safe_open(path, framework="pt") as f:
        meta = f.metadata() or {}
    if meta.get("format") != "snapy_amd restart":
        raise RestartError(f"{path}: not a restart part")
    return load_file(path)


# ---------------------------------------------------------------------------
# bundles
# ---------------------------------------------------------------------------
def make_bundle(bundle: str, parts: list) -> str:
    """One bundle of the part files `parts`, in their order
    (make_restart_bundle, restart.cpp:150-187)."""
    sizes = [os.path.getsize(p) for p in parts]
    tmp = bundle + ".tmp"
    with open(tmp, "wb") as out:
        out.write(f"{BUNDLE_MAGIC}\n{len(parts)}\n".encode())
        for p, s in zip(parts, sizes):
            out.write(f"{os.path.basename(p)}\t{s}\n".encode())
        out.write(b"\n")
        for p in parts:
            with open(p, "rb") as f:
                while True:
                    chunk = f.read(1 << 20)
                    if not chunk:
                        break
                    out.write(chunk)
    os.replace(tmp, bundle)
    return bundle


def combine_parts(output_dir: str, basename: str, number: int | None = None) -> str:
    """RestartOutput::combine_blocks (restart.cpp:189-243): the parts of
    `number` become `<basename>.<number>.restart`, a lone part renamed, more
    than one bundled (in rank order) and removed."""
    import glob

    tag = "final" if number is None else f"{number:05d}"
    pattern = os.path.join(output_dir, f"{glob.escape(basename)}.block*.{tag}.part")
    parts = sorted(glob.glob(pattern), key=lambda p: part_rank(os.path.basename(p)))
    if not parts:
        raise RestartError(f"combine_parts: no part matches {pattern}")
    out = os.path.join(output_dir, restart_name(basename, number))
    if os.path.exists(out):
        os.remove(out)
    if len(parts) == 1:
        os.replace(parts[0], out)
    else:
        make_bundle(out, parts)
        for p in parts:
            os.remove(p)
    return out


def load_restart(path: str, block_rank: int = 0) -> dict:
    """The tensors of block `block_rank`'s part, on the CPU, from a part, a
    combined restart file or a bundle (load_restart, restart.cpp:270-321);
    raises RestartError when the file is missing or holds no part of the
    block."""
    if not os.path.exists(path):
        raise RestartError(f"{path}: failed to open restart file")
    with open(path, "rb") as f:
        head = f.read(len(BUNDLE_MAGIC) + 1)
        if head == (BUNDLE_MAGIC + "\n").encode():
            line = f.readline()
            if not line.strip():
                raise RestartError(f"{path}: restart bundle missing entry count")
            entries = []
            for _ in range(int(line)):
                line = f.readline().decode().rstrip("\n")
                if "\t" not in line:
                    raise RestartError(f"{path}: restart bundle index line missing tab")
                name, size = line.split("\t", 1)
                entries.append((name, int(size)))
            if f.readline() != b"\n":
                raise RestartError(f"{path}: restart bundle header missing terminator")
            for name, size in entries:
                if part_rank(name) != block_rank:
                    f.seek(size, os.SEEK_CUR)
                    continue
                payload = f.read(size)
                if len(payload) != size:
                    raise RestartError(f"{path}: truncated restart bundle payload")
                return load_tensors(payload)
            raise RestartError(f"{path}: no part for block {block_rank} in restart bundle")
    if head[:2] == b"PK":
        try:
            return load_tensors(path)
        except Exception as e:  # a zip that is not a TorchScript archive
            raise RestartError(f"{path}: not a readable restart part ({e})") from e
    try:
        return _load_safetensors_part(path)
    except RestartError:
        raise
    except Exception as e:  # a truncated or foreign file
        raise RestartError(f"{path}: not a readable restart part ({e})") from e


# ---------------------------------------------------------------------------
# the part's tensors
# ---------------------------------------------------------------------------
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


def write_restart(block, vars, path: str, time: float, extra: dict | None = None) -> str:
    """Write the part file of `block` to `path` (its directory is made) and
    return the path; `extra` adds tensors (the output schedule and the budget
    carry)."""
    data = restart_variables(block, vars, time)
    if extra:
        data.update(extra)
    save_tensors(data, path)
    return path


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


# ---------------------------------------------------------------------------
# the output schedule and the budget carry
# ---------------------------------------------------------------------------
_FNV_OFFSET = 14695981039346656037
_FNV_PRIME = 1099511628211
_MASK = (1 << 64) - 1


def _fnv(data: bytes, h: int = _FNV_OFFSET) -> int:
    for c in data:
        h = ((h ^ c) * _FNV_PRIME) & _MASK
    return h


def _signed(h: int) -> int:
    return h - (1 << 64) if h >= (1 << 63) else h


def schedule_key(o) -> int:
    """OutputType::schedule_key (output_type.cpp:14-20): FNV-1a of
    `<type>|<dt as %f>|<variable>...`."""
    s = f"{o.file_type}|{o.dt:f}" + "".join("|" + v for v in o.variables)
    return _signed(_fnv(s.encode()))


def schedule_key_v2(o) -> int:
    """OutputType::schedule_key_v2 (output_type.cpp:22-42): FNV-1a of the
    type, a zero byte, the eight little-endian bytes of dt's bits, then per
    variable a zero byte and its name."""
    h = _fnv(o.file_type.encode())
    h = _fnv(b"\0" + struct.pack("<d", float(o.dt)), h)
    for v in o.variables:
        h = _fnv(b"\0" + v.encode(), h)
    return _signed(h)


def schedule_tensors(outputs, file_number, next_time) -> dict:
    """`file_number`, `next_time` and `output_key` (the legacy keys, then the
    exact ones) of a block's outputs (restart.cpp:91-121)."""
    keys = [schedule_key(o) for o in outputs] + [schedule_key_v2(o) for o in outputs]
    return {
        "file_number": torch.tensor([int(n) for n in file_number], dtype=torch.int64),
        "next_time": torch.tensor([float(t) for t in next_time], dtype=torch.float64),
        "output_key": torch.tensor(keys, dtype=torch.int64),
    }


def restore_schedule(data: dict, outputs, current_time: float):
    """(file_number, next_time) of `outputs` from a restart's schedule
    (MeshBlockImpl::_init_from_restart, meshblock_restart.cpp:61-123): each
    output takes its saved slot when the key still matches, else the first
    unclaimed slot of its key, else its own position; the counter follows the
    position; a saved next time more than dt past the resume is pulled back
    to one dt; a new output joins the grid of its dt."""
    n_out = len(outputs)
    if "file_number" not in data or "next_time" not in data:
        raise RestartError("restart file is missing required variable: file_number")
    fn = data["file_number"].tolist()
    nt = data["next_time"].tolist()
    nsaved = len(fn)
    keys = data.get("output_key")
    keys = keys.tolist() if keys is not None and keys.dim() == 1 else None
    precise = keys is not None and len(keys) == 2 * nsaved
    off = nsaved if precise else 0
    cur = [(schedule_key_v2(o) if precise else schedule_key(o)) for o in outputs]
    claimed = [False] * nsaved
    match = [-1] * n_out
    if keys is not None:
        for n in range(min(nsaved, n_out)):
            if keys[off + n] == cur[n]:
                match[n] = n
                claimed[n] = True
        for n in range(n_out):
            if match[n] >= 0:
                continue
            for k in range(nsaved):
                if not claimed[k] and keys[off + k] == cur[n]:
                    match[n] = k
                    claimed[k] = True
                    break
    for n in range(n_out):
        if match[n] < 0 and n < nsaved and not claimed[n]:
            match[n] = n
            claimed[n] = True
    file_number, next_time = [0] * n_out, [0.0] * n_out
    for n, o in enumerate(outputs):
        if n < nsaved:
            file_number[n] = int(fn[n])
        m = match[n]
        if m >= 0:
            saved = float(nt[m])
            next_time[n] = current_time + o.dt if (o.dt > 0.0 and saved > current_time + o.dt) else saved
        else:
            next_time[n] = current_time - math.fmod(current_time, o.dt) if o.dt > 0.0 else current_time
    return file_number, next_time


def pack_names(names) -> torch.Tensor:
    """The row names as one int8 tensor, each newline-terminated
    (pack_names, restart.cpp:360-367)."""
    b = "".join(n + "\n" for n in names).encode()
    return torch.tensor(list(b), dtype=torch.uint8).to(torch.int8)


def unpack_names(t: torch.Tensor) -> list:
    b = bytes(t.to(torch.int8).view(torch.uint8).tolist())
    return b.decode().splitlines()


def carry_tensors(rows, total_prev, pending, interval: float) -> dict:
    """write_budget_carry (restart.cpp:382-392): the open interval's start
    totals, its boundary transfer and its length (no monitor counters or
    minima are kept here: empty tensors)."""
    p = CARRY_PREFIX
    return {
        p + "rows": pack_names(rows),
        p + "total_prev": total_prev.detach().to("cpu", torch.float64),
        p + "pending": pending.detach().to("cpu", torch.float64),
        p + "counter": torch.zeros(0, dtype=torch.int64),
        p + "minimum": torch.zeros(0, dtype=torch.float64),
        p + "interval": torch.tensor([float(interval)], dtype=torch.float64),
    }


def read_carry(data: dict, rows):
    """read_budget_carry (restart.cpp:394-407): (total_prev, pending,
    interval), or None when the file holds none or one with other rows."""
    p = CARRY_PREFIX
    for f in ("rows", "total_prev", "pending", "counter", "minimum", "interval"):
        if p + f not in data:
            return None
    if unpack_names(data[p + "rows"]) != list(rows):
        return None
    return (data[p + "total_prev"].to(torch.float64), data[p + "pending"].to(torch.float64),
            float(data[p + "interval"][0]))
