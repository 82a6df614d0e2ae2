"""Restart files on the CPU: the part name, the TorchScript tensor archive
(kintera::save_tensors' format), the bundle of several parts
(restart.cpp:150-321), the metadata checks of check_restart_metadata
(restart.cpp:323-358), the output schedule's keys and its restore by key
(output_type.cpp:14-42, meshblock_restart.cpp:61-123), and the budget carry
(restart.cpp:360-407)."""

import pytest
import torch

from snapy_amd import restart


class _Coord:
    def __init__(self, n=(8, 6, 1)):
        self.f = {k: torch.linspace(0.0, 1.0, m + 1, dtype=torch.float64) for k, m in zip(("x1f", "x2f", "x3f"), n)}

    def buffer(self, k):
        return self.f[k]


class _State:
    def __init__(self, comp):
        self._c = comp

    def comp(self):
        return self._c


class _Block:
    """What restart_variables reads of a block."""

    def cycle(self):
        return self._cycle

    def __init__(self, fp32):
        torch.manual_seed(0)
        self._cycle = 17
        self._coord = _Coord()
        dt = torch.float32 if fp32 else torch.float64
        self.u = torch.rand((5, 1, 12, 14), dtype=torch.float64).to(dt)
        self._state = _State(torch.rand_like(self.u) * 1e-8 if fp32 else None)

    def total_u(self, vars):
        c = self._state.comp()
        return self.u.double() + c.double() if c is not None else self.u.clone()


def test_part_name():
    assert restart.part_name("run", 3, 12) == "run.block3.00012.part"
    assert restart.part_name("run") == "run.block0.final.part"


@pytest.mark.parametrize("fp32", [False, True])
def test_round_trip_and_checks(tmp_path, fp32):
    blk = _Block(fp32)
    w = torch.rand((5, 1, 12, 14), dtype=torch.float64)
    vars = {"hydro_u": blk.u, "hydro_w": w}
    path = str(tmp_path / restart.part_name("run", 0, 4))
    restart.write_restart(blk, vars, path, 12.5)
    data = restart.load_restart(path)
    assert torch.equal(data["hydro_w"], w)
    assert torch.equal(data["hydro_u_stored"], blk.u)
    assert torch.equal(data["hydro_u"], blk.total_u(vars))
    assert data["hydro_u"].dtype == torch.float64
    assert float(data["last_time"][0]) == 12.5 and int(data["last_cycle"][0]) == 17
    assert ("hydro_u_comp" in data) == fp32
    faces = tuple(blk._coord.buffer(k) for k in ("x1f", "x2f", "x3f"))
    restart.check_restart_metadata(data, faces)
    # each check refuses its own mistake
    bad = dict(data)
    del bad["last_cycle"]
    with pytest.raises(restart.RestartError, match="last_cycle"):
        restart.check_restart_metadata(bad, faces)
    bad = dict(data)
    bad[restart.METADATA] = torch.tensor([restart.VERSION + 1, int(data[restart.METADATA][1])])
    with pytest.raises(restart.RestartError, match="version"):
        restart.check_restart_metadata(bad, faces)
    bad = dict(data)
    bad[restart.METADATA] = torch.tensor([restart.VERSION, 96 - int(data[restart.METADATA][1])])
    with pytest.raises(restart.RestartError, match="dtype/shape"):
        restart.check_restart_metadata(bad, faces)
    other = (faces[0] * 2.0, faces[1], faces[2])
    with pytest.raises(restart.RestartError, match="grid mismatch at hydro_ref_x1f"):
        restart.check_restart_metadata(data, other)
    if fp32:
        bad = dict(data)
        del bad["hydro_u_comp"]
        with pytest.raises(restart.RestartError, match="compensation"):
            restart.check_restart_metadata(bad, faces)


def test_refuses_missing_and_foreign_files(tmp_path):
    with pytest.raises(restart.RestartError, match="failed to open"):
        restart.load_restart(str(tmp_path / "nope.part"))
    p = tmp_path / "junk.part"
    p.write_bytes(b"not an archive")
    with pytest.raises(restart.RestartError):
        restart.load_restart(str(p))
    from safetensors.torch import save_file

    q = str(tmp_path / "other.part")
    save_file({"x": torch.zeros(2)}, q)
    with pytest.raises(restart.RestartError, match="not a restart part"):
        restart.load_restart(q)


def test_archive_is_the_reference_tensor_archive(tmp_path):
    """The part is a TorchScript module whose buffers are the tensors, as the
    reference's tests/test_fp32_restart.py:22-29 writes one: torch.jit.load
    reads it back by name, dotted keys included, and a part written the
    reference's way (a traced module with buffers) loads here."""
    blk = _Block(True)
    vars = {"hydro_u": blk.u, "hydro_w": torch.rand((5, 1, 12, 14), dtype=torch.float64)}
    path = str(tmp_path / restart.part_name("run", 0, 2))
    extra = restart.carry_tensors(["mass", "energy"], torch.tensor([1.0, 2.0], dtype=torch.float64),
                                  torch.tensor([0.5, -0.5], dtype=torch.float64), 3.0)
    restart.write_restart(blk, vars, path, 1.0, extra=extra)
    with open(path, "rb") as f:
        assert f.read(2) == b"PK"
    mod = torch.jit.load(path, map_location="cpu")
    bufs = dict(mod.named_buffers())
    assert torch.equal(bufs["hydro_u_stored"], blk.u)
    assert torch.equal(bufs["budget_carry.pending"], extra["budget_carry.pending"])

    class Archive(torch.nn.Module):
        def __init__(self, data):
            super().__init__()
            for name, tensor in data.items():
                self.register_buffer(name, tensor.clone())

        def forward(self, x):
            return x

    theirs = str(tmp_path / "theirs.part")
    data = restart.restart_variables(blk, vars, 4.0)
    torch.jit.trace(Archive(data), torch.zeros(1), check_trace=False).save(theirs)
    back = restart.load_restart(theirs)
    assert set(back) == set(data)
    for k in data:
        assert torch.equal(back[k], data[k]), k


def test_bundle_of_parts(tmp_path):
    """Three block parts combine into one bundle in rank order
    (make_restart_bundle): the magic line, the count, name<TAB>bytes lines and
    an empty line; each block's part reads back from it, a missing rank is
    refused, and a lone part is renamed, not bundled."""
    d = str(tmp_path)
    for rank in (2, 0, 1):
        restart.save_tensors({"hydro_u": torch.full((3,), float(rank), dtype=torch.float64)},
                             f"{d}/{restart.part_name('run', rank, 7)}")
    sizes = [__import__("os").path.getsize(f"{d}/{restart.part_name('run', r, 7)}") for r in range(3)]
    out = restart.combine_parts(d, "run", 7)
    assert out.endswith("run.00007.restart")
    with open(out, "rb") as f:
        head = f.read().split(b"\n", 5)
    assert head[0] == b"SNAPY_RESTART_BUNDLE_V1" and head[1] == b"3"
    assert head[2:5] == [f"run.block{r}.00007.part\t{sizes[r]}".encode() for r in range(3)]
    assert head[5].startswith(b"\nPK")
    for rank in range(3):
        assert float(restart.load_restart(out, rank)["hydro_u"][0]) == rank
    with pytest.raises(restart.RestartError, match="no part for block 5"):
        restart.load_restart(out, 5)
    assert not list(tmp_path.glob("*.part"))
    restart.save_tensors({"hydro_u": torch.zeros(2)}, f"{d}/{restart.part_name('run', 0)}")
    lone = restart.combine_parts(d, "run", None)
    assert lone.endswith("run.final.restart") and float(restart.load_restart(lone)["hydro_u"].sum()) == 0.0


def test_schedule_keys_are_fnv1a():
    """FNV-1a 64 (the published test vectors pin the hash), over the
    reference's key strings."""
    assert restart._fnv(b"") == 0xCBF29CE484222325
    assert restart._fnv(b"a") == 0xAF63DC4C8601EC8C
    assert restart._fnv(b"foobar") == 0x85944171F73967E8
    o = output_options("netcdf", 300.0, ["prim", "uov"])
    want = restart._signed(restart._fnv(b"netcdf|300.000000|prim|uov"))
    assert restart.schedule_key(o) == want
    import struct

    raw = b"netcdf\0" + struct.pack("<d", 300.0) + b"\0prim\0uov"
    assert restart.schedule_key_v2(o) == restart._signed(restart._fnv(raw))
    assert restart.schedule_key_v2(o) != restart.schedule_key_v2(output_options("netcdf", 300.0, ["prim"]))


def output_options(t, dt, variables=()):
    from snapy_amd.output import OutputOptions

    return OutputOptions(file_type=t, dt=dt, variables=list(variables))


def test_schedule_restored_by_key():
    """Saved slots match by exact key whatever the order; the file counter
    follows the position; a next time more than dt past the resume is pulled
    back; a new output joins the grid of its dt; legacy keys are used when
    the file holds only those."""
    a = output_options("netcdf", 10.0, ["prim"])
    b = output_options("restart", 100.0)
    c = output_options("budget", 5.0)
    data = restart.schedule_tensors([a, b], [3, 1], [40.0, 1.0e9])
    fn, nt = restart.restore_schedule(data, [b, a], 35.0)
    assert fn == [3, 1]
    assert nt == [135.0, 40.0]  # b's 1e9 pulled back to 35 + 100; a keeps its slot's 40
    fn, nt = restart.restore_schedule(data, [a, b, c], 37.0)
    assert fn == [3, 1, 0] and nt[:2] == [40.0, 137.0] and nt[2] == 35.0
    legacy = dict(data)
    legacy["output_key"] = data["output_key"][:2]
    assert restart.restore_schedule(legacy, [b, a], 35.0)[1] == [135.0, 40.0]
    del legacy["output_key"]
    assert restart.restore_schedule(legacy, [a, b], 35.0)[1] == [40.0, 135.0]


def test_budget_carry_round_trip():
    rows = ["mass", "momentum1", "energy"]
    c = restart.carry_tensors(rows, torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64),
                              torch.tensor([0.1, 0.2, 0.3], dtype=torch.float64), 12.5)
    assert restart.unpack_names(c["budget_carry.rows"]) == rows
    tp, pend, interval = restart.read_carry(c, rows)
    assert tp.tolist() == [1.0, 2.0, 3.0] and pend.tolist() == [0.1, 0.2, 0.3] and interval == 12.5
    assert restart.read_carry(c, rows[:2]) is None
    bad = dict(c)
    del bad["budget_carry.minimum"]
    assert restart.read_carry(bad, rows) is None
