"""Resume from a restart part on the GPU: a block stepped N + M steps and a
block restarted from the part written after N steps, then stepped M, hold
the same bits (FP64 and the FP32 state form, a captured and an outflow
deck); the cycle and time come back; an FP64 part resumes in the FP32 form."""

import os

import pytest
import torch

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _convection(precision, execution="hip-graph"):
    from snapy_amd import MeshBlock, MeshBlockOptions, problems

    op = MeshBlockOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", "convection3d.yaml"))
    op.coord.nx1, op.coord.nx2, op.coord.nx3 = 32, 24, 16
    op.coord.x1max, op.coord.x2max, op.coord.x3max = 1600.0, 2400.0, 1600.0
    op.intg.execution = execution
    op.hydro.precision = precision
    blk = MeshBlock(op)
    blk.to("cuda")
    return blk, (lambda b: problems.convection3d(b))


def _sod():
    from snapy_amd import serving
    from snapy_amd.meshblock import MeshBlock

    op = serving._options("sod", {"n": 300})
    blk = MeshBlock(op)
    blk.to("cuda")
    return blk, (lambda b: serving._initial_state("sod", b, {"n": 300}, 4))


def _run(blk, vars, n, t):
    for _ in range(n):
        dt = blk.max_time_step(vars)
        for s in range(3):
            blk.forward(vars, dt, s)
        blk.inc_cycle()
        t += dt
    torch.cuda.synchronize()
    return t


@pytest.mark.parametrize("case", ["convection-f64", "convection-f32", "sod"])
def test_resume_is_bit_exact(case, tmp_path):
    from snapy_amd import restart

    make = {"convection-f64": lambda: _convection("float64"), "convection-f32": lambda: _convection("float32"),
            "sod": _sod}[case]
    print(f"restart {case}", flush=True)
    a, init = make()
    va, _ = a.initialize({"hydro_w": init(a).cuda()})
    t = _run(a, va, 2, 0.0)
    path = str(tmp_path / restart.part_name("run", 0, 1))
    a.write_restart(va, path, t)
    t_a = _run(a, va, 2, t)
    b, _ = make()
    vb, t_b = b.initialize_from_restart(path)
    assert t_b == t and b.cycle() == 2
    t_b = _run(b, vb, 2, t_b)
    assert t_b == t_a and b.cycle() == a.cycle() == 4
    assert torch.equal(a.total_u(va), b.total_u(vb))
    assert torch.equal(va["hydro_w"], vb["hydro_w"])


def test_fp64_part_resumes_in_the_fp32_form(tmp_path):
    from snapy_amd import restart

    a, init = _convection("float64")
    va, _ = a.initialize({"hydro_w": init(a).cuda()})
    t = _run(a, va, 1, 0.0)
    path = str(tmp_path / restart.part_name("run", 0, None))
    a.write_restart(va, path, t)
    b, _ = _convection("float32")
    vb, tb = b.initialize_from_restart(path)
    # the pair holds the FP64 totals to ~2^-48
    assert torch.allclose(b.total_u(vb), a.total_u(va), rtol=1e-13, atol=0.0)
    _run(a, va, 2, t)
    _run(b, vb, 2, tb)
    ua, ub = a.total_u(va), b.total_u(vb)
    err = ((ua - ub).norm() / ua.norm()).item()
    print("fp32 resume rel err", err, flush=True)
    assert err < 1e-5


def test_mismatched_grid_is_refused(tmp_path):
    from snapy_amd import restart

    a, init = _convection("float64", "eager")
    va, _ = a.initialize({"hydro_w": init(a).cuda()})
    path = str(tmp_path / "a.part")
    a.write_restart(va, path, 0.0)
    from snapy_amd import MeshBlock, MeshBlockOptions

    op = MeshBlockOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", "convection3d.yaml"))
    op.coord.nx1, op.coord.nx2, op.coord.nx3 = 32, 24, 16
    op.coord.x1max, op.coord.x2max, op.coord.x3max = 3200.0, 2400.0, 1600.0  # same shape, other grid
    b = MeshBlock(op)
    b.to("cuda")
    with pytest.raises(restart.RestartError, match="grid mismatch"):
        b.initialize_from_restart(path)
