"""The C-ABI driven from a C++ host with no Python and no torch
(tests/cxx/sod_host.cpp, built by __graft_entry__.build()): its Sod run,
made of the library's entry points alone on buffers it allocates itself,
equals MeshBlock's run of the same deck bit for bit."""

import os
import subprocess

import numpy as np
import pytest
import torch

from snapy_amd import MeshBlock, MeshBlockOptions, problems

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cxx", "sod_host")


def test_cxx_host_equals_meshblock(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cxx/sod_host is not built: run __graft_entry__.build()")
    steps = 20
    out = tmp_path / "u.bin"
    print("C++ host: Sod 256, 20 steps", flush=True)
    r = subprocess.run([EXE, str(steps), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(str(out), dtype=np.float64).reshape(5, -1)

    op = MeshBlockOptions.from_yaml(os.path.join(ROOT, "snapy_amd", "decks", "sod.yaml"))
    op.intg.execution = "eager"
    block = MeshBlock(op)
    block.to("cuda")
    vars, _ = block.initialize({"hydro_w": problems.sod(block).cuda()})
    for _ in range(steps):
        dt = block.device_time_step(vars)
        for s in range(3):
            block.forward(vars, dt, s)
    torch.cuda.synchronize()
    ref = vars["hydro_u"].cpu().numpy().reshape(5, -1)
    assert got.shape == ref.shape
    assert np.isfinite(got).all()
    assert np.array_equal(got, ref)
