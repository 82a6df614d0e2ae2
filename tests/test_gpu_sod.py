"""C1: the whole Sod run (256 cells, t_end = 0.25) through the MeshBlock API
on the GPU against the CPU FP64 oracle: 1e-12 relative L2 per row, the same
step count, and closeness to the exact Riemann solution."""

import os

import numpy as np
import pytest
import torch

from oracle import fv
import snapy_amd
from snapy_amd import MeshBlock, MeshBlockOptions, kIDN, kIPR, kIVX

pytestmark = pytest.mark.gpu

DECK = os.path.join(os.path.dirname(snapy_amd.__file__), "decks", "sod.yaml")


def run_gpu(tlim=0.25):
    op = MeshBlockOptions.from_yaml(DECK)
    block = MeshBlock(op)
    block.to("cuda")
    x1v = block.module("coord").buffer("x1v").view(1, 1, -1)
    w = torch.zeros(block.shape(), dtype=torch.float64)
    w[kIDN] = torch.where(x1v < 0.5, torch.tensor(1.0, dtype=torch.float64), torch.tensor(0.125, dtype=torch.float64))
    w[kIPR] = torch.where(x1v < 0.5, torch.tensor(1.0, dtype=torch.float64), torch.tensor(0.1, dtype=torch.float64))
    vars, _ = block.initialize({"hydro_w": w.cuda()})
    t, n, dts = 0.0, 0, []
    while t < tlim:
        dt = min(block.max_time_step(vars), tlim - t)
        dts.append(dt)
        for s in range(3):
            block.forward(vars, dt, s)
        assert block.check_redo(vars) == 0
        t += dt
        n += 1
    torch.cuda.synchronize()
    return vars, n, dts


def run_oracle(tlim=0.25):
    d = fv.sod_deck(256)
    b = fv.Block(d)
    b.initialize(fv.sod_initial(d.grid))
    t, n, dts = 0.0, 0, []
    while t < tlim:
        dt = min(b.max_time_step(), tlim - t)
        dts.append(dt)
        for s in range(3):
            b.forward(dt, s)
        t += dt
        n += 1
    return b, n, dts


def test_sod_matches_oracle_and_exact():
    print("C1 Sod 256 cells FP64: GPU run", flush=True)
    vars, n, dts = run_gpu()
    b, n_ref, dts_ref = run_oracle()
    assert n == n_ref
    np.testing.assert_allclose(dts, dts_ref, rtol=1e-12)
    u = vars["hydro_u"].cpu().numpy()
    sub = (slice(None), 0, 0, slice(3, -3))
    for rows in ([0], [1, 2, 3], [4]):
        a, r = u[sub][rows], b.u[sub][rows]
        err = np.linalg.norm(a - r) / np.linalg.norm(r)
        print("rows", rows, "relL2", err, flush=True)
        assert err < 1e-12
    w = fv.cons2prim(u, 1.4)[sub]
    x = 0.5 * (np.arange(256) + 0.5) / 128
    ex = fv.sod_exact(x, 0.25)
    assert np.abs(w[kIDN] - ex[0]).mean() < 5e-3
    assert np.abs(w[kIVX] - ex[1]).mean() < 1e-2
    assert np.abs(w[kIPR] - ex[2]).mean() < 5e-3
