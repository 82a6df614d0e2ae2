"""snapy_budget_noalloc (the allocation-free budget, partial sums in the
library's static device array) against snapy_budget (the same reduction
with a caller workspace), through the C-ABI: bit for bit on two calls in a
row, the second on a smaller block (the array keeps the first call's
contents), in both dtypes, and a row count beyond the array refused before
any launch."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import fv
from snapy_amd import _capi
from tests._gpu_util import dev, grid_desc, host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_budget_noalloc_matches_budget_on_two_calls(dtype):
    lib = _capi.load()
    rng = np.random.default_rng(17)
    for shape in ((40, 24, 12), (9, 5, 1)):
        g = fv.Grid(*shape, x1max=3.0, x2max=2.0, x3max=1.0)
        print(f"budget_noalloc {shape} {dtype}", flush=True)
        gd = grid_desc(g)
        u = dev(rng.random((5,) + g.shape) + 0.5, dtype)
        ws = torch.empty(int(lib.snapy_budget_workspace(ctypes.byref(gd))), dtype=torch.float64, device="cuda")
        ref = torch.empty(5, dtype=torch.float64, device="cuda")
        got = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
        code = _capi.dtype_code(dtype)
        assert lib.snapy_budget(ctypes.byref(gd), code, u.data_ptr(), ws.data_ptr(), ref.data_ptr(), None) == 0
        rc = lib.snapy_budget_noalloc(ctypes.byref(gd), code, u.data_ptr(), got.data_ptr(), None)
        assert rc == 0, lib.snapy_last_error()
        a, b = host(got), host(ref)
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), (a, b)
        vol = g.dx(3) * (g.dx(2) * g.dx(1))
        un = host(u).astype(np.float64)
        np.testing.assert_allclose(a, (un[(slice(None),) + g.interior()] * vol).reshape(5, -1).sum(1), rtol=1e-12)


def test_budget_noalloc_refuses_rows_beyond_its_scratch():
    lib = _capi.load()
    g = fv.Grid(8, 4, 1)
    gd = grid_desc(g, nvar=9)
    u = torch.zeros((9,) + g.shape, dtype=torch.float64, device="cuda")
    out = torch.zeros(9, dtype=torch.float64, device="cuda")
    rc = lib.snapy_budget_noalloc(ctypes.byref(gd), 0, u.data_ptr(), out.data_ptr(), None)
    assert rc != 0
    assert b"exceeds" in lib.snapy_last_error()
