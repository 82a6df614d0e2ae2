"""Kernel-level parity of the HIP stage against the CPU FP64 oracle, through
the C-ABI (include/snapy_amd.h).  FP64 bar: 1e-12 relative L2 (the tolerance
for the plain FP64 form); copies (boundaries) bit-exact."""

import ctypes
import itertools

import numpy as np
import pytest
import torch

from oracle import fv
from snapy_amd import _capi
from tests._gpu_util import dev, grid_desc, host, phys_desc, random_state, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _faces_slice(g: fv.Grid, axis):
    """interior faces along axis (lo..lo+nx), interior cells across it"""
    k, j, i = g.interior()
    sl = [slice(None), k, j, i]
    ax = {1: 3, 2: 2, 3: 1}[axis]
    s = sl[ax]
    sl[ax] = slice(s.start, s.stop + 1)
    return tuple(sl)


CASES = [
    # (nx1, nx2, nx3)
    (40, 1, 1),
    (24, 20, 1),
    (16, 12, 10),
]


@pytest.mark.parametrize("shape,riemann,shock", list(itertools.product(CASES, ["hllc", "lmars"], [True, False])))
def test_sweep_matches_oracle(shape, riemann, shock):
    lib = _capi.load()
    g = fv.Grid(*shape, x1min=0.0, x1max=2.0, x2min=-1.0, x2max=1.0, x3min=0.0, x3max=3.0)
    w = random_state(g, seed=sum(shape))
    gd, pd = grid_desc(g), phys_desc(riemann=riemann, shock_x1=shock, shock_x23=shock)
    wd = dev(w)
    for axis, tdim in ((1, 3), (2, 2), (3, 1)):
        if g.nc(axis) == 1:
            continue
        print(f"sweep x{axis} {shape} {riemann} shock={shock}", flush=True)
        flx = torch.zeros_like(wd)
        fp = torch.zeros_like(wd[0])
        rc = lib.snapy_sweep(axis, ctypes.byref(gd), ctypes.byref(pd), 0, wd.data_ptr(), flx.data_ptr(),
                             fp.data_ptr() if axis == 1 else None, None)
        assert rc == 0, lib.snapy_last_error()
        wlr = fv.reconstruct(w, tdim, shock)
        solver = fv.hllc if riemann == "hllc" else fv.lmars
        ref, pref = solver(wlr[fv.ILT], wlr[fv.IRT], 1.4, tdim)
        sl = _faces_slice(g, axis)
        got = host(flx)
        for v in range(5):
            assert rel_l2(got[v][sl[1:]], ref[v][sl[1:]]) < TOL, (axis, v)
        if axis == 1:
            assert rel_l2(host(fp)[sl[1:]], pref[sl[1:]]) < TOL


COV_CASES = [
    # (nx1, nx2, nx3, walls): several faces per lane along the sweep, runs
    # past the last face, one to eight x1 lines per workgroup, two tiles
    # along a line
    (24, 20, 1, False),
    (16, 12, 10, False),
    (13, 7, 9, True),
    (8, 50, 3, True),
    (126, 70, 3, False),
]


@pytest.mark.parametrize("case,riemann,shock,dtype",
                         list(itertools.product(COV_CASES, ["hllc", "lmars"], [False, True], [0, 1])))
def test_cov_sweep_matches_oracle(case, riemann, shock, dtype):
    """The x2/x3 sweep with the x1 flux covariance (snapy_sweep_cov) against
    the oracle's face fluxes (reconstruction, Riemann solver, covariance,
    wall prescription): FP64 within 1e-12, FP32 (one face per lane in FP64,
    runs of faces per lane in FP32) within 1e-5 relative L2 per row.  The
    FP32 LMARS fluxes carry the normal momentum row about the line's gauge,
    the pressure of its first interior cell along the sweep, which the test
    adds back."""
    nx1, nx2, nx3, walls = case
    lib = _capi.load()
    g = fv.Grid(nx1, nx2, nx3, x1min=0.0, x1max=2.0, x2min=-1.0, x2max=1.0, x3min=0.0, x3max=3.0)
    w = random_state(g, seed=nx1 + 3 * nx2 + 7 * nx3)
    bcs = ("reflecting",) * 6 if walls else ("outflow",) * 6
    gd = grid_desc(g)
    pd = phys_desc(riemann=riemann, shock_x1=shock, shock_x23=shock, bcs=bcs)
    pd.flux_covariance = 1
    wd = dev(w)
    tol = TOL if dtype == 0 else 1e-5
    for axis, tdim in ((2, 2), (3, 1)):
        if g.nc(axis) == 1:
            continue
        print(f"cov sweep x{axis} {case} {riemann} shock={shock} dtype={dtype}", flush=True)
        flx = torch.zeros(wd.shape, dtype=torch.float64 if dtype == 0 else torch.float32, device="cuda")
        rc = lib.snapy_sweep_cov(axis, ctypes.byref(gd), ctypes.byref(pd), dtype, wd.data_ptr(),
                                 dev(g.centers(1)).data_ptr(), flx.data_ptr(), None)
        assert rc == 0, lib.snapy_last_error()
        wlr = fv.reconstruct(w, tdim, shock)
        solver = fv.hllc if riemann == "hllc" else fv.lmars
        ref, _ = solver(wlr[fv.ILT], wlr[fv.IRT], 1.4, tdim)
        ref[fv.IPR][..., 1 : g.nc(1) - 1] += fv.flux_covariance(wlr[fv.ILT], wlr[fv.IRT], tdim, g, 1.4)
        fv.prescribe_walls(ref, tdim, bcs, axis, g)
        got = host(flx).astype(np.float64)
        if dtype == 1 and riemann == "lmars":
            # the gauge of each line: its first interior cell along the sweep
            lo = g.nghost
            gauge = w[fv.IPR][:, lo : lo + 1, :] if axis == 2 else w[fv.IPR][lo : lo + 1, :, :]
            got[fv.IVX + axis - 1] += gauge
        sl = _faces_slice(g, axis)
        for v in range(5):
            err = rel_l2(got[v][sl[1:]], ref[v][sl[1:]])
            assert err < tol, (axis, v, err)


@pytest.mark.parametrize("shape", CASES)
def test_update_matches_oracle(shape):
    lib = _capi.load()
    g = fv.Grid(*shape, x1min=0.0, x1max=2.0, x2min=-1.0, x2max=1.0, x3min=0.0, x3max=3.0)
    rng = np.random.default_rng(3)
    shp = (5,) + g.shape
    u0 = rng.random(shp) + 1.0
    u = rng.random(shp) + 1.0
    flux = {a: rng.random(shp) - 0.5 for a in (1, 2, 3) if g.nc(a) > 1}
    dt = 0.0123
    for stage, (w0, w1, w2) in enumerate(fv.RK3_WEIGHTS):
        print(f"update {shape} stage {stage}", flush=True)
        div = fv.divergence(g, flux)
        du = np.zeros_like(div)
        sub = (slice(None),) + g.interior()
        du[sub] = -dt * div[sub]
        ref = fv.rk_combine(u0, u, du, stage)
        ud = dev(u)
        keep = [dev(flux[a]) for a in flux]
        fl = (ctypes.c_void_p * 3)(*[keep[list(flux).index(a)].data_ptr() if a in flux else None for a in (1, 2, 3)])
        u0d = dev(u0)
        rc = lib.snapy_update(ctypes.byref(grid_desc(g)), 0, fl, u0d.data_ptr(), ud.data_ptr(), None, dt, None, w0, w1, w2, None)
        assert rc == 0, lib.snapy_last_error()
        got = host(ud)
        assert rel_l2(got[sub], ref[sub]) < TOL
        # ghosts untouched
        mask = np.ones(shp, bool)
        mask[sub] = False
        assert np.array_equal(got[mask], u[mask])


BC_KINDS = ["reflecting", "periodic", "extrapolation", "outflow"]


@pytest.mark.parametrize("shape,kinds", list(itertools.product(CASES, [
    ("reflecting",) * 6,
    ("periodic",) * 6,
    ("extrapolation",) * 6,
    ("outflow",) * 6,
    ("outflow", "reflecting", "periodic", "periodic", "extrapolation", "outflow"),
    ("reflecting", "outflow", "outflow", "extrapolation", "periodic", "periodic"),
])))
def test_boundaries_bit_exact(shape, kinds):
    lib = _capi.load()
    g = fv.Grid(*shape)
    w = random_state(g, seed=7, amp=0.4)
    bref = random_state(g, seed=8, amp=0.4)
    u = fv.prim2cons(w, 1.4)
    ref = u.copy()
    fv.apply_boundaries(ref, kinds, 1.4, bref)
    ud = dev(u)
    bd = dev(bref)
    scratch = torch.empty_like(ud)
    rc = lib.snapy_apply_boundaries(ctypes.byref(grid_desc(g)), ctypes.byref(phys_desc(bcs=kinds)), 0,
                                    ud.data_ptr(), bd.data_ptr(), scratch.data_ptr(), None)
    assert rc == 0, lib.snapy_last_error()
    got = host(ud)
    if "outflow" in kinds:
        # the radiating formula is arithmetic: FP64 round-off; the copies exact
        assert rel_l2(got, ref) < TOL
    else:
        assert np.array_equal(got, ref)
    # primitive form
    wp = w.copy()
    fv.apply_boundaries(wp, kinds, 1.4, bref, primitive=True)
    wd = dev(w)
    rc = lib.snapy_apply_boundaries_prim(ctypes.byref(grid_desc(g)), ctypes.byref(phys_desc(bcs=kinds)), 0,
                                         wd.data_ptr(), bd.data_ptr(), None)
    assert rc == 0, lib.snapy_last_error()
    if "outflow" in kinds:
        assert rel_l2(host(wd), wp) < TOL
    else:
        assert np.array_equal(host(wd), wp)


@pytest.mark.parametrize("shape", CASES)
def test_cfl_and_eos(shape):
    lib = _capi.load()
    g = fv.Grid(*shape, x1min=0.0, x1max=2.0, x2min=-1.0, x2max=1.0, x3min=0.0, x3max=3.0)
    w = random_state(g, seed=11, amp=0.5)
    u = fv.prim2cons(w, 1.4)
    ud, wd = dev(u), torch.zeros(w.shape, dtype=torch.float64, device="cuda")
    gd, pd = grid_desc(g), phys_desc()
    assert lib.snapy_cons2prim(ctypes.byref(gd), ctypes.byref(pd), 0, ud.data_ptr(), wd.data_ptr(), None) == 0
    assert rel_l2(host(wd), fv.cons2prim(u, 1.4)) < TOL
    u2 = torch.zeros_like(ud)
    wd2 = dev(w)  # referenced until the results are read back
    assert lib.snapy_prim2cons(ctypes.byref(gd), ctypes.byref(pd), 0, wd2.data_ptr(), u2.data_ptr(), None) == 0
    assert rel_l2(host(u2), u) < TOL
    out = torch.zeros((), dtype=torch.float64, device="cuda")
    assert lib.snapy_cfl(ctypes.byref(gd), ctypes.byref(pd), 0, wd2.data_ptr(), 1.0, out.data_ptr(), None) == 0
    # the min of the same per-cell quotients: exact
    assert float(host(out)) == fv.max_time_step(g, w, 1.4)


def test_budget_fp64():
    lib = _capi.load()
    g = fv.Grid(30, 20, 10, x1max=3.0, x2max=2.0, x3max=1.0)
    rng = np.random.default_rng(5)
    u = rng.random((5,) + g.shape)
    gd = grid_desc(g)
    n = lib.snapy_budget_workspace(ctypes.byref(gd))
    ws = torch.empty(int(n), dtype=torch.float64, device="cuda")
    out = torch.empty(5, dtype=torch.float64, device="cuda")
    ud = dev(u)  # referenced until the results are read back
    assert lib.snapy_budget(ctypes.byref(gd), 0, ud.data_ptr(), ws.data_ptr(), out.data_ptr(), None) == 0
    vol = g.dx(3) * (g.dx(2) * g.dx(1))
    ref = (u[(slice(None),) + g.interior()] * vol).reshape(5, -1).sum(1)
    np.testing.assert_allclose(host(out), ref, rtol=1e-13)
    # deterministic: the same bits twice
    out2 = torch.empty_like(out)
    assert lib.snapy_budget(ctypes.byref(gd), 0, ud.data_ptr(), ws.data_ptr(), out2.data_ptr(), None) == 0
    assert np.array_equal(host(out), host(out2))


def test_invalid_arguments_are_rejected_on_host():
    lib = _capi.load()
    g = fv.Grid(16, 1, 1)
    gd = grid_desc(g)
    bad = grid_desc(g)
    bad.nc1 = 17
    pd = phys_desc()
    w = torch.zeros((5,) + g.shape, dtype=torch.float64, device="cuda")
    assert lib.snapy_sweep(1, ctypes.byref(bad), ctypes.byref(pd), 0, w.data_ptr(), w.data_ptr(), None, None) < 0
    assert lib.snapy_sweep(2, ctypes.byref(gd), ctypes.byref(pd), 0, w.data_ptr(), w.data_ptr(), None, None) < 0
    assert lib.snapy_sweep(1, ctypes.byref(gd), ctypes.byref(pd), 7, w.data_ptr(), w.data_ptr(), None, None) < 0
    assert b"dtype" in lib.snapy_last_error()
