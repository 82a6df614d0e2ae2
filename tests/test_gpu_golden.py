"""The HIP path against the frozen golden fixtures (tests/golden/*.npz,
written from the CPU FP64 oracle by tests/golden/make_golden.py): the Sod
run, one x1 and one x2 sweep, one vertical implicit call per scheme and the
cubed-sphere ghost fill and seam fluxes.  Bars: 1e-12 relative L2 for the
FP64 stage (1e-11 for the implicit kernel alone, as its own test), bit-exact
for the ghost fill and the seam fluxes."""

import ctypes
import os

import numpy as np
import pytest
import torch

from oracle import fv
from snapy_amd import _capi
from snapy_amd.geometry import vic_table, x1_gravity_table
from tests._gpu_util import dev, grid_desc, host, phys_desc, rel_l2
from tests.golden import make_golden as mg

pytestmark = pytest.mark.gpu


def test_sod_against_fixture():
    from tests.test_gpu_sod import run_gpu

    print("golden: Sod 256 cells to t = 0.25", flush=True)
    z = mg.load("sod_c1")
    vars, n, dts = run_gpu()
    assert n == len(z["dts"])
    np.testing.assert_allclose(dts, z["dts"], rtol=1e-12)
    u = host(vars["hydro_u"])
    sub = (slice(None), 0, 0, slice(3, -3))
    for rows in ([0], [1, 2, 3], [4]):
        assert rel_l2(u[sub][rows], z["u"][sub][rows]) < 1e-12, rows


@pytest.mark.parametrize("riemann,shock", mg.SWEEP_CASES)
def test_sweeps_against_fixture(riemann, shock):
    lib = _capi.load()
    z = mg.load("sweeps")
    g = mg.sweep_grid()
    gd, pd = grid_desc(g), phys_desc(riemann=riemann, shock_x1=shock, shock_x23=shock)
    wd = dev(z["w"])
    k, j, i = g.interior()
    for axis in (1, 2):
        print(f"golden: sweep x{axis} {riemann}", flush=True)
        flx = torch.zeros_like(wd)
        fp = torch.zeros_like(wd[0])
        rc = lib.snapy_sweep(axis, ctypes.byref(gd), ctypes.byref(pd), 0, wd.data_ptr(), flx.data_ptr(),
                             fp.data_ptr() if axis == 1 else None, None)
        assert rc == 0, lib.snapy_last_error()
        if axis == 1:
            sl = (k, j, slice(i.start, i.stop + 1))
        else:
            sl = (k, slice(j.start, j.stop + 1), i)
        got = host(flx)
        ref = z[f"{riemann}_flux_x{axis}"]
        for v in range(5):
            assert rel_l2(got[v][sl], ref[v][sl]) < 1e-12, (axis, v)
        if axis == 1:
            assert rel_l2(host(fp)[sl], z[f"{riemann}_pface_x1"][sl]) < 1e-12


@pytest.mark.parametrize("scheme", [9, 1])
def test_vic_against_fixture(scheme):
    lib = _capi.load()
    z = mg.load("vic_columns")
    g = mg.vic_grid()
    pd = phys_desc(gammad=mg.GAMMA_DRY, riemann="lmars", shock_x1=False, shock_x23=False, bcs=("reflecting",) * 6,
                   grav1=mg.VIC_GRAV)
    pd.implicit_scheme = scheme
    gd = grid_desc(g)
    x1f, x1v = torch.from_numpy(g.faces(1)), torch.from_numpy(g.centers(1))
    a1, vol = g.dx(3) * g.dx(2), g.dx(3) * (g.dx(2) * g.dx(1))
    tab, t0, t1 = vic_table(x1f, x1v, 3, g.nx1, mg.VIC_GRAV, a1, vol)
    xt = x1_gravity_table(x1f, x1v, 3, g.nx1, mg.VIC_GRAV)
    wsize = lib.snapy_vic_partial_workspace if scheme == 1 else lib.snapy_vic_workspace
    entry = lib.snapy_vic_partial if scheme == 1 else lib.snapy_vic_full
    ws = torch.empty(int(wsize(ctypes.byref(gd))), dtype=torch.float64, device="cuda")
    st = torch.zeros(g.nx2 * g.nx3, dtype=torch.int32, device="cuda")
    u0 = torch.zeros(z["w"].shape, dtype=torch.float64, device="cuda")
    u = torch.zeros_like(u0)
    tabd, xtd, wd, dud = tab.cuda(), xt.cuda(), dev(z["w"]), dev(z["du"])
    print(f"golden: vic scheme {scheme}", flush=True)
    rc = entry(ctypes.byref(gd), ctypes.byref(pd), wd.data_ptr(), dud.data_ptr(), u0.data_ptr(), u.data_ptr(),
               tabd.data_ptr(), xtd.data_ptr(), t0, t1, mg.VIC_DT, None, 1.0, 0.0, 0.0, 1.0, ws.data_ptr(),
               st.data_ptr(), None, None)
    assert rc == 0, lib.snapy_last_error()
    assert int(host(st).max()) == 0
    got = host(u)
    k, j, _ = g.interior()
    n, ng = g.nx1, g.nghost
    ncol = (k.stop - k.start) * (j.stop - j.start)
    gc = np.moveaxis(got[:, k, j, ng : ng + n], 0, -1).reshape(ncol, n, 5)
    ref = z[f"delta_scheme{scheme}"]
    for v in range(5):
        assert rel_l2(gc[..., v], ref[..., v]) < 1e-11, v


@pytest.mark.parametrize("order", [2, 4])
def test_swe_ghosts_against_fixture(order):
    from tests.test_gpu_cubed import _mesh, _not_corner

    z = mg.load("swe_seams")
    print(f"golden: cubed-sphere ghosts order {order}", flush=True)
    m = _mesh(mg.SWE_N, order)
    q = torch.from_numpy(z[f"ghost_in_order{order}"]).cuda().contiguous()
    lib = _capi.load()
    _capi.check(lib.snapy_cs_ghosts(ctypes.byref(m._desc), 1, _capi.ptr(q), _capi.stream_of(q.device)), "cs_ghosts")
    out = q.cpu().numpy()
    ref = z[f"ghost_out_order{order}"]
    from oracle import cubed as oc

    mask = _not_corner(oc.Grid(mg.SWE_N, interp_order=order))
    for f in range(6):
        np.testing.assert_array_equal(out[f][:, mask], ref[f][:, mask])


def test_swe_seam_fluxes_against_fixture():
    from tests.test_gpu_cubed import _mesh

    z = mg.load("swe_seams")
    print("golden: cubed-sphere seam fluxes", flush=True)
    n = mg.SWE_N
    m = _mesh(n)
    m.initialize([{"hydro_w": torch.from_numpy(w)} for w in z["w"]])
    for name, fn in m._stage_calls(0)[:4]:
        _capi.check(fn(), name)
    got2, got3 = m._f2.cpu().numpy(), m._f3.cpu().numpy()
    ng = 3
    i, fa = slice(ng, ng + n), slice(ng, ng + n + 1)
    for f in range(6):
        np.testing.assert_array_equal(got2[f][:, i, fa], z["flux2"][f][:, i, fa])
        np.testing.assert_array_equal(got3[f][:, fa, i], z["flux3"][f][:, fa, i])


def test_tracer_block_against_fixture():
    """Two RK3 steps of the tracer block at the fixture's dts: the hydro
    state and both tracers to 1e-12."""
    from snapy_amd import MeshBlock
    from snapy_amd.options import (CoordinateOptions, HydroOptions, IntegratorOptions, MeshBlockOptions,
                                   ScalarOptions)

    z = mg.load("tracer_block")
    print("golden: tracer block", flush=True)
    a = mg.TRACER_GRID
    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", a["x1min"], a["x1max"], a["x2min"], a["x2max"], a["x3min"],
                                 a["x3max"], a["nx1"], a["nx2"], a["nx3"], 3)
    op.hydro = HydroOptions(riemann="lmars", grav1=mg.VIC_GRAV, flux_covariance=True, nu_iso=mg.TRACER_NU,
                            kappa_iso=mg.TRACER_NU)
    op.hydro.eos.gammad = mg.GAMMA_DRY
    op.hydro.eos.weight = fv.Deck.weight
    op.hydro.recon1.shock = False
    op.hydro.recon23.shock = False
    op.intg = IntegratorOptions(cfl=0.9, implicit_scheme=9)
    op.bcnames = ["reflecting"] * 6
    op.scalar = ScalarOptions(nvar=2, names=["slab", "ramp"])
    block = MeshBlock(op)
    block.to("cuda")
    vars, _ = block.initialize({"hydro_w": dev(z["w"]), "scalar_r": dev(z["r"])})
    for dt in z["dts"]:
        for s in range(3):
            block.forward(vars, float(dt), s)
    g = mg.tracer_grid()
    sub = (slice(None),) + g.interior()
    u, s_ = host(vars["hydro_u"]), host(vars["scalar_s"])
    for rows in ([0], [1, 2, 3], [4]):
        assert rel_l2(u[sub][rows], z["u"][sub][rows]) < 1e-12
    for n in range(2):
        assert rel_l2(s_[n][sub[1:]], z["s"][n][sub[1:]]) < 1e-12, n
