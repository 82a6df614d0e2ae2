"""Grouped launches on the GPU: density-current blocks of different widths
and heights (one deck) stepped by one launch per kernel give each block the
same bits as its own steps, in both state forms and both implicit schemes;
plans that do not match are refused."""

import ctypes

import pytest
import torch

from snapy_amd import _capi, problems, serving
from snapy_amd.group import BlockGroup, group_key
from snapy_amd.meshblock import MeshBlock

pytestmark = pytest.mark.gpu

SIZES = [(48, 16, -15.0), (96, 32, -9.0), (64, 24, -19.0)]


def _block(nx, nz, dT, precision, scheme):
    op = serving._options("density_current", {"nx": nx, "nz": nz, "dT": dT, "precision": precision,
                                             "execution": "eager"})
    op.intg.implicit_scheme = scheme
    blk = MeshBlock(op)
    blk.to("cuda")
    vars, _ = blk.initialize({"hydro_w": problems.density_current(blk, dT=dT, grav=serving.GRAV).cuda()})
    return blk, vars


def _state(blk, vars):
    out = [vars["hydro_u"].clone()]
    if blk._state.comp() is not None:
        out.append(blk._state.comp().clone())
    return out


@pytest.mark.parametrize("precision,scheme,capture", [("float64", 9, False), ("float32", 9, False),
                                                    ("float64", 1, False), ("float32", 1, False),
                                                    ("float32", 9, True), ("float64", 1, True)])
def test_grouped_steps_equal_each_blocks_own(precision, scheme, capture):
    print(f"group of {SIZES}, {precision}, scheme {scheme}, captured {capture}", flush=True)
    nsteps = 4
    solo = []
    for nx, nz, dT in SIZES:
        blk, vars = _block(nx, nz, dT, precision, scheme)
        for _ in range(nsteps):
            dt = blk.device_time_step(vars)
            for s in range(3):
                blk.forward(vars, dt, s)
        torch.cuda.synchronize()
        solo.append(_state(blk, vars))
        blk.release()
    members = [_block(nx, nz, dT, precision, scheme) for nx, nz, dT in SIZES]
    grp = BlockGroup(members, capture=capture)
    assert grp.launches(0) > 5
    for _ in range(nsteps):
        grp.step()
    torch.cuda.synchronize()
    for (blk, vars), ref in zip(members, solo):
        got = _state(blk, vars)
        assert torch.isfinite(got[0]).all()
        for a, b in zip(got, ref):
            assert torch.equal(a, b), (blk.options.coord.nx2, precision, scheme)
    grp.release()


def test_groups_refuse_blocks_that_launch_differently():
    a = _block(48, 16, -15.0, "float64", 9)
    b = _block(48, 16, -15.0, "float64", 1)  # the other implicit scheme: other kernels
    assert group_key(a[0]) != group_key(b[0])
    with pytest.raises(ValueError, match="share a deck"):
        BlockGroup([a, b])
    # the library refuses the plans themselves as well
    lib = _capi.load()
    plans = []
    for blk, vars in (a, b):
        assert lib.snapy_record_begin() == 0
        for name, fn in blk._stage_calls(vars, 0):
            if name != "vic_status_clear":
                _capi.check(fn(), name)
        p = ctypes.c_void_p()
        assert lib.snapy_record_end(ctypes.byref(p)) == 0
        assert lib.snapy_plan_launches(p) > 5
        plans.append(p)
    arr = (ctypes.c_void_p * 2)(*[p.value for p in plans])
    g = ctypes.c_void_p()
    rc = lib.snapy_group_build(arr, 2, ctypes.byref(g))
    for p in plans:
        lib.snapy_plan_free(p)
    assert rc == -21
    assert b"differ" in lib.snapy_last_error()


def _explicit_block(nx1, nx2, nx3, precision, seed):
    """A shock-capturing block without gravity (the plain x1 sweep and the
    divergence + combine kernels): HLLC, outflow x1, periodic x2/x3."""
    import numpy as np

    from snapy_amd.options import CoordinateOptions, HydroOptions, IntegratorOptions, MeshBlockOptions

    op = MeshBlockOptions()
    op.coord = CoordinateOptions("cartesian", 0.0, 1.0, 0.0, 0.5, 0.0, 0.5, nx1, nx2, nx3, 3)
    op.hydro = HydroOptions(riemann="hllc", precision=precision)
    op.hydro.eos.gammad = 1.4
    op.intg = IntegratorOptions(cfl=0.8, execution="eager")
    op.bcnames = ["outflow", "outflow", "periodic", "periodic", "periodic", "periodic"]
    blk = MeshBlock(op)
    blk.to("cuda")
    rng = np.random.default_rng(seed)
    w = np.zeros(blk.shape())
    w[0] = 1.0 + 0.2 * rng.random(w[0].shape)
    w[1] = 0.1 * rng.standard_normal(w[0].shape)
    w[4] = 1.0 + 0.2 * rng.random(w[0].shape)
    vars, _ = blk.initialize({"hydro_w": torch.from_numpy(w).cuda()})
    return blk, vars


def _convection_block(nx1, nx2, nx3, precision, seed):
    """A small 3D convection block (the convection deck's physics: gravity,
    the covariance sweeps, bot-heat, implicit scheme 1)."""
    import os

    from snapy_amd.options import MeshBlockOptions

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    op = MeshBlockOptions.from_yaml(os.path.join(root, "snapy_amd", "decks", "convection3d.yaml"))
    op.coord.nx1, op.coord.nx2, op.coord.nx3 = nx1, nx2, nx3
    op.coord.x1max = 50.0 * nx1
    op.coord.x2max = 100.0 * nx2
    op.coord.x3max = 100.0 * nx3
    op.hydro.precision = precision
    op.intg.execution = "eager"
    blk = MeshBlock(op)
    blk.to("cuda")
    vars, _ = blk.initialize({"hydro_w": problems.convection3d(blk, seed=seed, amp=2.0, layers=4).cuda()})
    return blk, vars


@pytest.mark.parametrize("kind,precision", [("explicit", "float64"), ("explicit", "float32"),
                                            ("convection", "float64"), ("convection", "float32")])
def test_grouped_3d_blocks_equal_each_blocks_own(kind, precision):
    shapes = [(16, 8, 8), (24, 6, 10), (16, 12, 4)]
    make = _explicit_block if kind == "explicit" else _convection_block
    print(f"group of {kind} blocks {shapes}, {precision}", flush=True)
    nsteps = 3
    solo = []
    for n, shp in enumerate(shapes):
        blk, vars = make(*shp, precision, 40 + n)
        for _ in range(nsteps):
            dt = blk.device_time_step(vars)
            for s in range(3):
                blk.forward(vars, dt, s)
        torch.cuda.synchronize()
        solo.append(_state(blk, vars))
        blk.release()
    members = [make(*shp, precision, 40 + n) for n, shp in enumerate(shapes)]
    grp = BlockGroup(members)
    for _ in range(nsteps):
        grp.step()
    torch.cuda.synchronize()
    for (blk, vars), ref in zip(members, solo):
        got = _state(blk, vars)
        assert torch.isfinite(got[0]).all()
        for a, b in zip(got, ref):
            assert torch.equal(a, b), (kind, precision, blk.options.coord.nx1)
    grp.release()


@pytest.mark.parametrize("precision,scheme", [("float64", 9), ("float32", 1)])
def test_grouped_tracer_blocks_equal_each_blocks_own(precision, scheme):
    """Blocks with two tracers (the covariance, the positivity limiter and
    the donor-capped implicit transfer) of different sizes, grouped."""
    from oracle import fv
    from tests.test_gpu_scalar import _pair

    grids = [fv.Grid(16, 12, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0),
             fv.Grid(24, 20, 1, 0.0, 6400.0, -6400.0, 6400.0, 0.0, 100.0),
             fv.Grid(16, 16, 1, 0.0, 4800.0, -4800.0, 4800.0, 0.0, 100.0)]
    print(f"group of tracer blocks, {precision}, scheme {scheme}", flush=True)
    nsteps = 3

    def tracer_state(blk, vars):
        return _state(blk, vars) + [vars["scalar_s"].clone(), vars["scalar_r"].clone()]

    solo = []
    for g in grids:
        _, blk, vars = _pair(g, ("reflecting",) * 6, scheme, True, precision=precision)
        for _ in range(nsteps):
            dt = blk.device_time_step(vars)
            for s in range(3):
                blk.forward(vars, dt, s)
        torch.cuda.synchronize()
        solo.append(tracer_state(blk, vars))
        blk.release()
    members = [_pair(g, ("reflecting",) * 6, scheme, True, precision=precision)[1:] for g in grids]
    grp = BlockGroup(members)
    for _ in range(nsteps):
        grp.step()
    torch.cuda.synchronize()
    for (blk, vars), ref in zip(members, solo):
        for a, b in zip(tracer_state(blk, vars), ref):
            assert torch.equal(a, b), (precision, scheme, blk.options.coord.nx2)
    grp.release()


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_grouped_cubed_spheres_equal_each_meshes_own(precision):
    """Cubed-sphere shallow-water meshes of different panel sizes (and
    member counts) grouped: each ends with its own bits."""
    import os

    from snapy_amd.cubed import CubedSphereMesh
    from snapy_amd.options import MeshOptions

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    specs = [(16, 1, 11), (24, 2, 12), (20, 1, 13)]

    def make(n, members, seed):
        mo = MeshOptions.from_yaml(os.path.join(root, "snapy_amd", "decks", "swe_cubed.yaml"))
        mo.block.coord.nx2 = mo.block.coord.nx3 = n
        mo.block.intg.execution = "eager"
        mo.block.hydro.precision = precision
        m = CubedSphereMesh(mo, members=members).to("cuda")
        vars = [{"hydro_w": w} for k in range(members) for w in problems.swe_hill(m, seed + k)]
        m.initialize(vars)
        return m, vars

    print(f"group of cubed spheres {specs}, {precision}", flush=True)
    nsteps = 3
    solo = []
    for n, members, seed in specs:
        m, vars = make(n, members, seed)
        for _ in range(nsteps):
            dt = m.device_time_step(vars)
            for s in range(3):
                m.forward(vars, dt, s)
        torch.cuda.synchronize()
        solo.append(m.total_u().clone())
        m.release()
    meshes = [make(n, members, seed) for n, members, seed in specs]
    grp = BlockGroup(meshes)
    for _ in range(nsteps):
        grp.step()
    torch.cuda.synchronize()
    for (m, _), ref in zip(meshes, solo):
        got = m.total_u()
        assert torch.isfinite(got).all()
        assert torch.equal(got, ref), (m.n, precision)
    grp.release()
