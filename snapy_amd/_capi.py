"""ctypes binding of the C-ABI in include/snapy_amd.h.

The library is loaded after ``import torch`` so it shares torch's HIP runtime
(both resolve the soname libamdhip64.so.7).  A missing library is an error:
there is no CPU fallback behind this module.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (load torch's HIP runtime first)

LIB_PATH = Path(__file__).resolve().parent / "_snapy_amd.so"

F64, F32 = 0, 1
RIEMANN = {"hllc": 0, "lmars": 1, "shallow-roe": 2}
BC = {
    "custom": 0,
    "none": 0,
    "reflecting": 1,
    "periodic": 2,
    "extrapolation": 3,
    "outflow": 4,
}


class Grid(ctypes.Structure):
    _fields_ = [
        ("nvar", ctypes.c_int32),
        ("nghost", ctypes.c_int32),
        ("nx1", ctypes.c_int32),
        ("nx2", ctypes.c_int32),
        ("nx3", ctypes.c_int32),
        ("nc1", ctypes.c_int32),
        ("nc2", ctypes.c_int32),
        ("nc3", ctypes.c_int32),
        ("dx1", ctypes.c_double),
        ("dx2", ctypes.c_double),
        ("dx3", ctypes.c_double),
        ("x1min", ctypes.c_double),
        ("x2min", ctypes.c_double),
        ("x3min", ctypes.c_double),
        ("members3", ctypes.c_int32),
        ("pad3", ctypes.c_int32),
    ]


class Physics(ctypes.Structure):
    _fields_ = [
        ("gammad", ctypes.c_double),
        ("grav1", ctypes.c_double),
        ("density_floor", ctypes.c_double),
        ("pressure_floor", ctypes.c_double),
        ("riemann", ctypes.c_int32),
        ("shock_x1", ctypes.c_int32),
        ("shock_x23", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 6),
        ("implicit_scheme", ctypes.c_int32),
        ("advection", ctypes.c_double),
        ("bot_heat_flux", ctypes.c_double),
        ("bot_heat_depth", ctypes.c_int32),
        ("flux_covariance", ctypes.c_int32),
        ("nu_iso", ctypes.c_double),
        ("kappa_iso", ctypes.c_double),
        ("gas_r", ctypes.c_double),
        ("eos", ctypes.c_int32),
    ]


class CsMesh(ctypes.Structure):
    """snapy_cs_mesh: the cubed-sphere descriptor and its device tables."""

    _fields_ = [
        ("n", ctypes.c_int32),
        ("nghost", ctypes.c_int32),
        ("members", ctypes.c_int32),
        ("interp_order", ctypes.c_int32),
        ("shock", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("radius", ctypes.c_double),
        ("geom", ctypes.c_void_p),
        ("ghost_idx", ctypes.c_void_p),
        ("ghost_w", ctypes.c_void_p),
        ("ghost_mat_c", ctypes.c_void_p),
        ("ghost_mat_p", ctypes.c_void_p),
        ("nghost_entries", ctypes.c_int64),
        ("seam_idx", ctypes.c_void_p),
        ("seam_d", ctypes.c_void_p),
    ]


class WlBlock(ctypes.Structure):
    """snapy_wl_block: one x1 block of a work list (sizes and device buffers)."""

    _fields_ = [
        ("nx1", ctypes.c_int32),
        ("bc_lo", ctypes.c_int32),
        ("bc_hi", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("dx1", ctypes.c_double),
        ("dx2", ctypes.c_double),
        ("dx3", ctypes.c_double),
        ("scale", ctypes.c_double),
        ("u", ctypes.c_void_p),
        ("u0", ctypes.c_void_p),
        ("w", ctypes.c_void_p),
        ("flux", ctypes.c_void_p),
        ("bref", ctypes.c_void_p),
        ("dt", ctypes.c_void_p),
    ]


class WorkList(ctypes.Structure):
    """snapy_work_list: the device descriptor array and prefix counts."""

    _fields_ = [
        ("blocks", ctypes.c_void_p),
        ("prefix", ctypes.c_void_p),
        ("nblocks", ctypes.c_int32),
        ("nghost", ctypes.c_int32),
        ("total", ctypes.c_int64 * 4),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_GP = ctypes.POINTER(Grid)
_WP = ctypes.POINTER(WorkList)
_PP = ctypes.POINTER(Physics)
_MP = ctypes.POINTER(CsMesh)

# name -> (restype, argtypes); every name here must be declared in
# include/snapy_amd.h (tests/test_capi_load.py checks both ways)
SIGNATURES = {
    "snapy_last_error": (ctypes.c_char_p, []),
    "snapy_kernel_count": (_I, []),
    "snapy_cons2prim": (_I, [_GP, _PP, _I, _P, _P, _P]),
    "snapy_prim2cons": (_I, [_GP, _PP, _I, _P, _P, _P]),
    "snapy_sweep": (_I, [_I, _GP, _PP, _I, _P, _P, _P, _P]),
    "snapy_sweep_cov": (_I, [_I, _GP, _PP, _I, _P, _P, _P, _P]),
    "snapy_update": (_I, [_GP, _I, _P * 3, _P, _P, _P, _D, _P, _D, _D, _D, _P]),
    "snapy_step_field": (_I, [_GP, _I, _P, _P, _D, _D, _D, _P, _P]),
    "snapy_cfl": (_I, [_GP, _PP, _I, _P, _D, _P, _P]),
    "snapy_apply_boundaries": (_I, [_GP, _PP, _I, _P, _P, _P, _P]),
    "snapy_apply_boundaries_prim": (_I, [_GP, _PP, _I, _P, _P, _P]),
    "snapy_hydro_ref_x1": (_I, [_GP, _PP, _P, _P, _P]),
    "snapy_sweep_x1_wb": (_I, [_GP, _PP, _I, _P, _P, _P, _P, _P]),
    "snapy_sweep_x1_wb_ref": (_I, [_GP, _PP, _I, _P, _P, _P, _P, _P]),
    "snapy_sweep_x1_wb_inc_f32": (_I, [_GP, _PP, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P]),
    "snapy_update_gravity": (_I, [_GP, _PP, _I, _P * 3, _P, _P, _P, _P, _P, _D, _P, _D, _D, _D, _P]),
    "snapy_vic_workspace": (ctypes.c_int64, [_GP]),
    "snapy_diffusion_workspace": (ctypes.c_int64, [_GP]),
    "snapy_cons2prim_f32": (_I, [_GP, _PP, _P, _P, _P, _P]),
    "snapy_prim2cons_f32": (_I, [_GP, _PP, _P, _P, _P, _P]),
    "snapy_update_f32": (_I, [_GP, _P * 3, _P, _P, _P, _P, _P, _D, _P, _D, _D, _D, _P]),
    "snapy_update_gravity_f32": (
        _I,
        [_GP, _PP, _P * 3, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _D, _D, _D, _P],
    ),
    "snapy_vic_full_f32": (
        _I,
        [_GP, _PP, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _D, _D, _D, _D, _P, _P, _P, _P],
    ),
    "snapy_vic_partial_workspace": (ctypes.c_int64, [_GP]),
    "snapy_vic_partial": (
        _I,
        [_GP, _PP, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _D, _D, _D, _D, _P, _P, _P, _P],
    ),
    "snapy_vic_partial_f32": (
        _I,
        [_GP, _PP, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _D, _D, _D, _D, _P, _P, _P, _P],
    ),
    "snapy_vic_partial_prim_f32": (
        _I,
        [_GP, _PP, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _D, _D, _D, _D, _P, _P, _P, _P],
    ),
    "snapy_diffusion": (_I, [_GP, _PP, _I, _P, _P * 3, _P, _P, _D, _P, _P]),
    "snapy_vic_full": (_I, [_GP, _PP, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _D, _D, _D, _D, _P, _P, _P, _P]),
    "snapy_vic_full_serial": (_I, [_GP, _PP, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _D, _D, _D, _D, _P, _P, _P, _P]),
    "snapy_scalar_flux": (_I, [_I, _GP, _PP, _I, _P, _P, _I, _P, _P, _P, _P]),
    "snapy_scalar_theta": (_I, [_GP, _PP, _I, _P, _P, _I, _P * 3, _P * 3, _D, _P, _P, _P]),
    "snapy_scalar_update": (_I, [_GP, _PP, _I, _P * 3, _P, _P, _P, _P, _P, _P, _D, _P, _D, _D, _D, _P]),
    "snapy_scalar_prim": (_I, [_GP, _PP, _I, _I, _P, _P, _P, _P, _P]),
    "snapy_scalar_boundaries": (_I, [_GP, _PP, _I, _P, _P]),
    "snapy_budget_workspace": (ctypes.c_int64, [_GP]),
    "snapy_budget": (_I, [_GP, _I, _P, _P, _P, _P]),
    "snapy_budget_noalloc": (_I, [_GP, _I, _P, _P, _P]),
    "snapy_record_begin": (_I, []),
    "snapy_record_end": (_I, [ctypes.POINTER(ctypes.c_void_p)]),
    "snapy_plan_launches": (_I, [_P]),
    "snapy_plan_free": (None, [_P]),
    "snapy_group_build": (_I, [ctypes.POINTER(ctypes.c_void_p), _I, ctypes.POINTER(ctypes.c_void_p)]),
    "snapy_group_launch": (_I, [_P, _P]),
    "snapy_group_launches": (_I, [_P]),
    "snapy_group_free": (None, [_P]),
    "snapy_cs_geom_size": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32]),
    "snapy_cs_edge_size": (ctypes.c_int64, [_MP]),
    "snapy_cs_cons2prim": (_I, [_MP, _P, _P, _P]),
    "snapy_cs_prim2cons": (_I, [_MP, _P, _P, _P]),
    "snapy_cs_ghosts": (_I, [_MP, _I, _P, _P]),
    "snapy_cs_sweep": (_I, [_MP, _I, _P, _P, _P, _P]),
    "snapy_cs_seams": (_I, [_MP, _P, _P, _P, _P]),
    "snapy_cs_update": (_I, [_MP, _P, _P, _P, _P, _P, _D, _P, _D, _D, _D, _P]),
    "snapy_cs_cfl": (_I, [_MP, _P, _D, _P, _P]),
    "snapy_cs_cons2prim_f32": (_I, [_MP, _P, _P, _P, _P]),
    "snapy_cs_prim2cons_f32": (_I, [_MP, _P, _P, _P, _P]),
    "snapy_cs_ghosts_f32": (_I, [_MP, _P, _P, _P]),
    "snapy_cs_sweep_f32": (_I, [_MP, _I, _P, _P, _P, _P]),
    "snapy_cs_seams_f32": (_I, [_MP, _P, _P, _P, _P]),
    "snapy_cs_update_f32": (_I, [_MP, _P, _P, _P, _P, _P, _P, _P, _D, _P, _D, _D, _D, _P]),
    "snapy_cs_cfl_f32": (_I, [_MP, _P, _D, _P, _P]),
    "snapy_tune_set": (_I, [_I, _GP, _I, _I]),
    "snapy_tune_get": (_I, [_I, _GP, _I]),
    "snapy_wl_counts": (_I, [ctypes.POINTER(WlBlock), ctypes.c_int32, ctypes.c_int32, _P, _P]),
    "snapy_wl_backup": (_I, [_WP, _PP, _P]),
    "snapy_wl_cons2prim": (_I, [_WP, _PP, _P]),
    "snapy_wl_sweep_x1": (_I, [_WP, _PP, _P]),
    "snapy_wl_update": (_I, [_WP, _PP, _D, _D, _D, _P]),
    "snapy_wl_boundaries": (_I, [_WP, _PP, _P]),
    "snapy_wl_cfl": (_I, [_WP, _PP, _P]),
    "snapy_wl_stage": (_I, [_WP, _PP, _I, _D, _D, _D, _P]),
}

_lib = None


class SnapyError(RuntimeError):
    pass


def load(path: Path | str | None = None):
    """Load the kernel library once; raise if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise SnapyError(
            f"{p} is not built: run `python -m snapy_amd.build` (or "
            "__graft_entry__.build()); the stage has no CPU fallback"
        )
    lib = ctypes.CDLL(str(p), mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().snapy_last_error().decode(errors="replace")
        raise SnapyError(f"{what} failed ({rc}): {msg}")


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise SnapyError("the stage kernels take HIP device tensors only")
    if not t.is_contiguous():
        raise SnapyError("the stage kernels take contiguous tensors only")
    return t.data_ptr()


def stream_of(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def dtype_code(dtype) -> int:
    if dtype == torch.float64:
        return F64
    if dtype == torch.float32:
        return F32
    raise SnapyError(f"unsupported dtype {dtype}")


def lib_path_in_use() -> str:
    return os.fspath(LIB_PATH)
