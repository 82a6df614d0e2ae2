"""The autotuner's host side on the CPU: the library's tuning table (set,
read back, default, refused values) and the JSON cache round trip."""

import ctypes
import threading

import pytest

from snapy_amd import _capi, autotune


def _grid(nx1, nx2, nx3):
    g = _capi.Grid()
    g.nvar, g.nghost = 5, 3
    g.nx1, g.nx2, g.nx3 = nx1, nx2, nx3
    g.nc1, g.nc2, g.nc3 = [n + 6 if n > 1 else 1 for n in (nx1, nx2, nx3)]
    return g


@pytest.mark.skipif(not _capi.LIB_PATH.exists(), reason="library not built")
def test_tuning_table():
    lib = _capi.load()
    g, h = _grid(128, 256, 256), _grid(64, 1, 1)
    gp, hp = ctypes.byref(g), ctypes.byref(h)
    assert lib.snapy_tune_get(autotune.TUNE_SWEEP, gp, 0) == 256
    assert lib.snapy_tune_set(autotune.TUNE_SWEEP, gp, 0, 128) == 0
    assert lib.snapy_tune_get(autotune.TUNE_SWEEP, gp, 0) == 128
    # per shape and per dtype
    assert lib.snapy_tune_get(autotune.TUNE_SWEEP, hp, 0) == 256
    assert lib.snapy_tune_get(autotune.TUNE_SWEEP, gp, 1) == 256
    assert lib.snapy_tune_set(autotune.TUNE_SWEEP, gp, 0, 0) == 0
    assert lib.snapy_tune_get(autotune.TUNE_SWEEP, gp, 0) == 256
    # the work list's grid is one value for every shape
    assert lib.snapy_tune_get(autotune.TUNE_WL_GRID, None, 0) == 2048
    assert lib.snapy_tune_set(autotune.TUNE_WL_GRID, None, 0, 512) == 0
    assert lib.snapy_tune_get(autotune.TUNE_WL_GRID, gp, 1) == 512
    assert lib.snapy_tune_set(autotune.TUNE_WL_GRID, None, 0, 0) == 0
    # values outside a tunable's range are refused (on a thread of its own:
    # the last-error message is per thread)
    out = []

    def bad():
        out.append(lib.snapy_tune_set(autotune.TUNE_EOS, gp, 0, 96))
        out.append(lib.snapy_tune_set(7, gp, 0, 64))
        out.append(lib.snapy_tune_set(autotune.TUNE_VIC_COLS, gp, 0, 3))
        out.append(lib.snapy_tune_set(autotune.TUNE_REF_COLS, gp, 0, 8))

    t = threading.Thread(target=bad)
    t.start()
    t.join()
    assert all(r < 0 for r in out) and len(out) == 4
    assert lib.snapy_tune_get(autotune.TUNE_EOS, gp, 0) == 256
    # the column kernels' columns per workgroup, with their own defaults
    assert lib.snapy_tune_get(autotune.TUNE_VIC_COLS, gp, 1) == 8
    assert lib.snapy_tune_get(autotune.TUNE_REF_COLS, gp, 1) == 4
    assert lib.snapy_tune_set(autotune.TUNE_VIC_COLS, gp, 1, 2) == 0
    assert lib.snapy_tune_get(autotune.TUNE_VIC_COLS, gp, 1) == 2
    assert lib.snapy_tune_set(autotune.TUNE_VIC_COLS, gp, 1, 0) == 0
    for fam, (t_, names, cands, dflt) in {**autotune.FAMILIES, **autotune.MANUAL}.items():
        assert dflt in cands
        assert lib.snapy_tune_get(t_, gp, 0) == dflt


def test_cache_round_trip(tmp_path):
    p = tmp_path / "tune.json"
    assert autotune.load(p) == {}
    autotune.save(p, "128x256x256:f32", {"eos": 128, "sweep": 256})
    autotune.save(p, "64x1x1:f64", {"update": 64})
    t = autotune.load(p)
    assert t["128x256x256:f32"] == {"eos": 128, "sweep": 256}
    assert t["64x1x1:f64"] == {"update": 64}
