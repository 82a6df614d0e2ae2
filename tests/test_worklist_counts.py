"""The work list's host-side prefix counts (snapy_wl_counts), on the CPU:
the four item counts of each block (cells incl. ghosts, x1 faces, interior
cells, ghost cells) and their running sums."""

import ctypes
import threading

import pytest

from snapy_amd import _capi


@pytest.mark.skipif(not _capi.LIB_PATH.exists(), reason="library not built")
def test_prefix_counts():
    lib = _capi.load()
    sizes = [37, 128, 4099, 3]
    descs = (_capi.WlBlock * len(sizes))()
    for d, n in zip(descs, sizes):
        d.nx1 = n
    n = len(sizes)
    prefix = (ctypes.c_int32 * (4 * (n + 1)))()
    total = (ctypes.c_int64 * 4)()
    assert lib.snapy_wl_counts(descs, n, 3, prefix, total) == 0
    per = [[s + 6, s + 1, s, 6] for s in sizes]
    for ph in range(4):
        run = 0
        assert prefix[ph * (n + 1)] == 0
        for b in range(n):
            run += per[b][ph]
            assert prefix[ph * (n + 1) + b + 1] == run
        assert total[ph] == run
    # a block below nghost cells is refused (on a thread of its own: the
    # library's last-error message is per thread)
    descs[3].nx1 = 2
    out = []
    t = threading.Thread(target=lambda: out.append(lib.snapy_wl_counts(descs, n, 3, prefix, total)))
    t.start()
    t.join()
    assert out[0] < 0
