"""The serving stream's request generator (CPU): seeded Poisson arrivals,
the 40/40/20 mix, step counts and grid ranges of the serving config."""

import numpy as np

from snapy_amd import serving


def test_mix_and_arrivals():
    reqs = serving.sample_requests(2000, seed=11, rate=20.0)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    # Poisson at 20/s: mean gap 0.05 s
    assert abs(np.mean(np.diff(t)) - 0.05) < 0.005
    kinds = [r.kind for r in reqs]
    frac = {k: kinds.count(k) / len(kinds) for k in ("sod", "density_current", "swe")}
    assert abs(frac["sod"] - 0.4) < 0.04 and abs(frac["density_current"] - 0.4) < 0.04
    assert abs(frac["swe"] - 0.2) < 0.03
    for r in reqs:
        assert 50 <= r.steps <= 500
        if r.kind == "sod":
            assert 128 <= r.params["n"] <= 8192
        elif r.kind == "density_current":
            assert r.params["nx"] in (256, 512, 1024) and r.params["nz"] in (32, 64, 128)
            assert -20.0 <= r.params["dT"] <= -5.0
            assert r.params["precision"] in ("float32", "float64")
        else:
            assert r.params["n"] in (32, 64, 96, 128, 192)


def test_stream_is_reproducible():
    a = serving.sample_requests(50, seed=3)
    b = serving.sample_requests(50, seed=3)
    assert [(r.kind, r.params, r.seed, r.steps, r.arrival) for r in a] == [
        (r.kind, r.params, r.seed, r.steps, r.arrival) for r in b
    ]


def test_group_key_groups_one_deck_on_any_grid():
    from snapy_amd import serving
    from snapy_amd.group import group_key, groupable
    from snapy_amd.meshblock import MeshBlock

    def key(nx, nz, prec):
        op = serving._options("density_current", {"nx": nx, "nz": nz, "dT": -5.0, "precision": prec})
        blk = MeshBlock(op)
        assert groupable(blk)
        return group_key(blk)

    assert key(256, 32, "float32") == key(1024, 32, "float32") == key(512, 128, "float32")
    assert key(256, 32, "float32") != key(256, 32, "float64")
