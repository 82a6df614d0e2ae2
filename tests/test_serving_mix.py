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


def test_member_key_groups_density_currents_of_one_grid_and_form():
    from snapy_amd import serving

    a = serving.Request(0, "density_current", {"nx": 256, "nz": 32, "dT": -5.0, "precision": "float32"}, 1, 10)
    b = serving.Request(1, "density_current", {"nx": 256, "nz": 32, "dT": -9.0, "precision": "float32"}, 2, 20)
    c = serving.Request(2, "density_current", {"nx": 256, "nz": 32, "dT": -9.0, "precision": "float64"}, 3, 20)
    d = serving.Request(3, "sod", {"n": 256}, 4, 20)
    assert serving.member_key(a) == serving.member_key(b)
    assert serving.member_key(a) != serving.member_key(c)
    assert serving.member_key(d) is None
