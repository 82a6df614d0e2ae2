"""The golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) against the CPU oracle as it stands: a change
of the oracle that moves any fixed number fails here, so the GPU tests that
compare against the fixtures keep meaning what they meant when the fixtures
were written."""

import numpy as np
import pytest

from tests.golden import make_golden as mg


@pytest.mark.parametrize("name", sorted(mg.CASES))
def test_oracle_reproduces_fixture(name):
    fixed = mg.load(name)
    now = mg.CASES[name]()
    assert sorted(fixed) == sorted(now)
    for key in fixed:
        assert fixed[key].shape == now[key].shape, key
        if name == "swe_seams" and key.startswith("ghost"):
            np.testing.assert_array_equal(now[key], fixed[key], err_msg=key)
        else:
            np.testing.assert_allclose(now[key], fixed[key], rtol=1e-14, atol=0.0, err_msg=key)


def test_fixtures_are_plain_arrays():
    for name in mg.CASES:
        with np.load(mg.path(name), allow_pickle=False) as z:
            for key in z.files:
                assert z[key].dtype in (np.float64, np.int64), (name, key)


def test_sod_fixture_is_near_the_exact_solution():
    """The frozen Sod state is the exact Riemann solution to within the
    scheme's error (Toro ch. 4)."""
    from oracle import fv

    z = mg.load("sod_c1")
    assert abs(z["dts"].sum() - 0.25) < 1e-14
    w = fv.cons2prim(z["u"], 1.4)[:, 0, 0, 3:-3]
    x = 0.5 * (np.arange(256) + 0.5) / 128
    ex = fv.sod_exact(x, 0.25)
    assert np.abs(w[0] - ex[0]).mean() < 5e-3
    assert np.abs(w[4] - ex[2]).mean() < 5e-3
