"""The deck boundary on the CPU: this build's decks parse; the reference's
own example decks parse or are refused with a ValueError naming the key
that asks for what the stage does not carry (skipped where the reference
tree is absent); sections that would change the state are never dropped in
silence; the reference's module name and Python names are exposed."""

import os
import textwrap

import pytest

import snapy
import snapy_amd
from snapy_amd.options import MeshBlockOptions, MeshOptions

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_EXAMPLES = "/root/reference/examples"

# deck -> None (parses) or the key names one of which the refusal must name
REFERENCE_DECKS = {
    "shock.yaml": None,
    "straka.yaml": ("distribute/nb2",),
    "shallow_xy.yaml": ("distribute/nb2", "scalar"),
    "shallow_splash.yaml": ("scalar",),
    "planar_convection.yaml": ("reconstruct/vertical/scale", "forcing/diffusion"),
}


def _load(path):
    import yaml

    with open(path) as f:
        cfg = yaml.safe_load(f) or {}
    if (cfg.get("distribute") or {}).get("layout") == "cubed-sphere":
        return MeshOptions.from_yaml(path)
    return MeshBlockOptions.from_yaml(path)


@pytest.mark.parametrize("name", sorted(os.listdir(os.path.join(ROOT, "snapy_amd", "decks"))))
def test_own_decks_parse(name):
    _load(os.path.join(ROOT, "snapy_amd", "decks", name))


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="the reference tree is not here")
@pytest.mark.parametrize("name,keys", sorted(REFERENCE_DECKS.items()))
def test_reference_decks_parse_or_are_refused_by_name(name, keys):
    path = os.path.join(REF_EXAMPLES, name)
    if keys is None:
        op = _load(path)
        assert op.coord.nx1 == 256 and op.hydro.riemann == "hllc"
        return
    with pytest.raises(ValueError) as e:
        _load(path)
    assert any(k in str(e.value) for k in keys), str(e.value)


def _deck(tmp_path, extra):
    base = textwrap.dedent(
        """
        geometry:
          type: cartesian
          bounds: {x1min: 0., x1max: 1., x2min: 0., x2max: 1., x3min: 0., x3max: 1.}
          cells: {nx1: 16, nx2: 1, nx3: 1, nghost: 3}
        dynamics:
          equation-of-state: {type: ideal-gas, gammad: 1.4}
          riemann-solver: {type: hllc}
        """
    )
    p = tmp_path / "deck.yaml"
    p.write_text(base + textwrap.dedent(extra))
    return str(p)


@pytest.mark.parametrize(
    "extra,key",
    [
        ("scalar: {nvar: 1, upper-bound: 1.}\n", "upper-bound"),
        ("scalar: {nvar: 1, riemann-solver: {type: hllc}}\n", "upwind"),
        ("scalar: {nvar: 1, reconstruct: {type: cp5}}\n", "scalar/reconstruct/type"),
        ("scalar: {nvar: 1, colour: red}\n", "colour"),
        ("scalar: {nvar: 1}\nboundary-condition: {external: {x1-inner: outflow, x1-outer: outflow}}\n",
         "outflow"),
        ("species:\n  - {name: dry}\n  - {name: H2O}\n", "species"),
        ("reactions:\n  - {equation: H2O => H2O(l)}\n", "reactions"),
        ("distribute: {layout: slab, nb2: 2}\n", "distribute/nb2"),
        ("distribute: {layout: slab, nb1: 2}\n", "distribute/nb1"),
    ],
)
def test_state_changing_sections_are_refused(tmp_path, extra, key):
    with pytest.raises(ValueError, match=key):
        MeshBlockOptions.from_yaml(_deck(tmp_path, extra))


def test_scalar_section_is_read(tmp_path):
    op = MeshBlockOptions.from_yaml(_deck(tmp_path, "scalar: {nvar: 2, names: [tracer, dye]}\n"))
    assert op.scalar.nvar == 2 and op.scalar.names == ["tracer", "dye"]
    assert op.scalar.recon_type == "weno5" and op.scalar.riemann == "upwind"
    op = MeshBlockOptions.from_yaml(_deck(tmp_path, "scalar: {nvar: 1}\n"))
    assert op.scalar.names == ["scalar0"]


def test_inert_sections_are_accepted(tmp_path):
    op = MeshBlockOptions.from_yaml(
        _deck(
            tmp_path,
            """
            scalar: {nvar: 0}
            species:
              - {name: dry, composition: {N: 1.56}, cv_R: 2.5}
            reference-state: {Tref: 0., Pref: 1.e5}
            distribute: {layout: slab, nb1: 1, nb2: 1, nb3: 1}
            outputs:
              - {type: restart, dt: 1.}
            problem: {p0: 1.e5}
            """,
        )
    )
    assert op.coord.nx1 == 16


def test_temperature_floor_is_read(tmp_path):
    p = _deck(tmp_path, "")
    text = open(p).read().replace("{type: ideal-gas, gammad: 1.4}", "{type: ideal-gas, gammad: 1.4, temperature-floor: 5.}")
    open(p, "w").write(text)
    assert MeshBlockOptions.from_yaml(p).hydro.eos.temperature_floor == 5.0


def test_reference_module_name_and_names():
    assert snapy.MeshBlock is snapy_amd.MeshBlock
    assert snapy.MeshBlockOptions is snapy_amd.MeshBlockOptions
    assert snapy.StageExecutor is snapy_amd.StageExecutor
    assert (snapy.kIDN, snapy.kIV1, snapy.kIV2, snapy.kIV3, snapy.kIPR, snapy.kICY) == (0, 1, 2, 3, 4, 5)
    assert (snapy.kInnerX1, snapy.kOuterX3) == (0, 5)
    for name in ("cycle", "inc_cycle", "forward", "initialize", "max_time_step", "check_redo", "part", "total_u",
                 "execution_report"):
        assert callable(getattr(snapy.MeshBlock, name)), name
