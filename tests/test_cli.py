"""The command-line front end (caroli_amd/cli.py) on the CPU: input parsing, parameters and
expressions, scans, number formats of the reference's tables, the systems it builds (through
cg_disordered_create) against the presets of the benchmark workloads, and refusals. The
transmission run itself is a GPU test (tests/test_gpu_cli.py)."""
import math

import numpy as np
import pytest

import caroli_amd as ca
from caroli_amd import cli

C2_LIKE = """
title = "BHZ ribbon with magnetic impurities"
[parameters]
J = 0.4
half = "J/2"
[model]
preset = "bhz"
[[disorder]]
type = "magnetic"
concentration = 0.2
exchange = "J"
moments = "polarized"
mz = 1.0
spin = "opposite-g"
[system]
size = [20, 8]
[[run]]
type = "transmission"
energies = { from = -0.16, to = 0.16, count = 5 }
configurations = 3
"""


def test_parameters_expressions_and_ranges():
    p = cli.resolve({"a": 2.0, "b": "a^3 - 1", "c": "2*pi/3", "d": "sqrt(b) + -a", "e": "b/0"}, {})
    assert p["b"] == 7.0 and p["c"] == pytest.approx(2 * math.pi / 3)
    assert p["d"] == pytest.approx(math.sqrt(7) - 2) and p["e"] == math.inf
    assert cli.resolve({"a": 1.0, "b": "2*a"}, {"a": 5.0})["b"] == 10.0  # --set overrides first
    with pytest.raises(cli.InputError):
        cli.resolve({"a": "b", "b": "a"}, {})  # a cycle
    with pytest.raises(cli.InputError):
        cli.resolve({"a": 1.0}, {"z": 2.0})  # unknown override
    with pytest.raises(cli.InputError):
        cli.number("__import__('os')", {})
    assert cli.values({"from": 0, "to": 1, "count": 3}, {}) == [0.0, 0.5, 1.0]
    assert cli.values({"from": 0, "to": 1, "step": 0.25}, {}) == [0.0, 0.25, 0.5, 0.75, 1.0]
    assert cli.values([1, "2*3"], {}) == [1.0, 6.0] and cli.values(4, {}) == [4.0]
    pts = cli.scan_points({"J": 1.0, "w": 0.5, "h": "J*w"}, {}, {"w": [1, 2], "J": [3, 4]})
    assert [v for v, _ in pts] == [[3.0, 1.0], [3.0, 2.0], [4.0, 1.0], [4.0, 2.0]]  # scan keys sorted
    assert [p["h"] for _, p in pts] == [3.0, 6.0, 4.0, 8.0]


def test_number_formats_of_the_tables():
    # Rust's `{}` (shortest round trip, no exponent) and `{:e}` of an f64.
    assert [cli.fmt_display(x) for x in (0.16, -0.16, 0.0, 1e-5, 2.0, 12345.5)] == \
        ["0.16", "-0.16", "0", "0.00001", "2", "12345.5"]
    assert [cli.fmt_exp(x) for x in (2.0, 1.5e-3, 0.0, 123.0, -2.5e10)] == ["2e0", "1.5e-3", "0e0", "1.23e2", "-2.5e10"]
    assert cli.fmt_exp(float("nan")) == "NaN"


def test_cli_systems_equal_the_presets():
    """The magnetic run of the equal-size batch (seed 1 + 1000 k + c for [[disorder]] k,
    configuration c, crates/caroli-cli/src/build.rs:281) and the Anderson one of the zigzag
    graphene batch build the same Hamiltonians as the presets."""
    data = cli.parse(C2_LIKE)
    setup = cli.Setup(data, cli.resolve(data["parameters"], {}))
    for c in range(3):
        a = setup.system(c).arrays()
        b = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [20, 8, 0], disorder=2, strength=0.4, concentration=0.2,
                                        seed=1 + c).arrays()
        for k in ("row_ptr", "cols", "vals"):
            assert np.array_equal(a[k], b[k]), (c, k)
    g = cli.parse("""
[model]
preset = "graphene"
[[disorder]]
type = "anderson"
strength = 1.0
[system]
size = [30, 6]
[[run]]
type = "transmission"
energies = [0.1]
""")
    setup = cli.Setup(g, {})
    a = setup.system(2).arrays()
    b = ca.NestedTwoTerminal.preset(ca.Model.Graphene, [30, 6, 0], disorder=1, strength=1.0, seed=3).arrays()
    assert np.array_equal(a["vals"], b["vals"])


def test_cli_disorder_restrictions():
    """`cells` and `sublattices` restrict the disorder to the listed cells and sublattices."""
    base = """
[model]
preset = "graphene"
[[disorder]]
type = "anderson"
strength = 2.0
{extra}
[system]
size = [10, 4]
[[run]]
type = "transmission"
energies = [0.1]
"""
    clean = ca.NestedTwoTerminal.preset(ca.Model.Graphene, [10, 4, 0]).arrays()

    def onsite(arr):
        rp, cl, vv = arr["row_ptr"], arr["cols"], arr["vals"]
        return np.array([sum(vv[e] for e in range(rp[r], rp[r + 1]) if cl[e] == r) for r in range(arr["n"])])

    for extra, pick in (("cells = { x = [3, 6] }", lambda cell, sub: 3 <= cell[0] < 6),
                        ("sublattices = [1]", lambda cell, sub: sub == 1)):
        data = cli.parse(base.format(extra=extra))
        arr = cli.Setup(data, {}).system(0).arrays()
        diff = onsite(arr) - onsite(clean)
        subs = np.tile([0, 1], len(arr["cells"]) // 2)
        for site, (cell, sub) in enumerate(zip(arr["cells"], subs)):
            assert (diff[arr["site_offsets"][site]] != 0) == pick(cell, sub), (extra, site)


@pytest.mark.parametrize("text,message", [
    ("[model]\npreset = 'qwz'\n[system]\nsize=[4,4]\n[[run]]\ntype='transmission'\nenergies=[0]", "qwz"),
    ("[model]\npreset = 'square'\nt = 2.0\n[system]\nsize=[4,4]\n[[run]]\ntype='transmission'\nenergies=[0]", "t = 2"),
    ("[model]\npreset = 'square'\n[system]\nsize=[4,4]\nleads=['-x']\n[[run]]\ntype='transmission'\n"
     "energies=[0]", "at least two leads"),
    ("[model]\npreset = 'square'\n[system]\nsize=[4,4]\nleads=['-x','+z']\n[[run]]\ntype='transmission'\n"
     "energies=[0]", "dimension 2"),
    ("[model]\npreset = 'bhz'\n[[disorder]]\ntype='magnetic'\nconcentration=0.1\nexchange=1\n[system]\n"
     "size=[4,4]\n[[run]]\ntype='transmission'\nenergies=[0]", "spin"),
    ("[model]\npreset = 'square'\n[system]\nsize=[4,4]\n[[run]]\ntype='dos'\nenergies=[0]", "dos"),
])
def test_cli_refuses_what_it_does_not_cover(text, message, tmp_path):
    path = tmp_path / "in.toml"
    path.write_text(text)
    with pytest.raises(cli.InputError) as err:
        data = cli.parse(text)
        if data["run"][0]["type"] != "transmission":
            cli.run_all(data, str(tmp_path / "out"), {}, None, ca.Precision.Double)
        cli.Setup(data, cli.resolve(data.get("parameters", {}), {}))
    assert message in str(err.value)


def test_cli_check_describes_the_sample(tmp_path, capsys):
    path = tmp_path / "c2.toml"
    path.write_text(C2_LIKE)
    assert cli.main(["check", str(path)]) == 0
    out = capsys.readouterr().out
    assert "sample: [20, 8] cells, 160 sites, 640 orbitals" in out and "run 0: 00-transmission.csv" in out
    assert "'half': 0.2" in out
    assert cli.main(["check", str(path), "--set", "nope=1"]) == 1


CROSS_LIKE = """
[model]
preset = "bhz"
[[disorder]]
type = "anderson"
strength = 0.5
[system]
size = [10, 8]
leads = ["-x", "+x", "-y", "+y"]
[[run]]
type = "transmission"
energies = [0.02, 0.3]
configurations = 2
[[run]]
type = "transmission"
energies = [0.3]
sectors = [0, 0, 1, 1]
"""


def test_cli_multi_terminal_systems():
    """Leads on four faces (build.rs:371-413): the engine's multi-terminal system of each
    configuration has the face slices as interfaces and the Hamiltonian of the two-terminal builder
    with the same seeded disorder."""
    data = cli.parse(CROSS_LIKE)
    setup = cli.Setup(data, {})
    assert setup.leads == "faces" and setup.faces == [(-1, 0), (1, 0), (-1, 1), (1, 1)]
    two = cli.Setup(cli.parse(CROSS_LIKE.replace('leads = ["-x", "+x", "-y", "+y"]', "")), {})
    assert two.leads == "axis"
    for c in (0, 1):
        m, s2 = setup.multi(c), two.system(c)
        a, b = m.arrays(), s2.arrays()
        assert np.array_equal(a["vals"], b["vals"]) and np.array_equal(a["cols"], b["cols"])
        assert m.lead_count() == 4
        for p, (sign, axis) in enumerate(setup.faces):
            edge = a["cells"][:, axis].min() if sign < 0 else a["cells"][:, axis].max()
            on = np.flatnonzero(a["cells"][:, axis] == edge)
            assert list(m.lead(p)["orbitals"]) == sorted(int(a["site_offsets"][i]) + o for i in on for o in range(4))
