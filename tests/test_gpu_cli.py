"""The command-line front end's transmission run on the GPU (caroli_amd/cli.py; the reference's
crates/caroli-cli/src/run.rs:70-190): its CSV against the oracle on the same configurations, with
a parameter scan, several configurations and both precisions. The reference's own CLI test checks
T(0) = 2 for the clean BHZ 16 x 10 ribbon (crates/caroli-cli/tests/cli.rs:55-98); so does this."""
import csv
import os

import numpy as np
import pytest

import caroli_amd as ca
from caroli_amd import cli
from helpers import oracle_of, say

pytestmark = pytest.mark.gpu

INPUT = """
[parameters]
J = 0.4
[model]
preset = "bhz"
[[disorder]]
type = "magnetic"
concentration = 0.2
exchange = "J"
moments = "polarized"
spin = "opposite-g"
[system]
size = [24, 10]
[[run]]
type = "transmission"
output = "t.csv"
energies = { from = -0.1, to = 0.1, count = 4 }
configurations = 3
scan = { J = [0.2, 0.4] }
[[run]]
type = "transmission"
energies = [0.0]
"""

CLEAN = """
[model]
preset = "bhz"
[system]
size = [16, 10]
[[run]]
type = "transmission"
energies = [0.0]
"""


def read(path):
    with open(path) as f:
        return list(csv.DictReader(f))


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_cli_transmission_run_matches_the_oracle(tmp_path, precision):
    path = tmp_path / "in.toml"
    path.write_text(INPUT)
    out = tmp_path / "out"
    say(f"[gpu] caroli run ({precision}): BHZ 24x10 magnetic, scan J = 0.2, 0.4, 3 configurations x 4 energies")
    assert cli.main(["run", str(path), "--out", str(out), "--precision", precision]) == 0
    rows = read(out / "t.csv")
    assert list(rows[0].keys()) == ["J", "config", "energy", "T"] and len(rows) == 2 * 3 * 4
    assert os.path.exists(out / "01-transmission.csv")
    worst = 0.0
    for j in (0.2, 0.4):
        for c in range(3):
            s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [24, 10, 0], disorder=2, strength=j, concentration=0.2,
                                            seed=1 + c)
            o = oracle_of(s)
            for r in rows:
                if float(r["J"]) != j or int(r["config"]) != c:
                    continue
                e = float(r["energy"])
                g = s.self_energies(e)
                t_ref, rc = o.transmission_with(e, [g.block(b)[:2] for b in range(s.block_count())])
                assert rc == 0
                worst = max(worst, abs(float(r["T"]) - t_ref))
    say(f"    max |T_cli - T_oracle| {worst:.1e}")
    assert worst <= (1e-10 if precision == "double" else 1e-8)


def test_cli_clean_bhz_ribbon_plateau(tmp_path):
    """crates/caroli-cli/tests/cli.rs:55-98: T(0) = 2 for the clean BHZ 16 x 10 ribbon."""
    path = tmp_path / "clean.toml"
    path.write_text(CLEAN)
    assert cli.main(["run", str(path), "--out", str(tmp_path / "o")]) == 0
    rows = read(tmp_path / "o" / "00-transmission.csv")
    assert rows[0]["config"] == "0" and rows[0]["energy"] == "0"
    assert abs(float(rows[0]["T"]) - 2.0) < 1e-8


CROSS = """
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
output = "t.csv"
energies = [0.02, 0.3]
configurations = 2
[[run]]
type = "transmission"
output = "s.csv"
energies = [0.3]
sectors = [0, 0, 1, 1]
"""


def test_cli_multi_terminal_run(tmp_path):
    """Leads on four faces go to the multi-terminal solver (run.rs:130-185): T_p_q columns for p != q
    against the dense solve of the same systems; with `sectors` the T_p.a_q.b columns, whose spin
    sectors add up to T_p_q."""
    path = tmp_path / "in.toml"
    path.write_text(CROSS)
    out = tmp_path / "out"
    assert cli.main(["run", str(path), "--out", str(out)]) == 0
    rows = read(out / "t.csv")
    pairs = [(p, q) for p in range(4) for q in range(4) if p != q]
    assert list(rows[0].keys()) == ["config", "energy"] + [f"T_{p}_{q}" for p, q in pairs] and len(rows) == 4
    setup = cli.Setup(cli.parse(CROSS), {})
    from helpers import dense_multi
    worst = 0.0
    for r in rows:
        m = setup.multi(int(r["config"]))
        e = float(r["energy"])
        d = dense_multi(m, m.self_energies(e), e)
        for p, q in pairs:
            worst = max(worst, abs(float(r[f"T_{p}_{q}"]) - d[p, q]))
    say(f"[gpu] caroli run, four leads: max |T_cli - T_dense| {worst:.1e}")
    assert worst < 1e-9
    sec = read(out / "s.csv")[0]
    keys = [(p, a) for p in range(4) for a in (0, 1)]
    assert list(sec.keys())[2:] == [f"T_{p}.{a}_{q}.{b}" for (p, a) in keys for (q, b) in keys if (p, a) != (q, b)]
    t0 = [r for r in rows if r["config"] == "0" and r["energy"] == "0.3"][0]
    for p, q in pairs:
        total = sum(float(sec[f"T_{p}.{a}_{q}.{b}"]) for a in (0, 1) for b in (0, 1))
        assert abs(total - float(t0[f"T_{p}_{q}"])) < 1e-10
