"""GPU parity: the engine's complex f64 path against the CPU oracle, through the C ABI.

The bar is |dT| <= 1e-10 (plain complex128 against the FP64 restatement of nested.rs), on the
systems the reference's own nested tests use (crates/caroli-green/tests/nested.rs) and on reduced
versions of the benchmark configurations.
"""
import numpy as np
import pytest

import caroli_amd as ca
from helpers import TOL, compare, hashed, oracle_of, say

pytestmark = pytest.mark.gpu


def test_device_present():
    assert ca.device_available(), "no HIP device: the engine has no CPU path"


def test_clean_bhz_16x10_33_energies():
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [16, 10, 0])
    t, tref = compare(s, np.linspace(-0.16, 0.16, 33), "BHZ 16x10 clean")
    # In the bulk gap the clean ribbon carries its two helical edge channels (or none where the
    # finite-width edge states hybridise), to 1e-8 (crates/caroli-validation/tests/ti_conductance.rs:29-38).
    assert np.all(np.abs(t - np.round(t)) < 1e-8)


def test_square_clean_with_delays():
    # E = 0: singular interior blocks of the clean lattice are delayed (nested.rs:52-59).
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [14, 9, 0])
    compare(s, [-3.1, -1.3, 0.0, 0.4, 2.2], "square 14x9 clean")


def test_square_disordered():
    s = hashed(ca.Model.Square, [14, 9, 0], 2.0)
    compare(s, [-2.7, -0.5, 0.0, 1.1, 3.0], "square 14x9 disordered")


def test_bhz_disordered_two_blocks():
    s = hashed(ca.Model.BHZ, [12, 7, 0], 0.6)
    assert s.block_count() == 2
    compare(s, [-0.4, -0.05, 0.0, 0.08, 0.5], "BHZ 12x7 disordered")


def test_graphene_ribbons():
    s = ca.NestedTwoTerminal.preset(ca.Model.Graphene, [10, 6, 0])
    compare(s, [-2.3, -0.7, 0.3, 1.6], "graphene 10x6 clean")
    s = hashed(ca.Model.Graphene, [10, 6, 0], 1.0, subs_per_cell=2)
    compare(s, [-1.1, 0.25, 0.9], "graphene 10x6 disordered")


def test_long_strip_and_short_devices():
    s = hashed(ca.Model.Square, [80, 5, 0], 1.5)
    compare(s, [-1.9, 0.3, 2.6], "strip 80x5")
    for length in (1, 2, 3):
        s = hashed(ca.Model.Square, [length, 6, 0], 1.5)
        compare(s, [-1.2, 0.7], f"short device {length}x6")


def test_cubic_samples():
    for extent, axis in (([14, 6, 5], 0), ([4, 9, 8], 0), ([7, 10, 6], 2)):
        s = hashed(ca.Model.Cubic, extent, 2.5, dim=3, axis=axis)
        compare(s, [-4.1, -1.3, 0.0, 2.2], f"cubic {extent} axis {axis}")


def test_large_fronts_bhz_magnetic():
    # A reduced benchmark system: fronts above the LDS size take the global-memory kernel.
    systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [48, 32, 0], disorder=2, strength=0.4, concentration=0.2,
                                           seed=1 + c) for c in range(2)]
    energies = [-0.12, 0.0, 0.1]
    sig = [systems[0].self_energies(e) for e in energies]
    pts_s, pts_g, tref = [], [], []
    for s in systems:
        o = oracle_of(s)
        for e, g in zip(energies, sig):
            arr = [g.block(b)[:2] for b in range(s.block_count())]
            t, rc = o.transmission_with(e, arr)
            assert rc == 0
            pts_s.append(s)
            pts_g.append(g)
            tref.append(t)
    say("[gpu] BHZ 48x32 magnetic, 2 configs x 3 energies")
    t, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Double)
    assert (st == 0).all()
    assert np.abs(t - np.array(tref)).max() <= TOL


def test_reused_self_energies():
    clean_sig = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [12, 9, 0])
    other = hashed(ca.Model.BHZ, [12, 9, 0], 1.3)
    for e in (-0.2, 0.05, 0.4):
        sig = clean_sig.self_energies(e)
        assert sig.energy() == e
        assert abs(other.transmission_with(sig) - other.transmission(e)) < 1e-12


def test_single_precision_fronts():
    # precision.rs / tests/nested.rs:233-275: c32 fronts follow c64 closely on these systems.
    s = hashed(ca.Model.Square, [48, 40, 0], 2.0)
    compare(s, [-1.7, 0.3, 1.9], "square 48x40 single", ca.Precision.Single, tol=1e-4)
    s = hashed(ca.Model.BHZ, [24, 20, 0], 0.6)
    compare(s, [-0.3, 0.05, 0.35], "BHZ 24x20 single", ca.Precision.Single, tol=5e-4)
