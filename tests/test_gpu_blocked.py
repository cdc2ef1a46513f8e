"""The blocked large-front path on whole systems, against the oracle, through the C ABI.

With CAROLI_FRONT_KERNEL=blocked every front above the LDS sizes, and every root, is factored by
the blocked launches (panel, row interchanges and U rows, trailing update per 64 columns) before
the MFMA Schur tiles, whatever its separator size. The small systems of the reference's nested
tests then exercise partial panels, roots with delayed variables (pivot test on the first ne1
columns only), both spin blocks, complex f32 fronts, the refined mixed mode and the adjoint (both
read the stored factors). Plans are built when a system's structure is first seen, so every test
here builds shapes that no other test uses, after setting the variable.
"""
import numpy as np
import pytest

import caroli_amd as ca
from helpers import compare, hashed, oracle_gradient, oracle_of, say

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def blocked(monkeypatch):
    monkeypatch.setenv("CAROLI_FRONT_KERNEL", "blocked")


def test_blocked_double():
    # E = 0: singular interior blocks of the clean lattice are delayed into their parents and the root.
    compare(ca.NestedTwoTerminal.preset(ca.Model.Square, [18, 9, 0]), [-1.3, 0.0, 0.4], "blocked square 18x9 clean")
    compare(hashed(ca.Model.BHZ, [13, 7, 0], 0.6), [-0.4, 0.0, 0.08], "blocked BHZ 13x7 disordered")
    compare(hashed(ca.Model.Cubic, [9, 7, 6], 2.5, dim=3), [-4.1, 0.0, 2.2], "blocked cubic 9x7x6")
    compare(hashed(ca.Model.Graphene, [11, 6, 0], 1.0, subs_per_cell=2), [-1.1, 0.25], "blocked graphene 11x6")


def test_blocked_bhz_magnetic_batch():
    systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [44, 28, 0], disorder=2, strength=0.4, concentration=0.2,
                                           seed=1 + c) for c in range(2)]
    energies = [-0.12, 0.03]
    sig = [systems[0].self_energies(e) for e in energies]
    pts_s, pts_g, ref = [], [], []
    for s in systems:
        o = oracle_of(s)
        for e, g in zip(energies, sig):
            t, rc = o.transmission_with(e, [g.block(b)[:2] for b in range(s.block_count())])
            assert rc == 0
            pts_s.append(s)
            pts_g.append(g)
            ref.append(t)
    say("[gpu] blocked BHZ 44x28 magnetic, 2 configs x 2 energies, double then mixed")
    t, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Double)
    assert (st == 0).all(), st
    assert np.abs(t - np.array(ref)).max() <= 1e-10
    tm, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Mixed)
    assert (st == 0).all(), st
    assert np.abs(tm - np.array(ref)).max() <= 1e-8


def test_blocked_single():
    compare(hashed(ca.Model.Square, [46, 38, 0], 2.0), [-1.7, 0.3], "blocked square 46x38 single", ca.Precision.Single,
            tol=1e-4)


def test_blocked_adjoint():
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [14, 8, 0], disorder=1, strength=40 / 72.9, seed=3)
    e = 0.01
    sig = s.self_energies(e)
    say("[gpu] blocked adjoint BHZ 14x8")
    t, grad = s.backward(sig)
    assert abs(t - s.transmission_with(sig)) < 1e-9
    ref = oracle_gradient(s, e)
    assert np.abs(grad - ref).max() <= 1e-8 * np.abs(ref).max()
