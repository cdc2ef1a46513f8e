"""The autotuner and the kernel choices, through the C ABI (cg_batch_autotune, cg_set_tuning).

cg_batch_autotune times eager sweeps of a resident batch under candidate kernel choices and leaves
the fastest in force; whatever it picks, the batch's results stay those of the oracle to 1e-10,
also on the graph replays after it. A choice set by hand that classifies fronts differently
re-plans systems that were planned before, and their results stay the oracle's too.
"""
import numpy as np
import pytest

import caroli_amd as ca
from helpers import TOL, oracle_of, say

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_tuning():
    yield
    ca.set_tuning(None)


def oracle_ts(systems, sigmas):
    ref = []
    for s, g in zip(systems, sigmas):
        t, rc = oracle_of(s).transmission_with(g.energy(), [g.block(b)[:2] for b in range(s.block_count())])
        assert rc == 0
        ref.append(t)
    return np.array(ref)


def test_autotuned_batch_keeps_parity(restore_tuning):
    # BHZ 60 x 24 (separators up to 96 orbitals: the f11 / trsm, border and blocked choices all
    # apply) in two configurations, and a cubic 16 x 8 x 8 bar: an uneven batch.
    bhz = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [60, 24, 0], disorder=2, strength=0.4, concentration=0.2,
                                       seed=5 + c) for c in range(2)]
    cubic = ca.NestedTwoTerminal.preset(ca.Model.Cubic, [16, 8, 8], disorder=1, strength=16.5, seed=1001)
    pts_s, pts_g = [], []
    for e in (-0.1, 0.03, 0.11):
        g = bhz[0].self_energies(e)
        for s in bhz:
            pts_s.append(s)
            pts_g.append(g)
    for e in (-1.0, 1.0 / 3.0):
        pts_s.append(cubic)
        pts_g.append(cubic.self_energies(e))
    ref = oracle_ts(pts_s, pts_g)
    batch = ca.ResidentBatch(pts_s, pts_g, ca.Precision.Double)
    say("[gpu] autotune: BHZ 60x24 x 6 points + cubic 16x8x8 x 2 points, before tuning")
    t, st = batch.run()
    assert (st == 0).all(), st
    assert np.abs(t - ref).max() <= TOL
    say("[gpu] autotune: tuning")
    choice, sec = batch.autotune()
    say(f"    choice {choice}, {sec * 1e3:.2f} ms per sweep")
    assert sec > 0 and ca.get_tuning() == choice
    for run in range(3):  # eager, captured, replayed
        say(f"[gpu] autotune: tuned batch, run {run}")
        t, st = batch.run()
        assert (st == 0).all(), st
        assert np.abs(t - ref).max() <= TOL, (run, np.abs(t - ref).max())
    # A hand-set choice that classifies the separators differently: the systems planned above are
    # re-planned at their next use.
    forced = {"small_fronts": 1, "split": 0, "blocked_min_ne": 64, "unblocked": 0,
              "schur_tile": 32 if choice["schur_tile"] != 32 else 64}
    if forced["split"] == choice["split"] and forced["blocked_min_ne"] == choice["blocked_min_ne"]:
        forced["blocked_min_ne"] = 1 << 30
    ca.set_tuning(forced)
    say(f"[gpu] autotune: forced {forced}")
    t2, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Double)
    assert (st == 0).all(), st
    assert np.abs(t2 - ref).max() <= TOL, np.abs(t2 - ref).max()


def test_autotune_mixed_precision_batch(restore_tuning):
    """Tuning a refined mixed-precision batch: every candidate run repeats the refinement; the
    tuned batch still meets 1e-8 against the oracle on every point."""
    systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [40, 20, 0], disorder=2, strength=0.4, concentration=0.2,
                                           seed=11 + c) for c in range(2)]
    pts_s, pts_g = [], []
    for e in (-0.08, 0.05):
        g = systems[0].self_energies(e)
        for s in systems:
            pts_s.append(s)
            pts_g.append(g)
    ref = oracle_ts(pts_s, pts_g)
    batch = ca.ResidentBatch(pts_s, pts_g, ca.Precision.Mixed)
    say("[gpu] autotune: mixed BHZ 40x20 x 4 points, tuning")
    choice, sec = batch.autotune()
    say(f"    choice {choice}, {sec * 1e3:.2f} ms per refined run")
    stats = ca.capi.BatchStats()
    t, st = batch.run(stats)
    assert (st == 0).all(), st
    assert np.abs(t - ref).max() <= 1e-8, np.abs(t - ref).max()
    assert stats.refinement_iters > 0


@pytest.mark.parametrize("precision", [ca.Precision.Double, ca.Precision.Mixed])
def test_schur_tile_shapes_agree(precision, restore_tuning):
    """32 x 32 Schur tiles (one wave) against 64 x 64 (four waves) on whole systems: BHZ 60 x 24
    (Schur blocks of 40 to 300 rows, partial tiles of both widths, the bordered root) and graphene
    zigzag 64 x 16 cells; complex128 to the oracle's bar, mixed to 1e-8, and the two tilings
    within rounding of each other."""
    bhz = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [60, 24, 0], disorder=2, strength=0.4, concentration=0.2, seed=9)
    gra = ca.NestedTwoTerminal.preset(ca.Model.Graphene, [64, 16, 0], disorder=1, strength=1.0, seed=4)
    pts_s, pts_g = [], []
    for s, es in ((bhz, (-0.1, 0.07)), (gra, (0.2, 0.45))):
        for e in es:
            pts_s.append(s)
            pts_g.append(s.self_energies(e))
    ref = oracle_ts(pts_s, pts_g)
    out = {}
    for tile in (64, 32, 0):
        ca.set_tuning(schur_tile=tile)
        say(f"[gpu] schur tiles {tile}: BHZ 60x24 + graphene 64x16, precision {precision}")
        t, st = ca.batch_transmission(pts_s, pts_g, precision)
        assert (st == 0).all(), st
        bar = TOL if precision == ca.Precision.Double else 1e-8
        assert np.abs(t - ref).max() <= bar, (tile, np.abs(t - ref).max())
        out[tile] = t
    if precision == ca.Precision.Double:
        assert np.abs(out[32] - out[64]).max() <= 1e-12 * (1 + np.abs(ref).max())
        assert np.abs(out[0] - out[64]).max() <= 1e-12 * (1 + np.abs(ref).max())
