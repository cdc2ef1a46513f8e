"""GPU mixed precision and adjoint, through the C ABI.

Mixed: complex f32 fronts with FP64-residual iterative refinement must reproduce the FP64 oracle
to |dT| <= 1e-8 on every point (the reference's unrefined f32 fronts miss by up to 0.46,
ARCHITECTURE.md:421-439).

Adjoint: dT/d(eps_a) for every orbital from one backward sweep, against an FP64 CPU gradient
built from the oracle's Green's-function columns (2 Re (G Gamma_R G^H Gamma_L G)_aa), and against
central finite differences of the GPU transmission (step 1e-6) on sampled sites, to 1e-6 relative.
"""
import numpy as np
import pytest

import caroli_amd as ca
from helpers import oracle_gradient, oracle_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def mixed_vs_oracle(systems, energies, label, tol=1e-8, max_recomputed=0.25):
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
    say(f"[gpu] mixed {label}: {len(pts_s)} points")
    stats = ca.capi.BatchStats()
    t, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Mixed, stats)
    assert (st == 0).all(), st
    d = np.abs(t - np.array(ref))
    say(f"    max |dT| {d.max():.2e}, refinement iterations {stats.refinement_iters}, "
        f"last correction {stats.max_refine_resid:.1e}, recomputed in f64 {stats.recomputed_points} "
        f"({stats.delayed_points} delayed)")
    assert d.max() <= tol, d
    # the f32 fronts plus refinement must carry the points, not the complex f64 recomputation
    assert stats.refinement_iters > 0
    assert stats.recomputed_points <= max_recomputed * len(pts_s), stats.recomputed_points


def test_mixed_bhz_magnetic():
    systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [40, 24, 0], disorder=2, strength=0.4, concentration=0.2,
                                           seed=1 + c) for c in range(2)]
    mixed_vs_oracle(systems, np.linspace(-0.16, 0.16, 5), "BHZ 40x24 magnetic")


def test_mixed_graphene_anderson():
    systems = [ca.NestedTwoTerminal.preset(ca.Model.Graphene, [64, 16, 0], disorder=1, strength=1.0, seed=1 + c)
               for c in range(2)]
    mixed_vs_oracle(systems, np.linspace(-0.5, 0.5, 4), "graphene 64x16 Anderson")


def test_mixed_square_with_delays():
    # The clean lattice at E = 0 has singular interior blocks: those points are recomputed in f64.
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [14, 9, 0])
    mixed_vs_oracle([s], [-1.3, 0.0, 0.4], "square 14x9 clean", max_recomputed=1.0)


def test_adjoint_matches_fp64_gradient_and_finite_differences():
    ext = [16, 8, 0]
    w = 40 / 72.9
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=w, seed=1)
    e = 0.01
    sig = s.self_energies(e)
    say("[gpu] adjoint BHZ 16x8")
    t, grad = s.backward(sig)
    assert abs(t - s.transmission_with(sig)) < 1e-9
    ref = oracle_gradient(s, e)
    scale = np.abs(ref).max()
    assert np.abs(grad - ref).max() <= 1e-8 * scale, (np.abs(grad - ref).max(), scale)
    # Central differences of the GPU transmission on sampled sites: V_i on the 4 orbitals of a site.
    rng = np.random.default_rng(2)
    nsites = ext[0] * ext[1]
    sites = rng.choice(nsites, 6, replace=False)
    h = 1e-6
    for i in sites:
        pot = np.zeros(nsites)
        pot[i] = h
        tp = ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=w, seed=1,
                                         site_potential=pot).transmission_with(sig)
        tm = ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=w, seed=1,
                                         site_potential=-pot).transmission_with(sig)
        fd = (tp - tm) / (2 * h)
        ad = grad[4 * i:4 * i + 4].sum()
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-3 * scale), (i, fd, ad)


@pytest.mark.parametrize("model,extent,energies,kw", [
    (ca.Model.BHZ, [60, 24, 0], [-0.16, -0.05, 0.02, 0.16], dict(disorder=2, strength=0.4, concentration=0.2)),
    (ca.Model.Graphene, [64, 16, 0], [-0.5, -0.1, 0.3], dict(disorder=1, strength=1.0)),
], ids=["BHZ 60x24 magnetic", "graphene 64x16 Anderson"])
def test_mixed_dual_corrected_refinement(monkeypatch, model, extent, energies, kw):
    """The dual-weighted acceptance (CAROLI_REFINE_GOAL=1: the transposed f32 solve, its FP64
    residual and the corrected amplitudes W_L^H x_L + phi^T (b - A x)) must give the same
    transmissions as the oracle to 1e-8, and accept points before their corrections fall below the
    relative tolerance."""
    monkeypatch.setenv("CAROLI_REFINE_GOAL", "1")
    systems = [ca.NestedTwoTerminal.preset(model, extent, seed=1 + c, **kw) for c in range(2)]
    sig = [systems[0].self_energies(e) for e in energies]
    ps = [s for s in systems for _ in energies]
    pg = [g for _ in systems for g in sig]
    ref = []
    for s in systems:
        o = oracle_of(s)
        for e, g in zip(energies, sig):
            t, rc = o.transmission_with(e, [g.block(b)[:2] for b in range(s.block_count())])
            assert rc == 0
            ref.append(t)
    stats = ca.capi.BatchStats()
    t, st = ca.batch_transmission(ps, pg, ca.Precision.Mixed, stats)
    d = np.abs(t - np.array(ref))
    say(f"[gpu] dual-corrected refinement: max |dT| {d.max():.2e}, iterations {stats.refinement_iters}, "
        f"accepted by the dual estimate {stats.goal_points}, recomputed {stats.recomputed_points}")
    assert (st == 0).all(), st
    assert d.max() <= 1e-8, d
    assert stats.goal_points > 0


def test_mixed_precision_is_bit_reproducible():
    """The forward solve of the refinement passes each front's boundary update to its parent in
    the order of the children (no atomics): repeated mixed runs give the same T bit for bit,
    through a fresh batch and a resident one (HIP graphs)."""
    systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [60, 24, 0], disorder=2, strength=0.4, concentration=0.2,
                                           seed=1 + c) for c in range(2)]
    energies = [-0.12, -0.03, 0.02, 0.1]
    sig = [systems[0].self_energies(e) for e in energies]
    ps = [s for s in systems for _ in energies]
    pg = [g for _ in systems for g in sig]
    stats = ca.capi.BatchStats()
    t1, st1 = ca.batch_transmission(ps, pg, ca.Precision.Mixed, stats)
    t2, st2 = ca.batch_transmission(ps, pg, ca.Precision.Mixed)
    batch = ca.ResidentBatch(ps, pg, ca.Precision.Mixed)
    runs = [batch.run()[0] for _ in range(3)]
    say(f"[gpu] mixed reproducibility: refinement iterations {stats.refinement_iters}, T = {np.round(t1, 6)}")
    assert stats.refinement_iters > 0
    assert (st1 == 0).all() and (st2 == 0).all()
    assert np.array_equal(t1, t2)
    for r in runs:
        assert np.array_equal(r, t1)
