"""GPU parity on the benchmark configurations at (or near) their real shapes, through the C ABI.

- Zigzag graphene ribbon 512 cells x 64 two-site cells (65,536 orbitals), Anderson w = 1:
  complex128 against the oracle to 1e-10, the refined mixed mode to 1e-8.
- BHZ 100 x 32 with Anderson w = 40/72.9 at E = 0.01: the adjoint dT/d(eps) of all 12,800
  orbitals against the FP64 gradient from the oracle's Green's-function columns (1e-8 of the
  largest entry), and central differences (h = 1e-6) of the GPU transmission on 16 sites drawn
  with a fixed seed (1e-6 relative above the rounding floor of the difference: T is ~2 here and
  the entries are ~1e-5, so the floor, 32 ulps of T over 2h, is what bounds small entries).
- Cubic Anderson model W = 16.5, cross-sections L = 16 and 24, length 2L: one uneven batch of both
  sizes against the oracle.
"""
import numpy as np
import pytest

import caroli_amd as ca
from helpers import oracle_gradient, oracle_of, sigmas_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def oracle_points(systems, energies):
    """Oracle T and the self-energy handles for every (system, energy) pair, system-major."""
    pts_s, pts_g, ref = [], [], []
    for s in systems:
        o = oracle_of(s)
        for e in energies:
            g, arr = sigmas_of(s, e)
            t, rc = o.transmission_with(e, arr)
            assert rc == 0
            pts_s.append(s)
            pts_g.append(g)
            ref.append(t)
    return pts_s, pts_g, np.array(ref)


def test_graphene_512x64_double_and_mixed():
    s = ca.NestedTwoTerminal.preset(ca.Model.Graphene, [512, 64, 0], disorder=1, strength=1.0, seed=1)
    energies = [-0.45, -0.1, 0.02, 0.3]
    pts_s, pts_g, ref = oracle_points([s], energies)
    say(f"[gpu] graphene 512x64: oracle T = {np.round(ref, 6)}")
    t, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Double)
    assert (st == 0).all(), st
    assert np.abs(t - ref).max() <= 1e-10, np.abs(t - ref).max()
    stats = ca.capi.BatchStats()
    tm, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Mixed, stats)
    assert (st == 0).all(), st
    say(f"    mixed max |dT| {np.abs(tm - ref).max():.2e}, refinement iterations {stats.refinement_iters}, "
        f"recomputed in f64 {stats.recomputed_points}")
    assert np.abs(tm - ref).max() <= 1e-8
    assert stats.refinement_iters > 0 and stats.recomputed_points <= 1


def test_adjoint_bhz_100x32():
    ext = [100, 32, 0]
    w = 40 / 72.9
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=w, seed=1)
    e = 0.01
    sig = s.self_energies(e)
    say("[gpu] adjoint BHZ 100x32")
    t, grad = s.backward(sig)
    assert abs(t - s.transmission_with(sig)) < 1e-9
    ref = oracle_gradient(s, e)
    scale = np.abs(ref).max()
    say(f"    T = {t:.6f}, max |grad - FP64 gradient| {np.abs(grad - ref).max():.2e} (scale {scale:.2e})")
    assert np.abs(grad - ref).max() <= 1e-8 * scale
    # The adjoint is deterministic: one thread per orbital, no atomics.
    t2, grad2 = s.backward(sig)
    assert t2 == t and np.array_equal(grad2, grad)
    # Central differences (h = 1e-6) on 64 sites drawn from seed 2, the 128 shifted systems in one
    # batched launch with the same self-energies.
    nsites = ext[0] * ext[1]
    sites = np.random.default_rng(2).choice(nsites, 64, replace=False)
    h = 1e-6
    shifted = []
    for i in sites:
        pot = np.zeros(nsites)
        pot[i] = h
        for sign in (1.0, -1.0):
            shifted.append(ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=w, seed=1,
                                                       site_potential=sign * pot))
    tt, st = ca.batch_transmission(shifted, [sig] * len(shifted), ca.Precision.Double)
    assert (st == 0).all()
    worst = 0.0
    for k, i in enumerate(sites):
        fd = (tt[2 * k] - tt[2 * k + 1]) / (2 * h)
        ad = grad[4 * i:4 * i + 4].sum()
        # Relative 1e-6, above the rounding floor of a central difference of T ~ 2 in double
        # precision (32 ulps of T over 2h): the gradient itself is pinned to 1e-8 above.
        floor = 32 * np.finfo(float).eps * abs(t) / (2 * h)
        worst = max(worst, abs(fd - ad) / (1e-6 * abs(ad) + floor))
        assert abs(fd - ad) <= 1e-6 * abs(ad) + floor, (i, fd, ad, floor)
    say(f"    64 sites: worst |FD - adjoint| / bound {worst:.2f}")


def test_adjoint_batch_over_energies_and_configurations():
    """cg_backward_batch over 3 configurations x 3 energies of BHZ 40 x 16 in one sweep equals the
    point-by-point adjoint bit for bit and the oracle's FP64 gradient to 1e-8."""
    ext = [40, 16, 0]
    systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=0.2, seed=1 + c) for c in range(3)]
    energies = [-0.14, 0.01, 0.05]
    sig = [systems[0].self_energies(e) for e in energies]
    ps = [s for s in systems for _ in energies]
    pg = [g for _ in systems for g in sig]
    say("[gpu] batched adjoint BHZ 40x16, 9 points")
    t, grad, st = ca.backward_batch(ps, pg)
    assert (st == 0).all(), st
    for k in range(len(ps)):
        t1, g1 = ps[k].backward(pg[k])
        assert t1 == t[k] and np.array_equal(g1, grad[k]), k
    say(f"    T = {np.round(t, 4)}")
    for k in (0, 4, 8):
        assert t[k] > 0.1  # conducting points: the gradient is of order T
        ref = oracle_gradient(ps[k], energies[k % 3])
        assert np.abs(grad[k] - ref).max() <= 1e-8 * np.abs(ref).max()
    tt, _ = ca.batch_transmission(ps, pg, ca.Precision.Double)
    assert np.abs(t - tt).max() <= 1e-10


def test_cubic_uneven_batch():
    systems = [ca.NestedTwoTerminal.preset(ca.Model.Cubic, [2 * L, L, L], disorder=1, strength=16.5, seed=1000 + k)
               for k, L in enumerate([16, 24, 16])]
    energies = [-1.0, 1.0 / 3.0]
    pts_s, pts_g, ref = [], [], []
    for s in systems:
        ps, pg, r = oracle_points([s], energies)
        pts_s += ps
        pts_g += pg
        ref = np.concatenate([ref, r])
    say(f"[gpu] cubic uneven batch (L = 16, 24, 16): oracle T = {np.round(ref, 6)}")
    t, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Double)
    assert (st == 0).all(), st
    assert np.abs(t - ref).max() <= 1e-10, np.abs(t - ref).max()
    # Mixed: many channels (several column chunks per solve), separators of 256 and 576 variables.
    stats = ca.capi.BatchStats()
    tm, st = ca.batch_transmission(pts_s, pts_g, ca.Precision.Mixed, stats)
    assert (st == 0).all(), st
    say(f"    mixed max |dT| {np.abs(tm - ref).max():.2e}, refinement iterations {stats.refinement_iters}, "
        f"recomputed in f64 {stats.recomputed_points}")
    assert np.abs(tm - ref).max() <= 1e-8
    assert stats.refinement_iters > 0 and stats.recomputed_points <= 1


def test_uneven_resident_batch_with_kernel_timing():
    """Three cubic structures of different sizes in one resident batch: three launch units on three
    lanes, each lane's buffers sized for its own unit. With the kernel timers on every unit runs on
    lane 0 in turn, which must grow to the largest unit first; then the concurrent run."""
    systems = [ca.NestedTwoTerminal.preset(ca.Model.Cubic, [2 * L, L, L], disorder=1, strength=16.5, seed=1000 + k)
               for k, L in enumerate([8, 16, 12])]
    energies = [-1.0, 1.0 / 3.0]
    pts_s, pts_g, ref = [], [], []
    for s in systems:
        ps, pg, r = oracle_points([s], energies)
        pts_s += ps
        pts_g += pg
        ref = np.concatenate([ref, r])
    batch = ca.ResidentBatch(pts_s, pts_g, ca.Precision.Double)
    say("[gpu] cubic resident batch L = 8, 16, 12 with kernel timing (units one after the other on lane 0)")
    ca.kernel_timing(True)
    try:
        t, st = batch.run()
    finally:
        ca.kernel_timing(False)
    assert (st == 0).all(), st
    assert np.abs(t - ref).max() <= 1e-10, np.abs(t - ref).max()
    say("[gpu] the same batch, units concurrent")
    t2, st = batch.run()
    assert (st == 0).all(), st
    assert np.abs(t2 - ref).max() <= 1e-10, np.abs(t2 - ref).max()
