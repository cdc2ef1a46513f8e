"""Multi-terminal devices on the GPU (NestedMultiTerminal, crates/caroli-green/src/nested_multi.rs):
every T[p][q] of the nested GPU solver against the dense direct solve with the same self-energies
(MultiTerminal::solve, multi.rs:89-110, as crates/caroli-green/tests/nested_multi.rs:64-85 does), on
a disordered square cross with four leads, a BHZ cross (two spin blocks), graphene with a narrow
lead (slanted bonds) and a clean square cross at E = 0; two leads at the ends of a ribbon against
the two-terminal GPU solver; spin-sector transmissions; complex f32 fronts; one batched launch
against one point at a time."""
import numpy as np
import pytest

import caroli_amd as ca
from helpers import cross, dense_multi, hash_disorder, multi_system, say
from test_multi_host import graphene_narrow_lead

pytestmark = pytest.mark.gpu

TOL = 1e-9  # nested_multi.rs tests: |dT| < 1e-9 (1 + |T|) against the dense solve


def check(s, energies, label, precision=ca.Precision.Double, tol=TOL):
    sig = [s.self_energies(e) for e in energies]
    say(f"[gpu] multi {label}: {s.lead_count()} leads, {s.block_count()} blocks, {len(energies)} energies")
    t, st = ca.multi_batch_transmissions([s] * len(energies), sig, precision)
    assert (st == 0).all(), st
    worst = 0.0
    for k, e in enumerate(energies):
        d = dense_multi(s, sig[k], e)
        err = np.abs(t[k] - d) / (1 + np.abs(d))
        worst = max(worst, err.max())
        assert err.max() < tol, f"{label} E = {e}:\n{t[k]}\n{d}"
    say(f"    max |dT| / (1 + |T|) {worst:.1e}")
    return t


def test_square_cross():
    sites, itf = cross(ca.Model.Square, 9, 7, 3)
    s = multi_system(ca.Model.Square, sites, itf, w=1.5)
    t = check(s, [-2.7, -0.4, 0.0, 1.3], "square cross 9x7, arms 3")
    assert np.abs(t - t.transpose(0, 2, 1)).max() < 1e-9  # time-reversal symmetric: T_pq = T_qp


def test_bhz_cross():
    sites, itf = cross(ca.Model.BHZ, 8, 6, 2)
    s = multi_system(ca.Model.BHZ, sites, itf, w=0.5)
    assert s.block_count() == 2
    check(s, [-0.2, 0.02, 0.3], "BHZ cross 8x6, arms 2")


def test_graphene_narrow_lead():
    sites, itf = graphene_narrow_lead()
    s = multi_system(ca.Model.Graphene, sites, itf, w=0.8)
    check(s, [-1.4, -0.3, 0.45, 1.3], "graphene narrow lead")


def test_clean_square_crosses_at_band_centre():
    """Clean crosses at E = 0, where the square domains of the leaves are singular: their separators
    delay to the parents (the re-planned path). Where the whole interior (the sites no lead couples
    to) is singular as well, the delayed block reaches the root singular and the point fails as
    the reference's root does (nested.rs:834-838)."""
    delayed = 0
    for nx, ny, arm in ((10, 7, 2), (9, 6, 3)):
        sites, itf = cross(ca.Model.Square, nx, ny, arm)
        s = multi_system(ca.Model.Square, sites, itf)
        stats = ca.capi.BatchStats()
        sig = [s.self_energies(e) for e in (0.0, 0.5)]
        t, st = ca.multi_batch_transmissions([s, s], sig, ca.Precision.Double, stats)
        assert (st == 0).all(), st
        for k, e in enumerate((0.0, 0.5)):
            d = dense_multi(s, sig[k], e)
            assert np.abs(t[k] - d).max() < TOL * (1 + np.abs(d).max()), (e, t[k], d)
        say(f"[gpu] clean square cross {nx}x{ny}, arms {arm}: delayed points {stats.delayed_points}")
        delayed += stats.delayed_points
    assert delayed > 0
    sites, itf = cross(ca.Model.Square, 6, 6, 2)  # singular interior at E = 0
    s = multi_system(ca.Model.Square, sites, itf)
    t, st = ca.multi_batch_transmissions([s], [s.self_energies(0.0)])
    assert st[0] == 1 and np.isnan(t[0]).all()


def test_two_leads_as_two_terminal():
    """Two leads at the ends of a disordered ribbon: T[1][0] = T[0][1] = the two-terminal T
    (nested_multi.rs tests two_leads_as_two_terminal, to 1e-10)."""
    sites = [(x, y, 0, 0) for x in range(14) for y in range(6)]
    pot = hash_disorder([s[:3] for s in sites], [0] * len(sites), 2.0)
    itf = [([(0, y, 0, 0) for y in range(6)], (-1, 0, 0)), ([(13, y, 0, 0) for y in range(6)], (1, 0, 0))]
    multi = ca.NestedMultiTerminal.from_model(ca.Model.Square, sites, itf, site_potential=pot)
    two = ca.NestedTwoTerminal.preset(ca.Model.Square, [14, 6, 0], site_potential=pot)
    for e in (-2.9, -1.1, 0.0, 0.7):
        t, t2 = multi.transmissions(e), two.transmission(e)
        assert abs(t[1, 0] - t2) < 1e-10 and abs(t[0, 1] - t2) < 1e-10, (e, t, t2)


def test_bhz_spin_sectors():
    """Spin-resolved transmissions of the BHZ cross (sectors [0, 0, 1, 1]: the spin blocks):
    the sectors of a lead pair add up to its T, and spin is conserved (no spin-flip terms)."""
    sites, itf = cross(ca.Model.BHZ, 8, 6, 2)
    s = multi_system(ca.Model.BHZ, sites, itf, w=0.5)
    for e in (-0.2, 0.3):
        sec = s.sector_transmissions(e, [[0, 0, 1, 1]] * 4)
        t = s.transmissions(e)
        assert sec.labels == [(p, a) for p in range(4) for a in (0, 1)]
        ts = sec.t.reshape(4, 2, 4, 2)
        assert np.abs(ts.sum(axis=(1, 3)) - t).max() < 1e-10 * (1 + np.abs(t).max())
        assert np.abs(ts[:, 0, :, 1]).max() < 1e-12 and np.abs(ts[:, 1, :, 0]).max() < 1e-12
        say(f"[gpu] BHZ cross spin sectors E = {e}: spin-up T_10 {ts[1, 0, 0, 0]:.6f}, down {ts[1, 1, 0, 1]:.6f}")


def test_single_precision_fronts():
    sites, itf = cross(ca.Model.Square, 9, 7, 3)
    s = multi_system(ca.Model.Square, sites, itf, w=1.5)
    check(s, [-0.4, 1.3], "square cross, complex f32 fronts", ca.Precision.Single, tol=1e-4)


def test_batch_equals_points():
    """One batched launch over three disorder configurations x three energies equals the points one
    at a time, bit for bit (no atomics in the labelled trace)."""
    sites, itf = cross(ca.Model.Square, 7, 5, 2)
    systems = [multi_system(ca.Model.Square, sites, itf, w=w) for w in (0.5, 1.0, 1.5)]
    pts = [(s, e) for s in systems for e in (-1.0, 0.2, 2.1)]
    sig = [s.self_energies(e) for s, e in pts]
    t, st = ca.multi_batch_transmissions([s for s, _ in pts], sig)
    assert (st == 0).all()
    for k, (s, e) in enumerate(pts):
        assert np.array_equal(t[k], s.transmissions_with(sig[k]))


@pytest.mark.parametrize("name", ["square", "bhz", "graphene"])
def test_smatrix_unitary_with_the_transmissions(name):
    """The scattering matrix in the basis of the lead modes is unitary and its blocks give the
    transmissions (crates/caroli-green/tests/nested_multi.rs:322-340)."""
    if name == "square":
        sites, itf = cross(ca.Model.Square, 9, 7, 3)
        s, energies = multi_system(ca.Model.Square, sites, itf, w=1.5), [-2.7, -0.4, 1.3]
    elif name == "bhz":
        sites, itf = cross(ca.Model.BHZ, 8, 6, 2)
        s, energies = multi_system(ca.Model.BHZ, sites, itf, w=0.5), [-0.2, 0.02, 0.3]
    else:
        sites, itf = graphene_narrow_lead()
        s, energies = multi_system(ca.Model.Graphene, sites, itf, w=0.8), [-1.4, 0.45, 1.3]
    for e in energies:
        sm = s.smatrix(e)
        full = sm.full()
        n = full.shape[1]
        assert full.shape[0] == n
        unitarity = np.abs(full.conj().T @ full - np.eye(n)).max()
        t = s.transmissions(e)
        worst = max(abs(sm.transmission(p, q) - t[p, q]) for p in range(len(t)) for q in range(len(t)) if p != q)
        say(f"[gpu] {name} S-matrix E = {e}: {n} modes, |S^H S - 1| {unitarity:.1e}, max |sum|S_pq|^2 - T_pq| {worst:.1e}")
        assert unitarity < 1e-9 and worst < 1e-9
        assert all(np.allclose(np.abs(l), 1.0) for l in sm.incoming + sm.outgoing)


def test_smatrix_clean_ribbon_sends_each_mode_through():
    """A clean ribbon: each column of S_RL has one entry of modulus 1, the others vanish, and
    nothing is reflected (nested_multi.rs:387-415)."""
    sites = [(x, y, 0, 0) for x in range(8) for y in range(9)]
    itf = [([(0, y, 0, 0) for y in range(9)], (-1, 0, 0)), ([(7, y, 0, 0) for y in range(9)], (1, 0, 0))]
    s = ca.NestedMultiTerminal.from_model(ca.Model.Square, sites, itf)
    for e in (-2.9, -1.1, 0.3, 1.7):
        sm = s.smatrix(e)
        t = np.abs(sm.blocks[1][0])
        assert ((np.abs(t - 1) < 1e-9).sum(axis=0) == 1).all() and ((t < 1e-9).sum(axis=0) == t.shape[0] - 1).all()
        assert np.abs(sm.blocks[0][0]).max() < 1e-9


def test_cross_bond_currents():
    """With lead q alone injecting, the current out through a cut across arm p is T_pq, into the
    device through arm q their sum, and Kirchhoff's law holds at every site no lead touches
    (crates/caroli-green/tests/nested_multi.rs:166-228)."""
    sites, itf = cross(ca.Model.Square, 9, 7, 4)
    s = multi_system(ca.Model.Square, sites, itf, w=1.5)
    lead_sites = {tuple(x) for sl, _ in itf for x in sl}
    for e in (-2.4, -0.6, 0.8):
        t = s.transmissions(e)
        states = s.scattering_states(s.self_energies(e))
        assert len(states.leads) == 4
        for q in range(4):
            f = np.zeros(4)
            f[q] = 1.0
            fr, to, cur = s.bond_currents(states, f)
            for p, (sl, d) in enumerate(itf):
                depth = lambda i: sites[i][0] * d[0] + sites[i][1] * d[1]
                edge = sl[0][0] * d[0] + sl[0][1] * d[1] - 1
                out = sum(c if (depth(a) < edge <= depth(b)) else -c if (depth(b) < edge <= depth(a)) else 0.0
                          for a, b, c in zip(fr, to, cur))
                expected = -sum(t[r, q] for r in range(4) if r != q) if p == q else t[p, q]
                assert abs(out - expected) < 1e-9, (e, q, p, out, expected)
            net = np.zeros(len(sites))
            np.add.at(net, fr, cur)
            np.add.at(net, to, -cur)
            inner = [i for i, x in enumerate(sites) if tuple(x) not in lead_sites]
            assert np.abs(net[inner]).max() < 1e-9


def test_multi_states_as_two_terminal():
    """Two leads at the ends of a ribbon: the multi-terminal states give the densities and bond
    currents of the two-terminal ones (nested_multi.rs:134-163)."""
    sites = [(x, y, 0, 0) for x in range(11) for y in range(6)]
    pot = hash_disorder([s[:3] for s in sites], [0] * len(sites), 2.0)
    itf = [([(0, y, 0, 0) for y in range(6)], (-1, 0, 0)), ([(10, y, 0, 0) for y in range(6)], (1, 0, 0))]
    multi = ca.NestedMultiTerminal.from_model(ca.Model.Square, sites, itf, site_potential=pot)
    two = ca.NestedTwoTerminal.preset(ca.Model.Square, [11, 6, 0], site_potential=pot)
    f = [0.9, 0.2]
    for e in (-2.2, 0.3, 1.6):
        sm = multi.scattering_states(multi.self_energies(e))
        st = two.scattering_states(two.self_energies(e))
        assert np.abs(multi.site_densities(sm, f) - two.site_densities(st, f)).max() < 1e-10
        a, b = multi.bond_currents(sm, f), two.bond_currents(st, f)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.abs(a[2] - b[2]).max() < 1e-10
