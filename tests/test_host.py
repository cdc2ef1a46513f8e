"""CPU tests: the oracle pinned against closed forms, dense inversion and the reference's golden
data; the host builders, symbolic phase and FP64 lead self-energies of the engine; the C ABI.

No GPU call is made here.
"""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import caroli_amd as ca
from caroli_amd import _capi as capi
from _oracle import OracleSystem, dense_transmission
from helpers import hashed, rebuilt, sigmas_of

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# -- random streams (crates/caroli-disorder/src/rng.rs:92-113, values printed by gfortran 16) ------

def uniforms(mode, seed, n, put=None):
    out = np.zeros(n)
    p = (C.c_int32 * 8)(*put) if put is not None else None
    assert capi.cg_uniforms(mode, seed, p, n, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return out


def test_rng_reproduces_gfortran_golden_draws():
    expected = [4.65636743897908767e-01, 1.00833111555351462e-01, 8.41914420882911130e-01,
                7.73831493534653414e-01, 1.80372332904081167e-01, 8.44873831312145440e-01]
    assert list(uniforms(1, 12345, 6)) == expected
    assert uniforms(1, 1237, 1)[0] == 4.72208155982870581e-01
    assert uniforms(2, 0, 1, put=[1, 2, 3, 4, 5, 6, 7, 8])[0] == 4.71070899360821249e-01


def xoshiro_seeded(seed, n):
    """Independent restatement of splitmix64 + xoshiro256** (Blackman & Vigna) in Python."""
    M = (1 << 64) - 1
    sm = seed

    def split():
        nonlocal sm
        sm = (sm + 0x9E3779B97F4A7C15) & M
        z = sm
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    s = [split() for _ in range(4)]
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M  # noqa: E731
    out = []
    for _ in range(n):
        r = (rotl((s[1] * 5) & M, 7) * 9) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        out.append((r >> 11) * (1.0 / (1 << 53)))
    return np.array(out)


def test_seed_from_u64_matches_independent_restatement():
    for seed in (1, 2, 5, 1000, 123456789):
        assert np.array_equal(uniforms(0, seed, 50), xoshiro_seeded(seed, 50))


# -- host builders ------------------------------------------------------------------------------

def dense_h(arr):
    n = arr["n"]
    h = np.zeros((n, n), complex)
    for r in range(n):
        for e in range(arr["row_ptr"][r], arr["row_ptr"][r + 1]):
            h[r, arr["cols"][e]] += arr["vals"][e]
    return h


def test_bhz_hamiltonian_entries():
    # crates/caroli-model/src/presets.rs:34-56 and gamma.rs, written out by hand.
    a, b, c, d, m = 1.0, -0.376406, 0.0, -0.280933, -0.137
    g0 = np.eye(4)
    g55 = np.diag([1, -1, 1, -1])
    g33 = np.array([[0, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, -1], [0, 0, -1, 0]])
    g44 = np.array([[0, -1j, 0, 0], [1j, 0, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]])
    onsite = (c - 4 * d) * g0 + (m - 4 * b) * g55
    tx = d * g0 + b * g55 - 0.5j * a * g33
    ty = d * g0 + b * g55 + 0.5j * a * g44
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [2, 2, 0])
    h = dense_h(s.arrays())
    # Sites (0,0),(0,1),(1,0),(1,1) at orbitals 0,4,8,12.
    assert np.allclose(h[0:4, 0:4], onsite)
    assert np.allclose(h[0:4, 8:12], tx)  # <(0,0)|H|(1,0)>
    assert np.allclose(h[0:4, 4:8], ty)  # <(0,0)|H|(0,1)>
    assert np.allclose(h, h.conj().T)
    assert s.block_count() == 2


def test_anderson_preset_uses_the_seeded_stream():
    # crates/caroli-cli/src/build.rs:283-295: w (u - 0.5) per site, sites in region order.
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [5, 4, 0], disorder=1, strength=2.0, seed=7)
    h = dense_h(s.arrays())
    assert np.allclose(np.diag(h).real, 2.0 * (xoshiro_seeded(7, 20) - 0.5), atol=0, rtol=0)


def test_magnetic_preset_polarized_opposite_g():
    # Polarized m_z = 1 moments with opposite g: J diag(1, -1, -1, 1) on floor(c N) sites
    # chosen by partial Fisher-Yates (crates/caroli-disorder/src/lib.rs:18-26).
    ext = [6, 5, 0]
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=2, strength=0.4, concentration=0.2, seed=3)
    clean = ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext)
    dh = dense_h(s.arrays()) - dense_h(clean.arrays())
    u = xoshiro_seeded(3, 6)
    n, k = 30, 6
    tab = list(range(n))
    for i in range(k):
        r = i + 1 + int((n - i) * u[i]) - 1
        tab[i], tab[r] = tab[r], tab[i]
    chosen = sorted(tab[:k])
    for site in range(n):
        blk = dh[4 * site:4 * site + 4, 4 * site:4 * site + 4]
        want = 0.4 * np.diag([1, -1, -1, 1]) if site in chosen else np.zeros((4, 4))
        assert np.allclose(blk, want), site
    assert np.allclose(dh - np.diag(np.diag(dh)), 0)


# -- symbolic phase: the engine's tree equals the oracle's --------------------------------------

@pytest.mark.parametrize("model,ext,kw", [
    (ca.Model.BHZ, [16, 10, 0], {}),
    (ca.Model.Graphene, [12, 6, 0], {}),
    (ca.Model.Cubic, [7, 5, 4], {}),
    (ca.Model.Square, [30, 9, 0], {}),
])
def test_tree_matches_oracle(model, ext, kw):
    s = ca.NestedTwoTerminal.preset(model, ext, **kw)
    o = OracleSystem(s.arrays())
    assert o.block_count() == s.block_count()
    for b in range(s.block_count()):
        nodes = s.block_sizes(b)[2]
        assert o.block(b)[2] == nodes
        for k in range(nodes):
            e1, d1 = s.tree_node(b, k)
            e2, d2 = o.node(b, k)
            assert np.array_equal(e1, e2) and np.array_equal(d1, d2), (b, k)


def test_benchmark_tree_statistics():
    # Node counts of the symbolic phase at the benchmark sizes (restated from nested.rs:1051-1122).
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [200, 64, 0], disorder=2, strength=0.4, concentration=0.2)
    assert s.block_count() == 2
    nl, nr, nodes, levels = s.block_sizes(0)
    assert (nl, nr, nodes) == (128, 128, 4109)
    sizes = [sum(len(x) for x in s.tree_node(0, k)) for k in range(nodes)]
    assert max(sizes) == 384 and sum(1 for x in sizes if x >= 64) == 297


# -- oracle against dense inversion and closed forms ----------------------------------------------

def blocks_of(o):
    return [tuple(x.astype(int) for x in o.block(b)[:2]) for b in range(o.block_count())]


@pytest.mark.parametrize("label,sysf,energies", [
    ("square clean", lambda: ca.NestedTwoTerminal.preset(ca.Model.Square, [14, 9, 0]), [-3.1, -1.3, 0.0, 0.4, 2.2]),
    ("square dirty", lambda: hashed(ca.Model.Square, [14, 9, 0], 2.0), [-2.7, -0.5, 0.0, 1.1, 3.0]),
    ("bhz dirty", lambda: hashed(ca.Model.BHZ, [12, 7, 0], 0.6), [-0.4, -0.05, 0.0, 0.08, 0.5]),
    ("graphene", lambda: ca.NestedTwoTerminal.preset(ca.Model.Graphene, [10, 6, 0]), [-2.3, -0.7, 0.3, 1.6]),
    ("graphene dirty", lambda: hashed(ca.Model.Graphene, [10, 6, 0], 1.0, subs_per_cell=2), [-1.1, 0.25, 0.9]),
    ("cubic", lambda: hashed(ca.Model.Cubic, [6, 4, 3], 2.5, dim=3), [-4.1, -1.3, 0.0, 2.2]),
    ("short", lambda: hashed(ca.Model.Square, [2, 6, 0], 1.5), [-1.2, 0.7]),
])
def test_oracle_matches_dense_inverse(label, sysf, energies):
    s = sysf()
    arr = s.arrays()
    o = OracleSystem(arr)
    for e in energies:
        _, sig = sigmas_of(s, e)
        t, rc = o.transmission_with(e, sig)
        assert rc == 0
        td = dense_transmission(arr, e, sig, blocks_of(o))
        assert abs(t - td) < 1e-9 * (1 + abs(td)), f"{label} E={e}: {t} vs {td}"


def test_oracle_delays_on_singular_blocks():
    # Clean square lattice at E = 0: singular interior blocks must be delayed, not divided by.
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [14, 9, 0])
    o = OracleSystem(s.arrays())
    _, sig = sigmas_of(s, 0.0)
    t, rc = o.transmission_with(0.0, sig)
    assert rc == 0 and abs(t - round(t)) < 1e-9


def has_fma():
    try:
        return " fma " in open("/proc/cpuinfo").read().replace("\n", " ")
    except OSError:
        return False


@pytest.mark.skipif(not has_fma(), reason="the contracted build of the oracle needs FMA instructions")
@pytest.mark.parametrize("config, k", [(0, 0), (5, 2)])
def test_c2_fixture_points_recompute(config, k):
    """Points of tests/golden/bhz200x64_magnetic.json recomputed on the CPU from the three oracle
    builds (double, long double, double with contracted multiply-adds): each value is reproduced
    exactly. On the second point (configuration 5, E = -0.14) the two double evaluations differ by
    about 1.5e-9, a rounding error of double precision alone, against which the long-double value
    is the reference."""
    fx = np.array(json.load(open(os.path.join(ROOT, "tests", "golden", "bhz200x64_magnetic.json")))["points"])
    row = fx[config * 33 + k]
    e = float(np.linspace(-0.16, 0.16, 33)[k])
    assert row[0] == config and row[1] == e
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [200, 64, 0], disorder=2, strength=0.4, concentration=0.2,
                                    seed=1 + config)
    arr = s.arrays()
    _, sig = sigmas_of(s, e)
    got = []
    for kw in (dict(), dict(extended=True), dict(fma=True)):
        t, rc = OracleSystem(arr, **kw).transmission_with(e, sig)
        assert rc == 0
        got.append(t)
    assert got == [row[2], row[3], row[4]]
    spread = abs(got[0] - got[2])
    assert (spread > 1e-9) == (config == 5), spread
    assert abs(got[0] - got[1]) < 2e-10 and abs(got[2] - got[1]) < 2e-9


def channel_count(w, e):
    """Open channels of a clean square ribbon of width w at energy e (t = 1): transverse modes
    eps_k = -2 cos(k pi / (w + 1)) with |e - eps_k| < 2."""
    eps = -2 * np.cos(np.arange(1, w + 1) * np.pi / (w + 1))
    return int(np.sum(np.abs(e - eps) < 2))


def test_clean_square_ribbon_counts_channels():
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [9, 7, 0])
    o = OracleSystem(s.arrays())
    for e in (-3.3, -2.1, -0.9, 0.2, 1.7, 3.5):
        _, sig = sigmas_of(s, e)
        t, rc = o.transmission_with(e, sig)
        assert rc == 0 and abs(t - channel_count(7, e)) < 1e-9, e


def test_chain_single_impurity_closed_form():
    # crates/caroli-green/src/two_terminal.rs:456-466: T = 1 / (1 + (u / (2 sin k))^2), E = -2 cos k.
    u = 0.7
    pot = np.array([0, 0, u, 0, 0], float)
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [5, 1, 0], site_potential=pot)
    o = OracleSystem(s.arrays())
    for e in (-1.5, -0.4, 0.3, 1.2):
        k = np.arccos(-e / 2)
        _, sig = sigmas_of(s, e)
        t, rc = o.transmission_with(e, sig)
        assert abs(t - 1 / (1 + (u / (2 * np.sin(k))) ** 2)) < 1e-12
    _, sig = sigmas_of(s, 2.5)  # outside the band
    assert o.transmission_with(2.5, sig)[0] < 1e-12


def test_clean_bhz_ribbon_matches_thesis_golden():
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "ti_cond2t_clean.json")))
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [16, 10, 0])
    o = OracleSystem(s.arrays())
    for e, g in zip(fx["energies"], fx["conductance"]):
        _, sig = sigmas_of(s, e)
        t, rc = o.transmission_with(e, sig)
        assert rc == 0
        assert abs(t - round(t)) < 1e-8, (e, t)
        assert abs(t - g) < fx["tolerance_vs_exact"], (e, t, g)


# -- lead self-energies --------------------------------------------------------------------------

def test_chain_self_energy_closed_form():
    # Surface Green's function of the t = -1 chain: g = (E - i sqrt(4 - E^2)) / 2 (two_terminal.rs:471-480).
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [3, 1, 0])
    for e in (-1.7, -0.3, 0.0, 0.9, 1.95):
        sl, sr, cl, cr = s.self_energies(e).block(0)
        g = (e - 1j * np.sqrt(4 - e * e)) / 2
        assert abs(sr[0, 0] - g) < 1e-12 and abs(sl[0, 0] - g) < 1e-12 and cl == cr == 1
    sl, sr, cl, cr = s.self_energies(2.6).block(0)
    g = (2.6 - np.sqrt(2.6 ** 2 - 4)) / 2
    assert abs(sr[0, 0] - g) < 1e-12 and cl == 0


@pytest.mark.parametrize("model,ext,energies", [
    (ca.Model.BHZ, [4, 10, 0], [-0.3, -0.16, 0.0, 0.07, 0.25]),
    (ca.Model.Graphene, [4, 8, 0], [-1.1, -0.45, 0.3, 0.77]),
    (ca.Model.Square, [4, 6, 0], [-2.5, -0.5, 1.1]),
    (ca.Model.Cubic, [3, 4, 3], [-3.3, 0.4, 2.1]),
])
def test_self_energies_solve_the_lead_dyson_equation(model, ext, energies):
    # The retarded Sigma_R = H01 (E - H00 - Sigma_R)^-1 H01^H, Sigma_L = H01^H (E - H00 - Sigma_L)^-1 H01,
    # with broadenings i(Sigma - Sigma^H) positive semidefinite (retarded branch).
    s = ca.NestedTwoTerminal.preset(model, ext)
    arr = s.arrays()
    o = OracleSystem(arr)
    for e in energies:
        sig = s.self_energies(e)
        for b in range(s.block_count()):
            left, right, _ = o.block(b)
            lpos = [list(np.nonzero(np.isin(np.sort(np.concatenate([left])), left))[0])]
            sl, sr, cl, cr = sig.block(b)
            # Restrict the lead cells to the block (left orbitals are the first slice's, in order).
            first = np.searchsorted(np.sort(left), left)
            allleft = np.sort(np.concatenate([o.block(q)[0] for q in range(o.block_count())]))
            kl = np.searchsorted(allleft, left)
            h00 = arr["left_h00"][np.ix_(kl, kl)]
            h01 = arr["left_h01"][np.ix_(kl, kl)]
            n = len(kl)
            rhs_r = h01 @ np.linalg.solve(e * np.eye(n) - h00 - sr, h01.conj().T)
            rhs_l = h01.conj().T @ np.linalg.solve(e * np.eye(n) - h00 - sl, h01)
            scale = 1 + np.abs(h01).max() ** 2
            assert np.abs(sr - rhs_r).max() < 1e-9 * scale, (model, e, b)
            assert np.abs(sl - rhs_l).max() < 1e-9 * scale, (model, e, b)
            for sg in (sl, sr):
                gam = 1j * (sg - sg.conj().T)
                assert np.linalg.eigvalsh((gam + gam.conj().T) / 2).min() > -1e-9
            del lpos, first


def test_reused_self_energies_are_checked():
    # Self-energies of a system with other leads are refused (nested.rs:121-129 panics).
    a = ca.NestedTwoTerminal.preset(ca.Model.Square, [6, 5, 0])
    b = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [6, 5, 0])
    sig = a.self_energies(0.4)
    with pytest.raises(ca.GreenError) as err:
        b.transmission_with(sig)
    assert err.value.code == 6


# -- the C ABI -----------------------------------------------------------------------------------

def test_library_exports_every_declared_function():
    header = open(os.path.join(ROOT, "include", "caroli_amd.h")).read()
    declared = set(re.findall(r"\b(cg_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(capi.lib, name), name
    assert declared == set(capi.EXPORTS)


def test_invalid_arguments_are_refused():
    h = C.c_void_p()
    ext = (C.c_uint64 * 3)(4, 4, 0)
    assert capi.cg_preset_create(9, ext, 0, 0, 0.0, 0.0, 1, None, C.byref(h)) == 6
    assert capi.cg_transmission_with(None, None, 0, None) == 6
    cols = (C.c_int64 * 2)()
    assert capi.cg_scattering_states(None, None, cols, None, None) == 6
    assert capi.cg_site_densities(None, None, None, None, None, None, 0, None) == 6
    assert capi.cg_bond_currents(None, None, None, None, None, None, None, None, None) == 6
    assert capi.cg_operator_currents(None, None, None, None, None, None, 0, None, None, None, None) == 6
    # A system's states checked against its self-energies: channel counts need no device.
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [8, 5, 0])
    sig = s.self_energies(0.3)
    assert capi.cg_scattering_states(s._h, sig._h, cols, None, None) == 0
    assert cols[0] == cols[1] and cols[0] > 0
    other = ca.NestedTwoTerminal.preset(ca.Model.Square, [8, 6, 0])
    assert capi.cg_scattering_states(other._h, sig._h, cols, None, None) == 6  # foreign self-energies


@pytest.mark.parametrize("model,ext,kw", [
    (ca.Model.BHZ, [20, 8, 0], dict(disorder=2, strength=0.4, concentration=0.2, seed=3)),
    (ca.Model.Graphene, [12, 6, 0], dict(disorder=1, strength=1.0, seed=2)),
    (ca.Model.Cubic, [6, 4, 4], dict(disorder=1, strength=16.5, seed=1000)),
])
def test_system_create_round_trip(model, ext, kw):
    """preset -> cg_system_arrays -> cg_system_create gives the same system: the same arrays, the
    same elimination trees node by node and bit-identical self-energies."""
    s = ca.NestedTwoTerminal.preset(model, ext, **kw)
    arr = s.arrays()
    r = rebuilt(arr)
    arr2 = r.arrays()
    for k, v in arr.items():
        assert np.array_equal(np.asarray(v), np.asarray(arr2[k])), k
    assert r.block_count() == s.block_count()
    for b in range(s.block_count()):
        assert r.block_sizes(b) == s.block_sizes(b)
        for k in range(s.block_sizes(b)[2]):
            e1, d1 = s.tree_node(b, k)
            e2, d2 = r.tree_node(b, k)
            assert np.array_equal(e1, e2) and np.array_equal(d1, d2), (b, k)
    g1, g2 = s.self_energies(0.07), r.self_energies(0.07)
    for b in range(s.block_count()):
        for m1, m2 in zip(g1.block(b), g2.block(b)):
            assert np.array_equal(np.asarray(m1), np.asarray(m2))
    # Self-energies of one are accepted by the other (same leads and blocks; the channel counts of
    # the states need no device).
    cols = (C.c_int64 * 2)()
    assert capi.cg_scattering_states(r._h, g1._h, cols, None, None) == 0


def test_system_create_refuses_malformed_arrays():
    """cg_system_create refuses malformed CSR and geometry with CG_ERR_INVALID (6)."""
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [6, 4, 0])
    arr = s.arrays()
    n = arr["n"]

    def refused(**changes):
        with pytest.raises(ca.GreenError) as err:
            rebuilt(arr, **changes)
        assert err.value.code == 6, changes

    rp = arr["row_ptr"].copy()
    rp[0] = 1
    refused(row_ptr=rp)                                  # CSR not starting at 0
    rp = arr["row_ptr"].copy()
    rp[3], rp[4] = rp[4], rp[3]
    refused(row_ptr=rp)                                  # decreasing row pointers
    cl = arr["cols"].copy()
    cl[0] = n
    refused(cols=cl)                                     # column out of range
    so = arr["site_orbitals"].copy()
    so[2] = -1
    refused(site_orbitals=so)                            # negative orbital count
    so = arr["site_orbitals"].copy()
    so[2] = 0
    refused(site_orbitals=so)                            # empty site
    sf = arr["site_offsets"].copy()
    sf[3] = sf[2]
    refused(site_offsets=sf)                             # overlapping sites (one orbital uncovered)
    sf = arr["site_offsets"].copy()
    sf[-1] = n
    refused(site_offsets=sf)                             # a site past the last orbital
    refused(left_h00=arr["left_h00"][:-1, :-1], left_h01=arr["left_h01"][:-1, :-1])  # lead cell != end slice
    refused(axis=2)                                      # axis outside the lattice dimension
    assert rebuilt(arr).sizes() == s.sizes()             # the untouched arrays are accepted


def test_tuning_defaults_variables_and_override(monkeypatch):
    """cg_get_tuning / cg_set_tuning: the defaults, the CAROLI_* variables, an override over them,
    refused values, and the way back (no device needed)."""
    for v in ("CAROLI_SMALL_FRONTS", "CAROLI_FRONT_KERNEL", "CAROLI_BLOCKED_MIN_NE", "CAROLI_SCHUR_TILE"):
        monkeypatch.delenv(v, raising=False)
    ca.set_tuning(None)
    try:
        assert ca.get_tuning() == {"small_fronts": 0, "split": 1, "blocked_min_ne": 128, "unblocked": 0,
                                   "schur_tile": 0}
        monkeypatch.setenv("CAROLI_FRONT_KERNEL", "border")
        monkeypatch.setenv("CAROLI_SMALL_FRONTS", "reg")
        monkeypatch.setenv("CAROLI_SCHUR_TILE", "32")
        assert ca.get_tuning() == {"small_fronts": 2, "split": 0, "blocked_min_ne": 1 << 30, "unblocked": 0,
                                   "schur_tile": 32}
        monkeypatch.delenv("CAROLI_SCHUR_TILE")
        monkeypatch.setenv("CAROLI_FRONT_KERNEL", "blocked")
        monkeypatch.setenv("CAROLI_BLOCKED_MIN_NE", "96")  # read after CAROLI_FRONT_KERNEL
        assert ca.get_tuning()["blocked_min_ne"] == 96
        ca.set_tuning(small_fronts=1, blocked_min_ne=256)
        assert ca.get_tuning() == {"small_fronts": 1, "split": 1, "blocked_min_ne": 256, "unblocked": 0,
                                   "schur_tile": 0}
        monkeypatch.setenv("CAROLI_SMALL_FRONTS", "lds")  # the override wins over the variables
        assert ca.get_tuning()["small_fronts"] == 1
        for bad in ({"small_fronts": 3}, {"split": 2}, {"blocked_min_ne": -1}, {"unblocked": -1}, {"schur_tile": 48}):
            with pytest.raises(ca.GreenError):
                ca.set_tuning(**bad)
        assert ca.get_tuning()["blocked_min_ne"] == 256
        assert capi.cg_get_tuning(None) == 6
        assert capi.cg_batch_autotune(None, None, None) == 6
    finally:
        ca.set_tuning(None)
    monkeypatch.delenv("CAROLI_BLOCKED_MIN_NE")
    assert ca.get_tuning() == {"small_fronts": 1, "split": 1, "blocked_min_ne": 0, "unblocked": 0, "schur_tile": 0}


def _oracle_states(s, e):
    """Scattering states psi_p = G[:, p] W_p from the oracle's G columns (W W^H = Gamma by an eigen-
    factorisation, a gauge of its own), with the oracle's T and the left-slice G columns."""
    arr = s.arrays()
    o = OracleSystem(arr)
    _, sig = sigmas_of(s, e)
    t, rc = o.transmission_with(e, sig)
    assert rc == 0

    def root(gam):
        w, v = np.linalg.eigh((gam + gam.conj().T) / 2)
        keep = w > 1e-12 * w.max()
        return v[:, keep] * np.sqrt(w[keep])

    left, right, gcols = [], [], []
    for b in range(s.block_count()):
        sl, sr = sig[b]
        gl_cols = o.green_columns(b, e, sl, sr, 0)
        left.append(gl_cols @ root(1j * (sl - sl.conj().T)))
        right.append(o.green_columns(b, e, sl, sr, 1) @ root(1j * (sr - sr.conj().T)))
        gcols.append((o.block(b)[0].astype(int), gl_cols))
    return ca.ScatteringStates(e, [np.hstack(left), np.hstack(right)]), t, gcols


def test_state_observables_conserve_current_and_sum_rule():
    """cg_bond_currents and cg_site_densities (host code) on states built from the oracle's G
    columns: with f = (1, 0) the current through every transverse cut is T (Landauer, and current
    conservation); with f = (1, 1) a site's density is its spectral density -Im G_aa / pi summed over
    its orbitals (G Gamma_L G^H + G Gamma_R G^H = i (G - G^H)), checked on the left-slice sites."""
    s = ca.NestedTwoTerminal.preset(ca.Model.Square, [14, 7, 0], disorder=1, strength=1.0, seed=3)
    states, t, gcols = _oracle_states(s, 0.4)
    arr = s.arrays()
    x = np.asarray(arr["cells"]).reshape(-1, 3)[:, 0]
    fr, to, cur = s.bond_currents(states, [1.0, 0.0])
    assert np.all(fr < to) and len(set(zip(fr, to))) == len(fr)
    for c in range(int(x.min()), int(x.max())):
        across = np.where((x[fr] <= c) & (c < x[to]), cur, 0.0) - np.where((x[to] <= c) & (c < x[fr]), cur, 0.0)
        assert abs(across.sum() - t) <= 1e-10 * t, (c, across.sum(), t)
    dens = s.site_densities(states, [1.0, 1.0])
    offs, norb = np.asarray(arr["site_offsets"]), np.asarray(arr["site_orbitals"])
    left, g = gcols[0]
    pos = {int(o): k for k, o in enumerate(left)}
    checked = 0
    for i in range(len(offs)):
        orbs = range(int(offs[i]), int(offs[i]) + int(norb[i]))
        if all(o in pos for o in orbs):
            ldos = sum(-g[o, pos[o]].imag / np.pi for o in orbs)
            assert abs(dens[i] - ldos) <= 1e-10 * abs(ldos), (i, dens[i], ldos)
            checked += 1
    assert checked > 0
    one = np.eye(int(norb[0]))
    assert np.allclose(s.site_densities(states, [0.3, 0.7], op=one), s.site_densities(states, [0.3, 0.7]), rtol=1e-13)
    with pytest.raises(ca.GreenError):
        s.site_densities(states, [1.0, 0.0], op=np.eye(int(norb[0]) + 1))


def test_operator_currents_resolve_the_decoupled_blocks():
    """cg_operator_currents (host code) on the clean BHZ ribbon, whose two blocks are decoupled
    sectors of each site's four orbitals: O = 1 gives the bond currents, and the projector on one
    block's orbitals carries that block's transmission through every cut (f = (1, 0))."""
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [20, 10, 0])
    e = 0.05
    states, t, _ = _oracle_states(s, e)
    arr = s.arrays()
    o = OracleSystem(arr)
    _, sig = sigmas_of(s, e)
    x = np.asarray(arr["cells"]).reshape(-1, 3)[:, 0]
    fr, to, cb = s.bond_currents(states, [1.0, 0.0])
    fr1, to1, c1 = s.operator_currents(states, [1.0, 0.0], np.eye(4))
    assert np.array_equal(fr, fr1) and np.array_equal(to, to1)
    assert np.abs(c1 - cb).max() <= 1e-13 * np.abs(cb).max()
    total = 0.0
    for b in range(s.block_count()):
        sl, sr = sig[b]
        left = o.block(b)[0].astype(int)
        g_r = o.green_columns(b, e, sl, sr, 1)
        local = sorted(set((np.nonzero(np.abs(g_r).max(axis=1) > 0)[0] % 4).tolist()))
        g_lr = g_r[left, :]
        t_b = np.trace(1j * (sl - sl.conj().T) @ g_lr @ (1j * (sr - sr.conj().T)) @ g_lr.conj().T).real
        total += t_b
        proj = np.diag([1.0 if k in local else 0.0 for k in range(4)])
        f3, t3, cp = s.operator_currents(states, [1.0, 0.0], proj)
        for c in range(int(x.min()), int(x.max())):
            across = np.where((x[f3] <= c) & (c < x[t3]), cp, 0.0) - np.where((x[t3] <= c) & (c < x[f3]), cp, 0.0)
            assert abs(across.sum() - t_b) <= 1e-10, (b, c, across.sum(), t_b)
    assert abs(total - t) <= 1e-12 * t
    with pytest.raises(ca.GreenError):
        s.operator_currents(states, [1.0, 0.0], np.eye(3))


def test_oracle_gradient_matches_finite_differences():
    # The FP64 gradient built from G columns (2 Re (G Gamma_R G^H Gamma_L G)_aa, the checker of the
    # GPU adjoint) against central differences of the oracle's own transmission.
    from helpers import oracle_gradient

    ext, w, e, h = [8, 5, 0], 40 / 72.9, 0.01, 1e-6
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=w, seed=1)
    sig = s.self_energies(e)
    sig_arr = [sig.block(b)[:2] for b in range(s.block_count())]
    grad = oracle_gradient(s, e)
    for i in (3, 17, 26):
        pot = np.zeros(ext[0] * ext[1])
        pot[i] = h
        tp = OracleSystem(ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=w, seed=1,
                                                      site_potential=pot).arrays()).transmission_with(e, sig_arr)[0]
        tm = OracleSystem(ca.NestedTwoTerminal.preset(ca.Model.BHZ, ext, disorder=1, strength=w, seed=1,
                                                      site_potential=-pot).arrays()).transmission_with(e, sig_arr)[0]
        fd = (tp - tm) / (2 * h)
        assert abs(fd - grad[4 * i:4 * i + 4].sum()) < 1e-6 * max(1.0, abs(fd)), (i, fd)


# -- checked reduced precision (host logic of cg_batch_transmission_checked) ----------------------

def test_checked_batch_sampling_failures_and_fallback(tmp_path):
    """caroli_amd/csrc/checked.hpp with the reference's fake runner (precision.rs:140-168): index 7
    off by 1e-4 in single precision, index 3 failing there."""
    import subprocess
    exe = str(tmp_path / "checked_driver")
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "native", "checked_driver.cpp"), "-o",
                    exe], check=True)
    lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = [[float(x) for x in ln.split()] for ln in lines if ln.strip()]
    exact = list(range(10))
    # every 5: index 7 not sampled (kept), index 3 failed (redone in double), 0 and 5 compared.
    rc, checked, failed, fell, prec, *rest = rows[0]
    t, dt = rest[:10], rest[10:20]
    assert (rc, checked, failed, fell, prec) == (0, 2, 1, 0, 1)
    assert t[3] == 3.0 and t[7] == 7.0 + 1e-4
    assert dt[0] == 0.0 and dt[5] == 0.0 and np.isnan(dt[7]) and np.isnan(dt[3])
    # every 7: index 7 sampled, 1e-4 above 1e-6: everything redone in double.
    rc, checked, failed, fell, prec, *rest = rows[1]
    assert (rc, fell, prec) == (0, 1, 0) and rest[:10] == exact
    assert rest[10 + 7] == pytest.approx(1e-4) and rest[21] == pytest.approx(1e-4)  # dt[7], max
    # double precision: nothing checked.
    rc, checked, failed, fell, prec, *rest = rows[2]
    assert (rc, checked, failed, fell, prec) == (0, 0, 0, 0, 0) and rest[:10] == exact
    # every point checked (the engine's report), tolerance above the error: no fallback; the
    # checked points return their double values, the differences are reported per point.
    rc, checked, failed, fell, prec, *rest = rows[3]
    assert (rc, checked, failed, fell, prec) == (0, 9, 1, 0, 1) and rest[:10] == exact
    dt = np.array(rest[10:20])
    assert dt[7] == pytest.approx(1e-4) and np.isnan(dt[3]) and np.nansum(dt) == pytest.approx(1e-4)


# -- magnetic impurities against the thesis golden data ---------------------------------------------

def thesis_magnetic(case, j):
    """The ti-cond2t magnetic setups (reference/baseline/cases/ti-cond2t-*/ti.ini) through the
    engine's builder (cg_bhz_magnetic_create): 18 x 10 BHZ ribbon, impurity region x in [5, 13)."""
    return ca.NestedTwoTerminal.bhz_magnetic([18, 10], seed=12345, x0=5, nx_imp=8, concentration=0.3, exchange=j,
                                             moments=case["moments"], moment_param=0.0, coupling=case["coupling"],
                                             seeding="gfortran")


@pytest.mark.parametrize("name", ["same-g", "isotropic", "opposite-g"])
def test_magnetic_impurities_reproduce_thesis_golden(name):
    """crates/caroli-validation/tests/ti_conductance_disorder.rs:25-58: with the thesis' own
    random configurations (Xoshiro256::gfortran(12345), the engine's choose_sites and moment draws),
    the leads' self-energies at E + 1e-5 i and the thesis formula Tr[Re(G_L) G Re(G_R) G^H], the
    45 golden conductances are reproduced to 1e-9. The exact Caroli transmission differs from them
    by up to ~0.2 (the thesis' real-part truncation), which the second bound records."""
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "ti_cond2t_magnetic.json")))
    case = fx["cases"][name]
    worst_thesis = worst_exact = 0.0
    for iu, j in enumerate(case["exchange"]):
        s = thesis_magnetic(case, j)
        assert s.block_count() == (1 if name == "isotropic" else 2)  # spin mixing only for isotropic moments
        o = OracleSystem(s.arrays())
        for ie, e in enumerate(case["energies"]):
            g = s.self_energies(e, eta=1e-5)
            blocks = [g.block(b)[:2] for b in range(s.block_count())]
            t, rc = o.transmission_thesis(e, blocks)
            assert rc == 0
            worst_thesis = max(worst_thesis, abs(t - case["conductance"][iu][ie]))
            t_exact, rc = o.transmission_with(e, blocks)
            worst_exact = max(worst_exact, abs(t_exact - case["conductance"][iu][ie]))
    print(f"{name}: |thesis formula - golden| <= {worst_thesis:.1e}, |exact - golden| <= {worst_exact:.2f}")
    assert worst_thesis <= fx["tolerance"]
    assert 1e-3 < worst_exact < 0.3


def test_magnetic_builder_matches_preset_for_the_cli_seeding():
    """The general builder with CLI seeding over the whole region in region order is the preset of
    the equal-size batch: cg_preset_create(BHZ, disorder 2) draws from seed_from_u64 over all
    sites (x slowest), the builder's candidate list lists x fastest; a one-column-wide ribbon has
    both orders equal, so the two must give the same Hamiltonian bit for bit."""
    a = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [12, 1, 0], disorder=2, strength=0.4, concentration=0.25, seed=3)
    b = ca.NestedTwoTerminal.bhz_magnetic([12, 1], seed=3, x0=0, nx_imp=12, concentration=0.25, exchange=0.4,
                                          moments="polarized", moment_param=1.0, coupling="opposite", seeding="u64")
    aa, bb = a.arrays(), b.arrays()
    for k in ("row_ptr", "cols", "vals"):
        assert np.array_equal(aa[k], bb[k]), k
    assert not np.array_equal(aa["vals"], ca.NestedTwoTerminal.preset(ca.Model.BHZ, [12, 1, 0]).arrays()["vals"])
