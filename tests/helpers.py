"""Shared helpers of the tests: systems, self-energies and the oracle's view of them."""
import numpy as np

import caroli_amd as ca
from _oracle import OracleSystem


TOL = 1e-10  # plain complex128 against the oracle: |dT| <= TOL (1 + |T|)


def say(*a):
    print(*a, flush=True)


def sigmas_of(sys_, energy):
    """The engine's host self-energies at `energy`, as the LeadSelfEnergies handle and per-block arrays."""
    sig = sys_.self_energies(energy)
    return sig, [sig.block(b)[:2] for b in range(sys_.block_count())]


def oracle_of(sys_):
    return OracleSystem(sys_.arrays())


def rebuilt(arrays, **changes):
    """A system created through cg_system_create (NestedTwoTerminal.from_arrays) from the arrays of
    another system (`arrays()`), with some of them replaced."""
    a = dict(arrays)
    a.update(changes)
    return ca.NestedTwoTerminal.from_arrays(a["n"], a["row_ptr"], a["cols"], a["vals"], a["cells"],
                                            a["site_orbitals"], a["site_offsets"], a["dim"], a["range"], a["axis"],
                                            a["left_h00"], a["left_h01"], a["right_h00"], a["right_h01"])


def compare(sys_, energies, label, precision=ca.Precision.Double, tol=TOL):
    """One batched launch over `energies` against the oracle's T, to tol (1 + max |T|)."""
    o = oracle_of(sys_)
    sig, tref = [], []
    for e in energies:
        g, arr = sigmas_of(sys_, e)
        t, rc = o.transmission_with(e, arr)
        assert rc == 0, f"{label} E={e}: oracle singular"
        sig.append(g)
        tref.append(t)
    say(f"[gpu] {label}: {len(energies)} points, precision {int(precision)}")
    t, st = ca.batch_transmission([sys_] * len(sig), sig, precision)
    assert (st == 0).all(), f"{label}: status {st}"
    d = np.abs(t - np.array(tref))
    assert d.max() <= tol * (1 + np.abs(tref).max()), f"{label}: max |dT| {d.max():.3e}\n{t}\n{tref}"
    return t, np.array(tref)


def hash_disorder(region_cells, subs, w):
    """The reproducible on-site disorder of crates/caroli-green/tests/nested.rs:30-39, per site."""
    out = []
    for (x, y, _), s in zip(region_cells, subs):
        k = (int(x) * 7919 + int(y) * 104_729 + int(s) * 31) & 0xFFFFFFFFFFFFFFFF
        u = (((k * 6_364_136_223_846_793_005 + 1_442_695_040_888_963_407) & 0xFFFFFFFFFFFFFFFF) >> 11) / float(1 << 53)
        out.append(w * (u - 0.5))
    return np.array(out)


def cuboid_cells(extent, subs_per_cell, dim):
    """Cells and sublattice of each site of Region::cuboid (first axis slowest, sublattice fastest)."""
    n = [extent[a] if a < dim else 1 for a in range(3)]
    cells, subs = [], []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                for s in range(subs_per_cell):
                    cells.append((i, j, k))
                    subs.append(s)
    return cells, subs


def hashed(model, extent, w, subs_per_cell=1, dim=2, axis=0):
    """`disordered(model, w, ...)` of the reference's nested tests, on a cuboid."""
    cells, subs = cuboid_cells(extent, subs_per_cell, dim)
    pot = hash_disorder(cells, subs, w)
    return ca.NestedTwoTerminal.preset(model, extent, axis=axis, site_potential=pot)


def oracle_gradient(s, energy):
    """2 Re (G Gamma_R G^H Gamma_L G)_aa from the oracle's G columns (right slice) and the columns of
    G^T (left slice, from the transposed system)."""
    arr = s.arrays()
    o, ot = OracleSystem(arr), OracleSystem(arr, transpose=True)
    sig = s.self_energies(energy)
    grad = np.zeros(arr["n"])
    for b in range(s.block_count()):
        sl, sr, _, _ = sig.block(b)
        left, right, _ = o.block(b)
        g_r = o.green_columns(b, energy, sl, sr, 1)              # G[:, R]
        gt_l = ot.green_columns(b, energy, sl.T, sr.T, 0)        # G^T[:, L] = G[L, :]^T
        gl = 1j * (sl - sl.conj().T)
        gr = 1j * (sr - sr.conj().T)
        g_lr = g_r[left.astype(int), :]
        m = gr @ g_lr.conj().T @ gl                               # Gamma_R G_LR^H Gamma_L
        grad += 2 * np.real(np.einsum("ac,cd,ad->a", g_r, m, gt_l))
    return grad
