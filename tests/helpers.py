"""Shared helpers of the tests: systems, self-energies and the oracle's view of them."""
import numpy as np

import caroli_amd as ca
from _oracle import OracleSystem


TOL = 1e-10  # plain complex128 against the oracle: |dT| <= TOL, absolute, point by point


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
    """One batched launch over `energies` against the oracle's T: |dT| <= tol absolute in complex128,
    tol (1 + max |T|) in the single-precision modes."""
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
    bar = tol if precision == ca.Precision.Double else tol * (1 + np.abs(tref).max())
    say(f"[gpu] {label}: max |dT| {d.max():.3e} (bar {bar:.1e})")
    assert d.max() <= bar, f"{label}: max |dT| {d.max():.3e}\n{t}\n{tref}"
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


# -- multi-terminal devices ---------------------------------------------------------------------

SUBS = {ca.Model.BHZ: 1, ca.Model.Square: 1, ca.Model.Graphene: 2, ca.Model.Cubic: 1}
ORBITALS = {ca.Model.BHZ: 4, ca.Model.Square: 1, ca.Model.Graphene: 1, ca.Model.Cubic: 1}


def cross(model, nx, ny, arm):
    """An nx x ny block with arms of length `arm` on its four edges, as wide as the edges, and a lead
    at the end of each arm (the geometry of crates/caroli-green/tests/nested_multi.rs:45-62):
    (sites as (x, y, z, sublattice), interfaces as (slice sites, outward step))."""
    ns = SUBS[model]
    cells = [(x, y) for x in range(-arm, nx + arm) for y in range(-arm, ny + arm) if 0 <= y < ny or 0 <= x < nx]
    sites = [(x, y, 0, s) for (x, y) in cells for s in range(ns)]
    column = lambda x: [(x, y, 0, s) for y in range(ny) for s in range(ns)]
    row = lambda y: [(x, y, 0, s) for x in range(nx) for s in range(ns)]
    interfaces = [(column(-arm), (-1, 0, 0)), (row(ny + arm - 1), (0, 1, 0)), (column(nx + arm - 1), (1, 0, 0)),
                  (row(-arm), (0, -1, 0))]
    return sites, interfaces


def multi_system(model, sites, interfaces, w=0.0):
    """NestedMultiTerminal of a preset model on `sites` with the hashed disorder of strength w."""
    pot = hash_disorder([s[:3] for s in sites], [s[3] for s in sites], w) if w else None
    return ca.NestedMultiTerminal.from_model(model, sites, interfaces, site_potential=pot)


def dense_hamiltonian(arr):
    n = arr["n"]
    h = np.zeros((n, n), np.complex128)
    rp, cl, vv = arr["row_ptr"], arr["cols"], arr["vals"]
    for r in range(n):
        h[r, cl[rp[r]:rp[r + 1]]] = vv[rp[r]:rp[r + 1]]
    return h


def dense_multi(sys_, sigma, energy):
    """T[p][q] = Tr[Gamma_p G_pq Gamma_q G_pq^H] from the whole device's Green's function
    G = (E - H - sum_p Sigma_p)^-1 by a dense inverse (MultiTerminal::solve, multi.rs:89-110), with
    the self-energies of `sigma`."""
    arr = sys_.arrays()
    n = arr["n"]
    a = energy * np.eye(n) - dense_hamiltonian(arr)
    leads = [sys_.lead(p)["orbitals"] for p in range(sys_.lead_count())]
    sig = [sigma.sigma(p) for p in range(len(leads))]
    for o, s in zip(leads, sig):
        a[np.ix_(o, o)] -= s
    g = np.linalg.inv(a)
    gam = [1j * (s - s.conj().T) for s in sig]
    P = len(leads)
    t = np.zeros((P, P))
    for p in range(P):
        for q in range(P):
            gpq = g[np.ix_(leads[p], leads[q])]
            t[p, q] = np.trace(gam[p] @ gpq @ gam[q] @ gpq.conj().T).real
    return t


def decimation_surface(h00, h01, z, tol=1e-15, iters=400):
    """Surface Green's function of semi-infinite cells h00 coupled outward by h01 at complex z, by
    the decimation of Lopez Sancho, Lopez Sancho and Rubio (J. Phys. F 15, 851 (1985))."""
    n = h00.shape[0]
    eye = np.eye(n)
    es, e, a, b = h00.copy(), h00.copy(), h01.copy(), h01.conj().T.copy()
    for _ in range(iters):
        g = np.linalg.inv(z * eye - e)
        agb, bga = a @ g @ b, b @ g @ a
        es = es + agb
        e = e + agb + bga
        a, b = a @ g @ a, b @ g @ b
        if max(np.abs(a).max(), np.abs(b).max()) < tol:
            break
    return np.linalg.inv(z * eye - es)
