"""Scattering states through the C ABI (cg_scattering_states): psi_p = G W_p on every orbital for
both leads, from the stored factors (nested.rs:851-895).

The states are compared with the oracle's G columns through what does not depend on the choice of
W_p (W W^H = Gamma fixes W only up to a unitary): psi_p psi_p^H = G Gamma_p G^H, seen through the
site densities, the bond currents and (sites of several orbitals) the currents of an on-site
operator (host code, itself pinned in tests/test_host.py), all to 1e-9 relative, on square,
two-block BHZ, graphene and cubic samples. The currents of the GPU states also carry T through
every cut.
"""
import numpy as np
import pytest

import caroli_amd as ca
from helpers import say
from test_host import _oracle_states

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("label,make,energy", [
    ("square 14x7 Anderson", lambda: ca.NestedTwoTerminal.preset(ca.Model.Square, [14, 7, 0], disorder=1,
                                                                   strength=1.0, seed=3), 0.4),
    ("BHZ 20x10 clean (two blocks)", lambda: ca.NestedTwoTerminal.preset(ca.Model.BHZ, [20, 10, 0]), 0.05),
    ("graphene 30x8 Anderson", lambda: ca.NestedTwoTerminal.preset(ca.Model.Graphene, [30, 8, 0], disorder=1,
                                                                     strength=0.5, seed=2), 0.3),
    ("cubic 8x4x4 Anderson", lambda: ca.NestedTwoTerminal.preset(ca.Model.Cubic, [8, 4, 4], disorder=1,
                                                                   strength=3.0, seed=7), 0.5),
])
def test_scattering_states_match_oracle(label, make, energy):
    s = make()
    ref, t, _ = _oracle_states(s, energy)
    say(f"[gpu] scattering states: {label}, E = {energy}, oracle T = {t:.6f}")
    sig = s.self_energies(energy)
    states = s.scattering_states(sig)
    n = s.sizes()[0]
    for p in range(2):
        assert states.leads[p].shape == (n, ref.leads[p].shape[1]), (p, states.leads[p].shape, ref.leads[p].shape)
    for f in ([1.0, 0.0], [0.0, 1.0]):
        d, dr = s.site_densities(states, f), s.site_densities(ref, f)
        assert np.abs(d - dr).max() <= 1e-9 * np.abs(dr).max(), (f, np.abs(d - dr).max())
        _, _, c = s.bond_currents(states, f)
        _, _, cr = s.bond_currents(ref, f)
        assert np.abs(c - cr).max() <= 1e-9 * max(np.abs(cr).max(), 1e-300), (f, np.abs(c - cr).max())
    norb = np.asarray(s.arrays()["site_orbitals"])
    if norb.min() == norb.max() > 1:  # an on-site operator (s_z-like) on every site
        op = np.diag([1.0 if k < norb[0] // 2 else -1.0 for k in range(int(norb[0]))])
        _, _, co = s.operator_currents(states, [1.0, 0.0], op)
        _, _, cor = s.operator_currents(ref, [1.0, 0.0], op)
        assert np.abs(co - cor).max() <= 1e-9 * max(np.abs(cor).max(), 1e-12), np.abs(co - cor).max()
    # Conservation: the GPU states carry T through every cut.
    arr = s.arrays()
    x = np.asarray(arr["cells"]).reshape(-1, 3)[:, 0]
    fr, to, cur = s.bond_currents(states, [1.0, 0.0])
    for c in range(int(x.min()), int(x.max())):
        across = np.where((x[fr] <= c) & (c < x[to]), cur, 0.0) - np.where((x[to] <= c) & (c < x[fr]), cur, 0.0)
        assert abs(across.sum() - t) <= 1e-9 * max(t, 1e-12), (c, across.sum(), t)
