"""Lead self-energies on the GPU (the scalar-coupling "cell" case of modes.rs:621-710) against the
host mode solver, and the uneven cubic batch at its large sizes against committed oracle values.

- Sigma_L, Sigma_R from the GPU (Y diag(d) Y^H on FP64 MFMA) equal the host's to 1e-12 of their
  size on cubic and square leads; Gamma = W W^H with W from the propagating eigenvectors; the
  channel counts agree; T through either self-energies agrees to 1e-12.
- Cells of 128 orbitals or more take the eigendecomposition on the GPU too (rocSOLVER zheevj):
  Sigma and T against a twin system whose leads stay on the host, to 1e-10.
- Cubic Anderson model W = 16.5, 2L x L x L at L = 32, 40, 48 (separators up to 2,304 variables,
  roots of 2L^2 = 4,608 kept orbitals): GPU complex128 T against the double oracle values in
  tests/golden/cubic_large.json (written on the CPU by tests/golden/make_transmission_fixtures.py,
  self-energies from the host mode solver), to 1e-10 absolute.
"""
import json
import os
import time

import numpy as np
import pytest

import caroli_amd as ca
from helpers import say

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def host_sigma(s, e):
    os.environ["CAROLI_SIGMA_HOST"] = "1"
    try:
        return s.self_energies(e)
    finally:
        del os.environ["CAROLI_SIGMA_HOST"]


@pytest.mark.parametrize("model,ext,energies", [
    (ca.Model.Cubic, [12, 8, 8], [-1.0, 0.37, 2.2]),
    (ca.Model.Cubic, [20, 10, 10], [-3.3, 0.0, 1.0 / 3.0]),
    (ca.Model.Square, [30, 40, 0], [-2.9, -0.4, 1.3]),
])
def test_gpu_cell_self_energies_match_host(model, ext, energies):
    s = ca.NestedTwoTerminal.preset(model, ext, disorder=1, strength=2.0, seed=7)
    for e in energies:
        g_gpu, g_host = s.self_energies(e), host_sigma(s, e)
        for b in range(s.block_count()):
            sl1, sr1, cl1, cr1 = g_gpu.block(b)
            sl2, sr2, cl2, cr2 = g_host.block(b)
            scale = max(np.abs(sl2).max(), np.abs(sr2).max())
            err = max(np.abs(sl1 - sl2).max(), np.abs(sr1 - sr2).max()) / scale
            say(f"[gpu] cell Sigma {model.name} {ext} E={e}: channels {cl1}/{cr1}, max |dSigma| / |Sigma| {err:.1e}")
            assert (cl1, cr1) == (cl2, cr2)
            assert err <= 1e-12
        t1 = s.transmission_with(g_gpu)
        t2 = s.transmission_with(g_host)
        assert abs(t1 - t2) <= 1e-12 * (1 + abs(t2)), (t1, t2)


@pytest.mark.parametrize("ext,energies", [([24, 12, 12], [-1.7, 0.45]), ([32, 16, 16], [0.0, 2.5])])
def test_gpu_cell_eigendecomposition_matches_host(ext, energies):
    """Cells of 144 and 256 orbitals take the GPU eigensolver (rocSOLVER zheevj); a twin system
    built under CAROLI_SIGMA_HOST=1 takes the host's (Householder + implicit QL) and computes Sigma
    on the host. The clean cubic cell has degenerate eigenvalues, so the two eigenvector bases
    differ; Sigma, Gamma, the channel counts and T do not."""
    s_gpu = ca.NestedTwoTerminal.preset(ca.Model.Cubic, ext, disorder=1, strength=2.0, seed=3)
    s_host = ca.NestedTwoTerminal.preset(ca.Model.Cubic, ext, disorder=1, strength=2.0, seed=3)
    for e in energies:
        t0 = time.perf_counter()
        g_gpu = s_gpu.self_energies(e)
        t_gpu = time.perf_counter() - t0
        g_host = host_sigma(s_host, e)
        sl1, sr1, cl1, cr1 = g_gpu.block(0)
        sl2, sr2, cl2, cr2 = g_host.block(0)
        scale = max(np.abs(sl2).max(), np.abs(sr2).max())
        err = max(np.abs(sl1 - sl2).max(), np.abs(sr1 - sr2).max()) / scale
        t1, t2 = s_gpu.transmission_with(g_gpu), s_host.transmission_with(g_host)
        say(f"[gpu] cell eigensolver cubic {ext} E={e}: channels {cl1}/{cr1}, |dSigma| / |Sigma| {err:.1e}, "
            f"|dT| {abs(t1 - t2):.1e}, GPU Sigma {t_gpu:.3f} s")
        assert (cl1, cr1) == (cl2, cr2)
        assert err <= 1e-10
        assert abs(t1 - t2) <= 1e-10 * (1 + abs(t2)), (t1, t2)


def test_gpu_cell_gamma_factor():
    """W W^H = Gamma = i (Sigma - Sigma^H) for the W the GPU path builds from the eigenvectors."""
    s = ca.NestedTwoTerminal.preset(ca.Model.Cubic, [16, 8, 8])
    g = s.self_energies(0.6)
    sl, sr, cl, cr = g.block(0)
    gam = 1j * (sr - sr.conj().T)
    ev = np.linalg.eigvalsh(gam)
    assert (ev > 1e-9 * ev.max()).sum() == cr  # rank of Gamma = channels
    t = s.transmission_with(g)
    assert abs(t - round(t)) < 1e-9  # clean cubic wire: an integer number of channels


def test_cubic_large_sizes_against_fixture():
    fx = json.load(open(os.path.join(HERE, "golden", "cubic_large.json")))
    systems = {}
    for s_idx, L, e, t_ref in fx["points"]:
        key = (int(s_idx), int(L))
        if key not in systems:  # one system (and one cell eigendecomposition) for its energies
            systems[key] = ca.NestedTwoTerminal.preset(ca.Model.Cubic, [2 * key[1], key[1], key[1]], disorder=1,
                                                       strength=16.5, seed=1000 + key[0])
        s = systems[key]
        t0 = time.perf_counter()
        g = s.self_energies(e)
        ts = time.perf_counter() - t0
        say(f"[gpu] cubic system {int(s_idx)} L = {int(L)} E = {e:+.4f}: self-energies {ts:.2f} s")
        t0 = time.perf_counter()
        t, st = ca.batch_transmission([s], [g], ca.Precision.Double)
        say(f"    T = {t[0]:.15f}, oracle {t_ref:.15f}, |dT| {abs(t[0] - t_ref):.1e} "
            f"({time.perf_counter() - t0:.2f} s)")
        assert st[0] == 0
        assert abs(t[0] - t_ref) <= 1e-10
