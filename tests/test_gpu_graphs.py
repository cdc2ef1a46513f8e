"""Sweeps replayed from HIP graphs, through the C ABI (cg_batch_create / cg_batch_run).

A resident batch runs its first sweep of every launch unit eagerly, captures the second into a
HIP graph and replays it afterwards. Every replay must give the eager result bit for bit (the same
kernels on the same inputs, in the same order) and the oracle's to 1e-10, with eager work of other
batches, which can move the lanes' buffers and force a new capture, and host synchronisations
between the replays; CAROLI_GRAPHS=0 keeps every sweep eager.
"""
import numpy as np
import pytest

import caroli_amd as ca
from helpers import TOL, hashed, oracle_of, say

pytestmark = pytest.mark.gpu


def oracle_ts(systems, sigmas):
    ref = []
    for s, g in zip(systems, sigmas):
        t, rc = oracle_of(s).transmission_with(g.energy(), [g.block(b)[:2] for b in range(s.block_count())])
        assert rc == 0
        ref.append(t)
    return np.array(ref)


@pytest.mark.parametrize("graphs", ["1", "0"])
def test_replayed_sweeps_match_eager(graphs, monkeypatch):
    monkeypatch.setenv("CAROLI_GRAPHS", graphs)
    systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [30, 14, 0], disorder=2, strength=0.4, concentration=0.2,
                                           seed=1 + c) for c in range(2)]
    energies = [-0.12, 0.0, 0.09]
    sig = [systems[0].self_energies(e) for e in energies]
    pts_s = [s for s in systems for _ in energies]
    pts_g = [g for _ in systems for g in sig]
    ref = oracle_ts(pts_s, pts_g)
    batch = ca.ResidentBatch(pts_s, pts_g, ca.Precision.Double)
    other = hashed(ca.Model.Square, [44, 31, 0], 2.0)  # eager work between the replays
    other_sig = [other.self_energies(e) for e in (-1.1, 0.6)]
    first = None
    for run in range(5):
        say(f"[gpu] graphs={graphs} resident BHZ 30x14 batch, run {run}")
        t, st = batch.run()
        assert (st == 0).all(), st
        assert np.abs(t - ref).max() <= TOL, (run, np.abs(t - ref).max())
        if first is None:
            first = t.copy()
        else:
            assert np.array_equal(t, first), (run, np.abs(t - first).max())
        if run % 2 == 1:
            to, so = ca.batch_transmission([other] * 2, other_sig, ca.Precision.Double)
            assert (so == 0).all()
