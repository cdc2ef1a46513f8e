"""Writes the transmission fixtures of the benchmark workloads from the oracle (run on the CPU).

The GPU tests compare the engine against these values at the workloads' full sizes, where the
oracle would take minutes to hours per point on the GPU box's host.

bhz200x64_magnetic.json (equal-size batch): BHZ 200 x 64 sites, polarized magnetic impurities
(concentration 0.2, J = 0.4, opposite g), configurations seeded 1 + c for c = 0..7 (the CLI's
default seeding, crates/caroli-cli/src/build.rs:281), 33 energies linspace(-0.16, 0.16). For
every point: T from the double oracle, from the long-double oracle (the same elimination with a
64-bit mantissa) and from the double oracle built with contracted multiply-adds (the same
elimination in double precision, rounded in another order), all with the engine's host
self-energies. The two double values, against the long-double one, sample the rounding error of a
double-precision nested dissection on that point.

cubic_large.json (uneven batch): cubic Anderson model W = 16.5, 2L x L x L, the systems s of the
512-system list (L_s = [16, 24, 32, 40, 48][floor(5 u_s)], u from xoshiro256** seeded 5) with
seed 1000 + s; for the first system of each L in `--cubic-l`, T at the energies `--cubic-e` from
the double oracle.

Usage: python tests/golden/make_transmission_fixtures.py [bhz|cubic] [--jobs N]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

BHZ_ENERGIES = np.linspace(-0.16, 0.16, 33)
CUBIC_ENERGIES = np.linspace(-1.0, 1.0, 4)


def bhz_system(c):
    import caroli_amd as ca

    return ca.NestedTwoTerminal.preset(ca.Model.BHZ, [200, 64, 0], disorder=2, strength=0.4, concentration=0.2,
                                       seed=1 + c)


def cubic_sizes():
    import caroli_amd as ca

    u = np.zeros(512)
    ca.capi.cg_uniforms(0, 5, None, 512, u.ctypes.data_as(C.POINTER(C.c_double)))
    return [[16, 24, 32, 40, 48][int(5 * x)] for x in u]


def cubic_system(s, L):
    import caroli_amd as ca

    return ca.NestedTwoTerminal.preset(ca.Model.Cubic, [2 * L, L, L], disorder=1, strength=16.5, seed=1000 + s)


def bhz_config(c):
    from _oracle import OracleSystem

    s = bhz_system(c)
    arr = s.arrays()
    o, oe, of = OracleSystem(arr), OracleSystem(arr, extended=True), OracleSystem(arr, fma=True)
    rows = []
    for e in BHZ_ENERGIES:
        sig = s.self_energies(float(e))
        blocks = [sig.block(b)[:2] for b in range(s.block_count())]
        td, rc = o.transmission_with(float(e), blocks)
        tl, rc2 = oe.transmission_with(float(e), blocks)
        tf, rc3 = of.transmission_with(float(e), blocks)
        assert rc == 0 and rc2 == 0 and rc3 == 0
        rows.append((float(e), td, tl, tf))
        print(f"bhz c={c} E={e:+.4f} T={td:.15f} |T - T_long|={abs(td - tl):.2e} (fma {abs(tf - tl):.2e})",
              flush=True)
    return c, rows


def cubic_point(args):
    s, L, e = args
    from _oracle import OracleSystem

    sysm = cubic_system(s, L)
    o = OracleSystem(sysm.arrays())
    sig = sysm.self_energies(float(e))
    t0 = time.time()
    t, rc = o.transmission_with(float(e), [sig.block(b)[:2] for b in range(sysm.block_count())])
    assert rc == 0
    print(f"cubic s={s} L={L} E={e:+.4f} T={t:.15f} ({time.time() - t0:.0f} s)", flush=True)
    return s, L, float(e), t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["bhz", "cubic"])
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--cubic-l", type=int, nargs="*", default=[32, 40, 48])
    ap.add_argument("--cubic-e", type=float, nargs="*", default=[-1.0, 1.0 / 3.0])
    a = ap.parse_args()
    if a.which == "bhz":
        with ProcessPoolExecutor(a.jobs) as ex:
            out = dict(ex.map(bhz_config, range(8)))
        fixture = {"workload": "BHZ 200x64 magnetic c=0.2 J=0.4 opposite-g, configurations seeded 1+c, "
                               "33 energies linspace(-0.16, 0.16)",
                   "columns": ["configuration", "energy", "T_double_oracle", "T_long_double_oracle",
                               "T_double_fma_oracle"],
                   "points": [[c, e, td, tl, tf] for c in range(8) for (e, td, tl, tf) in out[c]]}
        path = os.path.join(HERE, "bhz200x64_magnetic.json")
    else:
        sizes = cubic_sizes()
        tasks = []
        for L in a.cubic_l:
            s = sizes.index(L)
            tasks += [(s, L, e) for e in a.cubic_e]
        with ProcessPoolExecutor(a.jobs) as ex:
            res = list(ex.map(cubic_point, tasks))
        fixture = {"workload": "cubic Anderson W=16.5, 2L x L x L, system s of the 512-system list seeded 1000+s",
                   "columns": ["system", "L", "energy", "T_double_oracle"],
                   "points": [list(r) for r in res]}
        path = os.path.join(HERE, "cubic_large.json")
    with open(path, "w") as f:
        json.dump(fixture, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
