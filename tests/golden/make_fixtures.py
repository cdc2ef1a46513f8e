"""Writes the small fixtures of tests/golden/ from the reference's own data files.

ti_cond2t_magnetic.json: the thesis two-terminal conductance of the BHZ ribbon with magnetic
impurities, three cases (reference/baseline/golden/ti-cond2t-{same-g,isotropic,opposite-g}),
read the same way; cond(Nu, Ne) column-major with Nu = 3 exchange values and Ne = 5 energies.

ti_cond2t_clean.json: the thesis two-terminal conductance of the clean BHZ ribbon 16x10
(reference/baseline/golden/ti-cond2t-clean/{axes_00,cond_00}.dat: Fortran sequential unformatted,
int32 record markers, read as crates/caroli-validation/src/fortran.rs:9-42 does; the axes hold
the 3 values of U0 then the 13 energies, cond(Nu, Ne) is column-major).
"""
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/reference/baseline/golden"


def records(path):
    data = open(path, "rb").read()
    out, i = [], 0
    while i < len(data):
        n = struct.unpack("<i", data[i:i + 4])[0]
        payload = data[i + 4:i + 4 + n]
        end = struct.unpack("<i", data[i + 4 + n:i + 8 + n])[0]
        assert end == n, "not a sequential unformatted file"
        out.append(struct.unpack(f"<{n // 8}d", payload))
        i += 8 + n
    return out


def magnetic(src, case):
    """ti-cond2t-{same-g,isotropic,opposite-g}: U0 (the exchange J) x 5 energies of the 16 x 10
    ribbon with 24 magnetic impurities (crates/caroli-validation/tests/ti_conductance_disorder.rs)."""
    axes = records(os.path.join(src, case, "axes_00.dat"))[0]
    cond = records(os.path.join(src, case, "cond_00.dat"))[0]
    nu, ne = 3, 5
    assert len(axes) == nu + ne and len(cond) == nu * ne
    return {"exchange": list(axes[:nu]), "energies": list(axes[nu:]),
            "conductance": [[cond[iu + nu * ie] for ie in range(ne)] for iu in range(nu)]}


def main(src=SRC):
    # The three magnetic cases (reference/baseline/cases/ti-cond2t-*/ti.ini): Nx = 16, Ny = 10,
    # Nx_imp = 8, conc = 0.3, theta_0 = 0, eta = 1e-5, graine = 12345; spin = 1 (moments along z)
    # or 0 (isotropic); same g (ge = gh) or opposite g.
    mag = {"source": "reference/baseline/golden/ti-cond2t-{same-g,isotropic,opposite-g}",
           "system": "BHZ 18 x 10 (16 sample slices and one lead slice at each end), axis 0; impurity region "
                     "x in [5, 13), all y, x fastest; int(0.3 * 8 * 10) = 24 impurities from "
                     "Xoshiro256::gfortran(12345); leads at E + 1e-5 i; thesis formula Tr[Re(G_L) G Re(G_R) G^H]",
           "tolerance": 1e-9,
           "cases": {"same-g": dict(moments="cone", coupling="same", **magnetic(src, "ti-cond2t-same-g")),
                     "isotropic": dict(moments="isotropic", coupling="same", **magnetic(src, "ti-cond2t-isotropic")),
                     "opposite-g": dict(moments="cone", coupling="opposite",
                                        **magnetic(src, "ti-cond2t-opposite-g"))}}
    with open(os.path.join(HERE, "ti_cond2t_magnetic.json"), "w") as f:
        json.dump(mag, f, indent=1)
    axes = records(os.path.join(src, "ti-cond2t-clean", "axes_00.dat"))[0]
    cond = records(os.path.join(src, "ti-cond2t-clean", "cond_00.dat"))[0]
    nu, ne = 3, 13
    assert len(axes) == nu + ne and len(cond) == nu * ne
    fixture = {"source": "reference/baseline/golden/ti-cond2t-clean", "system": "BHZ 16x10 clean, axis 0",
               "energies": list(axes[nu:]), "conductance": [cond[ie * nu] for ie in range(ne)],
               "tolerance_vs_exact": 0.025}
    with open(os.path.join(HERE, "ti_cond2t_clean.json"), "w") as f:
        json.dump(fixture, f, indent=1)


if __name__ == "__main__":
    main(*sys.argv[1:])
