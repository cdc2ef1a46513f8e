"""Writes the small fixtures of tests/golden/ from the reference's own data files.

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


def main(src=SRC):
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
