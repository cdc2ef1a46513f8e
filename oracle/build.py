"""Builds the oracle (test infrastructure only) from oracle.cpp:
oracle/_ref/liboracle.so (complex double) and oracle/_ref/liboracle_ld.so (complex long double,
the same elimination with a 64-bit mantissa, for rounding-floor measurements)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref", "liboracle.so")
OUT_LD = os.path.join(HERE, "_ref", "liboracle_ld.so")


def _compile(out, defines):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    subprocess.run(["g++", "-O2", "-march=x86-64-v2", "-fPIC", "-shared", "-std=c++17", "-pthread", *defines,
                    os.path.join(HERE, "oracle.cpp"), "-o", tmp], check=True)
    os.replace(tmp, out)


def build(verbose=False):
    _compile(OUT, [])
    _compile(OUT_LD, ["-DORACLE_REAL=long double"])
    if verbose:
        print("built", OUT, "and", OUT_LD)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
