"""Builds the oracle (test infrastructure only) from oracle.cpp:
oracle/_ref/liboracle.so (complex double), oracle/_ref/liboracle_ld.so (complex long double, the
same elimination with a 64-bit mantissa) and oracle/_ref/liboracle_fma.so (complex double with
fused multiply-adds wherever the compiler can contract: a second double-precision evaluation of
the same elimination in another rounding order). The last two serve the rounding-floor
measurements; the FMA build runs only on a host with FMA instructions (x86-64-v3)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref", "liboracle.so")
OUT_LD = os.path.join(HERE, "_ref", "liboracle_ld.so")
OUT_FMA = os.path.join(HERE, "_ref", "liboracle_fma.so")


def _compile(out, flags):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    arch = [] if any(f.startswith("-march") for f in flags) else ["-march=x86-64-v2"]
    subprocess.run(["g++", "-O2", *arch, "-fPIC", "-shared", "-std=c++17", "-pthread", *flags,
                    os.path.join(HERE, "oracle.cpp"), "-o", tmp], check=True)
    os.replace(tmp, out)


def build(verbose=False):
    _compile(OUT, [])
    _compile(OUT_LD, ["-DORACLE_REAL=long double"])
    _compile(OUT_FMA, ["-march=x86-64-v3", "-ffp-contract=fast"])
    if verbose:
        print("built", OUT, OUT_LD, "and", OUT_FMA)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
