"""Builds the oracle (test infrastructure only): oracle/_ref/liboracle.so from oracle.cpp."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref", "liboracle.so")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    subprocess.run(["g++", "-O2", "-march=x86-64-v2", "-fPIC", "-shared", "-std=c++17", "-pthread",
                    os.path.join(HERE, "oracle.cpp"), "-o", tmp], check=True)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
