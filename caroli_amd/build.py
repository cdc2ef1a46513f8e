"""Builds the engine library `caroli_amd/libcaroli_amd.so` for gfx950 from the sources in csrc/.

Host C++ is compiled with g++, device code with hipcc (gfx950 only), then linked by hipcc.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcaroli_amd.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOST = ["model.cpp", "symbolic.cpp", "dense.cpp", "modes.cpp", "capi.cpp"]
DEVICE = ["fronts.hip", "solve.hip", "engine.hip", "leads.hip"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
             # packed FP32 ops off (compiler-generated v_pk_* FP32 forms are avoided)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + cmd[-1])
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for f in HOST:
        jobs.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, f), "-o", os.path.join(OBJ, f + ".o")])
    for f in DEVICE:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-x", "hip", os.path.join(CSRC, f), "-o", os.path.join(OBJ, f + ".o")])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, [j for j in jobs]))
    objs = [os.path.join(OBJ, f + ".o") for f in HOST + DEVICE]
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-L/opt/rocm/lib", "-lrocsolver",
          "-lrocblas", "-Wl,-rpath,/opt/rocm/lib", "-lpthread"])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
