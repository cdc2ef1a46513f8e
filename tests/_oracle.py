"""ctypes binding of the oracle (oracle/oracle.cpp): test infrastructure only.

The oracle is the checker of the engine: a CPU FP64 restatement of the reference's
nested-dissection path. Only tests, `__graft_entry__.smoke()` and the benchmark's CPU-baseline leg
use it.
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "liboracle.so")
LIB_LD = os.path.join(ROOT, "oracle", "_ref", "liboracle_ld.so")
LIB_FMA = os.path.join(ROOT, "oracle", "_ref", "liboracle_fma.so")


def _load(path=LIB):
    if not os.path.exists(path):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import build as oracle_build  # noqa: E402

        oracle_build.build()
    lib = C.CDLL(path)
    P = C.POINTER
    lib.oracle_create.restype = C.c_void_p
    lib.oracle_create.argtypes = [C.c_uint64, P(C.c_int64), P(C.c_uint32), P(C.c_double), C.c_uint64, P(C.c_int64),
                                  P(C.c_int32), P(C.c_uint64), C.c_int32, P(C.c_int64), C.c_int32, P(C.c_double),
                                  P(C.c_double), C.c_uint64, P(C.c_double), P(C.c_double), C.c_uint64]
    lib.oracle_destroy.argtypes = [C.c_void_p]
    lib.oracle_block_count.argtypes = [C.c_void_p]
    lib.oracle_block_info.argtypes = [C.c_void_p, C.c_int32, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64),
                                      P(C.c_uint64), P(C.c_uint64)]
    lib.oracle_node.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64),
                                P(C.c_uint64)]
    lib.oracle_transmission_with.argtypes = [C.c_void_p, C.c_double, P(C.c_double), P(C.c_double)]
    lib.oracle_transmission_thesis.argtypes = [C.c_void_p, C.c_double, P(C.c_double), P(C.c_double)]
    lib.oracle_green_columns.argtypes = [C.c_void_p, C.c_int32, C.c_double, P(C.c_double), P(C.c_double), C.c_int32,
                                         P(C.c_double)]
    lib.oracle_transmissions_parallel.argtypes = [P(C.c_void_p), P(C.c_double), P(P(C.c_double)), C.c_int64,
                                                  C.c_int32, P(C.c_double)]
    return lib


lib = _load()
_lib_ld = None
_lib_fma = None


def _ld():
    """The long-double build of the oracle (loaded on first use)."""
    global _lib_ld
    if _lib_ld is None:
        _lib_ld = _load(LIB_LD)
    return _lib_ld


def _fma():
    """The double build with contracted multiply-adds (loaded on first use; needs an FMA host)."""
    global _lib_fma
    if _lib_fma is None:
        _lib_fma = _load(LIB_FMA)
    return _lib_fma


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleSystem:
    """The oracle's copy of a system, built from the arrays of the engine's system (or any arrays)."""

    def __init__(self, arrays, transpose=False, extended=False, fma=False):
        """extended: the long-double build (64-bit mantissa) instead of the double one; fma: the
        double build with contracted multiply-adds (another rounding order of the same steps)."""
        self.lib = _ld() if extended else (_fma() if fma else lib)
        a = dict(arrays)
        vals = np.ascontiguousarray(a["vals"], np.complex128)
        mats = [np.asfortranarray(a[k], np.complex128) for k in ("left_h00", "left_h01", "right_h00", "right_h01")]
        if transpose:
            # A^T = z - H^T - Sigma^T: H^T = conj(H) for a Hermitian H (same pattern); leads transposed.
            vals = np.conj(vals)
            mats = [np.asfortranarray(m.T) for m in mats]
        self.n = int(a["n"])
        self._keep = [np.ascontiguousarray(a["row_ptr"], np.int64), np.ascontiguousarray(a["cols"], np.uint32), vals,
                      np.ascontiguousarray(a["cells"], np.int64).reshape(-1),
                      np.ascontiguousarray(a["site_orbitals"], np.int32),
                      np.ascontiguousarray(a["site_offsets"], np.uint64), np.ascontiguousarray(a["range"], np.int64),
                      *mats]
        k = self._keep
        self.h = self.lib.oracle_create(self.n, k[0].ctypes.data_as(C.POINTER(C.c_int64)),
                                   k[1].ctypes.data_as(C.POINTER(C.c_uint32)), _d(k[2]), len(k[4]),
                                   k[3].ctypes.data_as(C.POINTER(C.c_int64)), k[4].ctypes.data_as(C.POINTER(C.c_int32)),
                                   k[5].ctypes.data_as(C.POINTER(C.c_uint64)), int(a["dim"]),
                                   k[6].ctypes.data_as(C.POINTER(C.c_int64)), int(a["axis"]), _d(k[7]), _d(k[8]),
                                   k[7].shape[0], _d(k[9]), _d(k[10]), k[9].shape[0])
        if not self.h:
            raise ValueError("oracle_create refused the arrays")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_destroy(self.h)
            self.h = None

    def block_count(self):
        return self.lib.oracle_block_count(self.h)

    def block(self, b):
        nl, nr, nn = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.lib.oracle_block_info(self.h, b, C.byref(nl), C.byref(nr), C.byref(nn), None, None)
        left = np.zeros(nl.value, np.uint64)
        right = np.zeros(nr.value, np.uint64)
        self.lib.oracle_block_info(self.h, b, C.byref(nl), C.byref(nr), C.byref(nn),
                              left.ctypes.data_as(C.POINTER(C.c_uint64)), right.ctypes.data_as(C.POINTER(C.c_uint64)))
        return left, right, nn.value

    def node(self, b, k):
        ne, nb = C.c_uint64(), C.c_uint64()
        self.lib.oracle_node(self.h, b, k, C.byref(ne), C.byref(nb), None, None)
        e = np.zeros(ne.value, np.uint64)
        d = np.zeros(nb.value, np.uint64)
        self.lib.oracle_node(self.h, b, k, C.byref(ne), C.byref(nb), e.ctypes.data_as(C.POINTER(C.c_uint64)),
                        d.ctypes.data_as(C.POINTER(C.c_uint64)))
        return e, d

    @staticmethod
    def pack(sigmas):
        """Per block (Sigma_L, Sigma_R) -> the flat column-major array the oracle reads."""
        return np.concatenate([np.asfortranarray(m, np.complex128).reshape(-1, order="F")
                               for pair in sigmas for m in pair])

    def transmission_with(self, energy, sigmas):
        flat = self.pack(sigmas)
        t = C.c_double()
        rc = self.lib.oracle_transmission_with(self.h, float(energy), _d(flat), C.byref(t))
        return t.value, rc

    def transmission_thesis(self, energy, sigmas):
        """Tr[Re(Gamma_L) G Re(Gamma_R) G^H] (the thesis formula, thesis.rs:89-95)."""
        flat = self.pack(sigmas)
        t = C.c_double()
        rc = self.lib.oracle_transmission_thesis(self.h, float(energy), _d(flat), C.byref(t))
        return t.value, rc

    def green_columns(self, b, energy, sigma_l, sigma_r, which):
        """Columns of G on the right (which = 1) or left (0) slice of block b, on every orbital."""
        left, right, _ = self.block(b)
        m = len(right) if which else len(left)
        out = np.zeros((self.n, m), np.complex128, order="F")
        sl = np.asfortranarray(sigma_l, np.complex128)
        sr = np.asfortranarray(sigma_r, np.complex128)
        rc = self.lib.oracle_green_columns(self.h, b, float(energy), _d(sl), _d(sr), which, _d(out))
        if rc:
            raise RuntimeError("oracle: singular root")
        return out


def transmissions_parallel(systems, energies, sigma_flats, threads):
    n = len(systems)
    sh = (C.c_void_p * n)(*[s.h for s in systems])
    en = np.ascontiguousarray(energies, np.float64)
    ptrs = (C.POINTER(C.c_double) * n)(*[_d(f) for f in sigma_flats])
    out = np.zeros(n)
    failed = lib.oracle_transmissions_parallel(sh, _d(en), ptrs, n, threads, _d(out))
    return out, failed


def dense_transmission(arrays, energy, sigmas, blocks):
    """Brute force on tiny systems: G = (z - H - Sigma)^-1 densely (nested.rs:1-7), then the
    Caroli trace (crates/caroli-green/src/lib.rs:108-113) per block."""
    n = int(arrays["n"])
    h = np.zeros((n, n), np.complex128)
    rp, cl, vv = arrays["row_ptr"], arrays["cols"], arrays["vals"]
    for r in range(n):
        for e in range(rp[r], rp[r + 1]):
            h[r, cl[e]] += vv[e]
    a = energy * np.eye(n) - h
    for (left, right), (sl, sr) in zip(blocks, sigmas):
        a[np.ix_(left, left)] -= sl
        a[np.ix_(right, right)] -= sr
    g = np.linalg.inv(a)
    t = 0.0
    for (left, right), (sl, sr) in zip(blocks, sigmas):
        gl = 1j * (sl - sl.conj().T)
        gr = 1j * (sr - sr.conj().T)
        glr = g[np.ix_(left, right)]
        t += np.real(np.trace(gl @ glr @ gr @ glr.conj().T))
    return t
