"""cg_front_schur_noalloc (the allocation-free front update) against cg_front_schur.

Both run the same kernels on the same front, so the complex f64 and complex f32 results agree bit
for bit. Several calls in a row share the one static work area: a large front, then smaller ones
(the area keeps the previous call's contents), a front whose separator block is singular (status
5 in both), and a front above the work area's capacity (refused with CG_ERR_CAPACITY = 8 before
any device work).
"""
import ctypes as C

import numpy as np
import pytest

import caroli_amd as ca
from caroli_amd import _capi as capi
from helpers import say

pytestmark = pytest.mark.gpu


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def front(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    f += 4.0 * np.sqrt(n) * np.eye(n)  # well conditioned separator blocks
    return np.asfortranarray(f)


def schur_both(f, ne, precision):
    n = f.shape[0]
    nb = n - ne
    flat = np.ascontiguousarray(f.ravel(order="F")).view(np.float64)
    out = []
    for fn in (capi.cg_front_schur, capi.cg_front_schur_noalloc):
        s = np.zeros(nb * nb, np.complex128)
        st = C.c_int32(-1)
        rc = fn(_d(flat), n, ne, ne, int(precision), _d(s.view(np.float64)), C.byref(st))
        assert rc == 0, (fn.__name__, rc)
        out.append((s.reshape((nb, nb), order="F"), st.value))
    return out


@pytest.mark.parametrize("precision", [ca.Precision.Double, ca.Precision.Single])
def test_noalloc_front_schur_matches_front_schur(precision):
    # A blocked front (separator 200 > 128), then a smaller split one, then an LDS-sized one.
    for n, ne, seed in ((700, 200, 1), (300, 48, 2), (40, 12, 3)):
        f = front(n, seed)
        (s1, st1), (s2, st2) = schur_both(f, ne, precision)
        ref = f[ne:, ne:] - f[ne:, :ne] @ np.linalg.solve(f[:ne, :ne], f[:ne, ne:])
        err = np.abs(s2 - ref).max() / np.abs(ref).max()
        say(f"[gpu] noalloc front n={n} ne={ne} precision {int(precision)}: status {st1}/{st2}, "
            f"max |S_noalloc - S| {np.abs(s2 - s1).max():.1e}, relative error vs numpy {err:.1e}")
        assert st1 == 0 and st2 == 0
        assert np.array_equal(s1, s2)
        assert err <= (1e-12 if precision == ca.Precision.Double else 1e-4)


def test_noalloc_front_schur_refuses_and_delays():
    # Two equal rows in the separator block: refused in both (field.rs:199-208), status 5.
    f = front(120, 4)
    f[1, :] = f[0, :]
    (s1, st1), (s2, st2) = schur_both(f, 80, ca.Precision.Double)
    assert st1 == 5 and st2 == 5
    # Above the work area: CG_ERR_CAPACITY before any device work.
    n = 1025
    big = np.zeros(2 * n * n)
    st = C.c_int32(-1)
    out = np.zeros(2 * 24 * 24)
    assert capi.cg_front_schur_noalloc(_d(big), n, n - 24, n - 24, 0, _d(out), C.byref(st)) == 8
    # A smaller call after the refusal still works and matches.
    f = front(96, 5)
    (s1, st1), (s2, st2) = schur_both(f, 32, ca.Precision.Double)
    assert st1 == 0 and st2 == 0 and np.array_equal(s1, s2)
