"""GPU front update (Field::schur, crates/caroli-dense/src/field.rs:44-105) against numpy.

S = F22 - F21 F11^-1 F12 for random fronts of the sizes the trees produce, through every kernel
path: LDS-resident fronts (n <= 64), the separator LU in LDS with the triangular-solve tiles
(ne <= 64), the panel-by-panel border kernel, the blocked path (panel, row and update launches per
64 columns; the default above 128 separator variables) and the unblocked fallback; then the MFMA
Schur tiles (32 x 32 or 64 x 64). Complex f64 to a relative 1e-11, complex f32 to 1e-4; a singular leading block is refused
(field.rs:199-208).
"""
import ctypes as C
import os

import numpy as np
import pytest

from caroli_amd import _capi as capi

pytestmark = pytest.mark.gpu

SIZES = [(12, 5, 5), (60, 20, 20), (64, 64, 64), (90, 40, 40), (100, 37, 37), (200, 64, 64), (384, 128, 128),
         (60, 40, 0), (300, 256, 0), (700, 300, 300), (530, 200, 60), (260, 260, 260)]


def front(n, seed):
    """A front like field.rs:183-191: smooth entries plus a dominant diagonal."""
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return np.asfortranarray(f + 3.0 * np.sqrt(n) * np.eye(n))


def gpu_schur(f, ne, ne1, precision):
    n = f.shape[0]
    s = np.zeros((n - ne, n - ne), np.complex128, order="F")
    st = C.c_int32(-1)
    fd = np.asfortranarray(f, np.complex128)
    rc = capi.cg_front_schur(fd.ctypes.data_as(C.POINTER(C.c_double)), n, ne, ne1, precision,
                             s.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st))
    assert rc == 0
    return s, st.value


def numpy_schur(f, ne):
    return f[ne:, ne:] - f[ne:, :ne] @ np.linalg.solve(f[:ne, :ne], f[:ne, ne:])


@pytest.mark.parametrize("kernel", ["default", "border", "blocked", "unblocked"])
@pytest.mark.parametrize("n,ne,ne1", SIZES)
def test_front_schur_double(n, ne, ne1, kernel, monkeypatch):
    monkeypatch.setenv("CAROLI_FRONT_KERNEL", kernel)
    f = front(n, n + ne)
    print(f"[gpu] front n={n} ne={ne} ne1={ne1} {kernel}", flush=True)
    s, st = gpu_schur(f, ne, ne1, 0)
    assert st == 0
    if ne == n:
        return  # nothing left: only the pivot test ran
    ref = numpy_schur(f, ne)
    assert np.abs(s - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("tile", ["32", "64"])
@pytest.mark.parametrize("kernel", ["default", "blocked"])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n,ne", [(90, 40), (100, 37), (200, 64), (384, 128), (530, 200), (700, 300)])
def test_front_schur_tiles(n, ne, precision, kernel, tile, monkeypatch):
    """Both Schur tile shapes (32 x 32 with one wave, 64 x 64 with four) on Schur blocks of 50 to
    400 rows, whole and partial tiles."""
    monkeypatch.setenv("CAROLI_FRONT_KERNEL", kernel)
    monkeypatch.setenv("CAROLI_SCHUR_TILE", tile)
    f = front(n, 5 * n + ne)
    print(f"[gpu] front tile {tile} n={n} ne={ne} precision {precision} {kernel}", flush=True)
    s, st = gpu_schur(f, ne, ne, precision)
    assert st == 0
    ref = numpy_schur(f, ne)
    assert np.abs(s - ref).max() <= (1e-11 if precision == 0 else 1e-4) * np.abs(ref).max()


@pytest.mark.parametrize("n,ne", [(12, 5), (90, 40), (200, 64), (384, 128), (500, 200)])
def test_front_schur_single(n, ne):
    f = front(n, 7 * n + ne)
    print(f"[gpu] front c64 n={n} ne={ne}", flush=True)
    s, st = gpu_schur(f, ne, ne, 1)
    assert st == 0
    ref = numpy_schur(f, ne)
    assert np.abs(s - ref).max() <= 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("n,ne", [(10, 4), (120, 70), (150, 40), (400, 200)])
@pytest.mark.parametrize("precision", [0, 1])
def test_singular_block_is_refused(n, ne, precision):
    f = front(n, 3)
    f[1, :] = f[0, :]  # two equal rows in the leading block
    print(f"[gpu] singular front n={n} ne={ne} precision {precision}", flush=True)
    _, st = gpu_schur(f, ne, ne, precision)
    assert st == 5


@pytest.mark.parametrize("small", ["reg", "lds"])
def test_small_front_kernels_on_systems(small, monkeypatch):
    """Both small-front kernels (in LDS, the default, and register-resident) on whole systems
    against the oracle: delayed separators at E = 0 of the clean lattice, both BHZ spin blocks,
    complex f32 fronts."""
    import caroli_amd as ca
    from helpers import compare, hashed

    monkeypatch.setenv("CAROLI_SMALL_FRONTS", small)
    compare(ca.NestedTwoTerminal.preset(ca.Model.Square, [18, 9, 0]), [-1.3, 0.0, 0.4], f"{small} square 18x9 clean")
    compare(hashed(ca.Model.BHZ, [15, 9, 0], 0.6), [-0.1, 0.02], f"{small} BHZ 15x9 disordered")
    compare(hashed(ca.Model.Graphene, [13, 7, 0], 1.0, subs_per_cell=2), [-0.6, 0.3], f"{small} graphene 13x7")
    compare(hashed(ca.Model.Square, [41, 29, 0], 2.0), [-1.7, 0.3], f"{small} square 41x29 single",
            ca.Precision.Single, tol=1e-4)
