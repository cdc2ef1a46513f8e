"""MI355X nested-dissection transmission: the Python face of the engine's C ABI.

Mirrors the API of `NestedTwoTerminal` (crates/caroli-green/src/nested.rs:467-713), its
`LeadSelfEnergies` (nested.rs:439-459), `Precision`/`PrecisionOptions`/`PrecisionReport`
(crates/caroli-green/src/precision.rs:24-134) and `GreenError` (crates/caroli-green/src/lib.rs:35-47),
so that code written against the reference reads the same. `TwoTerminal` (two_terminal.rs:64-193)
is the same engine under the name the reference's README uses.

All numerical work runs in `libcaroli_amd.so`: the fronts, root and trace on the GPU, the symbolic
phase and the FP64 lead self-energies on the host.
"""
import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import _capi as capi

__all__ = [
    "GreenError", "Precision", "PrecisionOptions", "PrecisionReport", "LeadSelfEnergies", "NestedTwoTerminal",
    "TwoTerminal", "Model", "device_available", "batch_transmission", "batch_transmission_checked",
    "backward_batch",
]


class GreenError(RuntimeError):
    """GreenError of crates/caroli-green/src/lib.rs:35-47, by status code."""

    KINDS = {1: "Singular", 2: "Eigen", 3: "ModeCount", 4: "NoConvergence", 6: "Invalid", 7: "Device",
             8: "Capacity"}

    def __init__(self, code, what=""):
        self.code = int(code)
        self.kind = self.KINDS.get(self.code, f"status {code}")
        super().__init__(f"{self.kind}{': ' + what if what else ''}")


def _check(rc, what=""):
    if rc != 0:
        detail = (capi.cg_last_error() or b"").decode(errors="replace") if rc in (2, 6, 7) else ""
        raise GreenError(rc, "; ".join(x for x in (what, detail) if x))


class Precision(enum.IntEnum):
    """Front precision (precision.rs:24-32), plus the FP64-refined mixed mode."""

    Double = 0
    Single = 1
    Mixed = 2


@dataclass
class PrecisionOptions:
    """precision.rs:36-61: one point in `every` is recomputed in double precision and compared;
    the default checks every point (the engine's |dT| report for every point)."""

    precision: Precision = Precision.Mixed
    tolerance: float = 1e-8
    every: int = 1

    @classmethod
    def double(cls):
        """precision.rs:48-50: nothing to check."""
        return cls(Precision.Double, 0.0, 0)

    @classmethod
    def single(cls):
        """precision.rs:54-56: one point in fifty checked, differences up to 1e-2 accepted."""
        return cls(Precision.Single, 1e-2, 50)


@dataclass
class PrecisionReport:
    """precision.rs:65-78, with the per-point differences (NaN where a point was not compared)."""

    precision: Precision
    checked: int
    median: float
    max: float
    failed: int
    fallback: str
    differences: np.ndarray

    @property
    def fell_back(self) -> bool:
        return bool(self.fallback)


def device_available() -> bool:
    return bool(capi.cg_device_available())


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Model(enum.IntEnum):
    """Preset models of the engine's host builders (crates/caroli-model/src/presets.rs,
    crates/caroli-bench/src/main.rs:77-92)."""

    BHZ = 0
    Square = 1
    Graphene = 2
    Cubic = 3


class LeadSelfEnergies:
    """Lead self-energies of one energy, block by block (nested.rs:439-459)."""

    def __init__(self, handle, owner):
        self._h = C.c_void_p(handle)
        self._owner = owner

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            capi.cg_sigma_destroy(self._h)
            self._h = C.c_void_p(None)

    def energy(self) -> float:
        return capi.cg_sigma_energy(self._h)

    def block(self, b):
        """(Sigma_L, Sigma_R, channels_L, channels_R) of block b."""
        nl, nr = self._owner.block_sizes(b)[:2]
        sl = np.zeros((nl, nl), dtype=np.complex128, order="F")
        sr = np.zeros((nr, nr), dtype=np.complex128, order="F")
        cl, cr = C.c_int32(), C.c_int32()
        _check(capi.cg_sigma_block(self._h, b, _dptr(sl), _dptr(sr), C.byref(cl), C.byref(cr)))
        return sl, sr, cl.value, cr.value


class NestedTwoTerminal:
    """A two-terminal device solved by nested dissection (nested.rs:467-471)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            capi.cg_system_destroy(self._h)
            self._h = C.c_void_p(None)

    @classmethod
    def preset(cls, model, extent, axis=0, disorder=0, strength=0.0, concentration=0.0, seed=1, site_potential=None):
        """A preset model on `Region::cuboid(extent)` (see caroli_amd.h, cg_preset_create)."""
        ext = (C.c_uint64 * 3)(*[int(x) for x in (list(extent) + [0, 0, 0])[:3]])
        h = C.c_void_p()
        pot = None
        if site_potential is not None:
            site_potential = np.ascontiguousarray(site_potential, dtype=np.float64)
            pot = _dptr(site_potential)
        _check(capi.cg_preset_create(int(model), ext, axis, disorder, strength, concentration, seed, pot, C.byref(h)),
               "preset")
        return cls(h.value)

    @classmethod
    def bhz_magnetic(cls, extent, seed, x0, nx_imp, concentration, exchange, moments="cone", moment_param=0.0,
                     coupling="opposite", seeding="gfortran"):
        """BHZ ribbon with classical magnetic impurities (cg_bhz_magnetic_create): `moments`
        "isotropic", "cone" (half-angle moment_param, radians) or "polarized" (m_z = moment_param);
        `coupling` "opposite" or "same" g; `seeding` "gfortran" (Xoshiro256::gfortran) or "u64"
        (Xoshiro256::seed_from_u64)."""
        ext = (C.c_uint64 * 2)(int(extent[0]), int(extent[1]))
        h = C.c_void_p()
        _check(capi.cg_bhz_magnetic_create(
            ext, {"u64": 0, "gfortran": 1}[seeding], int(seed), int(x0), int(nx_imp), float(concentration),
            {"isotropic": 0, "cone": 1, "polarized": 2}[moments], float(moment_param),
            {"opposite": 0, "same": 1}[coupling], float(exchange), C.byref(h)), "bhz_magnetic")
        return cls(h.value)

    @classmethod
    def from_arrays(cls, n, row_ptr, cols, vals, cells, site_orbitals, site_offsets, dim, hop_range, axis,
                    left_h00, left_h01, right_h00, right_h01):
        """`from_model` from the arrays the models produce (cg_system_create)."""
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        cl = np.ascontiguousarray(cols, dtype=np.uint32)
        vv = np.ascontiguousarray(vals, dtype=np.complex128)
        ce = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1)
        so = np.ascontiguousarray(site_orbitals, dtype=np.int32)
        sf = np.ascontiguousarray(site_offsets, dtype=np.uint64)
        rg = np.ascontiguousarray(hop_range, dtype=np.int64)
        mats = [np.asfortranarray(m, dtype=np.complex128) for m in (left_h00, left_h01, right_h00, right_h01)]
        h = C.c_void_p()
        _check(capi.cg_system_create(
            n, rp.ctypes.data_as(C.POINTER(C.c_int64)), cl.ctypes.data_as(C.POINTER(C.c_uint32)), _dptr(vv),
            len(so), ce.ctypes.data_as(C.POINTER(C.c_int64)), so.ctypes.data_as(C.POINTER(C.c_int32)),
            sf.ctypes.data_as(C.POINTER(C.c_uint64)), dim, rg.ctypes.data_as(C.POINTER(C.c_int64)), axis,
            _dptr(mats[0]), _dptr(mats[1]), mats[0].shape[0], _dptr(mats[2]), _dptr(mats[3]), mats[2].shape[0],
            C.byref(h)), "from_arrays")
        return cls(h.value)

    # -- structure -------------------------------------------------------------------------
    def block_count(self) -> int:
        return capi.cg_block_count(self._h)

    def sizes(self):
        n, nnz, ns = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(capi.cg_system_sizes(self._h, C.byref(n), C.byref(nnz), C.byref(ns)))
        return n.value, nnz.value, ns.value

    def block_sizes(self, b):
        """(nl, nr, tree nodes, tree levels) of block b."""
        v = [C.c_uint64() for _ in range(4)]
        _check(capi.cg_block_sizes(self._h, b, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def arrays(self):
        """The system arrays: dict of CSR, geometry and lead cells."""
        n, nnz, ns = self.sizes()
        rp = np.zeros(n + 1, np.int64)
        cl = np.zeros(nnz, np.uint32)
        vv = np.zeros(nnz, np.complex128)
        ce = np.zeros((ns, 3), np.int64)
        so = np.zeros(ns, np.int32)
        sf = np.zeros(ns, np.uint64)
        dim, axis = C.c_int32(), C.c_int32()
        rg = np.zeros(3, np.int64)
        nl = self._lead_size(0)
        nr = self._lead_size(1)
        lh = [np.zeros((nl, nl), np.complex128, order="F") for _ in range(2)]
        rh = [np.zeros((nr, nr), np.complex128, order="F") for _ in range(2)]
        _check(capi.cg_system_arrays(
            self._h, rp.ctypes.data_as(C.POINTER(C.c_int64)), cl.ctypes.data_as(C.POINTER(C.c_uint32)), _dptr(vv),
            ce.ctypes.data_as(C.POINTER(C.c_int64)), so.ctypes.data_as(C.POINTER(C.c_int32)),
            sf.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(dim), rg.ctypes.data_as(C.POINTER(C.c_int64)),
            C.byref(axis), _dptr(lh[0]), _dptr(lh[1]), _dptr(rh[0]), _dptr(rh[1])))
        return dict(n=n, row_ptr=rp, cols=cl, vals=vv, cells=ce, site_orbitals=so, site_offsets=sf, dim=dim.value,
                    range=rg, axis=axis.value, left_h00=lh[0], left_h01=lh[1], right_h00=rh[0], right_h01=rh[1])

    def _lead_size(self, side):
        # The lead cells cover the whole end slices (all blocks).
        n, nnz, ns = self.sizes()
        ce = np.zeros((ns, 3), np.int64)
        so = np.zeros(ns, np.int32)
        axis = C.c_int32()
        _check(capi.cg_system_arrays(self._h, None, None, None, ce.ctypes.data_as(C.POINTER(C.c_int64)),
                                     so.ctypes.data_as(C.POINTER(C.c_int32)), None, None, None, C.byref(axis),
                                     None, None, None, None))
        c = ce[:, axis.value]
        end = c.min() if side == 0 else c.max()
        return int(so[c == end].sum())

    def tree_node(self, b, k):
        ne, nb = C.c_uint64(), C.c_uint64()
        _check(capi.cg_tree_node(self._h, b, k, C.byref(ne), C.byref(nb), None, None))
        e = np.zeros(ne.value, np.uint64)
        d = np.zeros(nb.value, np.uint64)
        _check(capi.cg_tree_node(self._h, b, k, C.byref(ne), C.byref(nb), e.ctypes.data_as(C.POINTER(C.c_uint64)),
                                 d.ctypes.data_as(C.POINTER(C.c_uint64))))
        return e, d

    def work(self):
        """(flops, bytes) of one transmission point (8 flops per complex MAC, 16-byte words)."""
        f, b = C.c_double(), C.c_double()
        _check(capi.cg_work(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value

    # -- self-energies -------------------------------------------------------------------------
    def self_energies(self, energy, eta=0.0) -> LeadSelfEnergies:
        """nested.rs:652-676 (FP64 mode matching; the scalar-coupling leads on the GPU when one is
        present). eta > 0: SelfEnergyMethod::ModesBroadened (the leads at E + i eta, the device
        at E; two_terminal.rs:157-172)."""
        g = C.c_void_p()
        if eta:
            _check(capi.cg_self_energies_broadened(self._h, float(energy), float(eta), C.byref(g)), f"E = {energy}")
        else:
            _check(capi.cg_self_energies(self._h, float(energy), C.byref(g)), f"E = {energy}")
        return LeadSelfEnergies(g.value, self)

    def sigma_from_arrays(self, energy, blocks) -> LeadSelfEnergies:
        """Self-energies given per block as (Sigma_L, Sigma_R)."""
        flat = np.concatenate([np.asfortranarray(m, dtype=np.complex128).reshape(-1, order="F")
                               for pair in blocks for m in pair])
        g = C.c_void_p()
        _check(capi.cg_sigma_from_arrays(self._h, float(energy), _dptr(flat), C.byref(g)))
        return LeadSelfEnergies(g.value, self)

    # -- transmission -------------------------------------------------------------------------
    def transmission(self, energy, precision=Precision.Double) -> float:
        """nested.rs:618-620."""
        t = C.c_double()
        _check(capi.cg_transmission(self._h, float(energy), int(precision), C.byref(t)), f"E = {energy}")
        return t.value

    def transmissions(self, energies, precision=Precision.Double):
        """nested.rs:623-626: one batched launch; NaN where a point failed."""
        e = np.ascontiguousarray(energies, dtype=np.float64)
        t = np.zeros_like(e)
        st = np.zeros(len(e), np.int32)
        _check(capi.cg_transmissions(self._h, _dptr(e), len(e), int(precision), _dptr(t),
                                     st.ctypes.data_as(C.POINTER(C.c_int32))))
        return t, st

    def transmission_with(self, sigma: LeadSelfEnergies) -> float:
        """nested.rs:682-684."""
        return self.transmission_with_precision(sigma, Precision.Double)

    def transmission_with_precision(self, sigma: LeadSelfEnergies, precision) -> float:
        """nested.rs:688-713."""
        t = C.c_double()
        _check(capi.cg_transmission_with(self._h, sigma._h, int(precision), C.byref(t)))
        return t.value

    def transmissions_checked(self, energies, options: PrecisionOptions):
        """nested.rs:630-641: (values, report) of `checked` (precision.rs:91-134) over the energies,
        through cg_batch_transmission_checked; NaN where a point failed."""
        sig = [self.self_energies(e) for e in energies]
        t, st, rep = batch_transmission_checked([self] * len(sig), sig, options)
        return t, rep

    def backward(self, sigma: LeadSelfEnergies):
        """T and dT/d(eps_a) for every orbital a (adjoint sweep over the stored fronts)."""
        n = self.sizes()[0]
        g = np.zeros(n)
        t = C.c_double()
        _check(capi.cg_backward(self._h, sigma._h, C.byref(t), _dptr(g)))
        return t.value, g

    def scattering_states(self, sigma: LeadSelfEnergies):
        """psi_p = G W_p on every orbital for the left and right lead (nested.rs:851-895):
        ScatteringStates with `.leads = [psi_left, psi_right]`, complex n x channels."""
        n = self.sizes()[0]
        cols = (C.c_int64 * 2)()
        _check(capi.cg_scattering_states(self._h, sigma._h, cols, None, None))
        left = np.zeros((int(cols[0]), n), complex)
        right = np.zeros((int(cols[1]), n), complex)
        _check(capi.cg_scattering_states(self._h, sigma._h, cols, _dptr(left.view(np.float64)),
                                         _dptr(right.view(np.float64))), "scattering states")
        return ScatteringStates(sigma.energy(), [left.T, right.T])

    def site_densities(self, states, f, op=None):
        """sum_p f_p psi_p^H O psi_p / 2 pi per site (nested.rs:994-1025), O = 1 without `op`."""
        left, right, cols = states._packed()
        out = np.zeros(len(self.arrays()["site_orbitals"]))
        fa = np.asarray(f, float)
        if op is None:
            rc = capi.cg_site_densities(self._h, _dptr(left), _dptr(right), cols, _dptr(fa), None, 0, _dptr(out))
        else:
            o = np.asfortranarray(np.asarray(op, complex))
            rc = capi.cg_site_densities(self._h, _dptr(left), _dptr(right), cols, _dptr(fa),
                                        _dptr(o.ravel(order="F").view(np.float64)), o.shape[0], _dptr(out))
        _check(rc, "site densities")
        return out

    def bond_currents(self, states, f):
        """2 Re Tr[H_ij G^<_ji] for every coupled pair of sites i < j (nested.rs:930-990):
        arrays (from, to, current), ordered by (from, to)."""
        left, right, cols = states._packed()
        fa = np.asarray(f, float)
        count = C.c_int64()
        _check(capi.cg_bond_currents(self._h, _dptr(left), _dptr(right), cols, _dptr(fa), C.byref(count), None, None,
                                     None))
        fr = np.zeros(count.value, np.int64)
        to = np.zeros(count.value, np.int64)
        cur = np.zeros(count.value)
        _check(capi.cg_bond_currents(self._h, _dptr(left), _dptr(right), cols, _dptr(fa), C.byref(count),
                                     fr.ctypes.data_as(C.POINTER(C.c_int64)), to.ctypes.data_as(C.POINTER(C.c_int64)),
                                     _dptr(cur)), "bond currents")
        return fr, to, cur

    def operator_currents(self, states, f, op):
        """Re Tr[(O H_ij + H_ij O) G^<_ji] for every coupled pair of sites i < j (nested.rs:925-985;
        a spin current for O = s_z): arrays (from, to, current)."""
        left, right, cols = states._packed()
        fa = np.asarray(f, float)
        o = np.asarray(op, complex)
        ob = np.ascontiguousarray(o.T).view(np.float64)  # column-major (re, im)
        count = C.c_int64()
        _check(capi.cg_operator_currents(self._h, _dptr(left), _dptr(right), cols, _dptr(fa), _dptr(ob), o.shape[0],
                                         C.byref(count), None, None, None), "operator currents")
        fr = np.zeros(count.value, np.int64)
        to = np.zeros(count.value, np.int64)
        cur = np.zeros(count.value)
        _check(capi.cg_operator_currents(self._h, _dptr(left), _dptr(right), cols, _dptr(fa), _dptr(ob), o.shape[0],
                                         C.byref(count), fr.ctypes.data_as(C.POINTER(C.c_int64)),
                                         to.ctypes.data_as(C.POINTER(C.c_int64)), _dptr(cur)), "operator currents")
        return fr, to, cur


class ScatteringStates:
    """ScatteringStates of nested.rs: the energy and, per lead, the states psi_p (n x channels)."""

    def __init__(self, energy, leads):
        self.energy = energy
        self.leads = leads

    def _packed(self):
        """Both leads as contiguous column-major (re, im) buffers and their channel counts."""
        bufs = [np.ascontiguousarray(np.asarray(m, complex).T).view(np.float64) for m in self.leads]
        cols = (C.c_int64 * 2)(*[m.shape[1] for m in self.leads])
        return bufs[0], bufs[1], cols


TwoTerminal = NestedTwoTerminal


def batch_transmission(systems, sigmas, precision=Precision.Double, stats=None):
    """One batched launch over points (systems[p], sigmas[p]) (cg_batch_transmission).
    Returns (T, status); `stats` (a capi.BatchStats) receives the timing and work counts."""
    n = len(systems)
    sh = (C.c_void_p * n)(*[s._h.value for s in systems])
    gh = (C.c_void_p * n)(*[g._h.value for g in sigmas])
    t = np.zeros(n)
    st = np.zeros(n, np.int32)
    bs = stats if stats is not None else capi.BatchStats()
    _check(capi.cg_batch_transmission(sh, gh, n, int(precision), _dptr(t), st.ctypes.data_as(C.POINTER(C.c_int32)),
                                      C.byref(bs)))
    return t, st


def batch_transmission_checked(systems, sigmas, options: PrecisionOptions, stats=None):
    """checked (precision.rs:91-134) over one batched launch (cg_batch_transmission_checked):
    returns (T, status, PrecisionReport) with the per-point |T - T_double| in report.differences."""
    n = len(systems)
    sh = (C.c_void_p * n)(*[s._h.value for s in systems])
    gh = (C.c_void_p * n)(*[g._h.value for g in sigmas])
    t = np.zeros(n)
    st = np.zeros(n, np.int32)
    dt = np.zeros(n)
    opt = capi.PrecisionOptionsC(int(options.precision), float(options.tolerance), int(options.every))
    rep = capi.PrecisionReportC()
    bs = stats if stats is not None else capi.BatchStats()
    _check(capi.cg_batch_transmission_checked(sh, gh, n, C.byref(opt), _dptr(t),
                                              st.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(dt), C.byref(rep),
                                              C.byref(bs)))
    report = PrecisionReport(Precision(rep.precision), rep.checked, rep.median, rep.max, rep.failed,
                             rep.fallback.decode(), dt)
    return np.where(st == 0, t, np.nan), st, report


def backward_batch(systems, sigmas):
    """T and dT/d(eps_a) of every orbital for many points of one structure in one sweep
    (cg_backward_batch): returns (T[count], grad[count, n], status[count])."""
    n = len(systems)
    nn = systems[0].sizes()[0] if n else 0
    sh = (C.c_void_p * n)(*[s._h.value for s in systems])
    gh = (C.c_void_p * n)(*[g._h.value for g in sigmas])
    t = np.zeros(n)
    grad = np.zeros((n, nn))
    st = np.zeros(n, np.int32)
    _check(capi.cg_backward_batch(sh, gh, n, _dptr(t), _dptr(grad), st.ctypes.data_as(C.POINTER(C.c_int32))),
           "backward batch")
    return t, grad, st


class ResidentBatch:
    """Points whose inputs stay in HBM between launches (cg_batch_create / cg_batch_run)."""

    def __init__(self, systems, sigmas, precision=Precision.Double):
        self._keep = (list(systems), list(sigmas))
        n = len(systems)
        self.n = n
        sh = (C.c_void_p * n)(*[s._h.value for s in systems])
        gh = (C.c_void_p * n)(*[g._h.value for g in sigmas])
        h = C.c_void_p()
        _check(capi.cg_batch_create(sh, gh, n, int(precision), C.byref(h)), "batch")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            capi.cg_batch_destroy(self._h)
            self._h = C.c_void_p(None)

    def run(self, stats=None):
        t = np.zeros(self.n)
        st = np.zeros(self.n, np.int32)
        bs = stats if stats is not None else capi.BatchStats()
        _check(capi.cg_batch_run(self._h, _dptr(t), st.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(bs)))
        return t, st

    def autotune(self):
        """Times eager sweeps of these points under candidate kernel choices (cg_batch_autotune),
        leaves the fastest in force and the batch planned with it. Returns (choice dict, seconds
        per sweep)."""
        tu = capi.Tuning()
        sec = C.c_double()
        _check(capi.cg_batch_autotune(self._h, C.byref(tu), C.byref(sec)), "autotune")
        return tu.as_dict(), sec.value


def get_tuning():
    """The kernel choices in force (cg_get_tuning) as a dict."""
    tu = capi.Tuning()
    _check(capi.cg_get_tuning(C.byref(tu)))
    return tu.as_dict()


def set_tuning(choice=None, **fields):
    """Sets the kernel choices (cg_set_tuning): a dict or keyword fields over the ones in force;
    set_tuning(None) with no fields goes back to the defaults and the CAROLI_* variables."""
    if choice is None and not fields:
        _check(capi.cg_set_tuning(None))
        return
    cur = get_tuning()
    cur.update(choice or {})
    cur.update(fields)
    _check(capi.cg_set_tuning(C.byref(capi.Tuning(**cur))), "tuning")


def kernel_timing(enable=True):
    _check(capi.cg_kernel_timing(1 if enable else 0))


def kernel_stats():
    """Per kernel family: dict name -> (launches, seconds, algorithmic flops, algorithmic bytes)."""
    buf = (capi.KernelStat * 32)()
    n = capi.cg_kernel_stats(buf, 32)
    return {buf[i].name.decode(): (buf[i].launches, buf[i].seconds, buf[i].flops, buf[i].bytes) for i in range(n)}


class MultiSelfEnergies:
    """Self-energies of every lead of a multi-terminal device at one energy, with the labels of the
    rows and columns of T: (lead, sector) pairs (nested_multi.rs:196-313)."""

    def __init__(self, handle, owner):
        self._h = C.c_void_p(handle)
        self._owner = owner

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            capi.cg_sigma_destroy(self._h)
            self._h = C.c_void_p(None)

    def energy(self) -> float:
        return capi.cg_sigma_energy(self._h)

    def sigma(self, p):
        """Sigma_p of lead p on its coupled orbitals (NestedMultiTerminal.lead(p)["orbitals"])."""
        k = len(self._owner.lead(p)["orbitals"])
        out = np.zeros((k, k), np.complex128, order="F")
        _check(capi.cg_multi_sigma(self._owner._h, self._h, p, _dptr(out)))
        return out

    def labels(self):
        n = C.c_int32()
        _check(capi.cg_multi_labels(self._h, C.byref(n), None, None))
        lead = np.zeros(n.value, np.int32)
        sector = np.zeros(n.value, np.int32)
        _check(capi.cg_multi_labels(self._h, C.byref(n), lead.ctypes.data_as(C.POINTER(C.c_int32)),
                                    sector.ctypes.data_as(C.POINTER(C.c_int32))))
        return [(int(a), int(b)) for a, b in zip(lead, sector)]


@dataclass
class SectorTransmissions:
    """nested_multi.rs:36-43: the (lead, sector) of each row and column, t[k][l] from labels[l]
    into labels[k]."""

    labels: list
    t: np.ndarray


@dataclass
class SMatrix:
    """smatrix.rs:22-34: blocks[p][q] the amplitudes into the outgoing modes of lead p from the
    incoming modes of lead q (unit current), and the Bloch factors of the modes."""

    energy: float
    blocks: list
    incoming: list
    outgoing: list

    def transmission(self, p, q) -> float:
        """sum |S_pq|^2 (a reflection for p = q)."""
        return float(np.sum(np.abs(self.blocks[p][q]) ** 2))

    def full(self):
        return np.block(self.blocks) if self.blocks else np.zeros((0, 0), complex)


class NestedMultiTerminal:
    """A device with any number of leads solved by nested dissection (nested_multi.rs:45-56): the
    sites the leads couple to stay at the root, whose border block gives G on all of their
    orbitals, and T[p][q] = Tr[Gamma_p G_pq Gamma_q G_pq^H] from lead q into lead p."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            capi.cg_system_destroy(self._h)

    @classmethod
    def from_model(cls, model, sites, interfaces, site_potential=None):
        """NestedMultiTerminal::from_model (nested_multi.rs:60) with a preset model as the clean lead:
        `sites` the region's (x, y, z, sublattice) in order, `interfaces` a list of (slice sites,
        outward step), site_potential an optional on-site potential per site."""
        st = np.ascontiguousarray(sites, dtype=np.int64).reshape(-1, 4)
        cells = np.ascontiguousarray(st[:, :3])
        subs = np.ascontiguousarray(st[:, 3], dtype=np.int32)
        keep = [cells, subs]
        itf = (capi.Interface * len(interfaces))()
        for p, (slice_sites, outward) in enumerate(interfaces):
            ss = np.ascontiguousarray(slice_sites, dtype=np.int64).reshape(-1, 4)
            sc = np.ascontiguousarray(ss[:, :3])
            sb = np.ascontiguousarray(ss[:, 3], dtype=np.int32)
            keep += [sc, sb]
            itf[p].sites = len(ss)
            itf[p].cells = sc.ctypes.data_as(C.POINTER(C.c_int64))
            itf[p].subs = sb.ctypes.data_as(C.POINTER(C.c_int32))
            for d in range(3):
                itf[p].outward[d] = int(outward[d])
        pot = None
        if site_potential is not None:
            pot = np.ascontiguousarray(site_potential, dtype=np.float64)
            if len(pot) != len(st):
                raise ValueError("one potential per site")
        h = C.c_void_p()
        _check(capi.cg_multi_model_create(int(model), len(st), cells.ctypes.data_as(C.POINTER(C.c_int64)),
                                          subs.ctypes.data_as(C.POINTER(C.c_int32)),
                                          _dptr(pot) if pot is not None else None, len(interfaces), itf,
                                          C.byref(h)), "multi-terminal system")
        return cls(h.value)

    @classmethod
    def from_arrays(cls, n, row_ptr, cols, vals, cells, site_orbitals, site_offsets, dim, hop_range, leads):
        """The device arrays as NestedTwoTerminal.from_arrays, and `leads` a list of dicts with h00,
        h01 (cell x cell), orbitals (the coupled device orbitals, increasing), coupling (coupled x
        cell) and optionally kinds (the index within its site of each cell orbital)."""
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        cl = np.ascontiguousarray(cols, dtype=np.uint32)
        vv = np.ascontiguousarray(vals, dtype=np.complex128)
        ce = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1, 3)
        so = np.ascontiguousarray(site_orbitals, dtype=np.int32)
        sf = np.ascontiguousarray(site_offsets, dtype=np.uint64)
        rg = np.ascontiguousarray(hop_range, dtype=np.int64)
        keep = []
        arr = (capi.Lead * len(leads))()
        for p, l in enumerate(leads):
            h00 = np.asfortranarray(l["h00"], dtype=np.complex128)
            h01 = np.asfortranarray(l["h01"], dtype=np.complex128)
            orb = np.ascontiguousarray(l["orbitals"], dtype=np.uint64)
            cp = np.asfortranarray(l["coupling"], dtype=np.complex128)
            kd = None if l.get("kinds") is None else np.ascontiguousarray(l["kinds"], dtype=np.int32)
            keep += [h00, h01, orb, cp, kd]
            arr[p].cell = h00.shape[0]
            arr[p].h00 = _dptr(h00)
            arr[p].h01 = _dptr(h01)
            arr[p].coupled = len(orb)
            arr[p].orbitals = orb.ctypes.data_as(C.POINTER(C.c_uint64))
            arr[p].coupling = _dptr(cp)
            arr[p].kinds = kd.ctypes.data_as(C.POINTER(C.c_int32)) if kd is not None else None
        h = C.c_void_p()
        _check(capi.cg_multi_create(int(n), rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                    cl.ctypes.data_as(C.POINTER(C.c_uint32)), _dptr(vv), len(so),
                                    ce.ctypes.data_as(C.POINTER(C.c_int64)), so.ctypes.data_as(C.POINTER(C.c_int32)),
                                    sf.ctypes.data_as(C.POINTER(C.c_uint64)), int(dim),
                                    rg.ctypes.data_as(C.POINTER(C.c_int64)), len(leads), arr, C.byref(h)),
               "multi-terminal system")
        return cls(h.value)

    def lead_count(self) -> int:
        return capi.cg_multi_lead_count(self._h)

    def block_count(self) -> int:
        return capi.cg_block_count(self._h)

    def sizes(self):
        n, nnz, ns = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(capi.cg_system_sizes(self._h, C.byref(n), C.byref(nnz), C.byref(ns)))
        return n.value, nnz.value, ns.value

    def arrays(self):
        """The device arrays (CSR and geometry)."""
        n, nnz, ns = self.sizes()
        rp = np.zeros(n + 1, np.int64)
        cl = np.zeros(nnz, np.uint32)
        vv = np.zeros(nnz, np.complex128)
        ce = np.zeros((ns, 3), np.int64)
        so = np.zeros(ns, np.int32)
        sf = np.zeros(ns, np.uint64)
        dim = C.c_int32()
        rg = np.zeros(3, np.int64)
        _check(capi.cg_system_arrays(
            self._h, rp.ctypes.data_as(C.POINTER(C.c_int64)), cl.ctypes.data_as(C.POINTER(C.c_uint32)), _dptr(vv),
            ce.ctypes.data_as(C.POINTER(C.c_int64)), so.ctypes.data_as(C.POINTER(C.c_int32)),
            sf.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(dim), rg.ctypes.data_as(C.POINTER(C.c_int64)),
            None, None, None, None, None))
        return dict(n=n, row_ptr=rp, cols=cl, vals=vv, cells=ce, site_orbitals=so, site_offsets=sf, dim=dim.value,
                    range=rg)

    def lead(self, p):
        """Lead p as attached: dict of h00, h01, orbitals, coupling (multi.rs:15-25)."""
        nc, nk = C.c_uint64(), C.c_uint64()
        _check(capi.cg_multi_lead(self._h, p, C.byref(nc), C.byref(nk), None, None, None, None))
        h00 = np.zeros((nc.value, nc.value), np.complex128, order="F")
        h01 = np.zeros_like(h00)
        orb = np.zeros(nk.value, np.uint64)
        cp = np.zeros((nk.value, nc.value), np.complex128, order="F")
        _check(capi.cg_multi_lead(self._h, p, C.byref(nc), C.byref(nk), _dptr(h00), _dptr(h01),
                                  orb.ctypes.data_as(C.POINTER(C.c_uint64)), _dptr(cp)))
        return dict(h00=h00, h01=h01, orbitals=orb.astype(np.int64), coupling=cp)

    def self_energies(self, energy, eta=0.0, sectors=None) -> MultiSelfEnergies:
        """Self-energies of the leads at E + i eta (nested_multi.rs:196-232); `sectors` as in
        sector_transmissions splits the labels."""
        g = C.c_void_p()
        if sectors is None:
            _check(capi.cg_multi_self_energies(self._h, float(energy), float(eta), None, None, C.byref(g)),
                   f"E = {energy}")
        else:
            if len(sectors) != self.lead_count():
                raise ValueError("one sector labelling per lead")
            ln = np.array([0 if s is None else len(s) for s in sectors], np.int32)
            tags = np.array([t for s in sectors if s is not None for t in s] or [0], np.int32)
            _check(capi.cg_multi_self_energies(self._h, float(energy), float(eta),
                                               ln.ctypes.data_as(C.POINTER(C.c_int32)),
                                               tags.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(g)),
                   f"E = {energy}")
        return MultiSelfEnergies(g.value, self)

    def transmissions(self, energy, precision=Precision.Double):
        """nested_multi.rs:174-176: T[p][q] from lead q into lead p."""
        n = self.lead_count()
        t = np.zeros((n, n))
        _check(capi.cg_multi_transmissions(self._h, float(energy), int(precision), _dptr(t)), f"E = {energy}")
        return t

    def transmissions_with(self, sigma: MultiSelfEnergies, precision=Precision.Double):
        """transmissions_with_precision (nested_multi.rs:180-194) with given self-energies."""
        t, st = multi_batch_transmissions([self], [sigma], precision)
        if st[0] != 0:
            raise GreenError(int(st[0]), f"E = {sigma.energy()}")
        return t[0]

    def smatrix(self, energy) -> "SMatrix":
        """nested_multi.rs:440-470: the scattering matrix in the basis of the propagating lead modes
        normalised to unit current (cg_multi_smatrix)."""
        P = self.lead_count()
        modes = np.zeros(2 * P, np.int32)
        mp = modes.ctypes.data_as(C.POINTER(C.c_int32))
        _check(capi.cg_multi_smatrix(self._h, float(energy), mp, None, None, None), f"E = {energy}")
        nout, nin = modes[:P].copy(), modes[P:].copy()
        full = np.zeros((int(nin.sum()), int(nout.sum())), complex)  # column-major = transposed C order
        lo = np.zeros(int(nout.sum()), complex)
        li = np.zeros(int(nin.sum()), complex)
        _check(capi.cg_multi_smatrix(self._h, float(energy), mp, _dptr(full.view(np.float64)),
                                     _dptr(lo.view(np.float64)), _dptr(li.view(np.float64))), f"E = {energy}")
        full = full.T
        ro = np.concatenate([[0], np.cumsum(nout)])
        co = np.concatenate([[0], np.cumsum(nin)])
        blocks = [[full[ro[p]:ro[p + 1], co[q]:co[q + 1]] for q in range(P)] for p in range(P)]
        return SMatrix(float(energy), blocks, [li[co[q]:co[q + 1]] for q in range(P)],
                       [lo[ro[p]:ro[p + 1]] for p in range(P)])

    def scattering_states(self, sigma: MultiSelfEnergies) -> ScatteringStates:
        """nested_multi.rs:354-413: psi = G W on every orbital for the channels of each label of
        `sigma` (each lead, or each sector of a lead): ScatteringStates with one n x channels array
        per label."""
        n = self.sizes()[0]
        L = len(sigma.labels())
        cols = (C.c_int64 * L)()
        _check(capi.cg_multi_scattering_states(self._h, sigma._h, cols, None))
        bufs = [np.zeros((int(cols[l]), n), complex) for l in range(L)]
        ptrs = (C.POINTER(C.c_double) * L)(*[_dptr(b.view(np.float64)) for b in bufs])
        _check(capi.cg_multi_scattering_states(self._h, sigma._h, cols, ptrs), "scattering states")
        return ScatteringStates(sigma.energy(), [b.T for b in bufs])

    @staticmethod
    def _states(states):
        bufs = [np.ascontiguousarray(np.asarray(m, complex).T).view(np.float64) for m in states.leads]
        L = len(bufs)
        ptrs = (C.POINTER(C.c_double) * L)(*[_dptr(b) for b in bufs])
        cols = (C.c_int64 * L)(*[m.shape[1] for m in states.leads])
        return bufs, ptrs, cols, L

    def site_densities(self, states, f, op=None):
        """sum_p f_p psi_p^H O psi_p / 2 pi per site (nested_multi.rs:510-513), O = 1 without `op`."""
        bufs, ptrs, cols, L = self._states(states)
        fa = np.asarray(f, float)
        out = np.zeros(self.sizes()[2])
        o = None if op is None else np.ascontiguousarray(np.asarray(op, complex).T).view(np.float64)
        _check(capi.cg_multi_site_densities(self._h, L, ptrs, cols, _dptr(fa), None if o is None else _dptr(o),
                                            0 if op is None else np.asarray(op).shape[0], _dptr(out)), "site densities")
        return out

    def bond_currents(self, states, f, op=None):
        """Bond currents 2 Re Tr[H_ij G^<_ji] (nested_multi.rs:497-500), or those of an on-site
        operator O (nested_multi.rs:503-505): arrays (from, to, current) ordered by (from, to)."""
        bufs, ptrs, cols, L = self._states(states)
        fa = np.asarray(f, float)
        o = None if op is None else np.ascontiguousarray(np.asarray(op, complex).T).view(np.float64)
        od = 0 if op is None else np.asarray(op).shape[0]
        count = C.c_int64()
        _check(capi.cg_multi_bond_currents(self._h, L, ptrs, cols, _dptr(fa), None if o is None else _dptr(o), od,
                                           C.byref(count), None, None, None), "bond currents")
        fr = np.zeros(count.value, np.int64)
        to = np.zeros(count.value, np.int64)
        cur = np.zeros(count.value)
        _check(capi.cg_multi_bond_currents(self._h, L, ptrs, cols, _dptr(fa), None if o is None else _dptr(o), od,
                                           C.byref(count), fr.ctypes.data_as(C.POINTER(C.c_int64)),
                                           to.ctypes.data_as(C.POINTER(C.c_int64)), _dptr(cur)), "bond currents")
        return fr, to, cur

    def sector_transmissions(self, energy, sectors, precision=Precision.Double) -> SectorTransmissions:
        """nested_multi.rs:246-313: `sectors[p]` tags the orbitals of each site of lead p (e.g.
        [0, 0, 1, 1] for the spin blocks of BHZ), or None for the lead as a whole."""
        g = self.self_energies(energy, sectors=sectors)
        return SectorTransmissions(g.labels(), self.transmissions_with(g, precision))


def multi_batch_transmissions(systems, sigmas, precision=Precision.Double, stats=None):
    """Every point (systems[p], sigmas[p]) of multi-terminal devices in one batched launch
    (cg_multi_batch_transmissions): T of shape (points, labels, labels) and the statuses."""
    n = len(systems)
    if n == 0:
        return np.zeros((0, 0, 0)), np.zeros(0, np.int32)
    L = len(sigmas[0].labels())
    sh = (C.c_void_p * n)(*[s._h.value for s in systems])
    gh = (C.c_void_p * n)(*[g._h.value for g in sigmas])
    t = np.zeros((n, L, L))
    st = np.zeros(n, np.int32)
    bs = stats if stats is not None else capi.BatchStats()
    _check(capi.cg_multi_batch_transmissions(sh, gh, n, int(precision), _dptr(t),
                                             st.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(bs)), "multi batch")
    return t, st
