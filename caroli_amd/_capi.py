"""ctypes binding of include/caroli_amd.h (the engine's C ABI).

Loads `caroli_amd/libcaroli_amd.so`, built in-tree by `caroli_amd/build.py`. There is no fallback:
a missing or unloadable library raises at import.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcaroli_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python caroli_amd/build.py`")
lib = C.CDLL(LIB_PATH)

c_sys = C.c_void_p
c_sig = C.c_void_p
P = C.POINTER


class BatchStats(C.Structure):
    _fields_ = [
        ("seconds", C.c_double),
        ("front_flops", C.c_double),
        ("front_bytes", C.c_double),
        ("delayed_points", C.c_int64),
        ("refinement_iters", C.c_int64),
        ("max_refine_resid", C.c_double),
        ("recomputed_points", C.c_int64),
        ("goal_points", C.c_int64),
    ]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


cg_system_create = _sig("cg_system_create", C.c_int32, C.c_uint64, P(C.c_int64), P(C.c_uint32), P(C.c_double),
                        C.c_uint64, P(C.c_int64), P(C.c_int32), P(C.c_uint64), C.c_int32, P(C.c_int64), C.c_int32,
                        P(C.c_double), P(C.c_double), C.c_uint64, P(C.c_double), P(C.c_double), C.c_uint64,
                        P(c_sys))
cg_bhz_magnetic_create = _sig("cg_bhz_magnetic_create", C.c_int32, P(C.c_uint64), C.c_int32, C.c_int64, C.c_int64,
                              C.c_int64, C.c_double, C.c_int32, C.c_double, C.c_int32, C.c_double, P(c_sys))
cg_self_energies_broadened = _sig("cg_self_energies_broadened", C.c_int32, c_sys, C.c_double, C.c_double, P(c_sig))
class Disorder(C.Structure):
    """cg_disorder: one [[disorder]] table of the CLI input."""
    _fields_ = [("kind", C.c_int32), ("strength", C.c_double), ("concentration", C.c_double),
                ("exchange", C.c_double), ("moments", C.c_int32), ("moment_param", C.c_double),
                ("spin", C.c_int32), ("has_cells", C.c_int32 * 3), ("cells", (C.c_int64 * 2) * 3),
                ("sublattices", C.c_uint64), ("seed", C.c_uint64)]


cg_disordered_create = _sig("cg_disordered_create", C.c_int32, C.c_int32, P(C.c_uint64), C.c_int32, C.c_int64,
                            P(Disorder), C.c_uint64, P(c_sys))
cg_preset_create = _sig("cg_preset_create", C.c_int32, C.c_int32, P(C.c_uint64), C.c_int32, C.c_int32, C.c_double,
                        C.c_double, C.c_uint64, P(C.c_double), P(c_sys))
cg_system_destroy = _sig("cg_system_destroy", None, c_sys)
cg_block_count = _sig("cg_block_count", C.c_int32, c_sys)
cg_system_sizes = _sig("cg_system_sizes", C.c_int32, c_sys, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64))
cg_block_sizes = _sig("cg_block_sizes", C.c_int32, c_sys, C.c_int32, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64),
                      P(C.c_uint64))
cg_system_arrays = _sig("cg_system_arrays", C.c_int32, c_sys, P(C.c_int64), P(C.c_uint32), P(C.c_double),
                        P(C.c_int64), P(C.c_int32), P(C.c_uint64), P(C.c_int32), P(C.c_int64), P(C.c_int32),
                        P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double))
cg_tree_node = _sig("cg_tree_node", C.c_int32, c_sys, C.c_int32, C.c_int32, P(C.c_uint64), P(C.c_uint64),
                    P(C.c_uint64), P(C.c_uint64))
cg_work = _sig("cg_work", C.c_int32, c_sys, P(C.c_double), P(C.c_double))
cg_self_energies = _sig("cg_self_energies", C.c_int32, c_sys, C.c_double, P(c_sig))
cg_sigma_from_arrays = _sig("cg_sigma_from_arrays", C.c_int32, c_sys, C.c_double, P(C.c_double), P(c_sig))
cg_sigma_block = _sig("cg_sigma_block", C.c_int32, c_sig, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_int32),
                      P(C.c_int32))
cg_sigma_energy = _sig("cg_sigma_energy", C.c_double, c_sig)
cg_sigma_destroy = _sig("cg_sigma_destroy", None, c_sig)
cg_transmission = _sig("cg_transmission", C.c_int32, c_sys, C.c_double, C.c_int32, P(C.c_double))
cg_transmission_with = _sig("cg_transmission_with", C.c_int32, c_sys, c_sig, C.c_int32, P(C.c_double))
cg_transmissions = _sig("cg_transmissions", C.c_int32, c_sys, P(C.c_double), C.c_int64, C.c_int32, P(C.c_double),
                        P(C.c_int32))
cg_batch_transmission = _sig("cg_batch_transmission", C.c_int32, P(c_sys), P(c_sig), C.c_int64, C.c_int32,
                             P(C.c_double), P(C.c_int32), P(BatchStats))
cg_backward = _sig("cg_backward", C.c_int32, c_sys, c_sig, P(C.c_double), P(C.c_double))
cg_backward_batch = _sig("cg_backward_batch", C.c_int32, P(c_sys), P(c_sig), C.c_int64, P(C.c_double),
                         P(C.c_double), P(C.c_int32))


class PrecisionOptionsC(C.Structure):
    """cg_precision_options."""
    _fields_ = [("precision", C.c_int32), ("tolerance", C.c_double), ("every", C.c_int64)]


class PrecisionReportC(C.Structure):
    """cg_precision_report."""
    _fields_ = [("precision", C.c_int32), ("checked", C.c_int64), ("median", C.c_double), ("max", C.c_double),
                ("failed", C.c_int64), ("fell_back", C.c_int32), ("fallback", C.c_char * 160)]


cg_batch_transmission_checked = _sig("cg_batch_transmission_checked", C.c_int32, P(c_sys), P(c_sig), C.c_int64,
                                     P(PrecisionOptionsC), P(C.c_double), P(C.c_int32), P(C.c_double),
                                     P(PrecisionReportC), P(BatchStats))
cg_device_available = _sig("cg_device_available", C.c_int32)

# Every function the header declares, for the export check.
EXPORTS = [
    "cg_system_create", "cg_preset_create", "cg_system_destroy", "cg_block_count", "cg_system_sizes",
    "cg_block_sizes", "cg_system_arrays", "cg_tree_node", "cg_work", "cg_self_energies", "cg_sigma_from_arrays",
    "cg_sigma_block", "cg_sigma_energy", "cg_sigma_destroy", "cg_transmission", "cg_transmission_with",
    "cg_transmissions", "cg_batch_transmission", "cg_backward", "cg_device_available",
    "cg_batch_transmission_checked", "cg_bhz_magnetic_create", "cg_self_energies_broadened",
    "cg_disordered_create", "cg_backward_batch",
]


class KernelStat(C.Structure):
    _fields_ = [
        ("name", C.c_char * 48),
        ("launches", C.c_int64),
        ("seconds", C.c_double),
        ("flops", C.c_double),
        ("bytes", C.c_double),
    ]


cg_batch_create = _sig("cg_batch_create", C.c_int32, P(c_sys), P(c_sig), C.c_int64, C.c_int32, P(C.c_void_p))
cg_batch_run = _sig("cg_batch_run", C.c_int32, C.c_void_p, P(C.c_double), P(C.c_int32), P(BatchStats))
cg_batch_destroy = _sig("cg_batch_destroy", None, C.c_void_p)
cg_kernel_timing = _sig("cg_kernel_timing", C.c_int32, C.c_int32)
cg_kernel_stats = _sig("cg_kernel_stats", C.c_int32, P(KernelStat), C.c_int32)
cg_front_schur = _sig("cg_front_schur", C.c_int32, P(C.c_double), C.c_int64, C.c_int64, C.c_int64, C.c_int32,
                      P(C.c_double), P(C.c_int32))
cg_front_schur_noalloc = _sig("cg_front_schur_noalloc", C.c_int32, P(C.c_double), C.c_int64, C.c_int64, C.c_int64,
                              C.c_int32, P(C.c_double), P(C.c_int32))
EXPORTS += ["cg_front_schur", "cg_front_schur_noalloc"]
cg_uniforms = _sig("cg_uniforms", C.c_int32, C.c_int32, C.c_uint64, P(C.c_int32), C.c_int64, P(C.c_double))
EXPORTS += ["cg_batch_create", "cg_batch_run", "cg_batch_destroy", "cg_kernel_timing", "cg_kernel_stats",
            "cg_uniforms"]


class Tuning(C.Structure):
    """cg_tuning: the kernel choices per front class."""
    _fields_ = [
        ("small_fronts", C.c_int32),
        ("split", C.c_int32),
        ("blocked_min_ne", C.c_int32),
        ("unblocked", C.c_int32),
        ("schur_tile", C.c_int32),
    ]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


cg_get_tuning = _sig("cg_get_tuning", C.c_int32, P(Tuning))
cg_set_tuning = _sig("cg_set_tuning", C.c_int32, P(Tuning))
cg_batch_autotune = _sig("cg_batch_autotune", C.c_int32, C.c_void_p, P(Tuning), P(C.c_double))
EXPORTS += ["cg_get_tuning", "cg_set_tuning", "cg_batch_autotune"]

cg_scattering_states = _sig("cg_scattering_states", C.c_int32, c_sys, c_sig, P(C.c_int64), P(C.c_double),
                            P(C.c_double))
cg_site_densities = _sig("cg_site_densities", C.c_int32, c_sys, P(C.c_double), P(C.c_double), P(C.c_int64),
                         P(C.c_double), P(C.c_double), C.c_int32, P(C.c_double))
cg_bond_currents = _sig("cg_bond_currents", C.c_int32, c_sys, P(C.c_double), P(C.c_double), P(C.c_int64),
                        P(C.c_double), P(C.c_int64), P(C.c_int64), P(C.c_int64), P(C.c_double))
cg_operator_currents = _sig("cg_operator_currents", C.c_int32, c_sys, P(C.c_double), P(C.c_double), P(C.c_int64),
                            P(C.c_double), P(C.c_double), C.c_int32, P(C.c_int64), P(C.c_int64), P(C.c_int64),
                            P(C.c_double))
EXPORTS += ["cg_scattering_states", "cg_site_densities", "cg_bond_currents", "cg_operator_currents"]


class Lead(C.Structure):
    """cg_lead: one lead as attached (cell, h00, h01, coupled orbitals, coupling, kinds)."""
    _fields_ = [
        ("cell", C.c_uint64),
        ("h00", P(C.c_double)),
        ("h01", P(C.c_double)),
        ("coupled", C.c_uint64),
        ("orbitals", P(C.c_uint64)),
        ("coupling", P(C.c_double)),
        ("kinds", P(C.c_int32)),
    ]


class Interface(C.Structure):
    """cg_interface: the device sites of one interface slice and its outward step."""
    _fields_ = [
        ("sites", C.c_uint64),
        ("cells", P(C.c_int64)),
        ("subs", P(C.c_int32)),
        ("outward", C.c_int64 * 3),
    ]


cg_multi_create = _sig("cg_multi_create", C.c_int32, C.c_uint64, P(C.c_int64), P(C.c_uint32), P(C.c_double),
                       C.c_uint64, P(C.c_int64), P(C.c_int32), P(C.c_uint64), C.c_int32, P(C.c_int64), C.c_int32,
                       P(Lead), P(c_sys))
cg_multi_model_create = _sig("cg_multi_model_create", C.c_int32, C.c_int32, C.c_uint64, P(C.c_int64), P(C.c_int32),
                             P(C.c_double), C.c_int32, P(Interface), P(c_sys))
cg_multi_lead_count = _sig("cg_multi_lead_count", C.c_int32, c_sys)
cg_multi_lead = _sig("cg_multi_lead", C.c_int32, c_sys, C.c_int32, P(C.c_uint64), P(C.c_uint64), P(C.c_double),
                     P(C.c_double), P(C.c_uint64), P(C.c_double))
cg_multi_self_energies = _sig("cg_multi_self_energies", C.c_int32, c_sys, C.c_double, C.c_double, P(C.c_int32),
                              P(C.c_int32), P(c_sig))
cg_multi_sigma = _sig("cg_multi_sigma", C.c_int32, c_sys, c_sig, C.c_int32, P(C.c_double))
cg_multi_labels = _sig("cg_multi_labels", C.c_int32, c_sig, P(C.c_int32), P(C.c_int32), P(C.c_int32))
cg_multi_batch_transmissions = _sig("cg_multi_batch_transmissions", C.c_int32, P(c_sys), P(c_sig), C.c_int64,
                                    C.c_int32, P(C.c_double), P(C.c_int32), P(BatchStats))
cg_multi_transmissions = _sig("cg_multi_transmissions", C.c_int32, c_sys, C.c_double, C.c_int32, P(C.c_double))
EXPORTS += ["cg_multi_create", "cg_multi_model_create", "cg_multi_lead_count", "cg_multi_lead",
            "cg_multi_self_energies", "cg_multi_sigma", "cg_multi_labels", "cg_multi_batch_transmissions", "cg_multi_transmissions"]
cg_multi_smatrix = _sig("cg_multi_smatrix", C.c_int32, c_sys, C.c_double, P(C.c_int32), P(C.c_double), P(C.c_double),
                        P(C.c_double))
cg_multi_scattering_states = _sig("cg_multi_scattering_states", C.c_int32, c_sys, c_sig, P(C.c_int64),
                                  P(P(C.c_double)))
cg_multi_site_densities = _sig("cg_multi_site_densities", C.c_int32, c_sys, C.c_int32, P(P(C.c_double)),
                               P(C.c_int64), P(C.c_double), P(C.c_double), C.c_int32, P(C.c_double))
cg_multi_bond_currents = _sig("cg_multi_bond_currents", C.c_int32, c_sys, C.c_int32, P(P(C.c_double)), P(C.c_int64),
                              P(C.c_double), P(C.c_double), C.c_int32, P(C.c_int64), P(C.c_int64), P(C.c_int64),
                              P(C.c_double))
EXPORTS += ["cg_multi_smatrix", "cg_multi_scattering_states", "cg_multi_site_densities", "cg_multi_bond_currents"]
cg_multi_disordered_create = _sig("cg_multi_disordered_create", C.c_int32, C.c_int32, P(C.c_uint64), C.c_int32,
                                  P(C.c_int32), P(C.c_int32), C.c_int64, P(Disorder), C.c_uint64, P(c_sys))
EXPORTS += ["cg_multi_disordered_create"]
cg_last_error = _sig("cg_last_error", C.c_char_p)
EXPORTS += ["cg_last_error"]
