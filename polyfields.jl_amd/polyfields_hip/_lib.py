"""ctypes binding of ``libpolyfields_hip.so`` (declarations in include/polyfields_hip.h).

The library is loaded from this package directory.  There is no fallback: if it is not
built, importing the binding raises, and every GPU entry point fails loudly.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpolyfields_hip.so")

_lib = None

P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I = C.c_int
D = C.c_double
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
S = C.c_char_p


class ScftOptions(C.Structure):
    _fields_ = [("nsteps", I), ("nsout", I), ("nsave", I), ("ftol", D), ("stol", D)]


class ScftResults(C.Structure):
    _fields_ = [("time", D), ("nsteps", I), ("converged", I), ("ferr", D), ("serr", D),
                ("fhelm", D)]


SaveFn = C.CFUNCTYPE(None, C.c_void_p, C.c_int)


SIGNATURES = {
    "pf_last_error": (C.c_char_p, []),
    "pf_version": (I, []),
    "pf_device_count": (I, [IP]),
    "pf_cell_geometry": (I, [I, S, DP, I, DP, DP, DP, DP]),
    "pf_chain_grid": (I, [I, DP, I, IP, DP, DP]),
    "pf_system_create": (I, [PP, I, I, I, I, S, DP, I, I, D, I, I]),
    "pf_system_destroy": (I, [P]),
    "pf_system_add_monomer": (I, [P, I, D, D]),
    "pf_system_add_multiblock": (I, [P, I, IP, IP, DP, I, D, IP, IP]),
    "pf_system_add_point": (I, [P, I, D, IP]),
    "pf_system_set_flory": (I, [P, I, I, D]),
    "pf_system_set_edwards": (I, [P, I, D]),
    "pf_system_set_field": (I, [P, I, DP]),
    "pf_system_fieldinit_seeded": (I, [P, C.c_uint64, D]),
    "pf_julia_randn": (I, [C.c_uint64, C.POINTER(C.c_int64), I, DP]),
    "pf_system_get_grid": (I, [P, I, I, DP]),
    "pf_system_set_density": (I, [P, I, DP]),
    "pf_uniform_fields": (I, [P]),
    "pf_system_get_species_density": (I, [P, I, I, DP]),
    "pf_system_get_propagator": (I, [P, I, I, I, DP]),
    "pf_system_species_scalars": (I, [P, I, DP, DP, DP]),
    "pf_system_cell": (I, [P, DP, I, I]),
    "pf_system_add_cell": (I, [P, I, S, DP, I]),
    "pf_system_synchronize": (I, [P]),
    "pf_system_set_precision": (I, [P, I]),
    "pf_system_set_mde": (I, [P, I]),
    "pf_validate": (I, [P]),
    "pf_density": (I, [P]),
    "pf_residuals": (I, [P]),
    "pf_field_error": (I, [P, DP]),
    "pf_free_energy": (I, [P, DP]),
    "pf_fhelm_grad": (I, [P, I, DP, DP, DP, I, DP, DP, DP, DP, I]),
    "pf_density_vjp": (I, [P, DP, DP, DP, DP, I, DP, I]),
    "pf_species_set_blocks": (I, [P, I, DP, I]),
    "pf_system_set_checkpoint": (I, [P, I]),
    "pf_free_energy_bulk": (I, [P, DP]),
    "pf_scfstress": (I, [P, DP, I]),
    "pf_scfstress_noalloc": (I, [P, DP, I]),
    "pf_updater_anderson": (I, [PP, I, I, D]),
    "pf_updater_euler": (I, [PP, I, D]),
    "pf_updater_momentum": (I, [PP, I, D, D, D, D]),
    "pf_updater_setup": (I, [P, P]),
    "pf_updater_step": (I, [P]),
    "pf_updater_destroy": (I, [P]),
    "pf_domain_create": (I, [PP, I, I, D, D, D]),
    "pf_domain_setup": (I, [P, P]),
    "pf_domain_set_fd_step": (I, [P, D]),
    "pf_domain_step": (I, [P, I, D]),
    "pf_domain_stress": (I, [P, I, DP, I, DP]),
    "pf_domain_destroy": (I, [P]),
    "pf_scft": (I, [P, P, P, C.POINTER(ScftOptions), C.POINTER(ScftResults)]),
    "pf_scft_save": (I, [P, P, P, C.POINTER(ScftOptions), C.POINTER(ScftResults), SaveFn, P]),
    "pf_batch_create": (I, [PP, C.POINTER(C.c_void_p), I]),
    "pf_batch_destroy": (I, [P]),
    "pf_batch_density": (I, [P]),
    "pf_batch_residuals": (I, [P]),
    "pf_batch_sweeps": (I, [P]),
    "pf_batch_set_precision": (I, [P, I]),
    "pf_batch_scft": (I, [P, I, I, D, I, D, IP, C.POINTER(ScftResults)]),
    "pf_batch_scft_domain": (I, [P, I, I, D, P, I, D, D, IP, C.POINTER(ScftResults)]),
    "pf_batch_scft_momentum": (I, [P, I, D, D, D, D, P, I, D, D, IP, C.POINTER(ScftResults)]),
    "pf_batch_scft_euler": (I, [P, I, D, P, I, D, D, IP, C.POINTER(ScftResults)]),
    "pf_batch_timing": (I, [P, I]),
    "pf_batch_set_graph": (I, [P, I]),
    "pf_system_autotune": (I, [P, I, IP, DP]),
    "pf_batch_autotune": (I, [P, I, IP, I, IP]),
    "pf_schedule_reset": (I, []),
    "pf_schedule_set": (I, [I, I, I, I, IP]),
    "pf_batch_timing_lanes": (I, [P, I, DP]),
    "pf_batch_timing_read": (I, [P, I, DP, IP]),
    "pf_system_timing": (I, [P, I]),
    "pf_system_timing_read": (I, [P, I, DP, IP]),
}


class PolyFieldsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """Load (once) and return the library with its prototypes declared."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m polyfields_hip._build` from "
                "polyfields.jl_amd/ (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Call ``name`` and raise PolyFieldsError on a negative status."""
    L = lib()
    rc = getattr(L, name)(*args)
    if isinstance(rc, int) and rc < 0:
        msg = L.pf_last_error().decode(errors="replace")
        if rc == -4:
            raise ZeroDivisionError(msg)
        raise PolyFieldsError(rc, msg)
    return rc


def dptr(a):
    return a.ctypes.data_as(DP)


def iptr(a):
    return a.ctypes.data_as(IP)
