"""ctypes binding of libipgpu.so (declarations mirror include/ipgpu.h).

The product path has no CPU fallback: if the library is missing or no GPU
is present, calls raise IPGError.
"""
import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libipgpu.so")

IPG_OK = 0
IPG_ERR_INVALID = -1
IPG_ERR_CAPACITY = -2
IPG_ERR_HIP = -3
IPG_ERR_UNSUPPORTED = -4
IPG_ERR_NOMEM = -5

IPG_LIST_HALF, IPG_LIST_FULL = 0, 1
IPG_PAIR_LJ, IPG_PAIR_MORSE, IPG_PAIR_ZBL, IPG_PAIR_BORN_MAYER, IPG_PAIR_COULOMB, IPG_PAIR_MOCK = range(6)
IPG_FP64, IPG_FP32 = 0, 1
IPG_SNAP_LIST_REFERENCE, IPG_SNAP_LIST_CUTOFF = 0, 1
IPG_SNAP_PLAIN_FP64, IPG_SNAP_ADJOINT_FP64, IPG_SNAP_ADJOINT_FP32 = 0, 1, 2

# Every symbol include/ipgpu.h declares (checked by the CPU tests).
EXPORTED = [
    "ipg_get_version", "ipg_last_error", "ipg_ctx_create", "ipg_ctx_destroy", "ipg_ctx_synchronize",
    "ipg_device_malloc", "ipg_device_free", "ipg_memcpy_h2d", "ipg_memcpy_d2h", "ipg_prof_enable",
    "ipg_prof_reset", "ipg_prof_query", "ipg_neighborlist", "ipg_nlist_build_dev", "ipg_nlist_info",
    "ipg_nlist_destroy", "ipg_pair_energy_forces", "ipg_pair_energy_forces_dev", "ipg_pair_eval_dev",
    "ipg_snap_create", "ipg_snap_destroy", "ipg_snap_info", "ipg_snap_neighborlist", "ipg_snap_local_descriptors",
    "ipg_snap_force_descriptors", "ipg_snap_all_descriptors", "ipg_snap_energy_forces",
    "ipg_snap_energy_forces_dev", "ipg_snap_force_vjp", "ipg_ctx_set_graphs", "ipg_snap_graph_info",
    "ipg_device_copy", "ipg_snap_batch_descriptors", "ipg_snap_batch_descriptors_dev", "ipg_snap_batch_info",
    "ipg_nlist_build_skin_dev", "ipg_nlist_update_dev", "ipg_prof_select", "ipg_snap_virial_stress",
    "ipg_snap_energy_forces_virial", "ipg_pair_eval_noalloc_dev",
]


class IPGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


class System(C.Structure):
    _fields_ = [
        ("natoms", C.c_int64),
        ("positions", C.c_void_p),
        ("species", C.c_void_p),
        ("box", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
    ]


class PairTerm(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("reserved", C.c_int32),
        ("coeff", C.c_double),
        ("rcut", C.c_double),
        ("p", C.c_double * 6),
        ("species_mask", C.c_uint64),
    ]


class SnapParams(C.Structure):
    _fields_ = [
        ("twojmax", C.c_int32),
        ("nspecies", C.c_int32),
        ("rcutfac", C.c_double),
        ("rmin0", C.c_double),
        ("rfac0", C.c_double),
        ("radii", C.POINTER(C.c_double)),
        ("weight", C.POINTER(C.c_double)),
        ("chem_flag", C.c_int32),
        ("bzero_flag", C.c_int32),
        ("bnorm_flag", C.c_int32),
        ("switch_flag", C.c_int32),
        ("wselfall_flag", C.c_int32),
    ]


_lib = None
_lock = threading.RLock()  # default_context() calls load() while holding it


def load():
    """Load libipgpu.so (raises IPGError if it was not built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise IPGError(IPG_ERR_INVALID, "libipgpu.so not built (run interatomicpotentials.jl_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        P, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        pS = C.POINTER(System)
        sig = {
            "ipg_get_version": (C.c_int, []),
            "ipg_last_error": (C.c_char_p, []),
            "ipg_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
            "ipg_ctx_destroy": (C.c_int, [P]),
            "ipg_ctx_synchronize": (C.c_int, [P]),
            "ipg_device_malloc": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
            "ipg_device_free": (C.c_int, [P, P]),
            "ipg_memcpy_h2d": (C.c_int, [P, P, P, C.c_size_t]),
            "ipg_memcpy_d2h": (C.c_int, [P, P, P, C.c_size_t]),
            "ipg_prof_enable": (C.c_int, [P, C.c_int]),
            "ipg_prof_reset": (C.c_int, [P]),
            "ipg_prof_select": (C.c_int, [P, C.c_char_p]),
            "ipg_prof_query": (C.c_int, [P, C.c_char_p, C.POINTER(i64), C.POINTER(d)]),
            "ipg_neighborlist": (C.c_int, [P, pS, d, C.c_int, P, P, P, i64, C.POINTER(i64)]),
            "ipg_nlist_build_dev": (C.c_int, [P, pS, d, C.c_int, C.POINTER(P)]),
            "ipg_nlist_info": (C.c_int, [P, C.POINTER(i64), C.POINTER(i64), C.POINTER(P), C.POINTER(P)]),
            "ipg_nlist_destroy": (C.c_int, [P]),
            "ipg_pair_energy_forces": (C.c_int, [P, pS, P, C.c_int, C.c_int, P, P, P]),
            "ipg_pair_energy_forces_dev": (C.c_int, [P, pS, P, C.c_int, C.c_int, P, P, P]),
            "ipg_pair_eval_dev": (C.c_int, [P, P, pS, P, C.c_int, C.c_int, P, P, P]),
            "ipg_pair_eval_noalloc_dev": (C.c_int, [P, P, pS, P, C.c_int, C.c_int, P, P, P]),
            "ipg_snap_create": (C.c_int, [P, C.POINTER(SnapParams), C.POINTER(P)]),
            "ipg_snap_destroy": (C.c_int, [P]),
            "ipg_snap_info": (C.c_int, [P, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
            "ipg_snap_neighborlist": (C.c_int, [P, pS, C.c_int, P, P, P, P, P, i64, C.POINTER(i64)]),
            "ipg_snap_local_descriptors": (C.c_int, [P, pS, C.c_int, P, P]),
            "ipg_snap_force_descriptors": (C.c_int, [P, pS, C.c_int, P, P]),
            "ipg_snap_all_descriptors": (C.c_int, [P, pS, C.c_int, P, P, P, P]),
            "ipg_snap_energy_forces": (C.c_int, [P, pS, C.c_int, P, P, i64, d, C.c_int, P, P]),
            "ipg_snap_energy_forces_dev": (C.c_int, [P, pS, C.c_int, P, d, C.c_int, P, P]),
            "ipg_snap_virial_stress": (C.c_int, [P, pS, C.c_int, P, P, i64, C.c_int, P]),
            "ipg_snap_energy_forces_virial": (C.c_int, [P, pS, C.c_int, P, P, i64, d, C.c_int, P, P, P]),
            "ipg_snap_force_vjp": (C.c_int, [P, pS, C.c_int, P, P, i64, P, C.c_int, P, P]),
            "ipg_ctx_set_graphs": (C.c_int, [P, C.c_int]),
            "ipg_device_copy": (C.c_int, [P, P, P, C.c_size_t]),
            "ipg_snap_graph_info": (C.c_int, [P, C.POINTER(i32), C.POINTER(i64), C.POINTER(i64), C.POINTER(i64),
                                              C.c_char_p, i64]),
            "ipg_snap_batch_descriptors": (C.c_int, [P, i64, pS, C.c_int, P, P, P, P]),
            "ipg_snap_batch_descriptors_dev": (C.c_int, [P, i64, P, P, P, P, P, C.c_int, P, P, P, P]),
            "ipg_nlist_build_skin_dev": (C.c_int, [P, pS, d, d, C.c_int, C.POINTER(P)]),
            "ipg_nlist_update_dev": (C.c_int, [P, pS, C.POINTER(i32)]),
            "ipg_snap_batch_info": (C.c_int, [P, C.c_int, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                              C.POINTER(i32), C.POINTER(i64), C.POINTER(i64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return _lib


def check(rc):
    if rc != IPG_OK:
        msg = load().ipg_last_error()
        raise IPGError(rc, msg.decode() if msg else "error")
    return rc


_ctx = None


def default_context():
    """Context on GPU 0, created on first use (one per process)."""
    global _ctx
    with _lock:
        if _ctx is None:
            L = load()
            h = C.c_void_p()
            rc = L.ipg_ctx_create(0, C.byref(h))
            if rc != IPG_OK:
                raise IPGError(rc, L.ipg_last_error().decode())
            _ctx = h
        return _ctx
