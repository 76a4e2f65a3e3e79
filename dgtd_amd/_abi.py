"""ctypes binding of libdgtd_hip.so (C ABI declared in include/dgtd_hip.h).

The library is the product: there is no Python or CPU fallback.  If it is
missing, importing this module raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# DGTD_LIB: another build of the same library (the sanitizer build of the host
# code, make -C dgtd_amd/csrc asan); by default the in-tree build.
LIB_PATH = os.environ.get("DGTD_LIB") or os.path.join(_HERE, "lib", "libdgtd_hip.so")

c_int32 = ctypes.c_int32
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i8p = ctypes.POINTER(ctypes.c_int8)
_i64p = ctypes.POINTER(ctypes.c_int64)

DGTD_OK = 0
DGTD_BC_NONE, DGTD_BC_PEC, DGTD_BC_PMC, DGTD_BC_SMA = 0, 1, 2, 3
DGTD_BC_INT_PEC, DGTD_BC_INT_PMC, DGTD_BC_INT_SMA, DGTD_BC_TFSF = 4, 5, 6, 7
DGTD_RK4, DGTD_LSERK4 = 0, 1
DGTD_KERNEL_AUTO, DGTD_KERNEL_VALU, DGTD_KERNEL_MFMA, DGTD_KERNEL_CSR = 0, 1, 2, 3


class DgtdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dgtd error {code}: {msg}")
        self.code = code


class MeshDesc(ctypes.Structure):
    _fields_ = [("K", c_int32), ("order", c_int32), ("vx", _dp), ("nv", c_int32),
                ("EToV", _i32p), ("face_bc", _i8p), ("eps_r", _dp), ("mu_r", _dp), ("dim", c_int32),
                ("sigma", _dp), ("nvert", c_int32)]


class Opts(ctypes.Structure):
    _fields_ = [("alpha", c_double), ("precision", c_int32), ("integrator", c_int32),
                ("device", c_int32), ("kernel", c_int32)]


class InitialField(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("field", c_int32), ("center", c_double * 3),
                ("polarization", c_double * 3), ("spread", c_double), ("dimension", c_int32),
                ("nmodes", c_int32), ("modes", c_int32 * 3)]


class PlaneWave(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("polarization", c_double * 3), ("propagation", c_double * 3),
                ("spread", c_double), ("mean", c_double), ("frequency", c_double)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("alpha", c_double), ("precision", c_int32), ("device", c_int32),
                ("max_instances", c_int32), ("reserved", c_int32), ("hbm_budget", ctypes.c_int64)]


class Tuning(ctypes.Structure):
    _fields_ = [("variant", c_int32), ("config", c_int32), ("waves_per_block", c_int32),
                ("gather_depth", c_int32), ("early_residual", c_int32), ("task_order", c_int32),
                ("from_cache", c_int32), ("ncandidates", c_int32), ("candidate_us", c_double * 8)]


class CaseInfo(ctypes.Structure):
    _fields_ = [("K", c_int32), ("nv", c_int32), ("order", c_int32), ("alpha", c_double),
                ("time_step", c_double), ("final_time", c_double), ("cfl", c_double),
                ("evolution_operator", c_int32), ("nsources", c_int32), ("nprobes", c_int32),
                ("exporter_steps", c_int32), ("refinement", c_int32), ("snapshot_steps", c_int32),
                ("snapshot_saves", c_int32), ("nplanewaves", c_int32), ("dim", c_int32),
                ("nfield_probes", c_int32)]


DGTD_E_CAPACITY = -6
DGTD_OPERATOR_MAXWELL, DGTD_OPERATOR_GLOBAL, DGTD_OPERATOR_HESTHAVEN = 0, 1, 2
DGTD_INST_FREE, DGTD_INST_RUNNING, DGTD_INST_DONE, DGTD_INST_FAILED, DGTD_INST_CANCELLED = 0, 1, 2, 3, 4

# name -> (restype, argtypes); every symbol declared in include/dgtd_hip.h.
SIGNATURES = {
    "dgtd_last_error": (ctypes.c_char_p, []),
    "dgtd_version": (ctypes.c_char_p, []),
    "dgtd_op_create": (c_int32, [ctypes.POINTER(MeshDesc), ctypes.POINTER(Opts),
                                 ctypes.POINTER(c_void_p)]),
    "dgtd_op_destroy": (None, [c_void_p]),
    "dgtd_op_ndofs": (c_int32, [c_void_p, _i64p]),
    "dgtd_op_np": (c_int32, [c_void_p, _i32p]),
    "dgtd_op_max_fscale": (c_int32, [c_void_p, _dp]),
    "dgtd_op_node_positions": (c_int32, [c_void_p, _dp]),
    "dgtd_op_mult": (c_int32, [c_void_p, c_double, _dp, _dp]),
    "dgtd_op_mult_device": (c_int32, [c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "dgtd_op_set_state": (c_int32, [c_void_p, _dp, c_double]),
    "dgtd_op_get_state": (c_int32, [c_void_p, _dp, _dp]),
    "dgtd_op_advance": (c_int32, [c_void_p, c_double, c_int32]),
    "dgtd_op_run": (c_int32, [c_void_p, c_double, c_double, _i32p]),
    "dgtd_op_time": (c_int32, [c_void_p, _dp]),
    "dgtd_op_step": (c_int32, [c_void_p, _dp, _dp, c_double, c_int32]),
    "dgtd_op_norms": (c_int32, [c_void_p, _dp]),
    "dgtd_op_unstable": (c_int32, [c_void_p, _i32p]),
    "dgtd_op_norms_noalloc": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "dgtd_op_set_planewaves": (c_int32, [c_void_p, ctypes.POINTER(PlaneWave), c_int32]),
    "dgtd_op_probe_set": (c_int32, [c_void_p, c_int32, _dp, _i32p]),
    "dgtd_op_probe_sample": (c_int32, [c_void_p, _dp]),
    "dgtd_op_set_timing": (c_int32, [c_void_p, c_int32]),
    "dgtd_op_last_kernel_us": (c_int32, [c_void_p, _dp, _i32p]),
    "dgtd_op_kernel_variant": (c_int32, [c_void_p, _i32p]),
    "dgtd_op_tuning": (c_int32, [c_void_p, ctypes.POINTER(Tuning)]),
    "dgtd_op_set_candidate": (c_int32, [c_void_p, c_int32]),
    "dgtd_op_work_per_element": (c_int32, [c_void_p, _dp, _dp]),
    "dgtd_cube_mesh": (c_int32, [c_int32, c_double, _i8p, _dp, _i32p, _i8p]),
    "dgtd_project_initial_fields": (c_int32, [ctypes.POINTER(MeshDesc), ctypes.POINTER(InitialField),
                                              c_int32, _dp]),
    "dgtd_case_load": (c_int32, [ctypes.c_char_p, ctypes.POINTER(c_void_p)]),
    "dgtd_case_destroy": (None, [c_void_p]),
    "dgtd_case_get_info": (c_int32, [c_void_p, ctypes.POINTER(CaseInfo)]),
    "dgtd_case_mesh": (c_int32, [c_void_p, ctypes.POINTER(MeshDesc)]),
    "dgtd_case_sources": (c_int32, [c_void_p, ctypes.POINTER(InitialField)]),
    "dgtd_case_probes": (c_int32, [c_void_p, _dp, _i32p, _i32p]),
    "dgtd_case_field_probes": (c_int32, [c_void_p, _dp, _i32p, _i32p, _i32p]),
    "dgtd_case_save_mesh": (c_int32, [c_void_p, ctypes.c_char_p]),
    "dgtd_case_time_step": (c_int32, [c_void_p, _dp]),
    "dgtd_op_estimate_time_step": (c_int32, [c_void_p, c_int32, c_double, _dp]),
    "dgtd_mesh_estimate_time_step": (c_int32, [ctypes.POINTER(MeshDesc), c_int32, c_double, _dp]),
    "dgtd_case_planewaves": (c_int32, [c_void_p, ctypes.POINTER(PlaneWave)]),
    "dgtd_save_fields": (c_int32, [ctypes.c_char_p, c_int32, ctypes.c_int64, _dp, c_double]),
    "dgtd_op_set_graphs": (c_int32, [c_void_p, c_int32]),
    "dgtd_case_save_fields": (c_int32, [c_void_p, ctypes.c_char_p, ctypes.c_int64, _dp, c_double]),
    "dgtd_batch_create": (c_int32, [ctypes.POINTER(BatchOpts), ctypes.POINTER(c_void_p)]),
    "dgtd_batch_destroy": (None, [c_void_p]),
    "dgtd_batch_footprint": (c_int32, [ctypes.POINTER(MeshDesc), c_int32, _i64p]),
    "dgtd_batch_add": (c_int32, [c_void_p, ctypes.POINTER(MeshDesc), _dp, c_double, c_double, _i32p]),
    "dgtd_batch_add_tfsf": (c_int32, [c_void_p, ctypes.POINTER(MeshDesc), _dp, c_double, c_double,
                                      ctypes.POINTER(PlaneWave), c_int32, _i32p]),
    "dgtd_batch_step": (c_int32, [c_void_p, c_int32]),
    "dgtd_batch_poll": (c_int32, [c_void_p, _i32p, c_int32, _i32p, _i32p]),
    "dgtd_batch_info": (c_int32, [c_void_p, c_int32, _dp, _i32p, _i32p, _dp]),
    "dgtd_batch_get_state": (c_int32, [c_void_p, c_int32, _dp]),
    "dgtd_batch_remove": (c_int32, [c_void_p, c_int32]),
    "dgtd_batch_cancel": (c_int32, [c_void_p, c_int32]),
    "dgtd_batch_stats": (c_int32, [c_void_p, _i64p, _i32p, _i32p, _i32p]),
    "dgtd_ref_element": (c_int32, [c_int32, _dp, _dp, _dp, _dp, _dp, _dp, _i32p]),
    "dgtd_mesh_vmaps": (c_int32, [ctypes.POINTER(MeshDesc), c_double, _i32p, _i32p, _i8p]),
    "dgtd_mesh_geometry": (c_int32, [ctypes.POINTER(MeshDesc), c_double, _dp]),
    "dgtd_global_assemble": (c_int32, [ctypes.POINTER(MeshDesc), c_double, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int64), _i32p, _dp]),
    "dgtd_mesh_locate": (c_int32, [ctypes.POINTER(MeshDesc), c_int32, _dp, _i32p, _dp]),
}

_lib = None


def lib():
    """Load libdgtd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                "g.build()' or make -C dgtd_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != DGTD_OK:
        raise DgtdError(rc, lib().dgtd_last_error().decode())
    return rc


def dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def i32ptr(a):
    return None if a is None else a.ctypes.data_as(_i32p)


def i8ptr(a):
    return None if a is None else a.ctypes.data_as(_i8p)
