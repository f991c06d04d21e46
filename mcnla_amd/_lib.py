"""ctypes binding of libmcnla_amd.so (the C ABI declared in include/mcnla_amd/isvd_capi.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C mcnla_amd/csrc``).  There
is no fallback: if the shared library is missing, importing any compute entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmcnla_amd.so")

# mcnla_status
OK, ERR_INVALID, ERR_NOMEM, ERR_HIP, ERR_NOT_CONVERGED, ERR_NUMERIC, ERR_STATE, ERR_NO_DEVICE = range(8)
# mcnla_dtype / mcnla_order
F64, F32 = 0, 1
COL_MAJOR, ROW_MAJOR = 0, 1
# mcnla_precision
(PREC_F64, PREC_MIXED_BF16, PREC_MIXED_BF16X3, PREC_MIXED_F16, PREC_MIXED_F16X2,
 PREC_MIXED_F16X3) = range(6)
# kinds
(SKETCHER_COL_BLOCK_GAUSSIAN, SKETCHER_ROW_BLOCK_GAUSSIAN, SKETCHER_GAUSSIAN,
 SKETCHER_COLUMN_SAMPLING, SKETCHER_ROW_BLOCK_COLUMN_SAMPLING) = range(5)
OMEGA_PHILOX, OMEGA_REFERENCE, OMEGA_PHILOX_F16 = 0, 1, 2
ROUTE_AUTO, ROUTE_SKETCH, ROUTE_GRAM_OF_A = 0, 1, 2
KERNEL_SKETCH, KERNEL_GRAM, KERNEL_FORMER_Z = 0, 1, 2
ORTH_GRAMIAN, ORTH_ROW_BLOCK_GRAMIAN = 0, 1
(INT_KN, INT_ROW_BLOCK_KN, INT_ROW_BLOCK_GRAMIAN_KN, INT_ROW_BLOCK_WY, INT_ROW_BLOCK_GRAMIAN_WY,
 INT_ROW_BLOCK_REDUCTION, INT_ROW_BLOCK_EXTRINSIC_MEAN) = range(7)
(FORM_SVD, FORM_GRAMIAN, FORM_COL_BLOCK_GRAMIAN, FORM_ROW_BLOCK_GRAMIAN,
 FORM_ROW_BLOCK_SYMMETRIC) = range(5)

# MCNLA_STATIC_WORKSPACE_BYTES: the static device array mcnla_solve_static carves its workspace from
STATIC_WORKSPACE_BYTES = 256 << 20
MAX_BATCH = 16

STATUS_NAMES = {
    OK: "OK", ERR_INVALID: "INVALID", ERR_NOMEM: "NOMEM", ERR_HIP: "HIP",
    ERR_NOT_CONVERGED: "NOT_CONVERGED", ERR_NUMERIC: "NUMERIC", ERR_STATE: "STATE",
    ERR_NO_DEVICE: "NO_DEVICE",
}


class McnlaError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [("nrow", ctypes.c_int64), ("ncol", ctypes.c_int64), ("rank", ctypes.c_int64),
                ("over_rank", ctypes.c_int64), ("num_sketch", ctypes.c_int64)]


class SolverConfig(ctypes.Structure):
    _fields_ = [("sketcher", ctypes.c_int), ("orthogonalizer", ctypes.c_int),
                ("integrator", ctypes.c_int), ("former", ctypes.c_int),
                ("precision", ctypes.c_int), ("jobv", ctypes.c_int),
                ("maxiter", ctypes.c_int64), ("tol", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("omega_mode", ctypes.c_int),
                ("omega_threads", ctypes.c_int), ("route", ctypes.c_int),
                ("graph", ctypes.c_int)]


class Tuning(ctypes.Structure):
    """Split-K workgroup targets of a solve's FP64 GEMMs and their kernels' tile shapes (a bit
    per GEMM in `kernels`: 1 gram, 2 former, 4 h_omega, 8 omega_p, 16 n_by_l, 32 qbar)."""
    _fields_ = [("gram", ctypes.c_int64), ("former", ctypes.c_int64), ("h_omega", ctypes.c_int64),
                ("omega_p", ctypes.c_int64), ("n_by_l", ctypes.c_int64), ("qbar", ctypes.c_int64),
                ("kernels", ctypes.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("solver", SolverConfig), ("slots", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("max_bypass", ctypes.c_int), ("batch_max", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("params", Params), ("A", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("a_dtype", ctypes.c_int), ("a_order", ctypes.c_int), ("a_on_host", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("ready_event", ctypes.c_void_p)]


class Served(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int64),
                ("S", ctypes.c_void_p), ("U", ctypes.c_void_p), ("V", ctypes.c_void_p),
                ("Qbar", ctypes.c_void_p), ("queue_s", ctypes.c_double),
                ("run_s", ctypes.c_double), ("latency_s", ctypes.c_double),
                ("error", ctypes.c_char_p)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("max_running", ctypes.c_int),
                ("peak_bytes", ctypes.c_size_t), ("budget_bytes", ctypes.c_size_t),
                ("p50_s", ctypes.c_double), ("p99_s", ctypes.c_double),
                ("mean_s", ctypes.c_double), ("max_s", ctypes.c_double),
                ("span_s", ctypes.c_double), ("batches", ctypes.c_uint64),
                ("kn_chunks", ctypes.c_uint64), ("kn_group_mean", ctypes.c_double),
                ("kn_chunk_mean_s", ctypes.c_double)]


# Every exported symbol and its signature; tests check the library exports exactly these.
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_d = ctypes.c_double
_sz = ctypes.c_size_t
_dp = ctypes.POINTER(ctypes.c_double)
_st = ctypes.c_int  # mcnla_status

SIGNATURES = {
    "mcnla_last_error": (ctypes.c_char_p, []),
    "mcnla_version": (ctypes.c_char_p, []),
    "mcnla_device_count": (_st, [ctypes.POINTER(ctypes.c_int)]),
    "mcnla_set_device": (_st, [_i]),
    "mcnla_malloc": (_st, [_pp, _sz]),
    "mcnla_free": (_st, [_vp]),
    "mcnla_memcpy_htod": (_st, [_vp, _vp, _sz]),
    "mcnla_memcpy_dtoh": (_st, [_vp, _vp, _sz]),
    "mcnla_memcpy_dtod": (_st, [_vp, _vp, _sz, _vp]),
    "mcnla_memset": (_st, [_vp, _i, _sz]),
    "mcnla_stream_create": (_st, [_pp]),
    "mcnla_stream_destroy": (_st, [_vp]),
    "mcnla_stream_synchronize": (_st, [_vp]),
    "mcnla_device_synchronize": (_st, []),
    "mcnla_event_create": (_st, [_pp]),
    "mcnla_event_destroy": (_st, [_vp]),
    "mcnla_event_record": (_st, [_vp, _vp]),
    "mcnla_event_synchronize": (_st, [_vp]),
    "mcnla_memcpy_htod_async": (_st, [_vp, _vp, _sz, _vp]),
    "mcnla_host_malloc": (_st, [_pp, _sz]),
    "mcnla_host_free": (_st, [_vp]),
    "mcnla_mem_info": (_st, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "mcnla_sketcher_create": (_st, [ctypes.POINTER(Params), _i, _i, _pp]),
    "mcnla_sketcher_set_seed": (_st, [_vp, _u64, _i, _i]),
    "mcnla_sketcher_run": (_st, [_vp, _vp, _i64, _i, _i, _vp, _vp]),
    "mcnla_sketcher_omega": (_st, [_vp, _pp]),
    "mcnla_sketcher_set_exponent": (_st, [_vp, _i64]),
    "mcnla_sketcher_indices": (_st, [_vp, _pp]),
    "mcnla_omega_reference": (_st, [_u64, _i, _i, _i64, _dp]),
    "mcnla_column_indices_reference": (_st, [_u64, _i, _i, _i64, _i64, ctypes.POINTER(_i64)]),
    "mcnla_orthogonalizer_create": (_st, [ctypes.POINTER(Params), _i, _i, _pp]),
    "mcnla_orthogonalizer_run": (_st, [_vp, _vp, _vp]),
    "mcnla_integrator_create": (_st, [ctypes.POINTER(Params), _i, _i, _i64, _d, _pp]),
    "mcnla_integrator_run": (_st, [_vp, _vp, _vp, _vp]),
    "mcnla_integrator_iteration": (_st, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "mcnla_former_create": (_st, [ctypes.POINTER(Params), _i, _i, _i, _pp]),
    "mcnla_former_run": (_st, [_vp, _vp, _i64, _i, _i, _vp, _vp]),
    "mcnla_former_results": (_st, [_vp, _pp, _pp, _pp]),
    "mcnla_stage_destroy": (_st, [_vp]),
    "mcnla_stage_time": (_st, [_vp, _dp, _dp, _i, ctypes.POINTER(ctypes.c_int)]),
    "mcnla_stage_name": (ctypes.c_char_p, [_vp]),
    "mcnla_stage_part_names": (ctypes.c_char_p, [_vp]),
    "mcnla_solver_config_default": (None, [ctypes.POINTER(SolverConfig)]),
    "mcnla_solver_create": (_st, [ctypes.POINTER(Params), ctypes.POINTER(SolverConfig), _pp]),
    "mcnla_solver_destroy": (_st, [_vp]),
    "mcnla_solver_workspace_bytes": (_st, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "mcnla_solver_set_seed": (_st, [_vp, _u64]),
    "mcnla_solver_solve": (_st, [_vp, _vp, _i64, _i, _i, _vp]),
    "mcnla_solver_solve_host": (_st, [_vp, _vp, _i64, _i, _i, _i64, _vp]),
    "mcnla_solver_solve_batch": (_st, [_i, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       _vp]),
    "mcnla_solver_results": (_st, [_vp, _pp, _pp, _pp, _pp]),
    "mcnla_solver_iteration": (_st, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "mcnla_solver_stage_times": (_st, [_vp, _dp, _dp]),
    "mcnla_solver_kernel_times": (_st, [_vp, _dp]),
    "mcnla_solver_kernel_time": (_st, [_vp, ctypes.c_int, _dp]),
    "mcnla_tuning_default": (None, [ctypes.POINTER(Tuning)]),
    "mcnla_solver_get_tuning": (_st, [_vp, ctypes.POINTER(Tuning)]),
    "mcnla_solver_autotune": (_st, [_vp, _vp, _i64, _i, _i, _i, _vp, ctypes.POINTER(Tuning)]),
    "mcnla_tune_set": (_st, [ctypes.POINTER(Params), ctypes.POINTER(SolverConfig),
                             ctypes.POINTER(Tuning)]),
    "mcnla_tune_get": (_st, [ctypes.POINTER(Params), ctypes.POINTER(SolverConfig),
                             ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_int)]),
    "mcnla_tune_clear": (None, []),
    "mcnla_tune_save": (_st, [ctypes.c_char_p]),
    "mcnla_tune_load": (_st, [ctypes.c_char_p]),
    "mcnla_solver_required_bytes": (_st, [ctypes.POINTER(Params), ctypes.POINTER(SolverConfig),
                                          ctypes.POINTER(ctypes.c_size_t)]),
    "mcnla_solve_static": (_st, [ctypes.POINTER(Params), ctypes.POINTER(SolverConfig), _vp, _i64,
                                 _i, _i, _vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_int64), _vp]),
    "mcnla_server_config_default": (None, [ctypes.POINTER(ServerConfig)]),
    "mcnla_server_create": (_st, [ctypes.POINTER(ServerConfig), _pp]),
    "mcnla_server_destroy": (_st, [_vp]),
    "mcnla_server_request_bytes": (_st, [_vp, ctypes.POINTER(Request),
                                         ctypes.POINTER(ctypes.c_size_t)]),
    "mcnla_server_submit": (_st, [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    "mcnla_server_wait": (_st, [_vp, _u64, _d, ctypes.POINTER(ctypes.c_int)]),
    "mcnla_server_next_done": (_st, [_vp, _d, ctypes.POINTER(ctypes.c_uint64)]),
    "mcnla_server_result": (_st, [_vp, _u64, ctypes.POINTER(Served)]),
    "mcnla_server_release": (_st, [_vp, _u64]),
    "mcnla_server_get_stats": (_st, [_vp, ctypes.POINTER(ServerStats)]),
    "mcnla_server_reset_stats": (_st, [_vp]),
    "mcnla_fill_gaussian": (_st, [_vp, _i64, _u64, _d, _vp]),
    "mcnla_synth_reflected": (_st, [_i64, _i64, _vp, _vp, _vp, _vp, _i64, _i, _i, _vp]),
    "mcnla_synth_lowrank": (_st, [_i64, _i64, _i64, _vp, _vp, _vp, _d, _u64, _vp, _i64, _i, _i,
                                  _vp]),
}

_LIB = None


def lib():
    """Loads libmcnla_amd.so once; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libmcnla_amd.so is not built (expected at %s); run __graft_entry__.build() or "
                "`make -C mcnla_amd/csrc`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(status, allow=()):
    """Raises McnlaError for a non-OK status not listed in `allow`; returns the status."""
    if status != OK and status not in allow:
        msg = lib().mcnla_last_error()
        raise McnlaError(status, msg.decode() if msg else "")
    return status
