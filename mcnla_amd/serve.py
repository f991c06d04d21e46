"""Serving loop: continuous batching of iSVD requests on one GPU (mcnla_server_* in the C ABI).

The server owns one device arena of ``hbm_budget`` bytes and ``slots`` workers (a host thread and
a HIP stream each).  ``submit`` never blocks: a request waits in a FIFO queue until a worker is free
and the arena has room for its workspaces and results, so requests that arrive while others run
are admitted as slots and memory free up, and admitted solves overlap on the device.  Results stay
in the arena until ``release``.  No counterpart in the reference, which runs one solve per job
(src/isvd_dense_cb.cpp:46-302).
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray
from .isvd import (ColBlockGaussianProjectionSketcher, RowBlockKolmogorovNagumoIntegrator,
                   SvdFormer)


class Served:
    """Outcome of one request: status, iteration count, timings and device views of the results
    (valid until the request is released)."""

    def __init__(self, params, former_kind, c):
        self.status = c.status
        self.iterations = c.iterations
        self.queue_s, self.run_s, self.latency_s = c.queue_s, c.run_s, c.latency_s
        self.error = c.error.decode() if c.error else ""
        m, n, k, l = params.nrow, params.ncol, params.rank, params.rank + params.over_rank
        u_order = "F" if former_kind == _lib.FORM_SVD else "C"
        self.S = DeviceArray.view(c.S, (k,)) if c.S else None
        self.U = DeviceArray.view(c.U, (m, k), order=u_order) if c.U else None
        self.V = DeviceArray.view(c.V, (n, k)) if c.V else None
        self.Qbar = DeviceArray.view(c.Qbar, (m, l)) if c.Qbar else None

    @property
    def ok(self):
        return self.status in (_lib.OK, _lib.ERR_NOT_CONVERGED)


class Server:
    def __init__(self, slots=8, hbm_budget=16 << 30, max_bypass=8, batch_max=None,
                 sketcher=ColBlockGaussianProjectionSketcher,
                 integrator=RowBlockKolmogorovNagumoIntegrator, former=SvdFormer,
                 precision=_lib.PREC_F64, max_iteration=256, tolerance=None, jobv=True,
                 omega_mode=_lib.OMEGA_PHILOX, omega_threads=1):
        cfg = _lib.ServerConfig()
        _lib.lib().mcnla_server_config_default(ctypes.byref(cfg))
        cfg.solver.sketcher = sketcher._kind
        cfg.solver.integrator = integrator._kind
        cfg.solver.former = former._kind
        cfg.solver.precision = precision
        cfg.solver.jobv = 1 if jobv else 0
        cfg.solver.maxiter = int(max_iteration)
        cfg.solver.tol = integrator._default_tol if tolerance is None else float(tolerance)
        cfg.solver.omega_mode = omega_mode
        cfg.solver.omega_threads = int(omega_threads)
        cfg.slots = int(slots)
        cfg.hbm_budget = int(hbm_budget)
        cfg.max_bypass = int(max_bypass)
        if batch_max is not None:
            cfg.batch_max = int(batch_max)
        self.config = cfg
        self.former_kind = former._kind
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().mcnla_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h.value
        self._params = {}
        self._hold = {}

    def _request(self, A, m, n, k, p, N, seed, lda=None, order=None, ready_event=None):
        r = _lib.Request()
        r.params = _lib.Params(int(m), int(n), int(k), int(p), int(N))
        r.seed = int(seed)
        r.ready_event = ready_event
        if isinstance(A, DeviceArray):
            r.A = A.ptr
            r.lda = int(lda if lda is not None else A.ld)
            r.a_dtype = A.dtype_code
            r.a_order = A.order_code
            r.a_on_host = 0
            host = None
        else:
            host = np.asarray(A)
            if host.dtype not in (np.float32, np.float64):
                host = host.astype(np.float64)
            fortran = host.flags["F_CONTIGUOUS"] and not host.flags["C_CONTIGUOUS"]
            if order is None:
                order = "F" if fortran else "C"
            host = np.require(host, requirements=[order])
            r.A = host.ctypes.data
            r.lda = int(host.shape[0] if order == "F" else host.shape[1])
            r.a_dtype = _lib.F32 if host.dtype == np.float32 else _lib.F64
            r.a_order = _lib.COL_MAJOR if order == "F" else _lib.ROW_MAJOR
            r.a_on_host = 1
        return r, host

    def request_bytes(self, A, m, n, k, p, N, lda=None):
        r, _ = self._request(A, m, n, k, p, N, 0, lda)
        b = ctypes.c_size_t()
        _lib.check(_lib.lib().mcnla_server_request_bytes(self._h, ctypes.byref(r), ctypes.byref(b)))
        return b.value

    def submit(self, A, m, n, k, p, N, seed=0, lda=None, ready_event=None):
        """Queues an m x n iSVD of A (a DeviceArray, of which the leading m x n block is used with
        leading dimension ``lda``, or a host array that is copied into the arena); returns an id.
        ``ready_event`` (an event handle): the solve waits for it on the device before reading
        A, so A may still be being written on another stream."""
        r, host = self._request(A, m, n, k, p, N, seed, lda, ready_event=ready_event)
        rid = ctypes.c_uint64()
        _lib.check(_lib.lib().mcnla_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        self._params[rid.value] = r.params
        if host is not None:
            self._hold[rid.value] = host  # the host copy is read when the request is admitted
        return rid.value

    def wait(self, rid, timeout=-1.0):
        done = ctypes.c_int()
        _lib.check(_lib.lib().mcnla_server_wait(self._h, rid, float(timeout), ctypes.byref(done)))
        return bool(done.value)

    def next_done(self, timeout=-1.0):
        """Id of the next finished request (completion order), or None on timeout."""
        rid = ctypes.c_uint64()
        _lib.check(_lib.lib().mcnla_server_next_done(self._h, float(timeout), ctypes.byref(rid)))
        return rid.value or None

    def result(self, rid):
        c = _lib.Served()
        _lib.check(_lib.lib().mcnla_server_result(self._h, rid, ctypes.byref(c)))
        self._hold.pop(rid, None)
        return Served(self._params[rid], self.former_kind, c)

    def release(self, rid):
        _lib.check(_lib.lib().mcnla_server_release(self._h, rid))
        self._params.pop(rid, None)
        self._hold.pop(rid, None)

    def stats(self):
        s = _lib.ServerStats()
        _lib.check(_lib.lib().mcnla_server_get_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in _lib.ServerStats._fields_}

    def reset_stats(self):
        _lib.check(_lib.lib().mcnla_server_reset_stats(self._h))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().mcnla_server_destroy(self._h)
            self._h = None
            self._hold.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def required_bytes(m, n, k, p, N, precision=_lib.PREC_F64,
                   integrator=RowBlockKolmogorovNagumoIntegrator, former=SvdFormer, jobv=True,
                   route=_lib.ROUTE_AUTO):
    """Device bytes of a solver of this shape (no device needed)."""
    cfg = _lib.SolverConfig()
    _lib.lib().mcnla_solver_config_default(ctypes.byref(cfg))
    cfg.route = int(route)
    cfg.integrator = integrator._kind
    cfg.former = former._kind
    cfg.precision = precision
    cfg.jobv = 1 if jobv else 0
    prm = _lib.Params(int(m), int(n), int(k), int(p), int(N))
    b = ctypes.c_size_t()
    _lib.check(_lib.lib().mcnla_solver_required_bytes(ctypes.byref(prm), ctypes.byref(cfg),
                                                      ctypes.byref(b)))
    return b.value
