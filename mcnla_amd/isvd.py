"""Python mirror of MCNLA's iSVD stage API over the C ABI.

Names, constructor arguments and call order follow the reference (emfomy/mcnla 0.7.2):
``Parameters`` (include/mcnla/isvd/core/parameters.hh:31-119), the stage classes selected by the
drivers' STYPE/OTYPE/ITYPE/FTYPE macros (src/isvd_dense_cb.cpp:20-34), the StageWrapper lifecycle
``initialize()`` / ``__call__`` / ``time()`` (include/mcnla/isvd/core/stage_wrapper.hh:35-101) and
the getters ``iteration()``, ``vectorS()``, ``matrixU()``, ``matrixV()``/``matrixVj()``.

All matrices are :class:`~mcnla_amd.device.DeviceArray` objects in HBM.  The compute always runs
on the GPU through libmcnla_amd; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray


class Parameters:
    """isvd::Parameters: sizes m, n, rank k, oversampling p, number of sketches N.
    On one GPU there is a single rank, so nrowRank() == nrow() and numSketchEach() == numSketch()."""

    def __init__(self, mpi_root=0, mpi_comm=None):
        self.mpi_root = mpi_root
        self.mpi_comm = mpi_comm
        self._m = self._n = self._k = 0
        self._p = 12
        self._N = 16
        self._synchronized = False

    def setSize(self, nrow, ncol=None):
        if ncol is None:
            nrow, ncol = nrow.shape
        self._m, self._n = int(nrow), int(ncol)
        self._synchronized = False
        return self

    def setRank(self, rank):
        self._k = int(rank)
        return self

    def setOverRank(self, over_rank):
        self._p = int(over_rank)
        return self

    def setNumSketch(self, num_sketch):
        self._N = int(num_sketch)
        return self

    def setNumSketchEach(self, num_sketch_each):
        return self.setNumSketch(num_sketch_each)

    def sync(self):
        self._synchronized = True

    def isSynchronized(self):
        return self._synchronized

    def nrow(self): return self._m
    def ncol(self): return self._n
    def nrowRank(self): return self._m
    def ncolRank(self): return self._n
    def nrowEach(self): return self._m
    def ncolEach(self): return self._n
    def nrowTotal(self): return self._m
    def ncolTotal(self): return self._n
    def rank(self): return self._k
    def overRank(self): return self._p
    def dimSketch(self): return self._k + self._p
    def dimSketchEach(self): return self._k + self._p
    def dimSketchTotal(self): return (self._k + self._p) * self._N
    def numSketch(self): return self._N
    def numSketchEach(self): return self._N

    def c_params(self):
        return _lib.Params(self._m, self._n, self._k, self._p, self._N)

    # parameters.hpp:288-363: buffers for the collection, Qbar, U, V (row-major)
    def createCollectionQ(self, dtype=np.float64):
        return DeviceArray((self._m, self.dimSketchTotal()), dtype)

    createCollectionQj = createCollectionQ
    createCollectionQjp = createCollectionQ

    def createMatrixQbar(self):
        return DeviceArray((self._m, self.dimSketch()))

    createMatrixQbarj = createMatrixQbar


class _Stage:
    _create = None

    def __init__(self, parameters, precision=_lib.PREC_F64):
        self.parameters = parameters
        self.precision = precision
        self._h = None
        self._computed = False

    def _make(self):
        raise NotImplementedError

    def initialize(self):
        if self._h is not None:
            _lib.lib().mcnla_stage_destroy(self._h)
        self._h = self._make()
        return self

    def _handle(self):
        if self._h is None:
            self.initialize()
        return self._h

    def time(self):
        tot = ctypes.c_double()
        _lib.check(_lib.lib().mcnla_stage_time(self._handle(), ctypes.byref(tot), None, 0, None))
        return tot.value

    def times(self):
        parts = (ctypes.c_double * 8)()
        n = ctypes.c_int()
        _lib.check(_lib.lib().mcnla_stage_time(self._handle(), None, parts, 8, ctypes.byref(n)))
        return list(parts[: n.value])

    def names(self):
        return _lib.lib().mcnla_stage_part_names(self._handle()).decode()

    def __str__(self):
        return _lib.lib().mcnla_stage_name(self._handle()).decode()

    def isInitialized(self):
        return self._h is not None

    def isComputed(self):
        return self._computed

    def __del__(self):
        try:
            if self._h is not None:
                _lib.lib().mcnla_stage_destroy(self._h)
        except Exception:
            pass


def _out_handle(fn, *args):
    h = ctypes.c_void_p()
    _lib.check(fn(*args, ctypes.byref(h)))
    return h.value


# ---------------------------------------------------------------------------------------------
# Sketchers
class _GaussianSketcher(_Stage):
    _kind = None

    def __init__(self, parameters, seed=0, exponent=0, precision=_lib.PREC_F64,
                 omega_mode=_lib.OMEGA_PHILOX, omega_threads=1):
        super().__init__(parameters, precision)
        self._seed = seed
        self._mode = omega_mode
        self._threads = omega_threads
        self._exponent = int(exponent)
        if self._exponent != 0:
            self._handle()  # the library checks the kind and precision now

    def _make(self):
        p = self.parameters.c_params()
        h = _out_handle(_lib.lib().mcnla_sketcher_create, ctypes.byref(p), self._kind,
                        self.precision)
        try:
            _lib.check(_lib.lib().mcnla_sketcher_set_seed(h, self._seed, self._mode,
                                                         self._threads))
            _lib.check(_lib.lib().mcnla_sketcher_set_exponent(h, self._exponent))
        except Exception:
            _lib.lib().mcnla_stage_destroy(h)
            raise
        return h

    def exponent(self):
        return self._exponent

    def setExponent(self, exponent):
        """Power iterations (gaussian_projection_sketcher.hh:87): Y = (A A')^q A Omega."""
        self._exponent = int(exponent)
        if self._h is not None:
            _lib.check(_lib.lib().mcnla_sketcher_set_exponent(self._h, self._exponent))
        return self

    def setSeed(self, seed, omega_mode=None, omega_threads=None):
        self._seed = int(seed)
        if omega_mode is not None:
            self._mode = omega_mode
        if omega_threads is not None:
            self._threads = omega_threads
        if self._h is not None:
            _lib.check(_lib.lib().mcnla_sketcher_set_seed(self._h, self._seed, self._mode,
                                                         self._threads))
        return self

    def seed(self):
        return self._seed

    def __call__(self, matrix_a, collection_q, stream=None):
        _lib.check(_lib.lib().mcnla_sketcher_run(self._handle(), matrix_a.ptr, matrix_a.ld,
                                                 matrix_a.dtype_code, matrix_a.order_code,
                                                 collection_q.ptr, stream))
        self._computed = True

    def omega(self):
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().mcnla_sketcher_omega(self._handle(), ctypes.byref(p)))
        P = self.parameters
        return DeviceArray.view(p.value, (P.ncol(), P.dimSketchTotal())).numpy()


class ColBlockGaussianProjectionSketcher(_GaussianSketcher):
    _kind = _lib.SKETCHER_COL_BLOCK_GAUSSIAN


class RowBlockGaussianProjectionSketcher(_GaussianSketcher):
    _kind = _lib.SKETCHER_ROW_BLOCK_GAUSSIAN


class GaussianProjectionSketcher(_GaussianSketcher):
    _kind = _lib.SKETCHER_GAUSSIAN


class _ColumnSamplingSketcher(_GaussianSketcher):
    """Y's columns are columns of A picked by the reference's seeded index stream
    (column_sampling_sketcher.hpp:84-96); FP64 only, no Omega."""

    def __init__(self, parameters, seed=0, omega_threads=1):
        super().__init__(parameters, seed=seed, precision=_lib.PREC_F64,
                         omega_mode=_lib.OMEGA_REFERENCE, omega_threads=omega_threads)

    def omega(self):
        raise AttributeError("a column-sampling sketcher has no Omega")

    def indices(self):
        """The column of A behind each column of Y (last run)."""
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().mcnla_sketcher_indices(self._handle(), ctypes.byref(p)))
        return DeviceArray.view(p.value, (self.parameters.dimSketchTotal(),), np.int64).numpy()


class ColumnSamplingSketcher(_ColumnSamplingSketcher):
    _kind = _lib.SKETCHER_COLUMN_SAMPLING


class RowBlockColumnSamplingSketcher(_ColumnSamplingSketcher):
    _kind = _lib.SKETCHER_ROW_BLOCK_COLUMN_SAMPLING


def column_indices_reference(seed, ncol, length, kind=ColumnSamplingSketcher, threads=1):
    """The column-sampling index stream on the host (no device)."""
    out = np.empty(int(length), np.int64)
    _lib.check(_lib.lib().mcnla_column_indices_reference(
        int(seed), kind._kind, int(threads), int(length), int(ncol),
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
    return out


# ---------------------------------------------------------------------------------------------
# Orthogonalizers
class _GramianOrthogonalizer(_Stage):
    _kind = None

    def _make(self):
        p = self.parameters.c_params()
        return _out_handle(_lib.lib().mcnla_orthogonalizer_create, ctypes.byref(p), self._kind,
                           self.precision)

    def __call__(self, collection_q, stream=None):
        _lib.check(_lib.lib().mcnla_orthogonalizer_run(self._handle(), collection_q.ptr, stream))
        self._computed = True


class GramianOrthogonalizer(_GramianOrthogonalizer):
    _kind = _lib.ORTH_GRAMIAN


class RowBlockGramianOrthogonalizer(_GramianOrthogonalizer):
    _kind = _lib.ORTH_ROW_BLOCK_GRAMIAN


# ---------------------------------------------------------------------------------------------
# Integrators
class _Integrator(_Stage):
    _kind = None
    _default_tol = 1e-4

    def __init__(self, parameters, max_iteration=256, tolerance=None, precision=_lib.PREC_F64):
        super().__init__(parameters, precision)
        self._maxiter = int(max_iteration)
        self._tol = self._default_tol if tolerance is None else float(tolerance)
        self._iteration = 0
        self.converged = None

    def _make(self):
        p = self.parameters.c_params()
        return _out_handle(_lib.lib().mcnla_integrator_create, ctypes.byref(p), self._kind,
                           self.precision, self._maxiter, self._tol)

    def setMaxIteration(self, maxiter):
        self._maxiter = int(maxiter)
        if self._h is not None:
            self.initialize()
        return self

    def setTolerance(self, tol):
        self._tol = float(tol)
        if self._h is not None:
            self.initialize()
        return self

    def maxIteration(self):
        return self._maxiter

    def tolerance(self):
        return self._tol

    def iteration(self):
        return self._iteration

    def __call__(self, collection_q, matrix_qbar, stream=None):
        st = _lib.check(_lib.lib().mcnla_integrator_run(self._handle(), collection_q.ptr,
                                                        matrix_qbar.ptr, stream),
                        allow=(_lib.ERR_NOT_CONVERGED,))
        it = ctypes.c_int64()
        _lib.check(_lib.lib().mcnla_integrator_iteration(self._h, ctypes.byref(it)))
        self._iteration = it.value
        self.converged = st == _lib.OK
        self._computed = True


class KolmogorovNagumoIntegrator(_Integrator):
    _kind = _lib.INT_KN


class RowBlockKolmogorovNagumoIntegrator(_Integrator):
    _kind = _lib.INT_ROW_BLOCK_KN


class RowBlockGramianKolmogorovNagumoIntegrator(_Integrator):
    _kind = _lib.INT_ROW_BLOCK_GRAMIAN_KN


class RowBlockWenYinIntegrator(_Integrator):
    _kind = _lib.INT_ROW_BLOCK_WY
    _default_tol = 1e-3


class RowBlockGramianWenYinIntegrator(_Integrator):
    _kind = _lib.INT_ROW_BLOCK_GRAMIAN_WY
    _default_tol = 1e-3


class _CombiningIntegrator(_Integrator):
    """Integrators that combine the sketches in one pass (no iteration, no tolerance); the
    reference constructs them from the parameters alone."""

    def __init__(self, parameters, precision=_lib.PREC_F64):
        super().__init__(parameters, 0, 0.0, precision)

    def __call__(self, collection_q, *args, stream=None):
        # the extrinsic mean takes (Qj, Q, Qbar) in the reference
        # (row_block_extrinsic_mean_integrator.hh:109-111); with one rank Qj and Q coincide
        return super().__call__(collection_q, args[-1], stream)


class RowBlockReductionIntegrator(_CombiningIntegrator):
    _kind = _lib.INT_ROW_BLOCK_REDUCTION


class RowBlockExtrinsicMeanIntegrator(_CombiningIntegrator):
    _kind = _lib.INT_ROW_BLOCK_EXTRINSIC_MEAN


# ---------------------------------------------------------------------------------------------
# Formers
class _Former(_Stage):
    _kind = None

    def __init__(self, parameters, jobv=True, precision=_lib.PREC_F64):
        super().__init__(parameters, precision)
        self.jobv = bool(jobv)

    def _make(self):
        p = self.parameters.c_params()
        return _out_handle(_lib.lib().mcnla_former_create, ctypes.byref(p), self._kind,
                           1 if self.jobv else 0, self.precision)

    def __call__(self, matrix_a, matrix_qbar, stream=None):
        _lib.check(_lib.lib().mcnla_former_run(self._handle(), matrix_a.ptr, matrix_a.ld,
                                               matrix_a.dtype_code, matrix_a.order_code,
                                               matrix_qbar.ptr, stream))
        self._computed = True

    def _results(self):
        s, u, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(_lib.lib().mcnla_former_results(self._handle(), ctypes.byref(s),
                                                   ctypes.byref(u), ctypes.byref(v)))
        return s.value, u.value, v.value

    def vectorS(self):
        s, _, _ = self._results()
        return DeviceArray.view(s, (self.parameters.rank(),)).numpy()

    def matrixU(self):
        _, u, _ = self._results()
        P = self.parameters
        order = "F" if self._kind == _lib.FORM_SVD else "C"
        return DeviceArray.view(u, (P.nrow(), P.rank()), order=order).numpy()

    matrixUj = matrixU

    def matrixV(self):
        if not self.jobv:
            raise RuntimeError("V was not requested (jobv=False)")
        _, _, v = self._results()
        P = self.parameters
        return DeviceArray.view(v, (P.ncol(), P.rank())).numpy()

    matrixVj = matrixV


class SvdFormer(_Former):
    _kind = _lib.FORM_SVD


class GramianFormer(_Former):
    _kind = _lib.FORM_GRAMIAN


class ColBlockGramianFormer(_Former):
    _kind = _lib.FORM_COL_BLOCK_GRAMIAN


class RowBlockGramianFormer(_Former):
    _kind = _lib.FORM_ROW_BLOCK_GRAMIAN


class RowBlockSymmetricFormer(_Former):
    """Square symmetric A: vectorS() holds the k algebraically largest Ritz values of
    Qbar' A Qbar in ascending order and matrixU() = Qbar W (row_block_symmetric_former.hpp)."""
    _kind = _lib.FORM_ROW_BLOCK_SYMMETRIC

    def matrixV(self):
        raise RuntimeError("the symmetric former forms U only")

    matrixVj = matrixV


# ---------------------------------------------------------------------------------------------
# Converters (include/mcnla/isvd/converter/*): they repartition data across MPI ranks.  With one
# GPU holding every row and every sketch they are the identity; a copy is made only when the
# caller hands over two different buffers.
class _Converter:
    def __init__(self, parameters):
        self.parameters = parameters

    def initialize(self):
        return self

    def __call__(self, src, dst, stream=None):
        if src.ptr != dst.ptr:
            if src.nbytes != dst.nbytes:
                raise ValueError("converter buffers differ in size")
            _lib.check(_lib.lib().mcnla_memcpy_dtod(dst.ptr, src.ptr, src.nbytes, stream))

    def time(self):
        return 0.0


class CollectionFromPartialSumToRowBlockConverter(_Converter): pass
class CollectionToRowBlockConverter(_Converter): pass
class CollectionFromRowBlockConverter(_Converter): pass
class MatrixFromRowBlockToAllConverter(_Converter): pass
class MatrixFromColBlockToAllConverter(_Converter): pass
class MatrixFromRowBlockConverter(_Converter): pass
class MatrixToRowBlockConverter(_Converter): pass
class DummyConverter(_Converter): pass


# ---------------------------------------------------------------------------------------------
class Solver:
    """Sketch -> orthogonalize -> integrate -> form as one resident plan (the isvd::Solver facade).

    ``integrator`` / ``former`` take the class objects above; ``precision`` is one of
    ``_lib.PREC_*``.  Results stay on the device until ``vectorS()/matrixU()/matrixV()``."""

    def __init__(self, parameters, sketcher=ColBlockGaussianProjectionSketcher,
                 integrator=RowBlockKolmogorovNagumoIntegrator, former=SvdFormer,
                 precision=_lib.PREC_F64, max_iteration=256, tolerance=None, seed=0,
                 omega_mode=_lib.OMEGA_PHILOX, omega_threads=1, jobv=True, route=_lib.ROUTE_AUTO,
                 graph=False):
        self.parameters = parameters
        cfg = _lib.SolverConfig()
        _lib.lib().mcnla_solver_config_default(ctypes.byref(cfg))
        cfg.sketcher = sketcher._kind
        cfg.integrator = integrator._kind
        cfg.former = former._kind
        cfg.precision = precision
        cfg.jobv = 1 if jobv else 0
        cfg.maxiter = int(max_iteration)
        cfg.tol = integrator._default_tol if tolerance is None else float(tolerance)
        cfg.seed = int(seed)
        cfg.omega_mode = omega_mode
        cfg.omega_threads = int(omega_threads)
        cfg.route = int(route)
        cfg.graph = 1 if graph else 0
        self.config = cfg
        self.former_kind = former._kind
        p = parameters.c_params()
        self._h = _out_handle(_lib.lib().mcnla_solver_create, ctypes.byref(p), ctypes.byref(cfg))
        self.converged = None

    def setSeed(self, seed):
        _lib.check(_lib.lib().mcnla_solver_set_seed(self._h, int(seed)))
        return self

    def solve(self, matrix_a, stream=None, allow_not_converged=True):
        allow = (_lib.ERR_NOT_CONVERGED,) if allow_not_converged else ()
        st = _lib.check(_lib.lib().mcnla_solver_solve(self._h, matrix_a.ptr, matrix_a.ld,
                                                      matrix_a.dtype_code, matrix_a.order_code,
                                                      stream), allow=allow)
        self.converged = st == _lib.OK
        return self

    __call__ = solve

    def solve_host(self, matrix_a, chunk=0, stream=None, allow_not_converged=True):
        """Solves a host numpy matrix (FP64 or FP32, C or F order), streamed to the device in
        chunks of ``chunk`` columns (F order) or rows (C order) beside the sketch of the chunks
        that have landed (``mcnla_solver_solve_host``)."""
        a = np.asarray(matrix_a)
        if a.dtype not in (np.float64, np.float32):
            raise TypeError("A must be float64 or float32")
        if a.flags.f_contiguous and not a.flags.c_contiguous:
            order, ld = _lib.COL_MAJOR, a.shape[0]
        else:
            a = np.ascontiguousarray(a)
            order, ld = _lib.ROW_MAJOR, a.shape[1]
        dt = _lib.F64 if a.dtype == np.float64 else _lib.F32
        allow = (_lib.ERR_NOT_CONVERGED,) if allow_not_converged else ()
        st = _lib.check(_lib.lib().mcnla_solver_solve_host(self._h, a.ctypes.data, ld, dt, order,
                                                           int(chunk), stream), allow=allow)
        self.converged = st == _lib.OK
        return self

    def iteration(self):
        it = ctypes.c_int64()
        _lib.check(_lib.lib().mcnla_solver_iteration(self._h, ctypes.byref(it)))
        return it.value

    def _ptrs(self):
        s, u, v, q = (ctypes.c_void_p() for _ in range(4))
        _lib.check(_lib.lib().mcnla_solver_results(self._h, ctypes.byref(s), ctypes.byref(u),
                                                   ctypes.byref(v), ctypes.byref(q)))
        return s.value, u.value, v.value, q.value

    def vectorS(self):
        return DeviceArray.view(self._ptrs()[0], (self.parameters.rank(),)).numpy()

    def matrixU(self):
        P = self.parameters
        order = "F" if self.former_kind == _lib.FORM_SVD else "C"
        return DeviceArray.view(self._ptrs()[1], (P.nrow(), P.rank()), order=order).numpy()

    def matrixV(self):
        P = self.parameters
        return DeviceArray.view(self._ptrs()[2], (P.ncol(), P.rank())).numpy()

    def matrixQbar(self):
        P = self.parameters
        return DeviceArray.view(self._ptrs()[3], (P.nrow(), P.dimSketch())).numpy()

    def stage_times(self):
        t = (ctypes.c_double * 4)()
        k = ctypes.c_double()
        _lib.check(_lib.lib().mcnla_solver_stage_times(self._h, t, ctypes.byref(k)))
        return list(t), k.value

    def kernel_times(self):
        """Seconds of single launches of the last solve: (sketch GEMM, Gram GEMM)."""
        t = (ctypes.c_double * 2)()
        _lib.check(_lib.lib().mcnla_solver_kernel_times(self._h, t))
        return t[0], t[1]

    def kernel_time(self, kernel):
        """Seconds of one launch of the last solve: _lib.KERNEL_SKETCH, KERNEL_GRAM or
        KERNEL_FORMER_Z (the former's Z = A'Qbar; 0 on the Gram-of-A route)."""
        t = ctypes.c_double()
        _lib.check(_lib.lib().mcnla_solver_kernel_time(self._h, int(kernel), ctypes.byref(t)))
        return t.value

    def workspace_bytes(self):
        b = ctypes.c_size_t()
        _lib.check(_lib.lib().mcnla_solver_workspace_bytes(self._h, ctypes.byref(b)))
        return b.value

    def tuning(self):
        """The split-K workgroup targets this solver's FP64 GEMMs run with."""
        t = _lib.Tuning()
        _lib.check(_lib.lib().mcnla_solver_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, matrix_a, reps=2, stream=None):
        """Times candidate split-K targets of each GEMM of this solver's route on ``matrix_a``,
        keeps the fastest (by more than 1 %), records it for solvers of the same shape and
        returns it; the solver then holds the results of a solve with the tuned plan."""
        t = _lib.Tuning()
        _lib.check(_lib.lib().mcnla_solver_autotune(self._h, matrix_a.ptr, matrix_a.ld,
                                                    matrix_a.dtype_code, matrix_a.order_code,
                                                    int(reps), stream, ctypes.byref(t)))
        return t.as_dict()

    def __del__(self):
        try:
            if self._h:
                _lib.lib().mcnla_solver_destroy(self._h)
        except Exception:
            pass


def solve_batch(solvers, matrices, stream=None, allow_not_converged=True):
    """Runs ``solvers[i].solve(matrices[i])`` for up to ``_lib.MAX_BATCH`` solvers as one batch
    (``mcnla_solver_solve_batch``: the KN iterations of all members in lockstep, one launch per
    kernel per iteration for the whole batch).  Each member's results are bit-identical to its
    solve alone.  Returns the members' status codes; raises on the first one that is an error
    (NOT_CONVERGED is allowed unless ``allow_not_converged`` is false)."""
    n = len(solvers)
    assert n == len(matrices) and 1 <= n <= _lib.MAX_BATCH
    hs = (ctypes.c_void_p * n)(*[s._h for s in solvers])
    As = (ctypes.c_void_p * n)(*[a.ptr for a in matrices])
    lds = (ctypes.c_int64 * n)(*[a.ld for a in matrices])
    dts = (ctypes.c_int * n)(*[a.dtype_code for a in matrices])
    ods = (ctypes.c_int * n)(*[a.order_code for a in matrices])
    sts = (ctypes.c_int * n)()
    _lib.check(_lib.lib().mcnla_solver_solve_batch(n, hs, As, lds, dts, ods, sts, stream))
    allow = (_lib.ERR_NOT_CONVERGED,) if allow_not_converged else ()
    for s, st in zip(solvers, sts):
        s.converged = st == _lib.OK
    for st in sts:
        _lib.check(st, allow=allow)
    return list(sts)


def _tune_args(parameters, precision, route):
    cfg = _lib.SolverConfig()
    _lib.lib().mcnla_solver_config_default(ctypes.byref(cfg))
    cfg.precision = precision
    cfg.route = int(route)
    return parameters.c_params(), cfg


def tuning_default():
    t = _lib.Tuning()
    _lib.lib().mcnla_tuning_default(ctypes.byref(t))
    return t.as_dict()


def tune_get(parameters, precision=_lib.PREC_F64, route=_lib.ROUTE_AUTO):
    """(targets, found) recorded for solvers of this shape, precision and route."""
    p, cfg = _tune_args(parameters, precision, route)
    t, found = _lib.Tuning(), ctypes.c_int()
    _lib.check(_lib.lib().mcnla_tune_get(ctypes.byref(p), ctypes.byref(cfg), ctypes.byref(t),
                                         ctypes.byref(found)))
    return t.as_dict(), bool(found.value)


def tune_set(parameters, targets, precision=_lib.PREC_F64, route=_lib.ROUTE_AUTO):
    p, cfg = _tune_args(parameters, precision, route)
    t = _lib.Tuning(**{k: int(v) for k, v in targets.items()})
    _lib.check(_lib.lib().mcnla_tune_set(ctypes.byref(p), ctypes.byref(cfg), ctypes.byref(t)))


def tune_clear():
    _lib.lib().mcnla_tune_clear()


def tune_save(path):
    _lib.check(_lib.lib().mcnla_tune_save(str(path).encode()))


def tune_load(path):
    _lib.check(_lib.lib().mcnla_tune_load(str(path).encode()))


def synth_lowrank(m, n, U, V, sigma, noise=0.0, seed=1, dtype=np.float64, order="F",
                  stream=None):
    """A = U diag(sigma) V' + noise * G on the device (G ~ N(0,1), Philox(seed)).  U (m x r) and
    V (n x r) are host arrays or row-major FP64 DeviceArrays; returns a DeviceArray."""
    r = len(sigma)

    def dev(X):
        if isinstance(X, DeviceArray):
            assert X.dtype == np.float64 and X.order == "C"
            return X
        return DeviceArray.from_numpy(np.ascontiguousarray(X, dtype=np.float64))
    dU, dV = dev(U), dev(V)
    ds = DeviceArray.from_numpy(np.ascontiguousarray(sigma, dtype=np.float64))
    A = DeviceArray((m, n), dtype, order)
    _lib.check(_lib.lib().mcnla_synth_lowrank(m, n, r, dU.ptr, dV.ptr, ds.ptr, float(noise),
                                              int(seed), A.ptr, A.ld, A.dtype_code,
                                              A.order_code, stream))
    _lib.check(_lib.lib().mcnla_stream_synchronize(stream))
    return A


def gaussian(shape, seed=0, scale=1.0, stream=None):
    """A row-major FP64 DeviceArray of scale * N(0,1) (the Philox stream of `seed`)."""
    out = DeviceArray(shape, np.float64, "C")
    count = int(np.prod(out.shape))
    _lib.check(_lib.lib().mcnla_fill_gaussian(out.ptr, count, int(seed), float(scale), stream))
    _lib.check(_lib.lib().mcnla_stream_synchronize(stream))
    return out
