"""CPU oracle for the iSVD hot path (TEST INFRASTRUCTURE ONLY).

A plain FP64 restatement of MCNLA 0.7.2's iSVD stages, compiled from ``isvd_oracle.cpp`` into
``liboracle.so``.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker.  The ``mcnla_amd``
product never imports it.

Every wrapper takes and returns numpy float64 arrays in the reference's layouts: the sketch
collection is one row-major ``m x (N*l)`` array (sketch ``i`` in columns ``[i*l, (i+1)*l)``),
``Qbar`` is row-major ``m x l``.

Parity pins (each enforced by a test under ``tests/``):
  * small dense kernels (syev / SVD / inverse): closed forms and numpy/LAPACK on random input;
  * orthogonalizer: ``data/yi.mtx -> data/qi.mtx`` projectors at 1e-8;
  * KN (row-block, X-form, Gramian): ``data/qi.mtx -> data/qb_kn.mtx``, 81 iterations, 1e-8;
  * WY (row-block, Gramian): ``data/qi.mtx -> data/qb_wy.mtx``, 21 iterations, 1e-8;
  * formers: ``data/a.mtx + data/q.mtx -> data/u.mtx`` projectors at 1e-8, sigma vs numpy;
  * reduction / extrinsic-mean integrators: ``data/qi.mtx -> data/qb_re.mtx / qb_em.mtx``
    projectors at 1e-8, plus the two-sketch midpoint and identical-sketch closed forms;
  * symmetric former: ``data/a_sym.mtx + data/q.mtx -> data/u_sym.mtx`` projectors at 1e-8,
    sigma vs the eigenvalues of ``Qbar' A Qbar``;
  * sketch: closed form ``A @ Omega``;
  * ``mt19937_raw``: the C++ standard's required 10000th draw of the default seed, and an
    independent restatement of MT19937 (``tests/stdrandom.py``) plus numpy's own MT19937;
  * ``seed_seq_generate``: a restatement of the standard's seed_seq::generate algorithm;
  * ``omega_ref``: bit-exact against a restatement of the Marsaglia-Bray polar method over
    two-draw ``generate_canonical<double, 53>`` (libstdc++'s ``normal_distribution``), of the
    thread chunking and of both seed chains.  The composed chain is still *parity unpinned*
    against the reference itself: no reference fixture stores Omega or its seed.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_syev.argtypes = [_i64, _d, _d, _d]
        L.orc_gesvd_tall.argtypes = [_i64, _i64, _d, _d, _d, _d]
        L.orc_inverse.argtypes = [_i64, _d]
        L.orc_inverse.restype = ctypes.c_int
        L.orc_omega_ref.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _i64, _d]
        L.orc_mt19937_raw.argtypes = [ctypes.c_uint32, _i64, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_seed_seq_generate.argtypes = [ctypes.POINTER(ctypes.c_uint32), _i64,
                                            ctypes.POINTER(ctypes.c_uint32), _i64]
        L.orc_sketch.argtypes = [_i64, _i64, _i64, _d, _i64, ctypes.c_int, _d, _d]
        L.orc_column_indices.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _i64, _i64,
                                         ctypes.POINTER(ctypes.c_int64)]
        L.orc_gramian_orthogonalize.argtypes = [_i64, _i64, _i64, _d]
        L.orc_gramian_orthogonalize.restype = ctypes.c_int
        for name in ("orc_kn_rowblock", "orc_kn_xform", "orc_kn_gramian", "orc_wy_rowblock",
                     "orc_wy_gramian"):
            f = getattr(L, name)
            f.argtypes = [_i64, _i64, _i64, _d, ctypes.c_int, ctypes.c_double, _d]
            f.restype = ctypes.c_int
        L.orc_svd_former.argtypes = [_i64, _i64, _i64, _i64, _d, _i64, ctypes.c_int, _d, _d, _d,
                                     _d]
        L.orc_gramian_former.argtypes = [_i64, _i64, _i64, _i64, _d, _i64, ctypes.c_int, _d, _d,
                                         _d, _d]
        L.orc_frerr.argtypes = [_i64, _i64, _i64, _d, _i64, ctypes.c_int, _d, _d, _d]
        L.orc_frerr.restype = ctypes.c_double
        L.orc_reduction.argtypes = [_i64, _i64, _i64, _d, _d]
        L.orc_extrinsic_mean.argtypes = [_i64, _i64, _i64, _d, _d]
        L.orc_symmetric_former.argtypes = [_i64, _i64, _i64, _d, _i64, ctypes.c_int, _d, _d, _d]
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_d)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def syev(a):
    """Symmetric eigen-decomposition: (w ascending, V with eigenvectors in columns)."""
    a = _c(a).copy()
    n = a.shape[0]
    w = np.empty(n)
    V = np.empty((n, n))
    lib().orc_syev(n, _p(a), _p(w), _p(V))
    return w, V


def svd_tall(x):
    """SVD of a tall matrix: (U r x c, s descending, V c x c with right vectors in columns)."""
    x = _c(x).copy()
    r, c = x.shape
    s = np.empty(c)
    U = np.empty((r, c))
    V = np.empty((c, c))
    lib().orc_gesvd_tall(r, c, _p(x), _p(s), _p(U), _p(V))
    return U, s, V


def inverse(a):
    a = _c(a).copy()
    rc = lib().orc_inverse(a.shape[0], _p(a))
    if rc != 0:
        raise np.linalg.LinAlgError("singular matrix")
    return a


def omega_ref(seed, n, Nl, col_block=True, nprocs=1, rank=0, threads=1):
    """The reference's seeded Omega (n x Nl row-major) for one rank."""
    out = np.empty((n, Nl))
    lib().orc_omega_ref(seed & 0xFFFFFFFF, 1 if col_block else 0, nprocs, rank, threads, n * Nl,
                        _p(out))
    return out


def mt19937_raw(seed, count):
    """`count` raw draws of std::mt19937 seeded with `seed` (uint32 array)."""
    out = np.empty(int(count), np.uint32)
    lib().orc_mt19937_raw(seed & 0xFFFFFFFF, int(count),
                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def seed_seq_generate(words, count):
    """std::seed_seq{words...}.generate() into `count` uint32 words."""
    w = np.ascontiguousarray(np.asarray(words, dtype=np.uint32))
    out = np.empty(int(count), np.uint32)
    lib().orc_seed_seq_generate(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), w.size,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), int(count))
    return out


def column_indices_ref(seed, ncol, length, col_block=True, nprocs=1, rank=0, threads=1):
    """The column-sampling sketchers' index stream for one rank (orc_column_indices)."""
    out = np.empty(int(length), np.int64)
    lib().orc_column_indices(seed & 0xFFFFFFFF, 1 if col_block else 0, nprocs, rank, threads,
                             int(length), int(ncol),
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return out


def column_sample(A, idx):
    """Y = the columns idx of A (column_sampling_sketcher.hpp:94-96)."""
    return np.ascontiguousarray(np.asarray(A)[:, np.asarray(idx)])


def power_sketch(A, omega, q, a_colmajor=True):
    """Y = A Omega, then q times Omega := A'Y, Y := A Omega (gaussian_projection_sketcher.hpp
    :96-100).  Returns (Y, final Omega)."""
    Y = sketch(A, omega, a_colmajor)
    for _ in range(q):
        omega = sketch(np.ascontiguousarray(np.asarray(A).T), Y, a_colmajor)
        Y = sketch(A, omega, a_colmajor)
    return Y, omega


def sketch(A, omega, a_colmajor=True):
    """Y = A @ Omega.  ``A`` is the logical m x n matrix; it is handed over in the requested
    storage order."""
    m, n = A.shape
    Nl = omega.shape[1]
    store = _c(A.T) if a_colmajor else _c(A)
    Y = np.empty((m, Nl))
    lib().orc_sketch(m, n, Nl, _p(store), m if a_colmajor else n, 1 if a_colmajor else 0,
                     _p(_c(omega)), _p(Y))
    return Y


def gramian_orthogonalize(Y, l, N):
    Q = _c(Y).copy()
    bad = lib().orc_gramian_orthogonalize(Q.shape[0], l, N, _p(Q))
    return Q, bad


def _integrate(name, Qs, l, N, maxiter, tol):
    Qs = _c(Qs)
    m = Qs.shape[0]
    Qbar = np.empty((m, l))
    it = getattr(lib(), name)(m, l, N, _p(Qs), maxiter, tol, _p(Qbar))
    return Qbar, it


def kn_rowblock(Qs, l, N, maxiter=256, tol=1e-4):
    return _integrate("orc_kn_rowblock", Qs, l, N, maxiter, tol)


def kn_xform(Qs, l, N, maxiter=256, tol=1e-4):
    return _integrate("orc_kn_xform", Qs, l, N, maxiter, tol)


def kn_gramian(Qs, l, N, maxiter=256, tol=1e-4):
    return _integrate("orc_kn_gramian", Qs, l, N, maxiter, tol)


def wy_rowblock(Qs, l, N, maxiter=256, tol=1e-3):
    return _integrate("orc_wy_rowblock", Qs, l, N, maxiter, tol)


def wy_gramian(Qs, l, N, maxiter=256, tol=1e-3):
    return _integrate("orc_wy_gramian", Qs, l, N, maxiter, tol)


def svd_former(A, Qbar, k, a_colmajor=True, jobv=True):
    """Returns (S[k], U m x k, V n x k)."""
    m, n = A.shape
    l = Qbar.shape[1]
    store = _c(A.T) if a_colmajor else _c(A)
    S = np.empty(k)
    Ucm = np.empty((k, m))  # col-major m x k == row-major k x m
    V = np.empty((n, k))
    lib().orc_svd_former(m, n, l, k, _p(store), m if a_colmajor else n, 1 if a_colmajor else 0,
                         _p(_c(Qbar)), _p(S), _p(Ucm), _p(V) if jobv else None)
    return S, Ucm.T.copy(), (V if jobv else None)


def gramian_former(A, Qbar, k, a_colmajor=True, jobv=True):
    """Returns (S[k], U m x k, V n x k)."""
    m, n = A.shape
    l = Qbar.shape[1]
    store = _c(A.T) if a_colmajor else _c(A)
    S = np.empty(k)
    U = np.empty((m, k))
    V = np.empty((n, k))
    lib().orc_gramian_former(m, n, l, k, _p(store), m if a_colmajor else n,
                             1 if a_colmajor else 0, _p(_c(Qbar)), _p(S), _p(U),
                             _p(V) if jobv else None)
    return S, U, (V if jobv else None)


def reduction(Qs, l, N):
    """Reduction integrator (pairwise tree); returns Qbar (m x l)."""
    Qs = _c(Qs)
    Qbar = np.empty((Qs.shape[0], l))
    lib().orc_reduction(Qs.shape[0], l, N, _p(Qs), _p(Qbar))
    return Qbar


def extrinsic_mean(Qs, l, N):
    """Extrinsic-mean integrator; returns Qbar (m x l, left singular vectors)."""
    Qs = _c(Qs)
    Qbar = np.empty((Qs.shape[0], l))
    lib().orc_extrinsic_mean(Qs.shape[0], l, N, _p(Qs), _p(Qbar))
    return Qbar


def symmetric_former(A, Qbar, k, a_colmajor=True):
    """Symmetric former for a square symmetric A: (S[k] ascending, U m x k)."""
    m = A.shape[0]
    l = Qbar.shape[1]
    store = _c(A.T) if a_colmajor else _c(A)
    S = np.empty(k)
    U = np.empty((m, k))
    lib().orc_symmetric_former(m, l, k, _p(store), m, 1 if a_colmajor else 0, _p(_c(Qbar)),
                               _p(S), _p(U))
    return S, U


def frerr(A, S, U, V):
    m, n = A.shape
    k = S.shape[0]
    store = _c(A.T)
    return lib().orc_frerr(m, n, k, _p(store), m, 1, _p(_c(S)), _p(_c(U.T)), _p(_c(V)))
