"""The reference's CPU pipeline with its large products on a multithreaded BLAS (TEST AND
BASELINE INFRASTRUCTURE ONLY: the CPU baseline that bench.py times, and a second statement of the
oracle that the tests hold against it).

The reference runs every contraction through BLAS (``la::mm`` -> dgemm, ``la::rk`` -> dsyrk,
include/mcnla/core/la/dense/routine/gemm.hpp:36-75) and the l x l factorizations through LAPACK;
this module does the same with numpy's BLAS (OpenBLAS, all threads it is given) for every m-sized
product and the oracle's own small kernels (``oracle.syev``, ``oracle.svd_tall``) for the l x l
and l x n factorizations:

* sketch: Y = A Omega (col_block_gaussian_projection_sketcher.hpp:99);
* Gramian orthogonalizer: W_i = Y_i' Y_i, eig, Q_i = Y_i V_i S_i^{-1/2}
  (row_block_gramian_orthogonalizer.hpp:79-101);
* row-block Kolmogorov-Nagumo integrator (row_block_kolmogorov_nagumo_integrator.hpp:104-205):
  Qc = Q_0, Bc = Qs'Qc; per iteration Gc = Qs Bc / N, Bgc = Qs'Gc, Dc = Bc'Bc/N,
  Z = Bc'Bgc/N - Dc^2, eig Z, s = sqrt(1/2 + sqrt(1/4 - lambda)), C, C^-1, F = C - Dc C^-1,
  Q+ = Qc F + Gc C^-1, B+ = Bc F + Bgc C^-1, stop when ||s - 1|| < tol;
* SVD former: W = Qbar' A, SVD of W, U = Qbar U_w[:, :k] (svd_former.hpp:86-113).

Each stage is timed separately, as the reference driver prints them
(src/isvd_dense_cb.cpp:252-270).
"""
import time

import numpy as np

from . import svd_tall, syev


def orthogonalize(Y, l, N):
    """In place: Q_i = Y_i V_i S_i^{-1/2} per sketch (descending eigenvalue order)."""
    for i in range(N):
        Yi = Y[:, i * l:(i + 1) * l]
        w, V = syev(Yi.T @ Yi)
        w, V = w[::-1], V[:, ::-1]
        Y[:, i * l:(i + 1) * l] = Yi @ (V / np.sqrt(w))
    return Y


def kn_small_step(Bc, Bgc, N):
    """The l x l part of one KN iteration: returns (F, C^-1, ||s - 1||_2)."""
    Dc = Bc.T @ Bc / N
    Z = Bc.T @ Bgc / N - Dc @ Dc
    Z = np.triu(Z) + np.triu(Z, 1).T  # the reference's syrk/syev read one triangle
    w, V = syev(Z)
    s = np.sqrt(0.5 + np.sqrt(0.25 - w))
    C = (V * s) @ V.T
    Ci = (V / s) @ V.T
    return C - Dc @ Ci, Ci, float(np.linalg.norm(s - 1.0))


def kn_rowblock(Qs, l, N, maxiter=256, tol=1e-4):
    """Row-block KN integrator; returns (Qbar, iterations executed)."""
    Qc = Qs[:, :l].copy()
    Bc = Qs.T @ Qc
    it, conv = 0, False
    while it < maxiter and not conv:
        Gc = Qs @ Bc / N
        Bgc = Qs.T @ Gc
        F, Ci, res = kn_small_step(Bc, Bgc, N)
        Qc = Qc @ F + Gc @ Ci
        Bc = Bc @ F + Bgc @ Ci
        it += 1
        conv = not (res >= tol)
    return Qc, it


def svd_former(A, Qbar, k):
    """(S, U) of the SVD former: W = Qbar' A, SVD via its transpose (n x l, tall)."""
    X = A.T @ Qbar  # = W'
    _, s, Vx = svd_tall(X)  # X = Ux S Vx'  =>  W = Vx S Ux'
    return s[:k], Qbar @ Vx[:, :k]


def solve(A, Omega, k, p, N, maxiter=256, tol=1e-4):
    """Whole pipeline; returns S, U, iterations and the per-stage seconds
    [sketch, orthogonalize, integrate, form]."""
    l = k + p
    t = [time.perf_counter()]
    Y = A @ Omega
    t.append(time.perf_counter())
    orthogonalize(Y, l, N)
    t.append(time.perf_counter())
    Qbar, it = kn_rowblock(Y, l, N, maxiter, tol)
    t.append(time.perf_counter())
    S, U = svd_former(A, Qbar, k)
    t.append(time.perf_counter())
    return S, U, it, [b - a for a, b in zip(t[:-1], t[1:])]
