"""Pins the BLAS-backed CPU pipeline (oracle/blas_pipeline.py, the CPU baseline that bench.py
times) against the reference's fixtures and against the C oracle, stage by stage."""
import numpy as np

import mtx
import oracle
from oracle import blas_pipeline as bp


def proj(Q):
    return Q @ Q.T


def test_kn_fixture():
    # check/isvd/integrator/kolmogorov_nagumo_integrator.cpp:41,61-66: 81 iterations, 1e-8
    Q, it = bp.kn_rowblock(mtx.load("qi.mtx"), 11, 12, 256, 1e-4)
    assert it == 81
    assert np.abs(Q - mtx.load("qb_kn.mtx")).max() < 1e-8


def test_orthogonalizer_fixture():
    # check/isvd/orthogonalizer/gramian_orthogonalizer.cpp:61-72 (projectors of each sketch)
    Y = mtx.load("yi.mtx").copy()
    qi = mtx.load("qi.mtx")
    Q = bp.orthogonalize(Y, 11, 12)
    for i in range(12):
        s = slice(i * 11, (i + 1) * 11)
        assert np.abs(proj(Q[:, s]) - proj(qi[:, s])).max() < 1e-8


def test_pipeline_matches_c_oracle():
    """The whole pipeline at m = 900, n = 500, k = 10, p = 6, N = 6 against the C oracle on the
    same Omega: equal iteration counts, Qbar to 1e-10, sigma to 1e-12, U projectors to 1e-10."""
    m, n, k, p, N = 900, 500, 10, 6, 6
    l = k + p
    rng = np.random.default_rng(3)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (U * 0.9 ** np.arange(n)) @ V.T
    om = oracle.omega_ref(4, n, N * l)
    S, Ub, it, times = bp.solve(A, om, k, p, N)
    Y = oracle.sketch(A, om)
    Q, _ = oracle.gramian_orthogonalize(Y, l, N)
    Qb_o, it_o = oracle.kn_rowblock(Q, l, N, 256, 1e-4)
    S_o, U_o, _ = oracle.svd_former(A, Qb_o, k)
    Qb, it_b = bp.kn_rowblock(bp.orthogonalize(A @ om, l, N), l, N)
    assert it == it_o == it_b
    assert np.abs(Qb - Qb_o).max() < 1e-10
    assert np.max(np.abs(S - S_o) / S_o) < 1e-12
    assert np.abs(proj(Ub) - proj(U_o)).max() < 1e-10
    assert len(times) == 4 and all(t >= 0 for t in times)
