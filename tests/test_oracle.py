"""Pins the CPU oracle (oracle/isvd_oracle.cpp) before anything is checked against it.

Every check comes from outside the oracle: the reference's golden fixtures and the iteration
counts they record (data/*.mtx, check/isvd/**), LAPACK via numpy, closed forms and invariants.
"""
import numpy as np
import pytest

import mtx
import oracle


def proj(Q):
    return Q @ Q.T


@pytest.fixture(scope="module")
def qi():
    return mtx.load("qi.mtx")


# --- small dense kernels -------------------------------------------------------------------------
def test_syev_matches_lapack():
    rng = np.random.default_rng(0)
    for n in (1, 2, 11, 32, 64):
        X = rng.standard_normal((n, n))
        A = X + X.T
        w, V = oracle.syev(A)
        w_ref = np.linalg.eigvalsh(A)
        assert np.allclose(w, w_ref, rtol=0, atol=1e-12 * np.abs(w_ref).max())
        assert np.allclose(V.T @ V, np.eye(n), atol=1e-13)
        assert np.allclose(V @ np.diag(w) @ V.T, A, atol=1e-12 * np.abs(A).max())


def test_svd_tall_matches_lapack():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((300, 20)) * (0.7 ** np.arange(20))
    U, s, V = oracle.svd_tall(X)
    assert np.allclose(s, np.linalg.svd(X, compute_uv=False), rtol=1e-13)
    assert np.allclose(U @ np.diag(s) @ V.T, X, atol=1e-13)
    assert np.allclose(U.T @ U, np.eye(20), atol=1e-13)


def test_inverse_closed_form():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((64, 64)) + 8 * np.eye(64)
    assert np.allclose(oracle.inverse(A) @ A, np.eye(64), atol=1e-12)
    P = np.eye(5)[[3, 0, 4, 1, 2]]  # a permutation needs pivoting; its inverse is its transpose
    assert np.array_equal(oracle.inverse(P), P.T)
    with pytest.raises(np.linalg.LinAlgError):
        oracle.inverse(np.zeros((3, 3)))


# --- stages against the reference fixtures ---------------------------------------------------------
def test_fixture_metadata():
    # data/qb_kn.mtx:3-4 and data/qb_wy.mtx:3-4 record the tolerance and the iteration count
    assert mtx.comment_value("qb_kn.mtx", "iter") == "81"
    assert mtx.comment_value("qb_kn.mtx", "tol") == "1e-4"
    assert mtx.comment_value("qb_wy.mtx", "iter") == "21"
    assert mtx.comment_value("qb_wy.mtx", "tol") == "1e-3"


@pytest.mark.parametrize("fn", [oracle.kn_rowblock, oracle.kn_xform, oracle.kn_gramian])
def test_kn_integrators_fixture(qi, fn):
    # check/isvd/integrator/kolmogorov_nagumo_integrator.cpp:41,61-66
    qb = mtx.load("qb_kn.mtx")
    Q, it = fn(qi, 11, 12, 256, 1e-4)
    assert it == 81
    assert np.abs(Q - qb).max() < 1e-8


@pytest.mark.parametrize("fn", [oracle.wy_rowblock, oracle.wy_gramian])
def test_wy_integrators_fixture(qi, fn):
    # check/isvd/integrator/row_block_wen_yin_integrator.cpp:41,69-74
    qb = mtx.load("qb_wy.mtx")
    Q, it = fn(qi, 11, 12, 256, 1e-3)
    assert it == 21
    assert np.abs(Q - qb).max() < 1e-8


def test_orthogonalizer_fixture(qi):
    # check/isvd/orthogonalizer/gramian_orthogonalizer.cpp:48-72 (projectors at 1e-8)
    yi = mtx.load("yi.mtx")
    Q, bad = oracle.gramian_orthogonalize(yi, 11, 12)
    assert bad == 0
    for i in range(12):
        b = slice(11 * i, 11 * (i + 1))
        assert np.abs(proj(Q[:, b]) - proj(qi[:, b])).max() < 1e-8
        assert np.allclose(Q[:, b].T @ Q[:, b], np.eye(11), atol=1e-12)


def test_reduction_and_extrinsic_fixtures(qi):
    # check/isvd/integrator/row_block_reduction_integrator.cpp:66-80 and
    # row_block_extrinsic_mean_integrator.cpp (Qbar Qbar' at 1e-8)
    for fn, name in ((oracle.reduction, "qb_re.mtx"), (oracle.extrinsic_mean, "qb_em.mtx")):
        Q = fn(qi, 11, 12)
        assert np.abs(proj(Q) - proj(mtx.load(name))).max() < 1e-8
        assert np.allclose(Q.T @ Q, np.eye(11), atol=1e-12)


def test_reduction_and_extrinsic_closed_forms():
    rng = np.random.default_rng(4)
    m, l = 60, 4
    Q0, _ = np.linalg.qr(rng.standard_normal((m, l)))
    # identical sketches: both integrators return that subspace
    same = np.hstack([Q0] * 4)
    for fn in (oracle.reduction, oracle.extrinsic_mean):
        assert np.abs(proj(fn(same, l, 4)) - proj(Q0)).max() < 1e-12
    # two sketches: the reduction is the geodesic midpoint -- principal vectors (Q0 w_k + Q1 t_k)
    # normalized, i.e. the principal angles to both inputs are half of the angles between them
    Q1, _ = np.linalg.qr(Q0 + 0.3 * rng.standard_normal((m, l)))
    Qm = oracle.reduction(np.hstack([Q0, Q1]), l, 2)
    th = np.arccos(np.clip(np.linalg.svd(Q0.T @ Q1, compute_uv=False), -1, 1))
    for Qx in (Q0, Q1):
        th_m = np.arccos(np.clip(np.linalg.svd(Qx.T @ Qm, compute_uv=False), -1, 1))
        assert np.allclose(np.sort(th_m), np.sort(th / 2), atol=1e-10)
    # the extrinsic mean of two subspaces spans the top-l eigenvectors of Q0Q0' + Q1Q1'
    Qe = oracle.extrinsic_mean(np.hstack([Q0, Q1]), l, 2)
    w, V = np.linalg.eigh(proj(Q0) + proj(Q1))
    assert np.abs(proj(Qe) - proj(V[:, -l:])).max() < 1e-10


def test_symmetric_former_fixture():
    # check/isvd/former/row_block_symmetric_former.cpp:45-66 (k = 5, p = 6; U U' at 1e-8)
    a, q, u = mtx.load("a_sym.mtx"), mtx.load("q.mtx"), mtx.load("u_sym.mtx")
    k = 5
    S, U = oracle.symmetric_former(a, q, k)
    assert np.abs(proj(U) - proj(u[:, :k])).max() < 1e-8
    w = np.linalg.eigvalsh(q.T @ a @ q)  # LAPACK, ascending
    assert np.allclose(S, w[-k:], rtol=1e-12)  # the k largest, ascending
    assert np.allclose(a @ U - U * S, (np.eye(a.shape[0]) - proj(q)) @ a @ U, atol=1e-9 * S[-1])


@pytest.mark.parametrize("colmajor", [True, False])
def test_formers_fixture(colmajor):
    # check/isvd/former/svd_former.cpp:41,52-69; col_block_gramian_former.cpp:59-63;
    # row_block_gramian_former.cpp:72-76 -- U projectors at 1e-8 (k = 5, p = 6)
    a, q, u = mtx.load("a.mtx"), mtx.load("q.mtx"), mtx.load("u.mtx")
    k = 5
    uk = u[:, :k]
    S1, U1, V1 = oracle.svd_former(a, q, k, a_colmajor=colmajor)
    S2, U2, V2 = oracle.gramian_former(a, q, k, a_colmajor=colmajor)
    for U in (U1, U2):
        assert np.abs(proj(U) - proj(uk)).max() < 1e-8
    s_ref = np.linalg.svd(q.T @ a, compute_uv=False)[:k]  # sigma of Q'A from LAPACK
    assert np.allclose(S1, s_ref, rtol=1e-13)
    assert np.allclose(S2, s_ref, rtol=1e-10)
    # V: right singular vectors of Q'A, and A'U = V S for the rotated basis
    for S, U, V in ((S1, U1, V1), (S2, U2, V2)):
        assert np.allclose(V.T @ V, np.eye(k), atol=1e-10)
        W = q.T @ a
        assert np.allclose(W @ V, (q.T @ U) * S, atol=1e-9 * S[0])


def test_sketch_closed_form():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((37, 23))
    Om = rng.standard_normal((23, 15))
    for cm in (True, False):
        assert np.allclose(oracle.sketch(A, Om, a_colmajor=cm), A @ Om, atol=1e-12)


def test_omega_ref_properties():
    # Determinism and N(0,1) moments; the bit stream's pieces (mt19937, seed_seq, the polar
    # method, chunking, both seed chains) are pinned in test_omega_pins.py.
    a = oracle.omega_ref(123, 500, 40)
    b = oracle.omega_ref(123, 500, 40)
    assert np.array_equal(a, b)
    assert abs(a.mean()) < 0.02 and abs(a.std() - 1) < 0.02
    c = oracle.omega_ref(124, 500, 40)
    assert not np.array_equal(a, c)
    rb = oracle.omega_ref(123, 500, 40, col_block=False)
    assert not np.array_equal(a, rb)
    t2 = oracle.omega_ref(123, 500, 40, threads=2)
    assert not np.array_equal(a, t2)


# --- end to end -----------------------------------------------------------------------------------
def test_pipeline_known_spectrum():
    """The whole oracle pipeline on A = U diag(0.95^i + 1e-6) V' (2000 x 1000 at reduced size):
    the top singular values agree with the generator's sigma to the randomized-SVD accuracy, and
    the KN result is a fixed point: (I - Qb Qb') Pbar Qb ~ 0."""
    rng = np.random.default_rng(4)
    m, n, k, p, N = 400, 200, 10, 10, 4
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sig = 0.95 ** np.arange(n) + 1e-6
    A = (U * sig) @ V.T
    l = k + p
    Y = oracle.sketch(A, oracle.omega_ref(9, n, N * l))
    Q, _ = oracle.gramian_orthogonalize(Y, l, N)
    tol = 1e-8
    Qb, it = oracle.kn_rowblock(Q, l, N, 5000, tol)
    assert it < 5000
    assert np.allclose(Qb.T @ Qb, np.eye(l), atol=1e-12)
    # The stopping test bounds ||s - 1|| ~ ||X'X|| / 2, so ||X|| is O(sqrt(2 tol)).
    Pbar = sum(proj(Q[:, i * l:(i + 1) * l]) for i in range(N)) / N
    X = (np.eye(m) - proj(Qb)) @ Pbar @ Qb
    assert np.linalg.norm(X, 2) < 2 * np.sqrt(2 * tol)
    S, Uo, Vo = oracle.svd_former(A, Qb, k)
    assert np.allclose(S, sig[:k], rtol=0.1)  # no power iterations: a few percent at 0.95^i
    assert np.all(S <= sig[:k] * (1 + 1e-12))  # Ritz values never exceed the true ones
    err = oracle.frerr(A, S, Uo, Vo)
    best = np.sqrt(np.sum(sig[k:] ** 2) / np.sum(sig ** 2))
    assert best <= err < 1.5 * best


def test_planted_mistake_is_caught(qi):
    """A plausible mistake in the KN restatement (a transposed Dc in F = C - Dc Cinv would need a
    rebuild; here the equivalent input-side slip: sketch 0 not first) changes the fixture result
    far beyond 1e-8, so the fixture check has teeth."""
    qb = mtx.load("qb_kn.mtx")
    swapped = np.concatenate([qi[:, 11:22], qi[:, :11], qi[:, 22:]], axis=1)
    Q, it = oracle.kn_rowblock(swapped, 11, 12, 256, 1e-4)
    assert np.abs(Q - qb).max() > 1e-3
    Q, it = oracle.kn_rowblock(qi, 11, 12, 256, 1e-3)  # wrong tolerance -> wrong count
    assert it != 81


def test_power_sketch_closed_form():
    """q power iterations (gaussian_projection_sketcher.hpp:96-100) give Y = (A A')^q A Omega and
    leave Omega = A' (A A')^(q-1) A Omega_0; the restatement is checked against those closed
    forms in both storage orders."""
    rng = np.random.default_rng(8)
    A = rng.standard_normal((60, 40)) / 8.0
    om = rng.standard_normal((40, 12))
    AAt = A @ A.T
    for q in (0, 1, 3):
        for cm in (True, False):
            Y, om_q = oracle.power_sketch(A, om, q, a_colmajor=cm)
            want = np.linalg.matrix_power(AAt, q) @ A @ om
            assert np.abs(Y - want).max() < 1e-13 * np.abs(want).max()
            if q:
                want_om = A.T @ np.linalg.matrix_power(AAt, q - 1) @ A @ om
                assert np.abs(om_q - want_om).max() < 1e-13 * np.abs(want_om).max()


def test_column_sample_is_a_gather():
    A = np.arange(30.0).reshape(5, 6)
    assert np.array_equal(oracle.column_sample(A, [5, 0, 0, 3]), A[:, [5, 0, 0, 3]])
