"""GPU parity of each stage, through the C ABI (libmcnla_amd), against the reference's golden
fixtures and the CPU oracle.  Tolerances are those of the reference's own tests (1e-8, element-wise
for the integrators, on projectors for the orthogonalizer and formers) unless stated."""
import sys

import numpy as np
import pytest

import mcnla_amd as mc
import mtx
import oracle

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def proj(Q):
    return Q @ Q.T


def params(m, n, k, p, N):
    return mc.Parameters().setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N)


# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(101, 1001, 11, 0, 12), (2000, 1000, 10, 10, 4), (257, 129, 3, 8, 3),
                                   (998, 333, 13, 4, 10), (1024, 4096, 20, 12, 16),
                                   (3001, 1000, 20, 12, 2), (1500, 600, 10, 10, 2)])
@pytest.mark.parametrize("order", ["F", "C"])
def test_sketch_f64(gpu, shape, order):
    m, n, k, p, N = shape
    note("sketch f64 %s order=%s" % (shape, order))
    rng = np.random.default_rng(1)
    A = rng.standard_normal((m, n))
    P = params(m, n, k, p, N)
    sk = mc.ColBlockGaussianProjectionSketcher(P, seed=5)
    sk.initialize()
    dA = mc.DeviceArray.from_numpy(np.asfortranarray(A) if order == "F" else np.ascontiguousarray(A))
    Y = P.createCollectionQ()
    sk(dA, Y)
    om = sk.omega()
    assert om.shape == (n, (k + p) * N)
    ref = A @ om
    assert np.abs(Y.numpy() - ref).max() < 1e-12 * np.abs(ref).max()
    # Philox Omega: N(0,1) moments and determinism per seed
    assert abs(om.mean()) < 0.05 and abs(om.std() - 1) < 0.05
    sk(dA, Y)
    assert np.array_equal(sk.omega(), om)


def test_sketch_f64_tail_split(gpu):
    """A column-major FP64 A whose row tiles do not fill whole rounds of the grid: 20 row tiles of
    256 x 16 column tiles of 128, of which 16 fill one round of a 256-CU device and the last 4
    (1000 rows, 64 workgroups) run as a second product with K split 4."""
    m, n, k, p, N = 5096, 4096, 20, 12, 64
    note("sketch f64 tail split %s" % ((m, n, k, p, N),))
    rng = np.random.default_rng(3)
    A = rng.standard_normal((m, n))
    P = params(m, n, k, p, N)
    sk = mc.ColBlockGaussianProjectionSketcher(P, seed=7)
    sk.initialize()
    dA = mc.DeviceArray.from_numpy(np.asfortranarray(A))
    Y = P.createCollectionQ()
    sk(dA, Y)
    ref = A @ sk.omega()
    got = Y.numpy()
    scale = np.abs(ref).max()
    assert np.abs(got[:4096] - ref[:4096]).max() < 1e-12 * scale
    assert np.abs(got[4096:] - ref[4096:]).max() < 1e-12 * scale
    sk(dA, Y)  # the same again, bit for bit (the partials are summed in a fixed order)
    assert np.array_equal(Y.numpy(), got)


@pytest.mark.parametrize("kind,cb", [(mc.ColBlockGaussianProjectionSketcher, True),
                                     (mc.RowBlockGaussianProjectionSketcher, False)])
def test_sketch_reference_omega(gpu, kind, cb):
    m, n, k, p, N = 300, 200, 5, 5, 3
    note("sketch reference omega %s" % kind.__name__)
    rng = np.random.default_rng(2)
    A = rng.standard_normal((m, n))
    P = params(m, n, k, p, N)
    sk = kind(P, seed=77, omega_mode=mc.OMEGA_REFERENCE, omega_threads=4)
    Y = P.createCollectionQ()
    order = "F" if cb else "C"
    sk(mc.DeviceArray.from_numpy(np.asarray(A, order=order)), Y)
    om_ref = oracle.omega_ref(77, n, N * (k + p), col_block=cb, threads=4)
    assert np.array_equal(sk.omega(), om_ref)  # bit-exact stream
    assert np.abs(Y.numpy() - oracle.sketch(A, om_ref, a_colmajor=cb)).max() < 1e-11


def bf16_round(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("order", ["F", "C"])
def test_sketch_mixed(gpu, order):
    m, n, k, p, N = 333, 517, 10, 6, 5
    note("sketch mixed order=%s" % order)
    rng = np.random.default_rng(3)
    A = rng.standard_normal((m, n)).astype(np.float32)
    P = params(m, n, k, p, N)
    dA = mc.DeviceArray.from_numpy(np.asarray(A, order=order))
    out = {}
    for prec in (mc.PREC_MIXED_BF16, mc.PREC_MIXED_BF16X3):
        sk = mc.ColBlockGaussianProjectionSketcher(P, seed=9, precision=prec)
        Y = P.createCollectionQ(np.float32)
        sk(dA, Y)
        om = sk.omega()
        out[prec] = (Y.numpy().astype(np.float64), om)
    Y1, om = out[mc.PREC_MIXED_BF16]
    exact = A.astype(np.float64) @ om
    scale = np.abs(exact).max()
    # one term: exactly the product of the BF16-rounded operands, up to FP32 accumulation
    pinned = bf16_round(A) @ bf16_round(om)
    assert np.abs(Y1 - pinned).max() < 2e-5 * scale
    assert np.linalg.norm(Y1 - exact) / np.linalg.norm(exact) < 6e-3
    Y3, om3 = out[mc.PREC_MIXED_BF16X3]
    assert np.array_equal(om3, om)
    assert np.linalg.norm(Y3 - exact) / np.linalg.norm(exact) < 5e-5


def f16_rows(A):
    """Per-row power-of-two exponents e (largest |a| 2^-e in [2^13, 2^14); 0 for a zero row) and
    the scaled rows A 2^-e in FP32."""
    mx = np.abs(A).max(axis=1).astype(np.float32)
    e = np.where(mx > 0, np.frexp(mx)[1] - 14, 0)
    return e, np.ldexp(A.astype(np.float32), -e[:, None]).astype(np.float32)


def f16_split(x):
    hi = x.astype(np.float16)
    lo = (x - hi.astype(np.float32)).astype(np.float16)
    return hi.astype(np.float64), lo.astype(np.float64)


@pytest.mark.parametrize("order", ["F", "C"])
def test_sketch_mixed_f16(gpu, order):
    """FP16 sketches on rows spanning 10^-30 .. 10^30 (and a zero row): each row of A is scaled by
    a power of two into FP16 range and the output row scaled back.  One term equals the product of
    the rounded scaled operands up to FP32 accumulation; the two-term split of A against an Omega
    that is exactly FP16 (OMEGA_PHILOX_F16), and the three-term split against any Omega, match
    the FP64 product row by row to FP32 accumulation accuracy."""
    m, n, k, p, N = 333, 517, 10, 6, 5
    note("sketch mixed fp16 order=%s" % order)
    rng = np.random.default_rng(4)
    A = (rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-30, 30, (m, 1))).astype(np.float32)
    A[7] = 0.0
    P = params(m, n, k, p, N)
    dA = mc.DeviceArray.from_numpy(np.asarray(A, order=order))
    out = {}
    for prec, omode in ((mc.PREC_MIXED_F16, mc.OMEGA_PHILOX), (mc.PREC_MIXED_F16X2, mc.OMEGA_PHILOX_F16),
                        (mc.PREC_MIXED_F16X3, mc.OMEGA_PHILOX)):
        sk = mc.ColBlockGaussianProjectionSketcher(P, seed=9, precision=prec, omega_mode=omode)
        Y = P.createCollectionQ(np.float32)
        sk(dA, Y)
        out[prec] = (Y.numpy().astype(np.float64), sk.omega())
    e, As = f16_rows(A)
    hi, lo = f16_split(As)

    def rowrel(Y, ref):
        scale = np.abs(ref).max(axis=1)
        scale[scale == 0] = 1.0
        return (np.abs(Y - ref).max(axis=1) / scale).max()

    Y1, om = out[mc.PREC_MIXED_F16]
    pinned = np.ldexp(hi @ om.astype(np.float32).astype(np.float16).astype(np.float64), e[:, None])
    assert np.all(Y1[7] == 0.0)
    assert rowrel(Y1, pinned) < 2e-5
    assert rowrel(Y1, A.astype(np.float64) @ om) < 3e-3
    Y2, om2 = out[mc.PREC_MIXED_F16X2]
    assert np.array_equal(om2.astype(np.float16).astype(np.float64), om2)  # Omega is FP16-exact
    assert np.abs(om2 - om).max() <= 2.0 ** -11 * np.abs(om).max()  # the same stream, rounded
    assert rowrel(Y2, A.astype(np.float64) @ om2) < 2e-6
    Y3, om3 = out[mc.PREC_MIXED_F16X3]
    assert np.array_equal(om3, om)
    assert rowrel(Y3, A.astype(np.float64) @ om3) < 2e-6


def test_sketch_mixed_tail_split(gpu):
    """16-bit sketches whose 256 x 256 tiles do not fill whole rounds of the grid: 36 row tiles
    of 8 column tiles, of which 32 fill one round of a 256-CU device and the last 4 (1000 rows)
    run with K split in 4 ranges whose FP32 partials are summed by a second kernel (FP16: then
    scaled by the row's power of two).  Every row equals the product of the rounded operands up
    to FP32 accumulation, and a second run is bit-identical."""
    m, n, k, p, N = 9192, 5000, 20, 12, 64
    note("sketch mixed tail split %s" % ((m, n, k, p, N),))
    rng = np.random.default_rng(5)
    A = (rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-3, 3, (m, 1))).astype(np.float32)
    P = params(m, n, k, p, N)
    dA = mc.DeviceArray.from_numpy(np.asfortranarray(A))
    e, As = f16_rows(A)
    for prec in (mc.PREC_MIXED_BF16, mc.PREC_MIXED_F16):
        sk = mc.ColBlockGaussianProjectionSketcher(P, seed=11, precision=prec)
        Y = P.createCollectionQ(np.float32)
        sk(dA, Y)
        got = Y.numpy().astype(np.float64)
        om = sk.omega()
        if prec == mc.PREC_MIXED_BF16:
            pinned = bf16_round(A) @ bf16_round(om)
        else:
            hi = As.astype(np.float16).astype(np.float64)
            pinned = np.ldexp(hi @ om.astype(np.float32).astype(np.float16).astype(np.float64),
                              e[:, None])
        scale = np.abs(pinned).max(axis=1)
        rel = np.abs(got - pinned).max(axis=1) / scale
        assert rel[:8192].max() < 2e-5 and rel[8192:].max() < 2e-5, (prec, rel[:8192].max(),
                                                                      rel[8192:].max())
        sk(dA, Y)
        assert np.array_equal(Y.numpy().astype(np.float64), got)


@pytest.mark.parametrize("kind,cb", [(mc.ColumnSamplingSketcher, True),
                                     (mc.RowBlockColumnSamplingSketcher, False)])
@pytest.mark.parametrize("dtype,order", [(np.float64, "F"), (np.float32, "C")])
def test_column_sampling_sketcher(gpu, kind, cb, dtype, order):
    """Column sampling (column_sampling_sketcher.hpp:84-96): Y's columns are exactly the columns
    of A at the oracle's index stream (3-thread chunking)."""
    m, n, k, p, N = 301, 257, 7, 5, 6
    note("column sampling %s %s %s" % (kind.__name__, np.dtype(dtype).name, order))
    A = np.random.default_rng(6).standard_normal((m, n)).astype(dtype)
    P = params(m, n, k, p, N)
    sk = kind(P, seed=31, omega_threads=3)
    Y = P.createCollectionQ()
    sk(mc.DeviceArray.from_numpy(np.asarray(A, order=order), order=order), Y)
    idx = sk.indices()
    assert np.array_equal(idx, oracle.column_indices_ref(31, n, N * (k + p), col_block=cb,
                                                         threads=3))
    assert np.array_equal(Y.numpy(), oracle.column_sample(A.astype(np.float64), idx))
    assert sk.time() > 0


@pytest.mark.parametrize("order", ["F", "C"])
def test_power_iteration_sketcher(gpu, order):
    """GaussianProjectionSketcher with exponent 2 (gaussian_projection_sketcher.hpp:96-100) on the
    reference Omega against the oracle: Y = (AA')^2 A Omega and the final Omega = A'(AA')A Omega."""
    m, n, k, p, N, q = 700, 300, 6, 4, 3, 2
    note("power iteration sketcher order=%s" % order)
    A = np.random.default_rng(12).standard_normal((m, n)) / np.sqrt(m)
    P = params(m, n, k, p, N)
    sk = mc.GaussianProjectionSketcher(P, seed=44, exponent=q, omega_mode=mc.OMEGA_REFERENCE)
    assert sk.exponent() == q
    Y = P.createCollectionQ()
    sk(mc.DeviceArray.from_numpy(np.asarray(A, order=order), order=order), Y)
    om0 = oracle.omega_ref(44, n, N * (k + p), col_block=True)
    Y_o, om_o = oracle.power_sketch(A, om0, q, a_colmajor=(order == "F"))
    assert np.abs(Y.numpy() - Y_o).max() < 1e-12 * np.abs(Y_o).max()
    assert np.abs(sk.omega() - om_o).max() < 1e-12 * np.abs(om_o).max()
    with pytest.raises(mc.McnlaError) as e:  # the col-block sketcher has no power iterations
        mc.ColBlockGaussianProjectionSketcher(P, seed=1, exponent=1)
    assert e.value.status == mc._lib.ERR_INVALID
    with pytest.raises(mc.McnlaError) as e:  # FP64 only
        mc.GaussianProjectionSketcher(P, seed=1, exponent=1, precision=mc.PREC_MIXED_BF16)
    assert e.value.status == mc._lib.ERR_INVALID


# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("kind", [mc.GramianOrthogonalizer, mc.RowBlockGramianOrthogonalizer])
def test_orthogonalizer_fixture(gpu, kind):
    # check/isvd/orthogonalizer/gramian_orthogonalizer.cpp:48-72
    note("orthogonalizer fixture %s" % kind.__name__)
    yi, qi = mtx.load("yi.mtx"), mtx.load("qi.mtx")
    P = params(101, 11, 11, 0, 12)
    o = kind(P)
    o.initialize()
    Q = mc.DeviceArray.from_numpy(yi)
    o(Q)
    q = Q.numpy()
    for i in range(12):
        b = slice(11 * i, 11 * (i + 1))
        assert np.abs(proj(q[:, b]) - proj(qi[:, b])).max() < 1e-8
    assert o.time() > 0


INTEGRATORS = [
    (mc.KolmogorovNagumoIntegrator, "qb_kn.mtx", 1e-4, 81),
    (mc.RowBlockKolmogorovNagumoIntegrator, "qb_kn.mtx", 1e-4, 81),
    (mc.RowBlockGramianKolmogorovNagumoIntegrator, "qb_kn.mtx", 1e-4, 81),
    (mc.RowBlockWenYinIntegrator, "qb_wy.mtx", 1e-3, 21),
    (mc.RowBlockGramianWenYinIntegrator, "qb_wy.mtx", 1e-3, 21),
]


@pytest.mark.parametrize("kind,fixture,tol,iters", INTEGRATORS)
def test_integrator_fixture(gpu, kind, fixture, tol, iters):
    # check/isvd/integrator/*.cpp: iteration() == 81 (KN) / 21 (WY), Qbar within 1e-8
    note("integrator fixture %s" % kind.__name__)
    qi, qb = mtx.load("qi.mtx"), mtx.load(fixture)
    P = params(101, 11, 11, 0, 12)
    it = kind(P, 256, tol)
    it.initialize()
    Qs = mc.DeviceArray.from_numpy(qi)
    Qbar = P.createMatrixQbar()
    it(Qs, Qbar)
    assert it.iteration() == iters
    assert it.converged
    assert np.abs(Qbar.numpy() - qb).max() < 1e-8


@pytest.mark.parametrize("kind,fixture,ofn", [
    (mc.RowBlockReductionIntegrator, "qb_re.mtx", oracle.reduction),
    (mc.RowBlockExtrinsicMeanIntegrator, "qb_em.mtx", oracle.extrinsic_mean),
])
def test_combining_integrator_fixture(gpu, kind, fixture, ofn):
    # check/isvd/integrator/row_block_{reduction,extrinsic_mean}_integrator.cpp: Qbar Qbar' at 1e-8
    note("integrator fixture %s" % kind.__name__)
    qi, qb = mtx.load("qi.mtx"), mtx.load(fixture)
    P = params(101, 11, 11, 0, 12)
    it = kind(P)
    Qbar = P.createMatrixQbar()
    it(mc.DeviceArray.from_numpy(qi), Qbar)
    Q = Qbar.numpy()
    assert it.iteration() == 0
    assert np.abs(proj(Q) - proj(qb)).max() < 1e-8
    assert np.allclose(Q.T @ Q, np.eye(11), atol=1e-12)
    # a larger case with an odd number of sketches (the tree re-reads entries of odd levels)
    m, l, N = 3000, 24, 13
    rng = np.random.default_rng(8)
    base, _ = np.linalg.qr(rng.standard_normal((m, l)))
    blocks = [np.linalg.qr(base + 0.2 * rng.standard_normal((m, l)))[0] for _ in range(N)]
    Qs = np.hstack(blocks)
    P2 = params(m, l, l, 0, N)
    it2 = kind(P2)
    Qbar2 = P2.createMatrixQbar()
    it2(mc.DeviceArray.from_numpy(Qs), Qbar2)
    Q2 = Qbar2.numpy()
    assert np.abs(proj(Q2) - proj(ofn(Qs, l, N))).max() < 1e-10
    assert np.allclose(Q2.T @ Q2, np.eye(l), atol=1e-12)


def test_symmetric_former_fixture(gpu):
    # check/isvd/former/row_block_symmetric_former.cpp: U U' at 1e-8 (k = 5, p = 6)
    note("former fixture RowBlockSymmetricFormer")
    a, q, u = mtx.load("a_sym.mtx"), mtx.load("q.mtx"), mtx.load("u_sym.mtx")
    k = 5
    P = params(101, 101, k, 6, 1)
    f = mc.RowBlockSymmetricFormer(P, True)
    f(mc.DeviceArray.from_numpy(np.asarray(a, order="C")), mc.DeviceArray.from_numpy(q))
    U, S = f.matrixU(), f.vectorS()
    assert np.abs(proj(U) - proj(u[:, :k])).max() < 1e-8
    S_o, U_o = oracle.symmetric_former(a, q, k)
    assert np.allclose(S, S_o, rtol=1e-12)
    assert np.all(np.diff(S) >= 0)  # ascending, as the reference returns them
    assert np.abs(proj(U) - proj(U_o)).max() < 1e-10
    with pytest.raises(mc.McnlaError):  # a non-square A is rejected
        mc.RowBlockSymmetricFormer(params(101, 102, k, 6, 1), True)(
            mc.DeviceArray.from_numpy(np.zeros((101, 102))), mc.DeviceArray.from_numpy(q))


def test_integrator_not_converged_returns_result(gpu):
    note("integrator maxiter exhaustion")
    qi = mtx.load("qi.mtx")
    P = params(101, 11, 11, 0, 12)
    it = mc.RowBlockKolmogorovNagumoIntegrator(P, 5, 1e-4)
    Qbar = P.createMatrixQbar()
    it(mc.DeviceArray.from_numpy(qi), Qbar)
    assert it.iteration() == 5 and it.converged is False
    Qo, ito = oracle.kn_rowblock(qi, 11, 12, 5, 1e-4)
    assert ito == 5
    assert np.abs(Qbar.numpy() - Qo).max() < 1e-10


FORMERS = [(mc.SvdFormer, "F"), (mc.GramianFormer, "F"), (mc.ColBlockGramianFormer, "F"),
           (mc.RowBlockGramianFormer, "C"), (mc.SvdFormer, "C")]


@pytest.mark.parametrize("kind,order", FORMERS)
def test_former_fixture(gpu, kind, order):
    # check/isvd/former/*.cpp: U projectors within 1e-8 (k = 5, p = 6)
    note("former fixture %s order=%s" % (kind.__name__, order))
    a, q, u = mtx.load("a.mtx"), mtx.load("q.mtx"), mtx.load("u.mtx")
    k = 5
    P = params(101, 1001, k, 6, 1)
    f = kind(P, True)
    f.initialize()
    f(mc.DeviceArray.from_numpy(np.asarray(a, order=order)), mc.DeviceArray.from_numpy(q))
    U, S, V = f.matrixU(), f.vectorS(), f.matrixV()
    assert np.abs(proj(U) - proj(u[:, :k])).max() < 1e-8
    S_ref, U_ref, V_ref = oracle.svd_former(a, q, k)
    assert np.allclose(S, S_ref, rtol=1e-12)
    assert np.abs(proj(V) - proj(V_ref)).max() < 1e-8
    assert np.allclose((q.T @ a) @ V, (q.T @ U) * S, atol=1e-9 * S[0])


def test_svd_former_wide_spectrum(gpu):
    """The SVD former on W = Qbar'A with singular values spanning 1 .. 1e-6 (known exactly:
    A = U diag(s) V', Qbar = U R with R orthogonal, so W = R' diag(s) V').  It takes the right
    vectors from eig(Z'Z), Z = A'Qbar, and sigma_j = ||Z w_j||, which keeps sigma to 1e-9 here
    (measured 4e-11 on the host with LAPACK's eigensolver; 4e-7 at 1e-7, and the squared form
    breaks down below about 1e-8, where a QR-based SVD of W would still give ~eps sigma_1)."""
    note("svd former wide spectrum")
    m, n, k, p = 2000, 900, 24, 8
    l = k + p
    rng = np.random.default_rng(14)
    s = np.logspace(0, -6, l)
    U, _ = np.linalg.qr(rng.standard_normal((m, l)))
    V, _ = np.linalg.qr(rng.standard_normal((n, l)))
    R, _ = np.linalg.qr(rng.standard_normal((l, l)))
    A = (U * s) @ V.T
    f = mc.SvdFormer(params(m, n, k, p, 1), True)
    f.initialize()
    f(mc.DeviceArray.from_numpy(np.asfortranarray(A)), mc.DeviceArray.from_numpy(U @ R))
    S = f.vectorS()
    note("  max rel sigma err %.3e" % np.max(np.abs(S - s[:k]) / s[:k]))
    assert np.max(np.abs(S - s[:k]) / s[:k]) < 1e-9
    assert np.abs(proj(f.matrixU()) - proj(U[:, :k])).max() < 1e-8


# ------------------------------------------------------------------------------------------------
def test_small_eigensolver_and_l64(gpu):
    """l = 64 (the largest supported): a pipeline through the stage API on a small problem,
    checked against the oracle stage by stage."""
    m, n, k, p, N = 700, 300, 50, 14, 2
    note("stage pipeline l=64")
    rng = np.random.default_rng(5)
    A = rng.standard_normal((m, n)) * (0.98 ** np.arange(n))
    P = params(m, n, k, p, N)
    sk = mc.RowBlockGaussianProjectionSketcher(P, seed=3, omega_mode=mc.OMEGA_REFERENCE)
    orth = mc.RowBlockGramianOrthogonalizer(P)
    integ = mc.RowBlockKolmogorovNagumoIntegrator(P, 300, 1e-4)
    form = mc.RowBlockGramianFormer(P, True)
    dA = mc.DeviceArray.from_numpy(np.ascontiguousarray(A))
    Y = P.createCollectionQ()
    sk(dA, Y)
    orth(Y)
    Qbar = P.createMatrixQbar()
    integ(Y, Qbar)
    form(dA, Qbar)
    l = k + p
    om = oracle.omega_ref(3, n, N * l, col_block=False)
    Yo = oracle.sketch(A, om, a_colmajor=False)
    Qo, _ = oracle.gramian_orthogonalize(Yo, l, N)
    for i in range(N):
        b = slice(l * i, l * (i + 1))
        assert np.abs(proj(Y.numpy()[:, b]) - proj(Qo[:, b])).max() < 1e-8
    Qbo, ito = oracle.kn_rowblock(Qo, l, N, 300, 1e-4)
    assert abs(integ.iteration() - ito) <= 1
    So, Uo, Vo = oracle.gramian_former(A, Qbo, k, a_colmajor=False)
    assert np.allclose(form.vectorS(), So, rtol=1e-8)
    assert np.abs(proj(form.matrixU()) - proj(Uo)).max() < 1e-6
