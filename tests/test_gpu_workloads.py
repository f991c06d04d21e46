"""The benchmark workloads as GPU tests, at full size (checked through properties the mathematics
fixes) and at reduced sizes with the same sketch shape (checked against the oracle pipeline).

* Headline: A 100000 x 10000 = U50 diag(0.9^i) V50' + E (Haar U50, V50; E = 1e-3 g / (sqrt(m) +
  sqrt(n))), k=20, p=12, N=64 (Y is 100000 x 2048), KN tol 1e-4, SVD former; FP64 (A FP64) and
  the BF16 mixed path (the same A in FP32) on one Philox Omega.
* Tall row-block: A 500000 x 2000 FP32 row-major with sigma_i = 1/(1+i)^2, k=50, p=14, N=128
  (Nl = 8192), row-block Wen-Yin tol 1e-3, row-block Gramian former: the Gram-of-A route against
  the sketch route.
* Serving: the first 256 requests of the serving stream (mixed m, n, k, N; sigma_i = rho^i), each
  result bit-identical alone and in an 8-slot mix, and the four cheapest against the oracle.

Tolerances are stated where they are asserted, with the reason.
"""
import sys

import numpy as np
import pytest

import mcnla_amd as mc
from mcnla_amd import workloads
import oracle

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def proj(Q):
    return Q @ Q.T


def rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def haar(rows, cols, seed):
    Q, R = np.linalg.qr(np.random.default_rng(seed).standard_normal((rows, cols)))
    return Q * np.sign(np.diag(R))


def params(m, n, k, p, N):
    return mc.Parameters().setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N)


# ------------------------------------------------------------------------------------------------
# Headline workload, full size.
HM, HN, HK, HP, HNS, HSEED = 100000, 10000, 20, 12, 64, 2


@pytest.fixture(scope="module")
def headline():
    U, V = haar(HM, 50, 1), haar(HN, 50, 11)
    sig = 0.9 ** np.arange(50)
    noise = 1e-3 / (np.sqrt(HM) + np.sqrt(HN))
    A64 = mc.synth_lowrank(HM, HN, U, V, sig, noise, seed=1, dtype=np.float64, order="F")
    A32 = mc.synth_lowrank(HM, HN, U, V, sig, noise, seed=1, dtype=np.float32, order="F")
    # ||E||_2 <= noise (sqrt(m) + sqrt(n) + t) except with probability exp(-t^2/2) (Gaussian
    # matrix norm concentration); t = 8.  FP32 storage adds at most 2^-24 ||A||_F < 1e-7.
    e_norm = noise * (np.sqrt(HM) + np.sqrt(HN) + 8.0) + 1e-7
    yield {"U": U, "V": V, "sig": sig, "A64": A64, "A32": A32, "e_norm": e_norm}
    A64.free()
    A32.free()


def test_headline_full_size(gpu, headline):
    """FP64 and mixed solves at 100000 x 10000 on the same Omega:
    * mixed sigma within 1e-5 of FP64's (the target), iteration counts within 1;
    * Qbar orthonormal to 1e-12 (both);
    * Ritz values against the generator: Qbar'A = (Qbar'U50) D V50' + Qbar'E, so by Weyl's
      inequality |S_i - sigma_i(Qbar'U50 D)| <= ||E||_2, and with eps = ||(I - Qbar Qbar')A||_2 <=
      ||(I - Qbar Qbar')U50 D||_2 + ||E||_2 the true values bracket them:
      sqrt((sigma_i - ||E||)^2 - eps^2) <= S_i <= sigma_i + ||E||."""
    note("headline full size: fp64 and mixed solves")
    P = params(HM, HN, HK, HP, HNS)
    f64 = mc.Solver(P, seed=HSEED).solve(headline["A64"])
    mix = mc.Solver(P, seed=HSEED, precision=mc.PREC_MIXED_BF16).solve(headline["A32"])
    S64, Smx = f64.vectorS(), mix.vectorS()
    note("  sigma rel err mixed vs fp64 %.3e, iterations %d / %d"
         % (rel(Smx, S64), mix.iteration(), f64.iteration()))
    assert rel(Smx, S64) <= 1e-5
    assert abs(mix.iteration() - f64.iteration()) <= 1
    U, sig, e = headline["U"], headline["sig"], headline["e_norm"]
    for name, s, S in (("fp64", f64, S64), ("mixed", mix, Smx)):
        Qb = s.matrixQbar()
        assert np.abs(Qb.T @ Qb - np.eye(HK + HP)).max() < 1e-12, name
        C = Qb.T @ U  # l x 50
        ritz_signal = np.linalg.svd(C * sig, compute_uv=False)[:HK]
        assert np.max(np.abs(S - ritz_signal)) <= e, name
        CD = C * sig
        r1 = np.sqrt(max(0.0, np.linalg.eigvalsh(np.diag(sig ** 2) - CD.T @ CD).max()))
        eps = r1 + e
        lo = np.sqrt(np.maximum(0.0, (sig[:HK] - e) ** 2 - eps ** 2))
        assert np.all(S <= sig[:HK] + e) and np.all(S >= lo), name
        note("  %s: max |S - sigma| / sigma vs generator %.3e, residual bound %.3e"
             % (name, rel(S, sig[:HK]), eps))


def test_headline_full_size_kn_fixed_point(gpu, headline):
    """The FP64 result is a KN fixed point of the collection's average projector
    P = (1/N) sum_i Q_i Q_i' (Q_s from the stage sketcher and orthogonalizer on the same Omega):
    X = (I - Qbar Qbar') P Qbar.  The integrator stops when ||s - 1||_2 < tol with
    s_i = sqrt(1/2 + sqrt(1/4 - lambda_i)), lambda_i the eigenvalues of X'X at the last iterate;
    1 - s >= lambda / 2 on [0, 1/4], so ||X||_2^2 < 2 tol there, and the returned (updated) basis is
    no worse: ||X||_2 <= sqrt(2 tol)."""
    note("headline full size: KN fixed point")
    P = params(HM, HN, HK, HP, HNS)
    s = mc.Solver(P, seed=HSEED).solve(headline["A64"])
    Qb = s.matrixQbar()
    sk = mc.ColBlockGaussianProjectionSketcher(P, HSEED)
    orth = mc.RowBlockGramianOrthogonalizer(P)
    for st in (sk, orth):
        st.initialize()
    Q = P.createCollectionQjp()
    sk(headline["A64"], Q)
    orth(Q)
    Qs = Q.numpy()
    PQ = Qs @ (Qs.T @ Qb) / HNS
    X = PQ - Qb @ (Qb.T @ PQ)
    res = np.linalg.norm(X, 2)
    note("  ||(I - Qbar Qbar') P Qbar||_2 = %.3e (bound %.3e)" % (res, np.sqrt(2e-4)))
    assert res <= np.sqrt(2e-4)


def test_headline_reduced_vs_oracle(gpu):
    """The headline spectrum and sketch shape (l = 32, N = 64, Nl = 2048) at 3000 x 2500, FP64 on
    the reference's Omega stream, against the oracle pipeline (sketch as the closed form A Omega,
    Gramian orthogonalizer, row-block KN, SVD former): sigma to 1e-9, U/V projectors to 1e-8, Qbar
    subspace distance to 1e-7 (KN is a contraction: results agree to rounding), iterations +-1."""
    m, n, k, p, N, seed = 3000, 2500, HK, HP, HNS, 2
    l = k + p
    note("headline reduced %dx%d l=%d N=%d vs oracle" % (m, n, l, N))
    U, V = haar(m, 50, 3), haar(n, 50, 4)
    rng = np.random.default_rng(5)
    A = (U * 0.9 ** np.arange(50)) @ V.T + 1e-3 * rng.standard_normal((m, n)) / (np.sqrt(m) + np.sqrt(n))
    om = oracle.omega_ref(seed, n, N * l, col_block=True)
    Q, _ = oracle.gramian_orthogonalize(A @ om, l, N)
    Qb_o, it_o = oracle.kn_rowblock(Q, l, N, 256, 1e-4)
    S_o, U_o, V_o = oracle.svd_former(A, Qb_o, k)
    s = mc.Solver(params(m, n, k, p, N), seed=seed, omega_mode=mc.OMEGA_REFERENCE,
                  route=mc.ROUTE_SKETCH)
    s.solve(mc.DeviceArray.from_numpy(np.asfortranarray(A)))
    note("  iterations gpu %d oracle %d, sigma rel err %.3e" % (s.iteration(), it_o, rel(s.vectorS(), S_o)))
    assert abs(s.iteration() - it_o) <= 1
    assert rel(s.vectorS(), S_o) < 1e-9
    assert np.abs(proj(s.matrixU()) - proj(U_o)).max() < 1e-8
    assert np.abs(proj(s.matrixV()) - proj(V_o)).max() < 1e-8
    assert np.linalg.norm(proj(s.matrixQbar()) - proj(Qb_o), 2) < 1e-7


# ------------------------------------------------------------------------------------------------
# Tall row-block workload.
TM, TN, TK, TP, TNS, TSEED = 500000, 2000, 50, 14, 128, 2


@pytest.fixture(scope="module")
def tall():
    """A = [diag(s_t) M]_t (250 row blocks of 2000, random signs s_t) with M = Q D V' / sqrt(250),
    Q, V Haar 2000 x 2000, D = diag(1/(1+i)^2), rounded to FP32.  The block stack of sign matrices
    over sqrt(250) has orthonormal columns, so sigma(A) = sqrt(250) sigma(fp32(M)) exactly."""
    blocks = TM // TN
    D = 1.0 / (1.0 + np.arange(TN)) ** 2
    M = ((haar(TN, TN, 42) * D) @ haar(TN, TN, 43).T / np.sqrt(blocks)).astype(np.float32)
    A, signs = workloads.sign_block_stack(blocks, M, seed=44)
    Mt = np.sqrt(blocks) * M.astype(np.float64)
    sig = np.linalg.svd(Mt, compute_uv=False)
    dA = mc.DeviceArray.from_numpy(A)
    del A
    yield {"A": dA, "signs": signs, "Mt": Mt, "sig": sig, "D": D, "blocks": blocks}
    dA.free()


def tall_kw():
    return dict(sketcher=mc.RowBlockGaussianProjectionSketcher,
                integrator=mc.RowBlockWenYinIntegrator, former=mc.RowBlockGramianFormer)


def test_tall_full_size_routes(gpu, tall):
    """500000 x 2000, N = 128: the Gram-of-A route (the automatic choice here: AUTO gives the same
    bits) against the sketch route (FP64 sketch and Gram of the 500000 x 8192 collection): equal
    iteration counts, sigma to 1e-9; Qbar orthonormal to 10 eps kappa^2; the returned sigma equal the
    singular values of Qbar'A formed on the host from the block structure (to 1e-8: the Gramian
    former takes sqrt of eig(Z'Z), which squares sigma_1/sigma_50 ~ 2600); and they lie between
    the true sigma and sqrt(sigma^2 - eps^2), eps = ||(I - Qbar Qbar')A||_2, also formed on the
    host.  The mixed precision on the automatic route keeps sigma to 1e-12 of FP64's."""
    P = params(TM, TN, TK, TP, TNS)
    A = tall["A"]
    out = {}
    for name, kw in (("auto", {}), ("gram_of_a", {"route": mc.ROUTE_GRAM_OF_A}),
                     ("sketch", {"route": mc.ROUTE_SKETCH}),
                     ("mixed", {"precision": mc.PREC_MIXED_BF16})):
        note("tall full size: %s" % name)
        s = mc.Solver(P, seed=TSEED, **tall_kw(), **kw).solve(A)
        out[name] = (s.vectorS(), s.matrixQbar(), s.iteration())
        del s
    S, Qb, it = out["gram_of_a"]
    assert np.array_equal(out["auto"][0], S) and out["auto"][2] == it
    Ss, Qs, its = out["sketch"]
    note("  iterations gram-of-A %d sketch %d, sigma rel err %.3e" % (it, its, rel(S, Ss)))
    assert it == its
    assert rel(S, Ss) < 1e-9
    Sm, _, itm = out["mixed"]
    assert itm == it and rel(Sm, S) < 1e-12
    B, n, l = tall["blocks"], TN, TK + TP
    # Both routes form Qbar from a Gram matrix (G = Y'Y, or Omega'(A'A)Omega), whose rounding is
    # amplified by kappa^2 in the directions of the smallest retained singular values (kappa =
    # sigma_1 / sigma_l of A = 65^2 here); the reference's Gramian orthogonalizer, which forms
    # Q_i = Y_i V_i S_i^{-1/2} from Y_i'Y_i, has the same property.
    kappa = tall["sig"][0] / tall["sig"][l - 1]
    for name, Q in (("gram_of_a", Qb), ("sketch", Qs)):
        orth = np.abs(Q.T @ Q - np.eye(l)).max()
        note("  %s: ||Qbar'Qbar - I||_max %.3e (bound 10 eps kappa^2 = %.3e)"
             % (name, orth, 10 * 2.2e-16 * kappa ** 2))
        assert orth < 10 * 2.2e-16 * kappa ** 2
    # Qbar'A = C Mt with C = Qbar' U, U = [diag(s_t)]_t / sqrt(B)
    Qr = Qb.reshape(B, n, l)
    C = np.einsum("tbl,tb->lb", Qr, tall["signs"].astype(np.float64)) / np.sqrt(B)
    Mt, sig = tall["Mt"], tall["sig"]
    ritz = np.linalg.svd(C @ Mt, compute_uv=False)[:TK]
    note("  sigma vs sigma(Qbar'A) %.3e, vs true sigma %.3e" % (rel(S, ritz), rel(S, sig[:TK])))
    assert rel(S, ritz) < 1e-8
    G = Mt.T @ Mt - (C @ Mt).T @ (C @ Mt)
    eps2 = max(0.0, np.linalg.eigvalsh((G + G.T) / 2).max())
    assert np.all(S <= sig[:TK] * (1 + 1e-9))
    assert np.all(S >= np.sqrt(np.maximum(0.0, sig[:TK] ** 2 - eps2)) * (1 - 1e-9))


def test_tall_reduced_vs_oracle(gpu):
    """The tall workload's spectrum, l = 64 and Wen-Yin at 3000 x 300 with N = 8 (Nl = 512 > n:
    the Gram-of-A route), A FP32 row-major, the reference's row-block Omega stream, against the
    oracle's row-block WY and row-block Gramian former.  WY's line search and alternating step
    amplify rounding (the reference's own row-block form moves 9e-9 in sigma under a 1e-15 input
    perturbation at the CPU-runnable workload), so sigma to 1e-6 and projectors to 1e-5;
    iterations +-1."""
    m, n, k, p, N, seed = 3000, 300, TK, TP, 8, 6
    l = k + p
    note("tall reduced %dx%d l=%d N=%d vs oracle" % (m, n, l, N))
    A = ((haar(m, n, 7) * (1.0 / (1.0 + np.arange(n)) ** 2)) @ haar(n, n, 8).T).astype(np.float32)
    A64 = A.astype(np.float64)
    om = oracle.omega_ref(seed, n, N * l, col_block=False)
    Q, _ = oracle.gramian_orthogonalize(A64 @ om, l, N)
    Qb_o, it_o = oracle.wy_rowblock(Q, l, N, 256, 1e-3)
    S_o, U_o, V_o = oracle.gramian_former(A64, Qb_o, k, a_colmajor=False)
    s = mc.Solver(params(m, n, k, p, N), seed=seed, omega_mode=mc.OMEGA_REFERENCE, **tall_kw())
    s.solve(mc.DeviceArray.from_numpy(np.ascontiguousarray(A)))
    note("  iterations gpu %d oracle %d, sigma rel err %.3e" % (s.iteration(), it_o, rel(s.vectorS(), S_o)))
    assert abs(s.iteration() - it_o) <= 1
    assert rel(s.vectorS(), S_o) < 1e-6
    assert np.abs(proj(s.matrixU()) - proj(U_o)).max() < 1e-5
    assert np.abs(proj(s.matrixV()) - proj(V_o)).max() < 1e-5


# ------------------------------------------------------------------------------------------------
# Serving stream.
SERVE_COUNT = 256


@pytest.fixture(scope="module")
def stream_mats():
    reqs = workloads.serving_stream(SERVE_COUNT)
    gen = workloads.ReflectedGenerator(50000, 5000)
    mats = [gen.make(r["m"], r["n"], workloads.stream_sigma(r), r["a_seed"]) for r in reqs]
    gen.free()
    yield reqs, mats
    for a in mats:
        a.free()


def test_reflected_generator(gpu):
    """The serving matrices: the device generator equals the host construction, and its singular
    values are exactly sigma (FP64, 1e-12 relative)."""
    gen = workloads.ReflectedGenerator(400, 200)
    sig = 0.9 ** np.arange(150)
    for order, dtype in (("F", np.float64), ("C", np.float64), ("F", np.float32)):
        A = gen.make(300, 150, sig, 9, dtype=dtype, order=order).numpy().astype(np.float64)
        H = workloads.ReflectedGenerator.host(300, 150, sig, 9)
        tol = 1e-14 if dtype == np.float64 else 1e-7
        assert np.abs(A - H).max() <= tol * np.abs(H).max()
        if dtype == np.float64:
            assert rel(np.linalg.svd(A, compute_uv=False), sig) < 1e-12
    gen.free()


def submit(srv, r, A):
    return srv.submit(A, r["m"], r["n"], r["k"], r["p"], r["N"], seed=r["omega_seed"])


def collect(srv, rid):
    res = srv.result(rid)
    assert res.ok, res.error
    out = (res.S.numpy(), res.U.numpy(), res.V.numpy(), res.Qbar.numpy(), res.iterations)
    srv.release(rid)
    return out


def test_serving_stream_alone_vs_mix(gpu, stream_mats):
    """256 requests of the serving stream on the mixed path: each result (S, U, V, Qbar, iteration
    count) is bit-identical when solved alone (one slot, one request at a time) and inside an
    8-slot mix with all 256 submitted at once (tolerance 0)."""
    reqs, mats = stream_mats
    kw = dict(hbm_budget=48 << 30, precision=mc.PREC_MIXED_BF16)
    note("serving stream: %d requests alone" % len(reqs))
    alone = []
    with mc.Server(slots=1, **kw) as srv:
        for r, A in zip(reqs, mats):
            rid = submit(srv, r, A)
            assert srv.wait(rid, 60.0)
            alone.append(collect(srv, rid))
    note("serving stream: %d requests in an 8-slot mix" % len(reqs))
    with mc.Server(slots=8, **kw) as srv:
        ids = [submit(srv, r, A) for r, A in zip(reqs, mats)]
        mix = {}
        for _ in ids:
            rid = srv.next_done(120.0)
            assert rid, "no request finished within 120 s"
            mix[rid] = collect(srv, rid)
        st = srv.stats()
    assert st["max_running"] > 1
    for i, rid in enumerate(ids):
        a, b = alone[i], mix[rid]
        for x, y in zip(a[:4], b[:4]):
            assert np.array_equal(x, y), "request %d differs alone vs mix" % i
        assert a[4] == b[4]


def test_serving_ready_event(gpu, stream_mats):
    """Requests submitted while their matrices are still being written on the generator's
    stream, each carrying the generator's event: the solve waits for it on the device and gives
    the same bits as the request whose matrix was written before submission."""
    reqs, mats = stream_mats
    picked = list(range(0, 64, 8))
    gen = workloads.ReflectedGenerator(50000, 5000, slots=4)  # slots wrap: write_async waits too
    fresh = [mc.DeviceArray((reqs[i]["m"], reqs[i]["n"]), np.float32, "F") for i in picked]
    with mc.Server(slots=4, hbm_budget=16 << 30, precision=mc.PREC_MIXED_BF16) as srv:
        ids = []
        for i, A in zip(picked, fresh):
            ev = gen.write_async(A, workloads.stream_sigma(reqs[i]), reqs[i]["a_seed"])
            ids.append(srv.submit(A, reqs[i]["m"], reqs[i]["n"], reqs[i]["k"], reqs[i]["p"],
                                  reqs[i]["N"], seed=reqs[i]["omega_seed"], ready_event=ev))
        got = []
        for rid in ids:
            assert srv.wait(rid, 60.0)
            got.append(collect(srv, rid))
        ref = []
        for i in picked:
            r = reqs[i]
            rid = submit(srv, r, mats[i])
            assert srv.wait(rid, 60.0)
            ref.append(collect(srv, rid))
    gen.free()
    for a, b in zip(got, ref):
        for x, y in zip(a[:4], b[:4]):
            assert np.array_equal(x, y)
        assert a[4] == b[4]
    for A in fresh:
        A.free()


def test_serving_stream_vs_oracle(gpu, stream_mats):
    """The four cheapest of those requests, served in FP64 alongside eight others, against the
    oracle pipeline on the same A (the FP32 matrix) and the same Philox Omega: sigma to 1e-9,
    U projector to 1e-8, iterations +-1."""
    reqs, mats = stream_mats
    order = sorted(range(len(reqs)),
                   key=lambda i: reqs[i]["m"] * (reqs[i]["k"] + 10) ** 2 * reqs[i]["N"])
    picked, others = order[:4], order[4:12]
    with mc.Server(slots=8, hbm_budget=16 << 30) as srv:
        ids = {i: submit(srv, reqs[i], mats[i]) for i in picked + others}
        got = {}
        for i in picked + others:
            assert srv.wait(ids[i], 60.0)
            got[i] = collect(srv, ids[i])
    for i in picked:
        r = reqs[i]
        m, n, k, N, l = r["m"], r["n"], r["k"], r["N"], r["k"] + r["p"]
        note("serving request %d (%dx%d k=%d N=%d) vs oracle" % (i, m, n, k, N))
        A = mats[i].numpy().astype(np.float64)
        om = mc.gaussian((n, N * l), seed=r["omega_seed"])
        Om = om.numpy()
        om.free()
        Q, _ = oracle.gramian_orthogonalize(A @ Om, l, N)
        Qb_o, it_o = oracle.kn_rowblock(Q, l, N, 256, 1e-4)
        S_o, U_o, _ = oracle.svd_former(A, Qb_o, k)
        S, U, V, Qb, it = got[i]
        assert abs(it - it_o) <= 1
        assert rel(S, S_o) < 1e-9
        assert np.abs(proj(U) - proj(U_o)).max() < 1e-8
