"""Batched solves (mcnla_solver_solve_batch, and the server's batch admission): every member's
results (S, U, V, Qbar), iteration count and status are bit-identical to its solve alone, across
shapes in both tile-width classes of the grouped Bs product (l <= 32 and l > 32, up to l = 64),
precisions, routes, tolerances, maxiter and integrators; a member that fails (NaN in A, maxiter
reached) leaves the others untouched; invalid batches are rejected before anything runs."""
import sys

import numpy as np
import pytest

import mcnla_amd as mc
from mcnla_amd import _lib

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def lowrank(m, n, rho, seed):
    rng = np.random.default_rng(seed)
    r = min(m, n)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return (U * rho ** np.arange(r)) @ V.T


KN = mc.RowBlockKolmogorovNagumoIntegrator
WY = mc.RowBlockWenYinIntegrator

# (m, n, k, p, N, rho, seed, options)
MEMBERS = [
    (900, 400, 5, 10, 8, 0.85, 1, {}),
    (1500, 700, 30, 10, 16, 0.9, 2, {}),                                    # l = 40: wide class
    (1200, 600, 12, 10, 12, 0.92, 3, {"precision": mc.PREC_MIXED_BF16}),
    (2000, 900, 44, 20, 6, 0.95, 4, {}),                                    # l = 64
    (800, 800, 10, 10, 5, 0.97, 5, {"max_iteration": 3}),                   # NOT_CONVERGED
    (1000, 500, 8, 10, 4, 0.88, 6, {"integrator": WY}),                     # not grouped
    (700, 350, 6, 10, 10, 0.8, 7, {"tolerance": 1e-8}),
    (600, 300, 5, 10, 4, 0.9, 8, {"nan": True}),                            # NUMERIC
    (3000, 150, 10, 10, 8, 0.9, 9, {"route": mc.ROUTE_GRAM_OF_A}),
    (1100, 520, 20, 10, 9, 0.93, 10, {"precision": mc.PREC_MIXED_F16}),
    (650, 640, 7, 10, 13, 0.86, 11, {}),                                    # odd Nl (unaligned rows)
]


def make(member):
    m, n, k, p, N, rho, seed, opt = member
    A = lowrank(m, n, rho, seed)
    if opt.get("nan"):
        A[17, 5] = np.nan
    prec = opt.get("precision", mc.PREC_F64)
    dA = mc.DeviceArray.from_numpy(np.asfortranarray(A.astype(np.float64 if prec == mc.PREC_F64
                                                              else np.float32)))
    kw = {k2: v for k2, v in opt.items() if k2 != "nan"}
    P = mc.Parameters().setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N)
    return P, kw, dA


def outputs(s):
    return (s.vectorS(), s.matrixU(), s.matrixV(), s.matrixQbar(), s.iteration())


def alone(P, kw, dA, seed):
    s = mc.Solver(P, seed=seed, **kw)
    try:
        s.solve(dA)
        return _lib.ERR_NOT_CONVERGED if not s.converged else _lib.OK, s
    except mc.McnlaError as e:
        return e.status, s


@pytest.fixture(scope="module")
def members():
    return [make(mb) for mb in MEMBERS]


def test_batch_equals_alone(gpu, members):
    note("batch: %d members (both tile classes, F64/BF16/F16, WY, Gram-of-A, NaN, maxiter) vs alone"
         % len(members))
    ref = []
    for i, (P, kw, dA) in enumerate(members):
        st, s = alone(P, kw, dA, 500 + i)
        ref.append((st, outputs(s) if st in (_lib.OK, _lib.ERR_NOT_CONVERGED) else None))
    assert ref[4][0] == _lib.ERR_NOT_CONVERGED and ref[7][0] == _lib.ERR_NUMERIC
    assert all(r[0] == _lib.OK for i, r in enumerate(ref) if i not in (4, 7))
    solvers = [mc.Solver(P, seed=500 + i, **kw) for i, (P, kw, _) in enumerate(members)]
    for rep in range(2):  # the same solvers twice: a batch leaves nothing behind
        sts = _batch_status(solvers, [dA for _, _, dA in members])
        for i, (st, out) in enumerate(ref):
            assert sts[i] == st, (rep, i, sts[i], st)
            if out is None:
                continue
            got = outputs(solvers[i])
            for a, b in zip(got[:4], out[:4]):
                assert a.shape == b.shape and np.array_equal(a, b), (rep, i)
            assert got[4] == out[4], (rep, i, got[4], out[4])
        note("  rep %d: %d members bit-identical, iterations %s" %
             (rep, len(ref), [solvers[i].iteration() for i in range(len(ref))]))


def _batch_status(solvers, mats):
    """Status codes of a batch with failing members (solve_batch raises on the first)."""
    import ctypes
    n = len(solvers)
    hs = (ctypes.c_void_p * n)(*[s._h for s in solvers])
    As = (ctypes.c_void_p * n)(*[a.ptr for a in mats])
    lds = (ctypes.c_int64 * n)(*[a.ld for a in mats])
    dts = (ctypes.c_int * n)(*[a.dtype_code for a in mats])
    ods = (ctypes.c_int * n)(*[a.order_code for a in mats])
    sts = (ctypes.c_int * n)()
    _lib.check(_lib.lib().mcnla_solver_solve_batch(n, hs, As, lds, dts, ods, sts, None))
    return list(sts)


def test_solve_batch_wrapper(gpu, members):
    note("batch: solve_batch of converging members, against alone")
    idx = [0, 1, 3, 6, 10]
    solvers = [mc.Solver(members[i][0], seed=900 + i, **members[i][1]) for i in idx]
    sts = mc.solve_batch(solvers, [members[i][2] for i in idx])
    assert sts == [_lib.OK] * len(idx)
    for j, i in enumerate(idx):
        P, kw, dA = members[i]
        s = mc.Solver(P, seed=900 + i, **kw).solve(dA)
        for a, b in zip(outputs(solvers[j]), outputs(s)):
            assert np.array_equal(a, b)
    with pytest.raises(mc.McnlaError) as e:  # NUMERIC member raises through the wrapper
        mc.solve_batch([mc.Solver(members[7][0], seed=1)], [members[7][2]])
    assert e.value.status == _lib.ERR_NUMERIC


def test_batch_rejects_invalid(gpu, members):
    note("batch: invalid batches fail before anything runs")
    import ctypes
    s0 = mc.Solver(members[0][0], seed=1)
    s1 = mc.Solver(members[1][0], seed=1)
    d0, d1 = members[0][2], members[1][2]

    def call(solvers, mats, count=None):
        n = len(solvers)
        hs = (ctypes.c_void_p * max(n, 1))(*[s._h for s in solvers])
        As = (ctypes.c_void_p * max(n, 1))(*[a.ptr for a in mats])
        lds = (ctypes.c_int64 * max(n, 1))(*[a.ld for a in mats])
        dts = (ctypes.c_int * max(n, 1))(*[a.dtype_code for a in mats])
        ods = (ctypes.c_int * max(n, 1))(*[a.order_code for a in mats])
        sts = (ctypes.c_int * max(n, 1))()
        return _lib.lib().mcnla_solver_solve_batch(n if count is None else count, hs, As, lds, dts,
                                                   ods, sts, None)

    assert call([], [], count=0) == _lib.ERR_INVALID
    assert call([s0] * 17, [d0] * 17) == _lib.ERR_INVALID  # more than MCNLA_MAX_BATCH
    assert call([s0, s0], [d0, d0]) == _lib.ERR_INVALID    # a solver twice
    assert call([s0, s1], [d0, d0]) == _lib.ERR_INVALID    # A of the wrong shape for s1
    # the solvers still work
    assert call([s0, s1], [d0, d1]) == _lib.OK


def test_server_batches_equal_alone(gpu, members):
    """One slot: requests submitted one at a time run alone; submitted together behind a long
    first request, the rest queue up and are admitted as one batch.  Results are the same."""
    note("serve: batch_max 8 == one at a time (1 slot)")
    idx = [3, 0, 1, 6, 10, 9]
    cases = [(members[i], MEMBERS[i]) for i in idx]
    # FP64 copies of the FP32 matrices, alive until the end (the server reads device A in place)
    mats = [dA if dA.dtype == np.float64 else
            mc.DeviceArray.from_numpy(np.asfortranarray(dA.numpy().astype(np.float64)))
            for (_, _, dA), _ in cases]
    for batch_max in (1, 8):
        with mc.Server(slots=1, hbm_budget=4 << 30, batch_max=batch_max,
                       precision=mc.PREC_F64) as srv:
            ids = {}
            for j, (_, mb) in enumerate(cases):
                m, n, k, p, N = mb[:5]
                rid = srv.submit(mats[j], m, n, k, p, N, seed=700 + j)
                ids[rid] = j
                if batch_max == 1:
                    assert srv.wait(rid, 60.0)
            res = {}
            for rid, j in ids.items():
                assert srv.wait(rid, 60.0)
                r = srv.result(rid)
                assert r.ok, r.error
                res[j] = (r.S.numpy(), r.U.numpy(), r.V.numpy(), r.Qbar.numpy(), r.iterations)
                srv.release(rid)
            st = srv.stats()
            note("  batch_max %d: %s" % (batch_max, st))
            assert st["completed"] == len(cases) and st["failed"] == 0
            if batch_max == 1:
                base = res
                assert st["batches"] == len(cases)
            else:
                assert st["batches"] < len(cases), "no request was admitted in a batch"
                for j in base:
                    for a, b in zip(res[j][:4], base[j][:4]):
                        assert np.array_equal(a, b), j
                    assert res[j][4] == base[j][4]
