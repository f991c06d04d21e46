"""The serving loop on the GPU: served results equal the stand-alone solver's and the oracle's,
each result is the same alone and in the mix (tolerance 0: bit-identical), admission respects the
HBM budget, and the request life cycle reports errors."""
import sys

import numpy as np
import pytest

import mcnla_amd as mc
from mcnla_amd import _lib
import oracle

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def lowrank(m, n, rho, seed, dtype=np.float64):
    """A = U diag(rho^i) V' with Haar U, V (the serving stream's matrices)."""
    rng = np.random.default_rng(seed)
    r = min(m, n)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return np.asfortranarray(((U * rho ** np.arange(r)) @ V.T).astype(dtype))


# A mixed stream of small requests: (m, n, k, p, N, rho, seed).
MIX = [(900, 400, 5, 10, 8, 0.85, 11), (1500, 700, 12, 10, 16, 0.9, 12),
       (600, 300, 8, 10, 4, 0.8, 13), (2100, 1000, 20, 10, 12, 0.93, 14),
       (1200, 1200, 10, 10, 8, 0.88, 15), (3000, 500, 6, 10, 20, 0.9, 16),
       (800, 800, 30, 10, 6, 0.95, 17), (2500, 900, 15, 10, 10, 0.86, 18),
       (700, 350, 5, 10, 32, 0.82, 19), (1800, 600, 25, 10, 8, 0.9, 20)]


@pytest.fixture(scope="module")
def mats():
    return [mc.DeviceArray.from_numpy(lowrank(m, n, rho, sd), order="F")
            for (m, n, k, p, N, rho, sd) in MIX]


def fetch(srv, rid):
    assert srv.wait(rid, 60.0), "request %d not finished within 60 s" % rid
    r = srv.result(rid)
    assert r.ok, r.error
    out = (r.S.numpy(), r.U.numpy(), r.V.numpy(), r.Qbar.numpy(), r.iterations)
    srv.release(rid)
    return out


def same(a, b):
    for x, y in zip(a[:4], b[:4]):
        assert x.shape == y.shape
        assert np.array_equal(x, y)
    assert a[4] == b[4]


def test_served_equals_solver(gpu, mats):
    note("serve: served == stand-alone solver (3 shapes, FP64 and mixed)")
    for prec in (mc.PREC_F64, mc.PREC_MIXED_BF16):
        with mc.Server(slots=2, hbm_budget=2 << 30, precision=prec) as srv:
            for i in (0, 3, 6):
                m, n, k, p, N, _, sd = MIX[i]
                P = mc.Parameters().setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N)
                A = mats[i]
                if prec != mc.PREC_F64:
                    A = mc.DeviceArray.from_numpy(A.numpy().astype(np.float32), order="F")
                s = mc.Solver(P, seed=100 + i, precision=prec).solve(A)
                ref = (s.vectorS(), s.matrixU(), s.matrixV(), s.matrixQbar(), s.iteration())
                got = fetch(srv, srv.submit(A, m, n, k, p, N, seed=100 + i))
                same(got, ref)


def test_served_matches_oracle(gpu):
    note("serve: reference-Omega request vs the oracle pipeline")
    m, n, k, p, N = 1000, 500, 8, 10, 4
    A = lowrank(m, n, 0.9, 5)
    l = k + p
    om = oracle.omega_ref(7, n, N * l, col_block=True)
    Q, _ = oracle.gramian_orthogonalize(oracle.sketch(A, om), l, N)
    Qb, it = oracle.kn_rowblock(Q, l, N, 256, 1e-4)
    S, Uo, Vo = oracle.svd_former(A, Qb, k)
    with mc.Server(slots=1, hbm_budget=1 << 30, omega_mode=mc.OMEGA_REFERENCE) as srv:
        rid = srv.submit(A, m, n, k, p, N, seed=7)  # host A: copied into the arena
        got_S, got_U, got_V, got_Q, got_it = fetch(srv, rid)
    assert abs(got_it - it) <= 1
    assert np.max(np.abs(got_S - S) / S) < 1e-10
    assert np.linalg.norm(got_U @ got_U.T - Uo @ Uo.T, 2) < 1e-8
    assert np.linalg.norm(got_Q @ got_Q.T - Qb @ Qb.T, 2) < 1e-8


def test_alone_vs_mix(gpu, mats):
    note("serve: each result alone (1 slot, one at a time) == in the mix (4 slots, all at once)")
    for prec in (mc.PREC_F64, mc.PREC_MIXED_BF16):
        alone = {}
        with mc.Server(slots=1, hbm_budget=4 << 30, precision=prec) as srv:
            for i, (m, n, k, p, N, _, _) in enumerate(MIX):
                rid = srv.submit(mats[i], m, n, k, p, N, seed=1000 + i)
                assert srv.wait(rid, 60.0)
                alone[i] = fetch(srv, rid)
            assert srv.stats()["max_running"] == 1
        with mc.Server(slots=4, hbm_budget=4 << 30, precision=prec) as srv:
            ids = {srv.submit(mats[i], m, n, k, p, N, seed=1000 + i): i
                   for i, (m, n, k, p, N, _, _) in enumerate(MIX)}
            seen = 0
            while seen < len(ids):
                rid = srv.next_done(60.0)
                assert rid is not None, "no completion within 60 s"
                same(fetch(srv, rid), alone[ids[rid]])
                seen += 1
            st = srv.stats()
            note("serve: mix stats %s" % st)
            assert st["completed"] == len(MIX) and st["failed"] == 0
            assert st["max_running"] > 1


def test_budget_admission(gpu, mats):
    note("serve: a budget that fits one request at a time")
    m, n, k, p, N, _, _ = MIX[3]
    with mc.Server(slots=4, hbm_budget=1 << 30) as probe:
        need = probe.request_bytes(mats[3], m, n, k, p, N)
    budget = need + need // 2  # room for one such request, not two
    with mc.Server(slots=4, hbm_budget=budget) as srv:
        ids = [srv.submit(mats[3], m, n, k, p, N, seed=s) for s in range(4)]
        outs = []
        for rid in ids:
            assert srv.wait(rid, 60.0)
            outs.append(fetch(srv, rid))
        st = srv.stats()
        assert st["max_running"] == 1, st
        assert st["peak_bytes"] <= budget
        # different seeds give different sketches; each recovers the leading sigma_i = rho^i
        rho = MIX[3][5]
        for o in outs:
            assert np.max(np.abs(o[0][:3] - rho ** np.arange(3)) / rho ** np.arange(3)) < 1e-2
        big = MIX[5]
        with pytest.raises(mc.McnlaError) as e:
            srv.submit(mats[5], big[0], big[1], 40, 20, 64)
        assert e.value.status == _lib.ERR_NOMEM


def test_lifecycle_errors(gpu, mats):
    note("serve: request life cycle")
    m, n, k, p, N, _, _ = MIX[0]
    with mc.Server(slots=1, hbm_budget=1 << 30) as srv:
        with pytest.raises(mc.McnlaError) as e:
            srv.submit(mats[0], m, n, 400, 10, 4)  # k + p > min(m, n)
        assert e.value.status == _lib.ERR_INVALID
        rid = srv.submit(mats[0], m, n, k, p, N, seed=1)
        assert srv.wait(rid, 60.0)
        srv.release(rid)
        with pytest.raises(mc.McnlaError) as e:
            srv.result(rid)
        assert e.value.status == _lib.ERR_INVALID
