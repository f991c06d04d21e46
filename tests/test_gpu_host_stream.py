"""The fused solve of a host matrix (mcnla_solver_solve_host): A streamed to the device in chunks
beside the sketch of the chunks that have landed.  In FP64 on the sketch route the chunked sums
only reorder Y = A Omega's K sum, so results equal the device-resident solve's to rounding (sigma
1e-12, projectors 1e-10, equal iterations); with a mixed precision or on the Gram-of-A route the
copy completes before the sketch and results are bit-identical.  Ragged chunks, the automatic
chunk size, a second call on the same solver and a bad chunk size are covered."""
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


def params(m, n, k, p, N):
    return mc.Parameters().setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N)


def outputs(s):
    return s.vectorS(), s.matrixU(), s.matrixV(), s.matrixQbar(), s.iteration()


def proj_err(U, V):
    return np.linalg.norm(U @ U.T - V @ V.T, 2)


@pytest.mark.parametrize("order,chunk", [("F", 37), ("C", 500), ("F", 0), ("C", 0)])
def test_host_stream_f64_matches_device(gpu, order, chunk):
    m, n, k, p, N = 3000, 1200, 10, 6, 8
    note("host stream FP64 %s-order chunk %d" % (order, chunk))
    A = lowrank(m, n, 0.9, 3)
    A = np.asfortranarray(A) if order == "F" else np.ascontiguousarray(A)
    dA = mc.DeviceArray.from_numpy(A, order=order)
    P = params(m, n, k, p, N)
    ref = outputs(mc.Solver(P, seed=11).solve(dA))
    s = mc.Solver(P, seed=11)
    for rep in range(2):  # the second call reuses the solver's device copy of A
        got = outputs(s.solve_host(A, chunk=chunk))
        rel = np.max(np.abs(got[0] - ref[0]) / ref[0])
        eu = proj_err(got[1], ref[1])
        eq = proj_err(got[3], ref[3])
        note("  rep %d: sigma rel %.2e, U projector %.2e, Qbar projector %.2e, iterations %d/%d"
             % (rep, rel, eu, eq, got[4], ref[4]))
        assert rel <= 1e-12 and eu <= 1e-10 and eq <= 1e-10
        assert got[4] == ref[4]


@pytest.mark.parametrize("case", ["bf16_colmajor", "gram_of_a_rowmajor"])
def test_host_stream_bit_identical(gpu, case):
    note("host stream %s" % case)
    if case == "bf16_colmajor":
        m, n, k, p, N = 2500, 900, 8, 8, 6
        A = np.asfortranarray(lowrank(m, n, 0.88, 4).astype(np.float32))
        kw = {"precision": mc.PREC_MIXED_BF16}
        dA = mc.DeviceArray.from_numpy(A, order="F")
    else:
        m, n, k, p, N = 3000, 150, 10, 10, 8
        A = np.ascontiguousarray(lowrank(m, n, 0.9, 5))
        kw = {"route": mc.ROUTE_GRAM_OF_A}
        dA = mc.DeviceArray.from_numpy(A, order="C")
    P = params(m, n, k, p, N)
    ref = outputs(mc.Solver(P, seed=7, **kw).solve(dA))
    got = outputs(mc.Solver(P, seed=7, **kw).solve_host(A, chunk=123))
    for a, b in zip(got[:4], ref[:4]):
        assert np.array_equal(a, b)
    assert got[4] == ref[4]


def test_host_stream_rejects_bad_chunk(gpu):
    m, n = 400, 200
    A = np.asfortranarray(lowrank(m, n, 0.9, 6))
    s = mc.Solver(params(m, n, 5, 5, 2), seed=1)
    with pytest.raises(mc.McnlaError) as e:
        s.solve_host(A, chunk=-1)
    assert e.value.status == _lib.ERR_INVALID
    s.solve_host(A, chunk=0)  # and the solver still works
    assert np.all(np.isfinite(s.vectorS()))
