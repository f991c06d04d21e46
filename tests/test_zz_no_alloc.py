"""mcnla_solve_static (the solve that allocates nothing: its workspace is a static device array)
against mcnla_solver_solve with the same configuration, through the C ABI.  Two calls in a row,
the second smaller (and on another precision and layout), each bit-identical to the allocating
solver's results; a problem larger than the static array and a graph-recording configuration are
refused without running."""
import ctypes
import sys

import numpy as np
import pytest

import mcnla_amd as mc
from mcnla_amd import _lib

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def config(precision, seed, integrator=_lib.INT_ROW_BLOCK_KN, tol=1e-4):
    cfg = _lib.SolverConfig()
    _lib.lib().mcnla_solver_config_default(ctypes.byref(cfg))
    cfg.precision = precision
    cfg.seed = seed
    cfg.integrator = integrator
    cfg.tol = tol
    return cfg


def lowrank(m, n, seed, dtype, order):
    rng = np.random.default_rng(seed)
    r = min(m, n)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    A = (U * 0.9 ** np.arange(r)) @ V.T
    return mc.DeviceArray.from_numpy(np.asarray(A.astype(dtype), order=order), order=order)


def allocating(params, cfg, A):
    """The existing entry points: mcnla_solver_create + solve + results."""
    h = ctypes.c_void_p()
    L = _lib.lib()
    _lib.check(L.mcnla_solver_create(ctypes.byref(params), ctypes.byref(cfg), ctypes.byref(h)))
    try:
        _lib.check(L.mcnla_solver_solve(h, A.ptr, A.ld, A.dtype_code, A.order_code, None),
                   allow=(_lib.ERR_NOT_CONVERGED,))
        ptrs = [ctypes.c_void_p() for _ in range(4)]
        _lib.check(L.mcnla_solver_results(h, *[ctypes.byref(p) for p in ptrs]))
        it = ctypes.c_int64()
        _lib.check(L.mcnla_solver_iteration(h, ctypes.byref(it)))
        m, n, k = params.nrow, params.ncol, params.rank
        l = k + params.over_rank
        S = mc.DeviceArray.view(ptrs[0].value, (k,)).numpy()
        U = mc.DeviceArray.view(ptrs[1].value, (k * m,)).numpy()
        V = mc.DeviceArray.view(ptrs[2].value, (n * k,)).numpy()
        Q = mc.DeviceArray.view(ptrs[3].value, (m * l,)).numpy()
        return S, U, V, Q, it.value
    finally:
        L.mcnla_solver_destroy(h)


def static(params, cfg, A):
    m, n, k = params.nrow, params.ncol, params.rank
    l = k + params.over_rank
    outs = [mc.DeviceArray((c,)) for c in (k, m * k, n * k, m * l)]
    for o in outs:  # a recognizable fill, so a skipped copy cannot pass
        _lib.check(_lib.lib().mcnla_memset(o.ptr, 0x7f, o.nbytes))
    it = ctypes.c_int64(-1)
    _lib.check(_lib.lib().mcnla_solve_static(
        ctypes.byref(params), ctypes.byref(cfg), A.ptr, A.ld, A.dtype_code, A.order_code,
        outs[0].ptr, outs[1].ptr, outs[2].ptr, outs[3].ptr, ctypes.byref(it), None),
        allow=(_lib.ERR_NOT_CONVERGED,))
    return tuple(o.numpy() for o in outs) + (it.value,)


CASES = [
    # (m, n, k, p, N, precision, dtype, order, seed)
    (3000, 1200, 10, 10, 8, _lib.PREC_F64, np.float64, "F", 5),
    (1500, 700, 8, 6, 4, _lib.PREC_MIXED_BF16, np.float32, "C", 6),
]


def test_static_matches_solver_twice(gpu):
    for (m, n, k, p, N, prec, dtype, order, seed) in CASES:
        note("static solve %dx%d k=%d p=%d N=%d precision %d" % (m, n, k, p, N, prec))
        params = _lib.Params(m, n, k, p, N)
        cfg = config(prec, seed)
        need = ctypes.c_size_t()
        _lib.check(_lib.lib().mcnla_solver_required_bytes(ctypes.byref(params), ctypes.byref(cfg),
                                                          ctypes.byref(need)))
        assert 0 < need.value <= _lib.STATIC_WORKSPACE_BYTES
        A = lowrank(m, n, seed, dtype, order)
        got = static(params, cfg, A)
        ref = allocating(params, cfg, A)
        for x, y in zip(got[:4], ref[:4]):
            assert x.shape == y.shape
            assert np.array_equal(x, y)
        assert got[4] == ref[4] > 0
        assert np.all(np.isfinite(got[0])) and np.all(np.diff(got[0]) <= 0)


def test_static_refuses_oversized_and_graph(gpu):
    note("static solve refusals")
    A = lowrank(200, 100, 1, np.float64, "F")
    out = mc.DeviceArray((64,))
    it = ctypes.c_int64(-1)
    # 400000 x 10000 with l = 64, N = 64: Y alone is 400000 * 4096 * 8 bytes
    big = _lib.Params(400000, 10000, 32, 32, 64)
    st = _lib.lib().mcnla_solve_static(ctypes.byref(big), ctypes.byref(config(_lib.PREC_F64, 1)),
                                       A.ptr, 400000, _lib.F64, _lib.COL_MAJOR, out.ptr, None,
                                       None, None, ctypes.byref(it), None)
    assert st == _lib.ERR_NOMEM
    assert b"static workspace" in _lib.lib().mcnla_last_error()
    cfg = config(_lib.PREC_F64, 1)
    cfg.graph = 1
    small = _lib.Params(200, 100, 5, 5, 2)
    st = _lib.lib().mcnla_solve_static(ctypes.byref(small), ctypes.byref(cfg), A.ptr, A.ld,
                                       A.dtype_code, A.order_code, out.ptr, None, None, None,
                                       ctypes.byref(it), None)
    assert st == _lib.ERR_INVALID
    assert it.value == -1
