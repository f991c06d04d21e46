"""The C-ABI boundary on the CPU: the library loads, exports every function include/*.h declares
(and nothing undeclared under the mcnla_ prefix), and rejects bad arguments before touching a
device.  No compute runs here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import mcnla_amd as mc
from mcnla_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include", "mcnla_amd")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mcnla_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def exported_functions():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if re.search(r" T mcnla_", ln)}


def test_library_loads():
    assert mc.lib() is not None
    assert b"gfx950" in mc.lib().mcnla_version()


def test_every_declared_symbol_is_exported():
    decl = declared_functions()
    assert len(decl) > 30
    exp = exported_functions()
    assert decl <= exp, sorted(decl - exp)
    assert exp <= decl, sorted(exp - decl)  # nothing exported that the header does not declare
    assert decl == set(_lib.SIGNATURES), sorted(decl ^ set(_lib.SIGNATURES))


def test_header_is_plain_c():
    # the boundary must compile as C (no torch/HIP/C++ types in signatures)
    hdr = os.path.join(ROOT, "include", "mcnla_amd", "isvd_capi.h")
    r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-x", "c", hdr], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


def _params(m=100, n=80, k=5, p=5, N=4):
    return _lib.Params(m, n, k, p, N)


@pytest.mark.parametrize("prm,msg", [
    (_params(k=60, p=10), "rank + over_rank"),
    (_params(m=10, k=8, p=4), "rank + over_rank"),
    (_params(N=0), "num_sketch"),
    (_params(k=0), "rank"),
    (_params(m=0), "matrix size"),
])
def test_invalid_parameters_rejected(prm, msg):
    h = ctypes.c_void_p()
    L = mc.lib()
    for st in (L.mcnla_sketcher_create(ctypes.byref(prm), 0, 0, ctypes.byref(h)),
               L.mcnla_orthogonalizer_create(ctypes.byref(prm), 0, 0, ctypes.byref(h)),
               L.mcnla_integrator_create(ctypes.byref(prm), 1, 0, 256, 1e-4, ctypes.byref(h)),
               L.mcnla_former_create(ctypes.byref(prm), 0, 1, 0, ctypes.byref(h))):
        assert st == _lib.ERR_INVALID
        assert msg in L.mcnla_last_error().decode()


def test_invalid_kinds_rejected():
    L = mc.lib()
    h = ctypes.c_void_p()
    prm = _params()
    assert L.mcnla_integrator_create(ctypes.byref(prm), 9, 0, 256, 1e-4, ctypes.byref(h)) == _lib.ERR_INVALID
    assert L.mcnla_integrator_create(ctypes.byref(prm), 1, 0, -1, 1e-4, ctypes.byref(h)) == _lib.ERR_INVALID
    assert L.mcnla_integrator_create(ctypes.byref(prm), 1, 0, 10, -1.0, ctypes.byref(h)) == _lib.ERR_INVALID
    assert L.mcnla_former_create(ctypes.byref(prm), 7, 1, 0, ctypes.byref(h)) == _lib.ERR_INVALID
    assert L.mcnla_sketcher_create(ctypes.byref(prm), 0, 9, ctypes.byref(h)) == _lib.ERR_INVALID
    cfg = _lib.SolverConfig()
    L.mcnla_solver_config_default(ctypes.byref(cfg))
    assert cfg.integrator == _lib.INT_ROW_BLOCK_KN and cfg.maxiter == 256 and cfg.tol == 1e-4
    cfg.omega_mode = 5
    assert L.mcnla_solver_create(ctypes.byref(prm), ctypes.byref(cfg), ctypes.byref(h)) == _lib.ERR_INVALID


def test_reference_omega_matches_oracle():
    import oracle
    L = mc.lib()
    for kind, cb in ((_lib.SKETCHER_COL_BLOCK_GAUSSIAN, True), (_lib.SKETCHER_ROW_BLOCK_GAUSSIAN, False),
                     (_lib.SKETCHER_GAUSSIAN, True)):
        for threads in (1, 3, 8):
            out = np.empty(1234)
            assert L.mcnla_omega_reference(2024, kind, threads, out.size,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
            ref = oracle.omega_ref(2024, 1, 1234, col_block=cb, threads=threads).ravel()
            assert np.array_equal(out, ref)


def test_python_mirror_names():
    # the reference's stage class names are all available (src/isvd_dense_cb.cpp:20-34 macros)
    for name in ("ColBlockGaussianProjectionSketcher", "RowBlockGaussianProjectionSketcher",
                 "GaussianProjectionSketcher", "GramianOrthogonalizer",
                 "RowBlockGramianOrthogonalizer", "KolmogorovNagumoIntegrator",
                 "RowBlockKolmogorovNagumoIntegrator", "RowBlockGramianKolmogorovNagumoIntegrator",
                 "RowBlockWenYinIntegrator", "RowBlockGramianWenYinIntegrator", "SvdFormer",
                 "GramianFormer", "ColBlockGramianFormer", "RowBlockGramianFormer",
                 "CollectionFromPartialSumToRowBlockConverter", "MatrixFromRowBlockToAllConverter",
                 "MatrixFromColBlockToAllConverter", "MatrixFromRowBlockConverter", "Parameters"):
        assert hasattr(mc, name), name
    P = mc.Parameters().setSize(2000, 1000).setRank(10).setOverRank(10).setNumSketch(4)
    assert P.dimSketch() == 20 and P.dimSketchTotal() == 80 and P.nrowRank() == 2000


def test_required_bytes_without_device():
    # the serving loop sizes a request's arena region from this; it needs no GPU
    m, n, k, p, N = 100000, 10000, 20, 12, 64
    nl = (k + p) * N
    f64 = mc.required_bytes(m, n, k, p, N)
    mixed = mc.required_bytes(m, n, k, p, N, precision=mc.PREC_MIXED_BF16)
    # Y (m x Nl FP64), G, Bs and a scratch Nl x Nl, Omega n x Nl, Qbar and U
    floor = 8 * (m * nl + 2 * nl * nl + n * nl + m * (k + p) + m * k)
    assert floor <= f64 < floor + 2 * 8 * 32 * nl * nl
    # the mixed sketch adds BF16 panels of A (m x n) and Omega
    assert mixed - f64 >= 2 * m * n
    assert mc.required_bytes(1000, 500, 5, 10, 8) < mc.required_bytes(2000, 500, 5, 10, 8)
    with pytest.raises(mc.McnlaError) as e:
        mc.required_bytes(10, 5, 5, 10, 8)
    assert e.value.status == _lib.ERR_INVALID


def test_route_choice_by_shape():
    # n small against Nl (500000 x 2000, Nl = 8192): the Gram-of-A route holds A'A and A as FP64
    # instead of the m x Nl collection; at 100000 x 10000, Nl = 2048 the sketch route is kept
    m, n, k, p, N = 500000, 2000, 50, 14, 128
    auto = mc.required_bytes(m, n, k, p, N, precision=mc.PREC_MIXED_BF16)
    via_a = mc.required_bytes(m, n, k, p, N, precision=mc.PREC_MIXED_BF16, route=mc.ROUTE_GRAM_OF_A)
    via_y = mc.required_bytes(m, n, k, p, N, precision=mc.PREC_MIXED_BF16, route=mc.ROUTE_SKETCH)
    assert auto == via_a and via_y > via_a + 8 * m * (k + p) * N - 8 * m * n - 16 * n * (k + p) * N
    m, n, k, p, N = 100000, 10000, 20, 12, 64
    assert mc.required_bytes(m, n, k, p, N) == mc.required_bytes(m, n, k, p, N, route=mc.ROUTE_SKETCH)


def test_device_view_leading_dimension():
    """A view of a block of a larger matrix carries the larger matrix's leading dimension (no
    device memory is touched)."""
    v = mc.DeviceArray.view(0x1000, (300, 40), np.float32, "F", ld=301)
    assert v.ld == 301 and v.order_code == _lib.COL_MAJOR and v.dtype_code == _lib.F32
    assert mc.DeviceArray.view(0x1000, (300, 40), np.float64, "C").ld == 40


def test_column_indices_match_oracle():
    """The column-sampling index stream of the library's host generator equals the oracle's for
    both seed chains and several thread chunkings; every index lies in [0, ncol) and the stream is
    uniform (chi-square over 50 bins).  The bit stream itself is parity unpinned, as Omega's: no
    reference fixture stores it."""
    import oracle
    for kind, cb in ((mc.ColumnSamplingSketcher, True), (mc.RowBlockColumnSamplingSketcher, False)):
        for threads in (1, 3, 8):
            got = mc.column_indices_reference(2024, 997, 1500, kind=kind, threads=threads)
            assert np.array_equal(got, oracle.column_indices_ref(2024, 997, 1500, col_block=cb,
                                                                 threads=threads))
            assert got.min() >= 0 and got.max() < 997
    big = mc.column_indices_reference(5, 50, 200000)
    counts = np.bincount(big, minlength=50)
    chi2 = ((counts - 4000.0) ** 2 / 4000.0).sum()
    assert chi2 < 100.0  # 49 degrees of freedom: p ~ 3e-5 at 100


def test_sketcher_variants_rejected_where_the_reference_has_none():
    L = mc.lib()
    h = ctypes.c_void_p()
    prm = _params()
    # column sampling copies A: no mixed precision
    assert L.mcnla_sketcher_create(ctypes.byref(prm), _lib.SKETCHER_COLUMN_SAMPLING,
                                   _lib.PREC_MIXED_BF16, ctypes.byref(h)) == _lib.ERR_INVALID
    assert L.mcnla_sketcher_create(ctypes.byref(prm), 5, 0, ctypes.byref(h)) == _lib.ERR_INVALID
    # the fused solver sketches with a Gaussian projection only
    with pytest.raises(mc.McnlaError) as e:
        cfg = _lib.SolverConfig()
        L.mcnla_solver_config_default(ctypes.byref(cfg))
        cfg.sketcher = _lib.SKETCHER_COLUMN_SAMPLING
        b = ctypes.c_size_t()
        _lib.check(L.mcnla_solver_required_bytes(ctypes.byref(prm), ctypes.byref(cfg), ctypes.byref(b)))
    assert e.value.status == _lib.ERR_INVALID


def _shape(m, n, k, p, N):
    return mc.Parameters().setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N)


def test_tune_table_round_trip(tmp_path):
    """The autotuner's plan table on the host: defaults, set/get per shape class, route AUTO
    resolved as the solver resolves it, save/load, rejection of bad targets and bad files."""
    mc.tune_clear()
    P = _shape(100000, 10000, 20, 12, 64)
    d = mc.tuning_default()
    assert d == {"gram": 2304, "former": 4096, "h_omega": 2304, "omega_p": 1024, "n_by_l": 512,
                 "qbar": 512, "kernels": 0}
    got, found = mc.tune_get(P, precision=_lib.PREC_MIXED_BF16)
    assert not found and got == d
    t = dict(d, gram=4608, former=2048)
    mc.tune_set(P, t, precision=_lib.PREC_MIXED_BF16)
    assert mc.tune_get(P, precision=_lib.PREC_MIXED_BF16) == (t, True)
    # AUTO resolves to the sketch route at this shape: the same entry
    assert mc.tune_get(P, precision=_lib.PREC_MIXED_BF16, route=_lib.ROUTE_SKETCH) == (t, True)
    assert not mc.tune_get(P, precision=_lib.PREC_MIXED_BF16, route=_lib.ROUTE_GRAM_OF_A)[1]
    assert not mc.tune_get(P, precision=_lib.PREC_F64)[1]  # precision is part of the key
    Q = _shape(500000, 2000, 50, 14, 128)  # resolves to the Gram-of-A route
    u = dict(d, omega_p=512, n_by_l=256, kernels=4 | 16)  # and the other tile shapes of two GEMMs
    mc.tune_set(Q, u)
    assert mc.tune_get(Q, route=_lib.ROUTE_GRAM_OF_A) == (u, True)
    path = tmp_path / "plans.txt"
    mc.tune_save(path)
    lines = path.read_text().split("\n")
    assert len([x for x in lines if x]) == 2
    mc.tune_clear()
    assert not mc.tune_get(P, precision=_lib.PREC_MIXED_BF16)[1]
    mc.tune_load(path)
    assert mc.tune_get(P, precision=_lib.PREC_MIXED_BF16) == (t, True)
    assert mc.tune_get(Q) == (u, True)
    with pytest.raises(mc.McnlaError):
        mc.tune_set(P, dict(d, gram=0))
    with pytest.raises(mc.McnlaError):
        mc.tune_set(P, dict(d, n_by_l=(1 << 20) + 1))
    with pytest.raises(mc.McnlaError):
        mc.tune_set(P, dict(d, kernels=64))  # six GEMMs, six bits
    # a line in the earlier 12-number form (no kernels column) loads with kernels = 0
    old = tmp_path / "old.txt"
    old.write_text("1 0 100 100 8 4 2304 4096 2304 1024 512 512\n")
    mc.tune_clear()
    mc.tune_load(old)
    assert mc.tune_get(_shape(100, 100, 4, 4, 4), route=_lib.ROUTE_SKETCH)[0]["kernels"] == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("1 0 100 100 8 4 2304 4096 2304 1024 512 512\n1 0 100 100 8 4 2304 4096\n")
    mc.tune_clear()
    with pytest.raises(mc.McnlaError):
        mc.tune_load(bad)
    assert not mc.tune_get(_shape(100, 100, 4, 4, 4))[1]  # nothing of a bad file loads
    with pytest.raises(mc.McnlaError):
        mc.tune_load(tmp_path / "missing.txt")
    mc.tune_clear()
