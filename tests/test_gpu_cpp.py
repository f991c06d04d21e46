"""The C++ host layer on the GPU: the reference's stage tests restated in C++
(tests/cpp/check_isvd.cpp) and the drop-in drivers run end to end on files, checked against the
oracle pipeline on the same reference-recipe Omega."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mcnla_amd", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(cmd, timeout=120):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def test_cpp_stage_checks(gpu):
    note("C++ stage checks")
    rc, out = run([os.path.join(ROOT, "tests", "cpp", "check_isvd"), GOLDEN])
    print(out)
    assert rc == 0, out
    assert out.count("PASS") == 11


def write_mtx(path, A):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n%d %d\n" % A.shape)
        np.savetxt(f, A.T.reshape(-1), fmt="%.17g")


def read_mtx(path):
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith("%")]
    m, n = (int(t) for t in lines[0].split())
    v = np.array(" ".join(lines[1:]).split(), dtype=np.float64)
    return v.reshape(n, m).T


@pytest.fixture(scope="module")
def problem(tmp_path_factory):
    d = tmp_path_factory.mktemp("driver")
    rng = np.random.default_rng(3)
    m, n = 600, 400
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (U * (0.93 ** np.arange(n) + 1e-6)) @ V.T
    write_mtx(str(d / "a.mtx"), A)
    return d, A


@pytest.mark.parametrize("prog,cb,integ,former", [
    ("isvd_dense_cb_kn_svd", True, "kn", "svd"),
    ("isvd_dense_cb", True, "wy", "gram"),
    ("isvd_dense_rb", False, "wy", "gram"),
])
def test_driver_end_to_end(gpu, problem, prog, cb, integ, former):
    note("driver %s" % prog)
    d, A = problem
    m, n = A.shape
    N, k, p, seed = 4, 8, 8, 17
    tol = 1e-4 if integ == "kn" else 1e-3
    args = [os.path.join(BIN, prog), str(d / "a.mtx"), str(d / "s.mtx"), str(d / "u.mtx"),
            str(d / "v.mtx"), str(N), str(k), str(p), str(tol), "256", "--seed", str(seed),
            "--omega", "reference"]
    rc, out = run(args)
    print(out)
    assert rc == 0, out
    S, U, Vv = read_mtx(str(d / "s.mtx"))[:, 0], read_mtx(str(d / "u.mtx")), read_mtx(str(d / "v.mtx"))
    l = k + p
    om = oracle.omega_ref(seed, n, N * l, col_block=cb)
    Q, _ = oracle.gramian_orthogonalize(oracle.sketch(A, om, a_colmajor=cb), l, N)
    fn = oracle.kn_rowblock if integ == "kn" else oracle.wy_rowblock
    Qb, it = fn(Q, l, N, 256, tol)
    if former == "svd":
        So, Uo, Vo = oracle.svd_former(A, Qb, k, a_colmajor=cb)
    else:
        So, Uo, Vo = oracle.gramian_former(A, Qb, k, a_colmajor=cb)
    tight = integ == "kn"  # WY amplifies rounding; see test_gpu_solver.py
    assert np.max(np.abs(S - So) / So) < (1e-9 if tight else 1e-6)
    assert np.abs(U @ U.T - Uo @ Uo.T).max() < (1e-8 if tight else 1e-5)
    assert np.abs(Vv @ Vv.T - Vo @ Vo.T).max() < (1e-8 if tight else 1e-5)
    line = [ln for ln in out.splitlines() if ln.startswith("Iteration")][0]
    assert abs(int(line.split("=")[1]) - it) <= 1
    err = [ln for ln in out.splitlines() if ln.startswith("Error")][0]
    assert abs(float(err.split("=")[1]) - oracle.frerr(A, So, Uo, Vo)) < (1e-8 if tight else 1e-6)


def test_driver_binary_input(gpu, problem):
    note("driver binary input")
    d, A = problem
    rc, out = run([os.path.join(BIN, "mtx2bin"), str(d / "a.mtx"), str(d / "a.bin")])
    assert rc == 0, out
    common = ["6", "8", "8", "1e-4", "256", "--seed", "5", "--omega", "reference"]
    rc, out_mtx = run([os.path.join(BIN, "isvd_dense_cb_kn_svd"), str(d / "a.mtx"), str(d / "s1.mtx"),
                       str(d / "u1.mtx"), str(d / "v1.mtx")] + common)
    assert rc == 0, out_mtx
    rc, out_bin = run([os.path.join(BIN, "isvd_dense_cb_kn_svd_bin"), str(d / "a.bin"),
                       str(d / "s2.mtx"), str(d / "u2.mtx"), str(d / "v2.mtx")] + common)
    assert rc == 0, out_bin
    assert np.array_equal(read_mtx(str(d / "s1.mtx")), read_mtx(str(d / "s2.mtx")))


def test_driver_usage_error(gpu):
    rc, out = run([os.path.join(BIN, "isvd_dense_cb"), "only-one-arg"])
    assert rc == 1 and "Usage" in out
