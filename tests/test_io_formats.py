"""The host file formats on the CPU: the MCNLA binary layout byte for byte, the size readers, the
column/row block loaders of both formats, and the bin2bin driver (tests/cpp/check_io and
mcnla_amd/bin/bin2bin exercise include/mcnla_amd/io.hpp; no device is used).

The expected header is built here from the reference's field definitions
(include/mcnla/core/io/binary/def.hpp:77-110: a 32-byte line "##MCNLABinaryFileFormat" left-justified
to 31 + '\\n', a 64-byte line "# <comment>" left-justified to 63 + '\\n', then the storage "DEN"
and the value type "R8" each left-justified to 8, 8 blanks for a dense file's index type, 7 blanks
and '\\n'), followed by int64 ndim = 2, the int64 storage dims (dim0 = the leading dimension) and
the values in storage order (binary/dense_save.hpp:81-90)."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "check_io")
BIN2BIN = os.path.join(ROOT, "mcnla_amd", "bin", "bin2bin")


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CHECK):
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "check_io"], check=True)
    if not os.path.exists(BIN2BIN):
        subprocess.run(["make", "-C", os.path.join(ROOT, "mcnla_amd", "drivers")], check=True)


def tool(*args):
    r = subprocess.run([CHECK] + [str(a) for a in args], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def parse(out):
    lines = out.strip().split("\n")
    m, n = (int(x) for x in lines[0].split())
    vals = np.array([[float(x) for x in line.split()] for line in lines[1:]]) if m else np.zeros((0, n))
    return vals.reshape(m, n)


def expected_header(comment, code=b"R8"):
    return (b"##MCNLABinaryFileFormat".ljust(31) + b"\n" + (b"# " + comment).ljust(63) + b"\n"
            + b"DEN".ljust(8) + code.ljust(8) + b" " * 8 + b" " * 7 + b"\n")


def full(m, n):
    return 1000.0 * np.arange(m)[:, None] + np.arange(n)[None, :]


@pytest.mark.parametrize("order", ["F", "C"])
def test_binary_layout(tmp_path, order):
    m, n = 5, 3
    f = tmp_path / "a.bin"
    tool("save", m, n, order, f, "layout check")
    raw = f.read_bytes()
    assert len(raw) == 128 + 3 * 8 + m * n * 8
    assert raw[:128] == expected_header(b"layout check")
    nd, d0, d1 = struct.unpack("<qqq", raw[128:152])
    A = full(m, n)
    if order == "F":  # column-major: dims (m, n), columns one after another
        assert (nd, d0, d1) == (2, m, n)
        data = A.T.ravel()
    else:  # row-major is stored as its transpose: dims (n, m), rows one after another
        assert (nd, d0, d1) == (2, n, m)
        data = A.ravel()
    assert np.array_equal(np.frombuffer(raw[152:], dtype="<f8"), data)


def test_binary_sizes_and_blocks(tmp_path):
    m, n = 7, 9
    cm, rm = tmp_path / "cm.bin", tmp_path / "rm.bin"
    tool("save", m, n, "F", cm, "cm")
    tool("save", m, n, "C", rm, "rm")
    assert tool("size", cm, "N").split() == [str(m), str(n)]
    assert tool("size", cm, "T").split() == [str(n), str(m)]
    assert tool("size", rm, "T").split() == [str(m), str(n)]  # a row-major file read as TRANS
    A = full(m, n)
    for j0, j1 in ((0, n), (2, 5), (8, 9), (3, 3)):
        assert np.array_equal(parse(tool("colblock", cm, j0, j1)), A[:, j0:j1])
    for i0, i1 in ((0, m), (1, 4), (6, 7)):
        assert np.array_equal(parse(tool("rowblock", rm, i0, i1)), A[i0:i1])
    r = subprocess.run([CHECK, "colblock", str(cm), "4", str(n + 1)], capture_output=True, text=True)
    assert r.returncode == 2 and "range" in r.stderr


def write_mtx(path, A):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n% a comment line\n")
        f.write("%d %d\n" % A.shape)
        for v in A.T.ravel():
            f.write("%.17g\n" % v)


def test_matrix_market_blocks(tmp_path):
    """Non-square, so that the row-block reader must skip m - end values per column."""
    rng = np.random.default_rng(2)
    A = rng.standard_normal((6, 4))
    f = tmp_path / "a.mtx"
    write_mtx(f, A)
    assert tool("mtxsize", f, "N").split() == ["6", "4"]
    assert tool("mtxsize", f, "T").split() == ["4", "6"]
    for j0, j1 in ((0, 4), (1, 3), (3, 4)):
        assert np.array_equal(parse(tool("mtxcol", f, j0, j1)), A[:, j0:j1])
    for i0, i1 in ((0, 6), (2, 5), (5, 6)):
        assert np.array_equal(parse(tool("mtxrow", f, i0, i1)), A[i0:i1])


def test_bin2bin(tmp_path):
    m, n = 4, 6
    src, dst = tmp_path / "cm.bin", tmp_path / "rm.bin"
    tool("save", m, n, "F", src, "source")
    r = subprocess.run([BIN2BIN, str(src), str(dst), "converted"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    raw = dst.read_bytes()
    assert raw[:128] == expected_header(b"converted")
    assert struct.unpack("<qqq", raw[128:152]) == (2, n, m)
    assert np.array_equal(np.frombuffer(raw[152:], dtype="<f8"), full(m, n).ravel())
    assert tool("size", dst, "T").split() == [str(m), str(n)]
