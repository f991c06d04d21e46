"""Reader for the MatrixMarket dense "array" fixtures under tests/golden/.

Format (include/mcnla/core/io/matrix_market/dense_load.hpp:87-121 and :198-235): header lines
start with '%', then "m n" (or "m n l" for a 3-D collection), then the values of each matrix in
column-major order, one matrix after another.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_header(path):
    comments = []
    with open(path) as f:
        for line in f:
            if line.startswith("%"):
                comments.append(line.rstrip("\n"))
            else:
                return comments, [int(t) for t in line.split()]
    raise ValueError("no size line in %s" % path)


def load(name):
    """Returns a 2-D array (m x n) or, for a 3-D collection, the reference's unfolded
    row-major m x (n*l) collection (matrix i occupies columns [i*n, (i+1)*n))."""
    path = os.path.join(GOLDEN, name)
    comments, dims = read_header(path)
    with open(path) as f:
        body = [ln for ln in f if not ln.startswith("%")]
    vals = np.array(" ".join(body[1:]).split(), dtype=np.float64)
    if len(dims) == 2:
        m, n = dims
        assert vals.size == m * n
        return vals.reshape(n, m).T.copy()
    m, n, l = dims
    assert vals.size == m * n * l
    mats = vals.reshape(l, n, m).transpose(0, 2, 1)  # l x m x n
    return np.concatenate(list(mats), axis=1).copy()


def comment_value(name, key):
    comments, _ = read_header(os.path.join(GOLDEN, name))
    for c in comments:
        for part in c.lstrip("% ").split(","):
            if "=" in part:
                k, v = part.split("=", 1)
                if k.strip() == key:
                    return v.strip()
    return None
