"""Matrix Market (coordinate) reader producing CRS with sorted columns.

Covers what the reference harness reads (examples/src/helpers/sparsemat.cpp:
110-280): real/integer/pattern, general/symmetric, gzip or plain text.
"""
import gzip

import numpy as np


def read_mm(path):
    op = gzip.open if str(path).endswith(".gz") else open
    with op(path, "rt") as f:
        header = f.readline().lower().split()
        if len(header) < 5 or header[0] != "%%matrixmarket" or header[2] != "coordinate":
            raise ValueError("only coordinate Matrix Market files are supported")
        field, symm = header[3], header[4]
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        nr, nc, nz = (int(v) for v in line.split())
        data = np.loadtxt(f, ndmin=2, max_rows=nz)
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    v = np.ones(len(r)) if field == "pattern" else data[:, 2].astype(np.float64)
    if symm in ("symmetric", "skew-symmetric", "hermitian"):
        off = r != c
        sgn = -1.0 if symm == "skew-symmetric" else 1.0
        r, c, v = np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, sgn * v[off]])
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rowPtr = np.zeros(nr + 1, np.int64)
    np.cumsum(np.bincount(r, minlength=nr), out=rowPtr[1:])
    return rowPtr.astype(np.int32), c.astype(np.int32), v
