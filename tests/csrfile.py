"""Test helper: read a reference .csr data file and normalise it to CSRx the
way ReadMat does (src/ReadData.cxx:221-357 ReadCSR -> src/MATUtil.cxx:439
SortCSRRow -> src/MATUtil.cxx:303 CSRtoMAT): 1-based indices shifted to 0,
columns sorted within each row (stable), a zero diagonal inserted in rows
that have none."""
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def read_csr(name: str):
    tok = (GOLDEN / name).read_text().split()
    n = int(tok[0])
    rp = np.array(tok[1:n + 2], dtype=np.int64)
    nnz = int(rp[-1] - rp[0])
    ci = np.array(tok[n + 2:n + 2 + nnz], dtype=np.int64)
    va = np.array(tok[n + 2 + nnz:n + 2 + 2 * nnz], dtype=np.float64)
    if rp[0] == 1:
        rp -= 1
        ci -= 1
    rows, cols, vals = [], [], []
    for i in range(n):
        s, e = rp[i], rp[i + 1]
        c, v = ci[s:e], va[s:e]
        order = np.argsort(c, kind="stable")
        c, v = c[order], v[order]
        if not np.any(c == i):
            k = np.searchsorted(c, i)
            c = np.insert(c, k, i)
            v = np.insert(v, k, 0.0)
        rows.append(np.full(c.shape[0], i))
        cols.append(c)
        vals.append(v)
    cols = np.concatenate(cols).astype(np.uint32)
    vals = np.concatenate(vals)
    counts = np.array([c.shape[0] for c in rows])
    rpn = np.zeros(n + 1, np.uint32)
    np.cumsum(counts, out=rpn[1:])
    return n, rpn, cols, vals
