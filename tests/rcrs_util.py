"""Test helpers: random canonical range-CRS matrices and dense conversions."""
import numpy as np

import diverg_amd as dv


def from_dense(m, size_bytes=8):
    """Dense Boolean matrix -> canonical RangeCRS (maximal runs per row);
    the Basic -> Range conversion of crs_matrix.hpp:197-242 on sorted rows."""
    m = np.asarray(m, dtype=bool)
    nrows, ncols = m.shape
    ent, rm = [], [0]
    for i in range(nrows):
        cols = np.flatnonzero(m[i])
        if len(cols):
            brk = np.flatnonzero(np.diff(cols) != 1)
            starts = np.concatenate([[cols[0]], cols[brk + 1]])
            ends = np.concatenate([cols[brk], [cols[-1]]])
            ent.extend(np.stack([starts, ends], 1).ravel().tolist())
        rm.append(len(ent))
    dt = np.uint32 if size_bytes == 4 else np.uint64
    return dv.RangeCRS(ncols, np.array(ent, np.uint32), np.array(rm, dt), nrows=nrows)


def to_dense(r):
    m = np.zeros((r.nrows, r.ncols), dtype=bool)
    rm = np.asarray(r.rowmap, np.int64)
    en = np.asarray(r.entries, np.int64)
    for i in range(r.nrows):
        for k in range(rm[i], rm[i + 1], 2):
            m[i, en[k]:en[k + 1] + 1] = True
    return m


def random_dense(rng, nrows, ncols, density=0.05, run=8):
    """Random Boolean matrix made of short runs (so rows hold real ranges)."""
    m = np.zeros((nrows, ncols), dtype=bool)
    nstarts = rng.binomial(ncols, density / max(1, run / 2), size=nrows)
    for i in range(nrows):
        s = rng.integers(0, ncols, size=nstarts[i])
        ln = rng.integers(1, run + 1, size=nstarts[i])
        for a, b in zip(s, ln):
            m[i, a:min(ncols, a + b)] = True
    return m


def bool_matmul(a, b):
    return (a.astype(np.int64) @ b.astype(np.int64)) > 0


def bool_power(a, k):
    n = a.shape[0]
    c = np.eye(n, dtype=bool)
    for _ in range(k):
        c = bool_matmul(c, a)
    return c


def walk_index_dense(a, dlo, dup):
    """D(i, j) = some walk i -> j has length in [dlo, dup] (dense layering)."""
    n = a.shape[0]
    cur = np.eye(n, dtype=bool)
    d = np.zeros((n, n), dtype=bool)
    for length in range(dup + 1):
        if length >= dlo:
            d |= cur
        cur = bool_matmul(cur, a)
    return d


def assert_identical(actual, expected, what=""):
    """check_identical (test_dindex.cpp:34-53): rows, cols, nnz, entries, rowmap."""
    assert actual.nrows == expected.nrows, what
    assert actual.ncols == expected.ncols, what
    ae = np.asarray(actual.entries, np.uint64)
    ee = np.asarray(expected.entries, np.uint64)
    assert len(ae) == len(ee), f"{what}: entry count {len(ae)} != {len(ee)}"
    if not np.array_equal(ae, ee):
        k = int(np.flatnonzero(ae != ee)[0])
        raise AssertionError(f"{what}: entries differ first at {k}: {ae[k]} != {ee[k]}")
    assert np.array_equal(np.asarray(actual.rowmap, np.uint64),
                          np.asarray(expected.rowmap, np.uint64)), f"{what}: rowmap"
    exp_nnz = expected.nnz
    if getattr(actual, "size_bytes", 8) == 4:
        exp_nnz %= 1 << 32
    assert actual.nnz == exp_nnz, f"{what}: nnz {actual.nnz} != {exp_nnz}"
