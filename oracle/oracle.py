"""ctypes binding of the CPU oracle (rcrs_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg; the product package never
imports it.  Matrices are passed as any object with `nrows`, `ncols`,
`rowmap` (uint32 or uint64 numpy array), `entries` (uint32 numpy array);
results come back as `OracleMatrix` with a uint64 rowmap and exact 64-bit nnz.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc_rcrs.so")


class _Rcrs(C.Structure):
    _fields_ = [
        ("nrows", C.c_uint64),
        ("ncols", C.c_uint64),
        ("rowmap", C.POINTER(C.c_uint64)),
        ("entries", C.POINTER(C.c_uint32)),
        ("nnz", C.c_uint64),
    ]


class OracleMatrix:
    """Host rCRS matrix as produced by the oracle."""

    def __init__(self, nrows, ncols, rowmap, entries, nnz):
        self.nrows = int(nrows)
        self.ncols = int(ncols)
        self.rowmap = rowmap
        self.entries = entries
        self.nnz = int(nnz)

    def __repr__(self):
        return (f"OracleMatrix({self.nrows}x{self.ncols}, "
                f"ranges={len(self.entries) // 2}, nnz={self.nnz})")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        l = C.CDLL(_LIB_PATH)
        P = C.POINTER(_Rcrs)
        for name, args in [
            ("orc_identity", [C.c_uint64, P]),
            ("orc_spadd", [P, P, P]),
            ("orc_spgemm", [P, P, P]),
            ("orc_power", [P, C.c_uint32, P]),
            ("orc_dindex", [P, C.c_uint32, C.c_uint32, P]),
            ("orc_query", [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                           C.POINTER(C.c_uint8), C.c_uint64]),
            ("orc_walk_rows", [P, C.POINTER(C.c_uint32), C.c_uint64,
                               C.c_uint32, C.c_uint32, P]),
            ("orc_adjacency", [C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                               C.POINTER(C.c_uint32), C.c_uint64, C.c_uint64, P]),
            ("orc_merge_stream", [P, P, P]),
        ]:
            f = getattr(l, name)
            f.argtypes = args
            f.restype = C.c_int
        l.orc_free.argtypes = [P]
        l.orc_free.restype = None
        l.orc_nnz.argtypes = [P]
        l.orc_nnz.restype = C.c_uint64
        l.orc_num_threads.restype = C.c_int
        l.orc_set_threads.argtypes = [C.c_int]
        _lib = l
    return _lib


class OracleError(RuntimeError):
    pass


def _in(m, keep):
    rm = np.ascontiguousarray(m.rowmap, dtype=np.uint64)
    en = np.ascontiguousarray(m.entries, dtype=np.uint32)
    if en.size == 0:
        en = np.zeros(1, dtype=np.uint32)
    keep.extend([rm, en])
    s = _Rcrs()
    s.nrows = m.nrows
    s.ncols = m.ncols
    s.rowmap = rm.ctypes.data_as(C.POINTER(C.c_uint64))
    s.entries = en.ctypes.data_as(C.POINTER(C.c_uint32))
    s.nnz = 0
    return s


def _out(s):
    n = int(s.nrows)
    rm = np.ctypeslib.as_array(s.rowmap, shape=(n + 1,)).copy()
    ne = int(rm[-1])
    if ne:
        en = np.ctypeslib.as_array(s.entries, shape=(ne,)).copy()
    else:
        en = np.zeros(0, dtype=np.uint32)
    m = OracleMatrix(n, s.ncols, rm, en, s.nnz)
    lib().orc_free(C.byref(s))
    return m


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with status {rc}")


def identity(n):
    s = _Rcrs()
    _check(lib().orc_identity(n, C.byref(s)), "identity")
    return _out(s)


def spadd(a, b):
    keep = []
    sa, sb, sc = _in(a, keep), _in(b, keep), _Rcrs()
    _check(lib().orc_spadd(C.byref(sa), C.byref(sb), C.byref(sc)), "spadd")
    return _out(sc)


def spgemm(a, b):
    keep = []
    sa, sb, sc = _in(a, keep), _in(b, keep), _Rcrs()
    _check(lib().orc_spgemm(C.byref(sa), C.byref(sb), C.byref(sc)), "spgemm")
    return _out(sc)


def power(a, k):
    keep = []
    sa, sc = _in(a, keep), _Rcrs()
    _check(lib().orc_power(C.byref(sa), k, C.byref(sc)), "power")
    return _out(sc)


def dindex(a, dlo, dup):
    keep = []
    sa, sc = _in(a, keep), _Rcrs()
    _check(lib().orc_dindex(C.byref(sa), dlo, dup, C.byref(sc)), "dindex")
    return _out(sc)


def query(m, qi, qj):
    keep = []
    sm = _in(m, keep)
    qi = np.ascontiguousarray(qi, dtype=np.uint32)
    qj = np.ascontiguousarray(qj, dtype=np.uint32)
    out = np.zeros(len(qi), dtype=np.uint8)
    rc = lib().orc_query(C.byref(sm), qi.ctypes.data_as(C.POINTER(C.c_uint32)),
                         qj.ctypes.data_as(C.POINTER(C.c_uint32)),
                         out.ctypes.data_as(C.POINTER(C.c_uint8)), len(qi))
    return out, rc


def walk_rows(a, rows, dlo, dup):
    keep = []
    sa, sc = _in(a, keep), _Rcrs()
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    _check(lib().orc_walk_rows(C.byref(sa),
                               rows.ctypes.data_as(C.POINTER(C.c_uint32)),
                               len(rows), dlo, dup, C.byref(sc)), "walk_rows")
    return _out(sc)


def num_threads():
    return lib().orc_num_threads()


def set_threads(t):
    lib().orc_set_threads(int(t))


def adjacency(node_len, edge_off, edge_to, lower=0, upper=0):
    """Character-level adjacency of nodes [lower, upper) (adjacency.hpp:170-228
    then crs_matrix.hpp:197-242)."""
    nl = np.ascontiguousarray(node_len, dtype=np.uint32)
    eo = np.ascontiguousarray(edge_off, dtype=np.uint64)
    et = np.ascontiguousarray(edge_to, dtype=np.uint32)
    if et.size == 0:
        et = np.zeros(1, dtype=np.uint32)
    s = _Rcrs()
    _check(lib().orc_adjacency(len(nl), nl.ctypes.data_as(C.POINTER(C.c_uint32)),
                               eo.ctypes.data_as(C.POINTER(C.c_uint64)),
                               et.ctypes.data_as(C.POINTER(C.c_uint32)), lower, upper,
                               C.byref(s)), "adjacency")
    return _out(s)


def merge_stream(a, b):
    """Streaming distance-index merge (dindex.hpp:290-420)."""
    keep = []
    s = _Rcrs()
    _check(lib().orc_merge_stream(C.byref(_in(a, keep)), C.byref(_in(b, keep)), C.byref(s)),
           "merge_stream")
    return _out(s)
