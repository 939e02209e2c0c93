"""MI355X (gfx950) range-CRS Boolean sparse engine -- Python host binding.

Mirrors DiVerG's operator interface for the distance-index path
(range_sparse.hpp:430-477, 2283-2368; dindex.hpp:46-189):

    ctx = Context()
    c = ctx.range_spgemm(a, b)         # RangeSpGEMM::spgemm
    c = ctx.range_spadd(a, b)          # RangeSpAdd::spadd
    c = ctx.range_power(a, k)          # RangePower::power
    d = ctx.create_distance_index(a, dlo, dup)   # util::create_distance_index
    hits = ctx.query(d, i, j)          # CRSMatrix::operator()(i, j), batched

Matrices are `RangeCRS` (host, numpy arrays) or `DeviceRCRS` (HBM-resident
handles returned with location="device").  Every call goes through the C ABI
in libdiverg_amd.so; there is no Python or CPU fallback -- a missing library
raises `NativeMissing`.
"""
import ctypes as C
import os

import numpy as np

from . import _native as N
from ._native import NativeMissing  # noqa: F401

__all__ = ["RangeCRS", "DeviceRCRS", "Context", "Server", "Graph", "DiVerGError",
           "generate_graph", "graph_preset", "identical", "NativeMissing"]


class DiVerGError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        super().__init__(f"{N.STATUS.get(status, status)}: {message}")


class RangeCRS:
    """Host range-CRS matrix: CRSMatrix<RangeDynamic, bool, uint32_t, S>.

    entries: uint32 array s0,e0,s1,e1,...; rowmap: uint32 or uint64 (S)
    array of nrows+1 offsets into entries; nnz: number of set cells.
    """

    def __init__(self, ncols, entries, rowmap, nnz=None, nrows=None):
        self.entries = np.ascontiguousarray(entries, dtype=np.uint32)
        rowmap = np.asarray(rowmap)
        if rowmap.dtype not in (np.uint32, np.uint64):
            rowmap = rowmap.astype(np.uint64)
        self.rowmap = np.ascontiguousarray(rowmap)
        self.nrows = len(self.rowmap) - 1 if nrows is None else int(nrows)
        self.ncols = int(ncols)
        if nnz is None:  # crs_matrix.hpp:175-189
            e = self.entries.astype(np.int64)
            nnz = int((e[1::2] - e[0::2] + 1).sum())
            if self.size_bytes == 4:
                nnz %= 1 << 32
        self.nnz = int(nnz)

    @property
    def size_bytes(self):
        return self.rowmap.dtype.itemsize

    def numRows(self):
        return self.nrows

    def numCols(self):
        return self.ncols

    def row(self, i):
        b, e = int(self.rowmap[i]), int(self.rowmap[i + 1])
        return self.entries[b:e]

    def as_size(self, size_bytes):
        dt = np.uint32 if size_bytes == 4 else np.uint64
        return RangeCRS(self.ncols, self.entries, self.rowmap.astype(dt),
                        self.nnz % (1 << 32) if size_bytes == 4 else self.nnz)

    def _desc(self, keep):
        d = N.dv_rcrs()
        d.nrows = self.nrows
        d.ncols = self.ncols
        en = self.entries if len(self.entries) else np.zeros(2, np.uint32)
        keep.extend([en, self.rowmap])
        d.rowmap = self.rowmap.ctypes.data
        d.entries = en.ctypes.data_as(C.POINTER(C.c_uint32))
        d.n_entries = len(self.entries)
        d.nnz = self.nnz
        d.size_bytes = self.size_bytes
        d.location = N.DV_HOST
        d.owner = None
        return d

    def __eq__(self, other):
        return identical(self, other)

    def __repr__(self):
        return (f"RangeCRS({self.nrows}x{self.ncols}, ranges={len(self.entries) // 2}, "
                f"nnz={self.nnz}, S=uint{8 * self.size_bytes})")


def identical(a, b):
    """check_identical of test_dindex.cpp:34-53: rows, cols, nnz, entries, rowmap."""
    return (a.nrows == b.nrows and a.ncols == b.ncols and a.nnz == b.nnz
            and np.array_equal(np.asarray(a.entries, np.uint64), np.asarray(b.entries, np.uint64))
            and np.array_equal(np.asarray(a.rowmap, np.uint64), np.asarray(b.rowmap, np.uint64)))


def _from_desc(d):
    """Copy a library-allocated host descriptor into a RangeCRS."""
    dt = np.uint32 if d.size_bytes == 4 else np.uint64
    n = int(d.nrows)
    rm = np.ctypeslib.as_array(C.cast(d.rowmap, C.POINTER(C.c_uint32 if d.size_bytes == 4
                                                           else C.c_uint64)),
                               shape=(n + 1,)).astype(dt, copy=True)
    ne = int(d.n_entries)
    en = (np.ctypeslib.as_array(d.entries, shape=(ne,)).copy() if ne
          else np.zeros(0, np.uint32))
    return RangeCRS(int(d.ncols), en, rm, nnz=int(d.nnz), nrows=n)


class DeviceRCRS:
    """An HBM-resident matrix owned by the library (freed with the handle)."""

    def __init__(self, ctx, desc):
        self._ctx = ctx
        self.desc = desc

    nrows = property(lambda self: int(self.desc.nrows))
    ncols = property(lambda self: int(self.desc.ncols))
    nnz = property(lambda self: int(self.desc.nnz))
    n_entries = property(lambda self: int(self.desc.n_entries))
    size_bytes = property(lambda self: int(self.desc.size_bytes))

    def _desc(self, keep):
        return self.desc

    def to_host(self):
        return self._ctx.to_host(self)

    def free(self):
        if self.desc is not None and self._ctx._h:
            N.lib().dv_free(self._ctx._h, C.byref(self.desc))
        self.desc = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __repr__(self):
        return (f"DeviceRCRS({self.nrows}x{self.ncols}, ranges={self.n_entries // 2}, "
                f"nnz={self.nnz})")


class BuildStats(dict):
    pass


class Context:
    """A device context (dv_init / dv_finalize): replaces Kokkos
    initialize/finalize (runtime.hpp:25-55)."""

    def __init__(self, device=0):
        self._h = None
        h = C.c_void_p()
        rc = N.lib().dv_init(device, C.byref(h))
        if rc:
            raise DiVerGError(rc, f"dv_init({device}) failed")
        self._h = h

    def close(self):
        if self._h:
            N.lib().dv_finalize(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers -------------------------------------------------------------
    def _check(self, rc, what):
        if rc:
            msg = N.lib().dv_last_error(self._h)
            raise DiVerGError(rc, f"{what}: {msg.decode() if msg else ''}")

    def _out(self, desc, location):
        if location == "device":
            return DeviceRCRS(self, desc)
        m = _from_desc(desc)
        N.lib().dv_free(self._h, C.byref(desc))
        return m

    @staticmethod
    def _loc(location):
        if location not in ("host", "device"):
            raise ValueError("location must be 'host' or 'device'")
        return N.DV_HOST if location == "host" else N.DV_DEVICE

    # -- tuning --------------------------------------------------------------
    def tuning(self):
        t = N.dv_tuning()
        self._check(N.lib().dv_get_tuning(self._h, C.byref(t)), "dv_get_tuning")
        return {f: getattr(t, f) for f, _ in N.dv_tuning._fields_}

    def last_tuning(self):
        """The launch shape the last index build ran with (autotuned or not)."""
        t = N.dv_tuning()
        self._check(N.lib().dv_last_tuning(self._h, C.byref(t)), "dv_last_tuning")
        return {f: getattr(t, f) for f, _ in N.dv_tuning._fields_}

    def set_tuning(self, **kw):
        """Update the named fields of the launch tuning (dv_tuning)."""
        t = N.dv_tuning()
        self._check(N.lib().dv_get_tuning(self._h, C.byref(t)), "dv_get_tuning")
        for k, v in kw.items():
            if k not in dict(N.dv_tuning._fields_):
                raise KeyError(k)
            setattr(t, k, int(v))
        self._check(N.lib().dv_set_tuning(self._h, C.byref(t)), "dv_set_tuning")

    # -- per-kernel timing ------------------------------------------------------
    def profile(self, enable=True):
        self._check(N.lib().dv_profile(self._h, 1 if enable else 0), "dv_profile")

    def profile_read(self):
        arr = (N.dv_kernel_time * len(N.KERNEL_KINDS))()
        self._check(N.lib().dv_profile_read(self._h, arr), "dv_profile_read")
        return {name: {f: getattr(arr[k], f) for f, _ in N.dv_kernel_time._fields_}
                for k, name in enumerate(N.KERNEL_KINDS)}

    # -- matrices ------------------------------------------------------------
    def to_device(self, m):
        keep = []
        out = N.dv_rcrs()
        self._check(N.lib().dv_copy(self._h, C.byref(m._desc(keep)), N.DV_DEVICE,
                                    C.byref(out)), "dv_copy")
        return DeviceRCRS(self, out)

    def to_host(self, m):
        keep = []
        out = N.dv_rcrs()
        self._check(N.lib().dv_copy(self._h, C.byref(m._desc(keep)), N.DV_HOST,
                                    C.byref(out)), "dv_copy")
        return self._out(out, "host")

    def identity(self, n, size_bytes=8, location="host"):
        out = N.dv_rcrs()
        self._check(N.lib().dv_range_identity(self._h, n, size_bytes, self._loc(location),
                                              C.byref(out)), "dv_range_identity")
        return self._out(out, location)

    def range_spadd(self, a, b, location="host"):
        keep = []
        out = N.dv_rcrs()
        self._check(N.lib().dv_range_spadd(self._h, C.byref(a._desc(keep)),
                                           C.byref(b._desc(keep)), self._loc(location),
                                           C.byref(out)), "dv_range_spadd")
        return self._out(out, location)

    def range_spgemm(self, a, b, location="host"):
        keep = []
        out = N.dv_rcrs()
        self._check(N.lib().dv_range_spgemm(self._h, C.byref(a._desc(keep)),
                                            C.byref(b._desc(keep)), self._loc(location),
                                            C.byref(out)), "dv_range_spgemm")
        return self._out(out, location)

    def range_power(self, a, k, location="host"):
        keep = []
        out = N.dv_rcrs()
        self._check(N.lib().dv_range_power(self._h, C.byref(a._desc(keep)), k,
                                           self._loc(location), C.byref(out)),
                    "dv_range_power")
        return self._out(out, location)

    def create_distance_index(self, a, dlo, dup, location="host", stats=False):
        keep = []
        out = N.dv_rcrs()
        st = N.dv_build_stats()
        self._check(N.lib().dv_create_distance_index(
            self._h, C.byref(a._desc(keep)), dlo, dup, self._loc(location), C.byref(out),
            C.byref(st)), "dv_create_distance_index")
        m = self._out(out, location)
        if stats:
            return m, BuildStats({f: getattr(st, f) for f, _ in N.dv_build_stats._fields_})
        return m

    def create_distance_index_batched(self, comps, dlo, dup, location="host", stats=False):
        """Indexes of independent components in one planned build over their
        block-diagonal assembly; outputs[c] equals create_distance_index(comps[c])."""
        keep = []
        n = len(comps)
        arr = (N.dv_rcrs * max(1, n))()
        for i, m in enumerate(comps):
            arr[i] = m._desc(keep)
        outs = (N.dv_rcrs * max(1, n))()
        st = N.dv_build_stats()
        self._check(N.lib().dv_create_distance_index_batched(
            self._h, arr, n, dlo, dup, self._loc(location), outs, C.byref(st)),
            "dv_create_distance_index_batched")
        res = [self._out(outs[i], location) for i in range(n)]
        if stats:
            return res, BuildStats({f: getattr(st, f) for f, _ in N.dv_build_stats._fields_})
        return res

    def merge_distance_index(self, d1, d2, location="host"):
        """util::merge_distance_index (dindex.hpp:435-456, 556-579): the union
        of two indices of one shape, through the device SpAdd."""
        keep = []
        out = N.dv_rcrs()
        self._check(N.lib().dv_merge_distance_index(self._h, C.byref(d1._desc(keep)),
                                                    C.byref(d2._desc(keep)),
                                                    self._loc(location), C.byref(out)),
                    "dv_merge_distance_index")
        return self._out(out, location)

    def save(self, m, path):
        """Write m in the reference's serialize() layout (crs_matrix.hpp:1392-1409,
        1986-2000)."""
        keep = []
        self._check(N.lib().dv_save(self._h, C.byref(m._desc(keep)), os.fsencode(path)),
                    "dv_save")

    def load(self, path, size_bytes=0, location="host"):
        """Read a matrix written by save() or by the reference's serialize()."""
        out = N.dv_rcrs()
        self._check(N.lib().dv_load(self._h, os.fsencode(path), size_bytes,
                                    self._loc(location), C.byref(out)), "dv_load")
        return self._out(out, location)

    def adjacency_matrix(self, graph, lower=0, upper=0, size_bytes=8, location="host"):
        """Character-level adjacency of nodes [lower, upper) of `graph` (upper 0:
        all), built on the device (adjacency.hpp:170-228, 249-262)."""
        out = N.dv_rcrs()
        self._check(N.lib().dv_adjacency_matrix(self._h, graph._h, lower, upper, size_bytes,
                                                self._loc(location), C.byref(out)),
                    "dv_adjacency_matrix")
        return self._out(out, location)

    def query_device(self, m, qi_ptr, qj_ptr, out_ptr, n):
        """Batched D(i, j) with the pairs and answers in device memory (raw
        device pointers); blocks until the answers are written."""
        keep = []
        rc = N.lib().dv_query_batch(self._h, C.byref(m._desc(keep)),
                                    C.cast(C.c_void_p(qi_ptr), C.POINTER(C.c_uint32)),
                                    C.cast(C.c_void_p(qj_ptr), C.POINTER(C.c_uint32)),
                                    C.cast(C.c_void_p(out_ptr), C.POINTER(C.c_uint8)), n,
                                    N.DV_DEVICE)
        self._check(rc, "dv_query_batch")

    def query(self, m, qi, qj):
        """Batched D(i, j); raises DiVerGError(DV_ERR_RANGE) for pairs outside."""
        keep = []
        qi = np.ascontiguousarray(qi, dtype=np.uint32)
        qj = np.ascontiguousarray(qj, dtype=np.uint32)
        out = np.zeros(len(qi), dtype=np.uint8)
        rc = N.lib().dv_query_batch(self._h, C.byref(m._desc(keep)),
                                    qi.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    qj.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    out.ctypes.data_as(C.POINTER(C.c_uint8)), len(qi),
                                    N.DV_HOST)
        self._check(rc, "dv_query_batch")
        return out


class Server:
    """Serving loop over HBM-resident indexes (dv_server_*): continuous
    batching of locus-pair queries, index builds admitted by HBM budget."""

    def __init__(self, device=0, hbm_budget=0, max_batch=0, max_inflight=0, query_cus=0):
        self._h = None
        o = N.dv_server_opts(hbm_budget, max_batch, max_inflight, query_cus, 0)
        h = C.c_void_p()
        rc = N.lib().dv_server_start(device, C.byref(o), C.byref(h))
        if rc:
            raise DiVerGError(rc, "dv_server_start failed")
        self._h = h

    def close(self):
        if self._h:
            N.lib().dv_server_stop(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _check(rc, what):
        if rc:
            raise DiVerGError(rc, what)

    def add_index(self, m):
        keep = []
        i = C.c_uint32()
        self._check(N.lib().dv_server_add_index(self._h, C.byref(m._desc(keep)), C.byref(i)),
                    "dv_server_add_index")
        return i.value

    def submit_queries(self, index_id, qi, qj):
        qi = np.ascontiguousarray(qi, dtype=np.uint32)
        qj = np.ascontiguousarray(qj, dtype=np.uint32)
        t = C.c_uint64()
        self._check(N.lib().dv_server_submit_queries(
            self._h, index_id, qi.ctypes.data_as(N.U32P), qj.ctypes.data_as(N.U32P), len(qi),
            C.byref(t)), "dv_server_submit_queries")
        return t.value

    def submit_build(self, a, dlo, dup):
        keep = []
        t = C.c_uint64()
        self._check(N.lib().dv_server_submit_build(self._h, C.byref(a._desc(keep)), dlo, dup,
                                                   C.byref(t)), "dv_server_submit_build")
        return t.value

    @staticmethod
    def _info(i):
        return {f: getattr(i, f) for f, _ in N.dv_request_info._fields_}

    def poll(self, ticket):
        i = N.dv_request_info()
        rc = N.lib().dv_server_poll(self._h, ticket, C.byref(i))
        if rc not in (0, 7):
            raise DiVerGError(rc, "dv_server_poll")
        return self._info(i) if rc == 0 else None

    def wait(self, ticket):
        i = N.dv_request_info()
        self._check(N.lib().dv_server_wait(self._h, ticket, C.byref(i)), "dv_server_wait")
        return self._info(i)

    def results(self, ticket, n):
        out = np.zeros(n, np.uint8)
        self._check(N.lib().dv_server_results(self._h, ticket,
                                              out.ctypes.data_as(C.POINTER(C.c_uint8)), n),
                    "dv_server_results")
        return out

    def get_index(self, index_id):
        d = N.dv_rcrs()
        self._check(N.lib().dv_server_get_index(self._h, index_id, C.byref(d)),
                    "dv_server_get_index")
        m = _from_desc(d)
        N.lib().dv_free(None, C.byref(d))
        return m

    def replay(self, qidx, qi, qj, arrival_us, builds=()):
        """Replay a query stream (and build requests) in real time.

        builds: iterable of (matrix, dlo, dup, arrival_us).  Returns answers,
        per-query latency (us), per-build completion time (us after its
        arrival) and produced index ids, and the elapsed wall time."""
        n = len(qi)
        qidx = np.ascontiguousarray(qidx, dtype=np.uint32)
        qi = np.ascontiguousarray(qi, dtype=np.uint32)
        qj = np.ascontiguousarray(qj, dtype=np.uint32)
        arr = np.ascontiguousarray(arrival_us, dtype=np.float64)
        ans = np.zeros(n, np.uint8)
        lat = np.zeros(n, np.float64)
        builds = list(builds)
        nb = len(builds)
        keep = []
        bd = (N.dv_rcrs * max(1, nb))()
        for k, (m, _, _, _) in enumerate(builds):
            bd[k] = m._desc(keep)
        blo = np.array([b[1] for b in builds] or [0], np.uint32)
        bup = np.array([b[2] for b in builds] or [0], np.uint32)
        bus = np.array([b[3] for b in builds] or [0], np.float64)
        bdone = np.zeros(max(1, nb), np.float64)
        bidx = np.zeros(max(1, nb), np.uint32)
        el = C.c_double()
        rc = N.lib().dv_server_replay(
            self._h, n, qidx.ctypes.data_as(N.U32P), qi.ctypes.data_as(N.U32P),
            qj.ctypes.data_as(N.U32P), arr.ctypes.data_as(N.DP),
            ans.ctypes.data_as(C.POINTER(C.c_uint8)), lat.ctypes.data_as(N.DP), nb, bd,
            blo.ctypes.data_as(N.U32P), bup.ctypes.data_as(N.U32P), bus.ctypes.data_as(N.DP),
            bdone.ctypes.data_as(N.DP), bidx.ctypes.data_as(N.U32P), C.byref(el))
        self._check(rc, "dv_server_replay")
        return {"answers": ans, "latency_us": lat, "build_done_us": bdone[:nb],
                "build_index": bidx[:nb], "elapsed_us": el.value}


class Graph:
    """A sequence graph (dv_graph): nodes with lengths in bases and directed
    edges.  Load GFA with Graph.load_gfa(path) or build from arrays."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def load_gfa(cls, path, version=0):
        h = C.c_void_p()
        rc = N.lib().dv_graph_load_gfa(os.fsencode(path), version, C.byref(h))
        if rc:
            raise DiVerGError(rc, f"dv_graph_load_gfa({path!r}) failed")
        return cls(h)

    @classmethod
    def from_arrays(cls, node_len, edges_from=(), edges_to=()):
        nl = np.ascontiguousarray(node_len, dtype=np.uint32)
        ef = np.ascontiguousarray(edges_from, dtype=np.uint32)
        et = np.ascontiguousarray(edges_to, dtype=np.uint32)
        if len(ef) != len(et):
            raise ValueError("edges_from and edges_to differ in length")
        h = C.c_void_p()
        u32 = C.POINTER(C.c_uint32)
        rc = N.lib().dv_graph_from_arrays(len(nl), nl.ctypes.data_as(u32), len(ef),
                                          ef.ctypes.data_as(u32), et.ctypes.data_as(u32),
                                          C.byref(h))
        if rc:
            raise DiVerGError(rc, "dv_graph_from_arrays failed")
        return cls(h)

    def close(self):
        if self._h:
            N.lib().dv_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = N.dv_graph_info()
        rc = N.lib().dv_graph_info_get(self._h, C.byref(i))
        if rc:
            raise DiVerGError(rc, "dv_graph_info_get failed")
        return {f: getattr(i, f) for f, _ in N.dv_graph_info._fields_}

    def arrays(self):
        """(node_len, charorder, edge_off, edge_to) as numpy arrays."""
        inf = self.info()
        n, e = inf["nodes"], inf["edges"]
        nl = np.zeros(n, np.uint32)
        co = np.zeros(n + 1, np.uint64)
        eo = np.zeros(n + 1, np.uint64)
        et = np.zeros(max(e, 1), np.uint32)
        u32, u64 = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        rc = N.lib().dv_graph_arrays(self._h, nl.ctypes.data_as(u32), co.ctypes.data_as(u64),
                                     eo.ctypes.data_as(u64), et.ctypes.data_as(u32))
        if rc:
            raise DiVerGError(rc, "dv_graph_arrays failed")
        return nl, co, eo, et[:e]

    def node_name(self, k):
        s = N.lib().dv_graph_node_name(self._h, k)
        return None if s is None else s.decode()

    def sort(self, order="min-breaks", seed=0):
        rc = N.lib().dv_graph_sort(self._h, N.DV_ORDERS[order], seed)
        if rc:
            raise DiVerGError(rc, "dv_graph_sort failed")

    def components(self):
        """Block boundaries [0, b1, ..., n] with no edge between blocks."""
        cnt = C.c_uint64()
        N.lib().dv_graph_components(self._h, None, 0, C.byref(cnt))
        arr = np.zeros(cnt.value + 1, np.uint64)
        rc = N.lib().dv_graph_components(self._h, arr.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         len(arr), C.byref(cnt))
        if rc:
            raise DiVerGError(rc, "dv_graph_components failed")
        return arr


def graph_preset(preset, component=0):
    p = N.dv_graph_params()
    rc = N.lib().dv_graph_preset(preset, component, C.byref(p))
    if rc:
        raise DiVerGError(rc, f"unknown preset {preset}")
    return p


def generate_graph(preset=None, component=0, size_bytes=8, params=None, with_vpos=False,
                   **overrides):
    """Seeded synthetic variation graph -> (RangeCRS A[, vpos])."""
    p = params if params is not None else graph_preset(preset, component)
    for k, v in overrides.items():
        setattr(p, k, v)
    out = N.dv_rcrs()
    vp = C.POINTER(C.c_uint32)()
    rc = N.lib().dv_generate_graph(C.byref(p), size_bytes, C.byref(out),
                                   C.byref(vp) if with_vpos else None)
    if rc:
        raise DiVerGError(rc, "dv_generate_graph failed")
    m = _from_desc(out)
    N.lib().dv_free(None, C.byref(out))
    if with_vpos:
        vpos = np.ctypeslib.as_array(vp, shape=(m.nrows,)).copy()
        N.lib().dv_free_host(C.cast(vp, C.c_void_p))
        return m, vpos
    return m
