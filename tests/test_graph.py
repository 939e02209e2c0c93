"""Sequence graphs on the host (CPU only): GFA loading, orderings, component
blocks, the oracle's adjacency and streaming-merge restatements, and the
on-disk matrix layout.

The GFA fixtures are the reference's own test data (test/data/small/x.gfa,
test/data/multi/multi.gfa, test/data/middle/m.gfa, the last gzipped).
"""
import ctypes as C
import gzip
import os
import shutil
import struct

import numpy as np
import pytest

import diverg_amd as dv
from diverg_amd import _native as N
from oracle import oracle as orc
from tests.rcrs_util import assert_identical, from_dense, random_dense, to_dense

GFA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gfa")


def gfa_path(name, tmp_path=None):
    p = os.path.join(GFA, name)
    if name.endswith(".gz"):
        out = os.path.join(str(tmp_path), name[:-3])
        with gzip.open(p, "rb") as f, open(out, "wb") as g:
            shutil.copyfileobj(f, g)
        return out
    return p


def parse_gfa_py(path):
    """Independent reading of the fixture: segment names and lengths in file
    order, and the set of (from, to) segment pairs of E / L records."""
    names, lens, edges = [], [], set()
    with open(path) as f:
        for line in f:
            t = line.rstrip("\n").split("\t")
            if t[0] == "S":
                names.append(t[1])
                lens.append(int(t[2]) if t[2].isdigit() and len(t) >= 4 else len(t[2]))
            elif t[0] == "E":
                edges.add((t[2].rstrip("+-"), t[3].rstrip("+-")))
            elif t[0] == "L":
                edges.add((t[1], t[3]))
    return names, lens, edges


def graph_edges_by_name(g):
    nl, co, eo, et = g.arrays()
    n = len(nl)
    names = [g.node_name(k) for k in range(n)]
    out = set()
    for u in range(n):
        for p in range(int(eo[u]), int(eo[u + 1])):
            out.add((names[u], names[int(et[p])]))
    return names, dict(zip(names, nl.tolist())), out


def dense_adjacency(node_len, edges, lower=0, upper=None):
    """The character-level adjacency from its definition (dense, small)."""
    upper = len(node_len) if upper is None else upper
    start = np.concatenate([[0], np.cumsum(node_len[lower:upper])]).astype(np.int64)
    n = int(start[-1])
    m = np.zeros((n, n), dtype=bool)
    for k in range(upper - lower):
        for j in range(int(node_len[lower + k]) - 1):
            m[start[k] + j, start[k] + j + 1] = True
    for u, v in edges:
        if lower <= u < upper:
            m[start[u - lower + 1] - 1, start[v - lower]] = True
    return m


@pytest.mark.parametrize("name", ["x.gfa", "multi.gfa", "m.gfa.gz"])
def test_gfa_load_matches_file(name, tmp_path):
    path = gfa_path(name, tmp_path)
    g = dv.Graph.load_gfa(path)
    names, lens, edges = parse_gfa_py(path)
    info = g.info()
    assert info["nodes"] == len(names)
    assert info["loci"] == sum(lens)
    assert info["edges"] == len(edges)
    got_names, got_len, got_edges = graph_edges_by_name(g)
    assert got_names == names  # S-line order
    assert got_len == dict(zip(names, lens))
    assert got_edges == edges


def test_gfa1_and_errors(tmp_path):
    p = tmp_path / "g1.gfa"
    p.write_text("H\tVN:Z:1.0\nS\ta\tACGT\nS\tb\t*\tLN:i:3\nS\tc\tG\n"
                 "L\ta\t+\tb\t-\t0M\nL\ta\t+\tc\t+\t0M\nL\ta\t+\tc\t-\t0M\nL\tb\t+\tc\t+\t0M\n")
    g = dv.Graph.load_gfa(str(p))
    assert g.info()["nodes"] == 3 and g.info()["loci"] == 8
    assert g.info()["edges"] == 3  # a->c twice (orientations differ) is one edge
    bad = tmp_path / "bad.gfa"
    bad.write_text("S\ta\tACGT\nL\ta\t+\tzz\t+\t0M\n")
    with pytest.raises(dv.DiVerGError) as e:
        dv.Graph.load_gfa(str(bad))
    assert e.value.status == 9
    empty = tmp_path / "zero.gfa"
    empty.write_text("S\ta\t*\n")
    with pytest.raises(dv.DiVerGError):
        dv.Graph.load_gfa(str(empty))
    with pytest.raises(dv.DiVerGError):
        dv.Graph.load_gfa(str(tmp_path / "missing.gfa"))


def weak_components_py(n, edges):
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    for u, v in edges:
        a, b = find(u), find(v)
        if a != b:
            parent[max(a, b)] = min(a, b)
    return [find(x) for x in range(n)]


def edge_list(g):
    nl, co, eo, et = g.arrays()
    return [(u, int(et[p])) for u in range(len(nl)) for p in range(int(eo[u]), int(eo[u + 1]))]


def test_components_multi():
    g = dv.Graph.load_gfa(gfa_path("multi.gfa"))
    b = g.components()
    n = g.info()["nodes"]
    comp = weak_components_py(n, edge_list(g))
    assert b[0] == 0 and b[-1] == n and np.all(np.diff(b.astype(np.int64)) > 0)
    # every block is exactly one weakly connected component here (the file
    # lists each component's segments together)
    assert len(b) - 1 == len(set(comp)) >= 2
    for lo, hi in zip(b[:-1], b[1:]):
        assert len({comp[k] for k in range(int(lo), int(hi))}) == 1


def test_components_interleaved_are_merged():
    # 0-2 and 1-3 interleave: the finest cut-free blocks are [0, 4) and [4, 5)
    g = dv.Graph.from_arrays([1, 1, 1, 1, 2], [0, 1], [2, 3])
    assert g.components().tolist() == [0, 4, 5]


@pytest.mark.parametrize("order", ["native", "topological", "min-breaks", "cuthill-mckee", "random"])
@pytest.mark.parametrize("name", ["x.gfa", "multi.gfa"])
def test_orderings_renumber_the_same_graph(order, name):
    g0 = dv.Graph.load_gfa(gfa_path(name))
    names0, len0, edges0 = graph_edges_by_name(g0)
    g = dv.Graph.load_gfa(gfa_path(name))
    g.sort(order, seed=7)
    names, lens, edges = graph_edges_by_name(g)
    assert sorted(names) == sorted(names0)
    assert lens == len0 and edges == edges0
    if order in ("topological", "min-breaks"):
        # the fixtures are DAGs: every edge goes forward
        assert all(u < v for u, v in edge_list(g))
    if order != "native" and order != "random":
        comp = weak_components_py(len(names), edge_list(g))
        # components contiguous
        seen, last = set(), None
        for c in comp:
            if c != last:
                assert c not in seen
                seen.add(c)
                last = c
    if order == "random":
        g2 = dv.Graph.load_gfa(gfa_path(name))
        g2.sort("random", seed=7)
        assert [g2.node_name(k) for k in range(len(names))] == names


def test_min_breaks_keeps_chains_consecutive():
    # a chain given in scrambled file order: min-breaks restores it
    perm = [3, 0, 4, 1, 2]
    # node k of the file is chain position perm[k]; edges along the chain
    pos_to_node = {p: k for k, p in enumerate(perm)}
    ef = [pos_to_node[p] for p in range(4)]
    et = [pos_to_node[p + 1] for p in range(4)]
    g = dv.Graph.from_arrays([2] * 5, ef, et)
    assert g.info()["breaks"] == 3  # only 3 -> 4 is consecutive in file order
    g.sort("min-breaks")
    assert g.info()["breaks"] == 0


def random_graph(rng, n_nodes, n_extra, maxlen=4):
    lens = rng.integers(1, maxlen + 1, size=n_nodes).astype(np.uint32)
    ef = list(range(n_nodes - 1))
    et = list(range(1, n_nodes))
    for _ in range(n_extra):
        u = int(rng.integers(0, n_nodes))
        v = int(rng.integers(max(0, u - 3), min(n_nodes, u + 8)))
        ef.append(u)
        et.append(v)
    return lens, ef, et


@pytest.mark.parametrize("seed", range(4))
def test_oracle_adjacency_matches_definition(seed):
    rng = np.random.default_rng(seed)
    lens, ef, et = random_graph(rng, 60, 40)
    g = dv.Graph.from_arrays(lens, ef, et)
    nl, co, eo, etg = g.arrays()
    a = orc.adjacency(nl, eo, etg)
    edges = set(zip(ef, et))
    want = from_dense(dense_adjacency(nl, edges))
    assert a.rowmap.tolist() == want.rowmap.tolist()
    assert a.entries.tolist() == want.entries.tolist()
    # adjacency.hpp:192-194
    assert a.nnz == int(nl.sum()) - len(nl) + len(edges)


def test_oracle_adjacency_block_diagonal_multi():
    """test_dindex.cpp:84-173: the whole-graph matrix equals the assembly of
    the per-component blocks, each shifted by its first locus."""
    g = dv.Graph.load_gfa(gfa_path("multi.gfa"))
    nl, co, eo, et = g.arrays()
    whole = orc.adjacency(nl, eo, et)
    b = g.components()
    rm, en = [0], []
    for lo, hi in zip(b[:-1], b[1:]):
        blk = orc.adjacency(nl, eo, et, int(lo), int(hi))
        shift = int(co[int(lo)])
        en.extend((blk.entries.astype(np.int64) + shift).tolist())
        rm.extend((blk.rowmap[1:].astype(np.int64) + rm[-1]).tolist())
    assert whole.rowmap.tolist() == rm
    assert whole.entries.tolist() == en


def test_oracle_adjacency_duplicate_columns_open_new_ranges():
    """crs_matrix.hpp:197-242: a repeated column does not extend the open
    range (hand trace: cols 1, 1, 2 -> [1,1], [1,2])."""
    a = orc.adjacency(np.array([1, 1, 1], np.uint32), np.array([0, 3, 3, 3], np.uint64),
                      np.array([1, 1, 2], np.uint32))
    assert a.entries.tolist() == [1, 1, 1, 2]
    assert a.nnz == 3


def random_canonical(rng, n, density):
    return from_dense(random_dense(rng, n, n, density, 6))


@pytest.mark.parametrize("seed", range(3))
def test_oracle_merge_stream_is_union(seed):
    rng = np.random.default_rng(100 + seed)
    a, b = random_canonical(rng, 200, 0.1), random_canonical(rng, 200, 0.06)
    m = orc.merge_stream(a, b)
    want = from_dense(to_dense(a) | to_dense(b))
    assert m.rowmap.tolist() == want.rowmap.tolist()
    assert m.entries.tolist() == want.entries.tolist()
    assert m.nnz == int((to_dense(a) | to_dense(b)).sum())
    s = orc.spadd(a, b)
    assert s.entries.tolist() == m.entries.tolist() and s.rowmap.tolist() == m.rowmap.tolist()


def test_oracle_merge_stream_reference_cases():
    """test_crsmatrix.cpp:925-1063: zero + zero, random + zero (both ways),
    triangular halves -> all ones."""
    n = 200
    z = from_dense(np.zeros((n, n), bool))
    assert orc.merge_stream(z, z).nnz == 0
    rng = np.random.default_rng(5)
    r = random_canonical(rng, n, 0.2)
    for x, y in [(r, z), (z, r)]:
        m = orc.merge_stream(x, y)
        assert m.entries.tolist() == r.entries.tolist() and m.nnz == int(to_dense(r).sum())
    i, j = np.indices((n, n))
    up, lo = from_dense(i + 20 < j), from_dense(~(i + 20 < j))
    m = orc.merge_stream(up, lo)
    assert to_dense(m).all() and m.entries.tolist() == [0, n - 1] * n


# -- on-disk layout --------------------------------------------------------

def save_host(m, path):
    keep = []
    rc = N.lib().dv_save(None, C.byref(m._desc(keep)), os.fsencode(path))
    assert rc == 0, rc


def load_host(path, size_bytes=0):
    out = N.dv_rcrs()
    rc = N.lib().dv_load(None, os.fsencode(path), size_bytes, N.DV_HOST, C.byref(out))
    if rc:
        return rc
    m = dv._from_desc(out)
    N.lib().dv_free(None, C.byref(out))
    return m


@pytest.mark.parametrize("size_bytes", [4, 8])
def test_save_layout_and_round_trip(tmp_path, size_bytes):
    m = dv.RangeCRS(10, np.array([1, 3, 0, 0, 5, 9], np.uint32),
                    np.array([0, 2, 2, 6], np.uint32 if size_bytes == 4 else np.uint64))
    p = tmp_path / "m.bin"
    save_host(m, p)
    raw = p.read_bytes()
    # crs_matrix.hpp:1392-1409 + :1986-2000 with utils.hpp's vector layout
    fmt = "<Q6IQ4" + ("I" if size_bytes == 4 else "Q") + "QQ"
    assert raw == struct.pack(fmt, 6, 1, 3, 0, 0, 5, 9, 4, 0, 2, 2, 6, 10, m.nnz)
    assert m.nnz == 3 + 1 + 5
    back = load_host(p)
    assert back.size_bytes == size_bytes
    assert dv.identical(back, m) and back.nnz == m.nnz


def test_load_rejects_malformed(tmp_path):
    good = struct.pack("<Q2IQ2QQQ", 2, 1, 2, 2, 0, 2, 5, 2)
    p = tmp_path / "ok.bin"
    p.write_bytes(good)
    assert load_host(p, 8).nrows == 1
    cases = {
        "truncated": good[:-4],
        "odd": struct.pack("<Q2IQ2QQQ", 2, 1, 2, 2, 0, 1, 5, 2),
        "outside": struct.pack("<Q2IQ2QQQ", 2, 1, 7, 2, 0, 2, 5, 7),
        "bad_end": struct.pack("<Q2IQ2QQQ", 2, 1, 2, 2, 0, 0, 5, 2),
    }
    for name, raw in cases.items():
        q = tmp_path / (name + ".bin")
        q.write_bytes(raw)
        assert load_host(q) == 9, name
    assert load_host(tmp_path / "none.bin") == 9
