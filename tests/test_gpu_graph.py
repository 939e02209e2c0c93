"""Sequence-graph path on the GPU: the device adjacency build against the
oracle's restatement of adjacency.hpp:170-228, distance indices of the
reference's own GFA fixtures (test_dindex.cpp:180-240), index merges
(test_crsmatrix.cpp:911-1165) and on-disk round trips of device indexes."""
import numpy as np
import pytest

import diverg_amd as dv
from oracle import oracle as orc
from tests.rcrs_util import assert_identical, from_dense, random_dense, to_dense
from tests.test_graph import gfa_path, random_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size_bytes", [4, 8])
@pytest.mark.parametrize("name", ["x.gfa", "multi.gfa", "m.gfa.gz"])
@pytest.mark.parametrize("order", ["native", "min-breaks", "cuthill-mckee"])
def test_adjacency_fixtures(ctx, name, size_bytes, order, tmp_path):
    print(f"adjacency {name} sb={size_bytes} order={order}", flush=True)
    g = dv.Graph.load_gfa(gfa_path(name, tmp_path))
    g.sort(order)
    nl, co, eo, et = g.arrays()
    want = orc.adjacency(nl, eo, et)
    got = ctx.adjacency_matrix(g, size_bytes=size_bytes)
    assert_identical(got, want, f"adjacency {name}")
    info = g.info()
    assert got.nnz == info["loci"] - info["nodes"] + info["edges"]  # adjacency.hpp:192-194


def test_adjacency_components_assemble_to_whole(ctx):
    """test_dindex.cpp:84-173: per-component blocks, shifted by their first
    locus, assemble to the whole-graph matrix."""
    g = dv.Graph.load_gfa(gfa_path("multi.gfa"))
    nl, co, eo, et = g.arrays()
    whole = ctx.adjacency_matrix(g)
    b = g.components()
    rm, en = [0], []
    for lo, hi in zip(b[:-1], b[1:]):
        blk = ctx.adjacency_matrix(g, int(lo), int(hi))
        assert_identical(blk, orc.adjacency(nl, eo, et, int(lo), int(hi)), "component block")
        shift = int(co[int(lo)])
        en.extend((blk.entries.astype(np.int64) + shift).tolist())
        rm.extend((blk.rowmap[1:].astype(np.int64) + rm[-1]).tolist())
    assert whole.rowmap.tolist() == rm
    assert whole.entries.tolist() == en


@pytest.mark.parametrize("seed", range(3))
def test_adjacency_random_graphs(ctx, seed):
    rng = np.random.default_rng(40 + seed)
    lens, ef, et = random_graph(rng, 3000, 2500, maxlen=40)
    g = dv.Graph.from_arrays(lens, ef, et)
    nl, co, eo, etg = g.arrays()
    for sb in (4, 8):
        assert_identical(ctx.adjacency_matrix(g, size_bytes=sb), orc.adjacency(nl, eo, etg),
                         f"random graph seed {seed}")


def test_adjacency_errors(ctx):
    g = dv.Graph.from_arrays([2, 3, 1], [0, 1], [1, 2])
    with pytest.raises(dv.DiVerGError):   # the edge 1 -> 2 leaves [0, 2)
        ctx.adjacency_matrix(g, 0, 2)
    with pytest.raises(dv.DiVerGError):
        ctx.adjacency_matrix(g, 2, 1)
    one = dv.Graph.from_arrays([5])
    m = ctx.adjacency_matrix(one)
    assert m.entries.tolist() == [1, 1, 2, 2, 3, 3, 4, 4] and m.nnz == 4


@pytest.mark.parametrize("size_bytes,dlo,dup", [(8, 1, 8), (4, 1, 3), (8, 0, 5), (8, 4, 4)])
def test_distance_index_multi_gfa(ctx, size_bytes, dlo, dup):
    """test_dindex.cpp:180-240 builds D of multi.gfa for [1, 8] (uint64
    rowmap) and [1, 3] (uint32); here against the oracle."""
    g = dv.Graph.load_gfa(gfa_path("multi.gfa"))
    a = ctx.adjacency_matrix(g, size_bytes=size_bytes)
    assert_identical(ctx.create_distance_index(a, dlo, dup), orc.dindex(a, dlo, dup),
                     f"multi.gfa D[{dlo},{dup}]")


def test_distance_index_m_gfa_from_device_adjacency(ctx, tmp_path):
    g = dv.Graph.load_gfa(gfa_path("m.gfa.gz", tmp_path))
    g.sort("min-breaks")
    a_dev = ctx.adjacency_matrix(g, location="device")
    d = ctx.create_distance_index(a_dev, 10, 60)
    a = a_dev.to_host()
    assert_identical(d, orc.dindex(a, 10, 60), "m.gfa D[10,60]")


# -- merge ------------------------------------------------------------------

def rnd(rng, n, density):
    return from_dense(random_dense(rng, n, n, density, 6))


def test_merge_reference_cases(ctx):
    """test_crsmatrix.cpp:925-1063."""
    n = 200
    z = from_dense(np.zeros((n, n), bool))
    assert ctx.merge_distance_index(z, z).nnz == 0
    rng = np.random.default_rng(11)
    r = rnd(rng, n, 0.2)
    for x, y in [(r, z), (z, r)]:
        m = ctx.merge_distance_index(x, y)
        assert dv.identical(m, r) and m.nnz == int(to_dense(r).sum())
    r1, r2 = rnd(rng, n, 0.15), rnd(rng, n, 0.1)
    m = ctx.merge_distance_index(r1, r2)
    assert_identical(m, orc.merge_stream(r1, r2), "random merge")
    assert (to_dense(m) == (to_dense(r1) | to_dense(r2))).all()
    i, j = np.indices((n, n))
    m = ctx.merge_distance_index(from_dense(i + 20 < j), from_dense(~(i + 20 < j)))
    assert to_dense(m).all()


def test_merge_large_dimensions(ctx):
    """test_crsmatrix.cpp:1067-1165: blocks at opposite corners of a matrix
    of order 32767."""
    n = 32767
    rng = np.random.default_rng(3)
    b1, b2 = random_dense(rng, 200, 200, 0.2, 5), random_dense(rng, 400, 400, 0.1, 5)

    def place(block, at):
        r = from_dense(block)
        rm = np.zeros(n + 1, np.uint64)
        rows = block.shape[0]
        rm[at + 1:at + rows + 1] = r.rowmap[1:]
        rm[at + rows + 1:] = r.rowmap[-1]
        return dv.RangeCRS(n, (r.entries.astype(np.int64) + at).astype(np.uint32), rm)
    m1, m2 = place(b1, 0), place(b2, n - 400)
    m = ctx.merge_distance_index(m1, m2)
    assert m.nnz == int(b1.sum()) + int(b2.sum())
    assert_identical(m, orc.merge_stream(m1, m2), "large merge")


def test_merge_of_two_indices(ctx):
    """Two indices of one graph over adjacent distance windows merge into the
    index of the joined window."""
    a = dv.generate_graph(1, backbone_bp=30000, seed=21)
    d1 = ctx.create_distance_index(a, 10, 40, location="device")
    d2 = ctx.create_distance_index(a, 41, 90, location="device")
    m = ctx.merge_distance_index(d1, d2)
    assert_identical(m, orc.dindex(a, 10, 90), "merged windows")


# -- persistence ------------------------------------------------------------

@pytest.mark.parametrize("size_bytes", [4, 8])
def test_save_load_device_round_trip(ctx, tmp_path, size_bytes):
    a = dv.generate_graph(1, backbone_bp=20000, seed=4, size_bytes=size_bytes)
    d = ctx.create_distance_index(a, 5, 50, location="device")
    p = str(tmp_path / "d.bin")
    ctx.save(d, p)
    back = ctx.load(p, location="device")
    assert_identical(back.to_host(), orc.dindex(a, 5, 50), "loaded index")
    hits = ctx.query(back, np.arange(100, dtype=np.uint32), np.arange(100, dtype=np.uint32) + 7)
    want = orc.query(orc.dindex(a, 5, 50), np.arange(100, dtype=np.uint32),
                     np.arange(100, dtype=np.uint32) + 7)[0]
    assert hits.tolist() == want.tolist()
