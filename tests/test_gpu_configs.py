"""GPU parity on the benchmark workloads' shapes at test sizes, the reference's
known-answer arrays through the HIP kernels, the 32-bit nnz wrap, and inputs
that do not suit the fast paths (non-canonical rows, a hub row of thousands
of ranges).  Every output is compared bit for bit (rowmap, entries, nnz)."""
import json
import os

import numpy as np
import pytest

import diverg_amd as dv
from oracle import oracle as orc
from tests.rcrs_util import assert_identical, from_dense, to_dense

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


@pytest.fixture(autouse=True)
def default_tuning(ctx):
    t = ctx.tuning()
    yield
    ctx.set_tuning(**t)


def long_edges(a, min_jump=50):
    """Edges i -> j with |j - i| >= min_jump (structural-variant edges)."""
    rm = a.rowmap.astype(np.int64)
    rows = np.repeat(np.arange(a.nrows), np.diff(rm) // 2)
    s = a.entries[0::2].astype(np.int64)
    return int(np.count_nonzero(np.abs(s - rows) >= min_jump))


@pytest.mark.parametrize("layout", [0, 1])
def test_chromosome_rates_u64_200_1000(ctx, layout):
    """The chromosome workload's generator at its default rates (SNP 0.01,
    indel 0.001, SVs at 1e-5 with 30% inversions) on a 1e6-bp backbone,
    uint64 rowmap, [200, 1000]."""
    ctx.set_tuning(layout=layout)
    a = dv.generate_graph(3, backbone_bp=1_000_000, size_bytes=8)
    assert a.rowmap.dtype == np.uint64
    assert long_edges(a) >= 5   # the backbone holds several SVs
    print(f"chromosome rates: {a.nrows} loci, [200,1000], layout {layout}", flush=True)
    d = ctx.create_distance_index(a, 200, 1000)
    assert_identical(d, orc.dindex(a, 200, 1000), "D[200,1000]")


@pytest.mark.parametrize("layout", [0, 1])
def test_wide_index_two_ranges_per_row(ctx, layout):
    """An index of about two ranges per row (the shape reported for real
    graphs): SNP 0.1, indel 0.05, SVs at 3e-3 (half inversions) on a 2e5-bp
    backbone, uint32 rowmap, [300, 600].  Two builds in a row: a slot build
    that runs out of its predicted side entries is repeated in the exact form
    and the next one runs with the prediction it learned; each build, in
    either layout, matches the oracle."""
    ctx.set_tuning(layout=layout)
    a = dv.generate_graph(2, size_bytes=4, backbone_bp=200_000, snp_rate=0.1, indel_rate=0.05,
                          sv_rate=3e-3, sv_inversion_frac=0.5, sv_deletion_frac=0.3)
    e = orc.dindex(a, 300, 600)
    assert len(e.entries) / 2 / e.nrows > 1.9
    for k in range(2):
        print(f"wide index: {a.nrows} loci, [300,600], layout {layout}, build {k}", flush=True)
        assert_identical(ctx.create_distance_index(a, 300, 600), e, f"D[300,600] build {k}")


@pytest.mark.parametrize("graphs", [0, 1])
def test_batched_preset4_components(ctx, graphs):
    """Components of the batched workload (preset 4: log-normal lengths,
    the rates of preset 2) at 1/100 of their lengths, batched at [300, 600]
    and checked one by one against the oracle."""
    ctx.set_tuning(graphs=graphs)
    comps = []
    for c in range(8):
        p = dv.graph_preset(4, c)
        p.backbone_bp = max(2000, p.backbone_bp // 100)
        comps.append(dv.generate_graph(params=p, size_bytes=8))
    print(f"batched preset-4 components {[a.nrows for a in comps]} graphs={graphs}", flush=True)
    outs = ctx.create_distance_index_batched(comps, 300, 600)
    for c, (a, d) in enumerate(zip(comps, outs)):
        assert_identical(d, orc.dindex(a, 300, 600), f"component {c}")


@pytest.mark.parametrize("plan_mode", [0, 1])
def test_power_a300_preset2(ctx, plan_mode):
    """A^d1 of the 5e6-bp workload's generator (d1 = 300) at 2e5 bp, uint32."""
    ctx.set_tuning(plan_mode=plan_mode)
    a = dv.generate_graph(2, backbone_bp=200_000, size_bytes=4)
    print(f"A^300 of {a.nrows} loci plan_mode={plan_mode}", flush=True)
    assert_identical(ctx.range_power(a, 300), orc.power(a, 300), "A^300")


def chain(n, size_bytes):
    """The path 0 -> 1 -> ... -> n-1."""
    rm_dt = np.uint32 if size_bytes == 4 else np.uint64
    en = np.repeat(np.arange(1, n, dtype=np.uint32), 2)
    rm = np.concatenate([np.arange(0, 2 * n - 1, 2), [2 * n - 2]]).astype(rm_dt)
    return dv.RangeCRS(n, en, rm)


@pytest.mark.parametrize("layout", [1, 2])
def test_u32_nnz_wraps_like_the_reference(ctx, layout):
    """A uint32 index whose true nnz is 4.8e9 > 2^32: the chain of 1e5 loci
    at [0, 80000] has rows [i, min(i + 80000, n - 1)]; nnz comes back modulo
    2^32, as the reference's size-typed counter (range_sparse.hpp:1152,
    crs_matrix.hpp:2003)."""
    ctx.set_tuning(layout=layout)
    n, dup = 100_000, 80_000
    a = chain(n, 4)
    print(f"u32 nnz wrap: chain {n} [0,{dup}] layout {layout}", flush=True)
    d = ctx.create_distance_index(a, 0, dup)
    i = np.arange(n, dtype=np.int64)
    ends = np.minimum(i + dup, n - 1)
    true_nnz = int(np.sum(ends - i + 1))
    assert true_nnz > 1 << 32
    assert d.rowmap.dtype == np.uint32
    assert np.array_equal(d.rowmap.astype(np.int64), 2 * np.arange(n + 1))
    assert np.array_equal(d.entries[0::2].astype(np.int64), i)
    assert np.array_equal(d.entries[1::2].astype(np.int64), ends)
    assert d.nnz == true_nnz % (1 << 32)
    d8 = ctx.create_distance_index(chain(n, 8), 0, dup)
    assert d8.nnz == true_nnz


@pytest.fixture
def kat_tiers(ctx):
    """Group tiers only, with one-word LDS windows (every row chunked)."""
    ctx.set_tuning(lane_slice_max=0, lds_words=1)


def test_hbitvector_kat_through_hip(ctx, kat_tiers):
    """test_hbitvector.cpp:569-661 as I . M on the GPU."""
    k = load("hbitvector_kat.json")
    m = dv.RangeCRS(k["ncols"], np.array(k["entries"], np.uint32),
                    np.array(k["rowmap"], np.uint64))
    c = ctx.range_spgemm(ctx.identity(5), m)
    assert c.rowmap.tolist() == k["expected_rowmap"]
    assert c.entries.tolist() == k["expected_entries"]
    assert c.ncols == k["ncols"]


def test_example26_through_hip(ctx, kat_tiers):
    """Manuscript Example 26: the word 0011110011001110 gathers to
    [2,5], [8,9], [12,14]."""
    k = load("example26.json")
    bits = [i for i, ch in enumerate(k["bits"]) if ch == "1"]
    m = dv.RangeCRS(16, np.repeat(np.array(bits, np.uint32), 2), np.array([0, 2 * len(bits)]))
    c = ctx.range_spgemm(ctx.identity(1), m)
    assert c.entries.tolist() == k["expected_entries"]
    assert c.nnz == len(bits)


def test_query_tiny_kat_through_hip(ctx):
    """test_crsmatrix.cpp:398-422: every cell of the 10 x 10 matrix."""
    k = load("crsmatrix_tiny.json")
    dense = np.zeros((k["nrows"], k["ncols"]), bool)
    for i, j in k["cells"]:
        dense[i, j] = True
    m = from_dense(dense)
    qi, qj = np.meshgrid(np.arange(10, dtype=np.uint32), np.arange(10, dtype=np.uint32),
                         indexing="ij")
    out = ctx.query(m, qi.ravel(), qj.ravel())
    assert np.array_equal(out.reshape(10, 10).astype(bool), dense)


def scrambled(a, rng):
    """A copy of `a` whose rows of several ranges are in reverse or shuffled
    order, plus rows with touching and overlapping ranges: non-canonical rows
    the reference accepts (its I + A merge then keeps a verbatim tail)."""
    rm = a.rowmap.astype(np.int64)
    rows = []
    for i in range(a.nrows):
        r = a.entries[rm[i]:rm[i + 1]].reshape(-1, 2).tolist()
        if len(r) > 1:
            r = r[::-1] if rng.random() < 0.5 else [r[k] for k in rng.permutation(len(r))]
        rows.append(r)
    n = a.nrows
    rows[7] = [[10, 12], [3, 5]]                # unsorted, i inside neither
    rows[20] = [[21, 22], [23, 25]]             # touching
    rows[40] = [[45, 50], [44, 46], [30, 30]]   # overlapping, unsorted
    rows[n - 3] = [[n - 1, n - 1], [0, 1]]
    en = np.array([x for r in rows for pair in r for x in pair], np.uint32)
    lens = np.array([2 * len(r) for r in rows], np.int64)
    rmo = np.concatenate([[0], np.cumsum(lens)]).astype(a.rowmap.dtype)
    return dv.RangeCRS(a.ncols, en, rmo)


@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("schedule", [0, 1])
def test_distance_index_noncanonical_input(ctx, layout, schedule):
    """Non-canonical input rows: every layout and schedule gives the
    reference's index (range CRS in the reference's order; the slot form and
    the paired schedule need canonical rows), on the first build and on the
    next builds of the same shape."""
    rng = np.random.default_rng(layout * 2 + schedule)
    a = scrambled(dv.generate_graph(3, backbone_bp=3000, seed=21, snp_rate=0.05,
                                    indel_rate=0.02, sv_rate=2e-3, sv_min=20, sv_max=200), rng)
    ctx.set_tuning(layout=layout, schedule=schedule)
    for dlo, dup in [(20, 40), (0, 30), (6, 6)]:
        want = orc.dindex(a, dlo, dup)
        print(f"non-canonical input [{dlo},{dup}] layout {layout} schedule {schedule}", flush=True)
        for build in range(2):
            d, st = ctx.create_distance_index(a, dlo, dup, stats=True)
            assert_identical(d, want, f"D[{dlo},{dup}] build {build}")
            assert st["layout"] == 1


@pytest.mark.parametrize("layout", [0, 2])
def test_hub_row_repeated_builds(ctx, layout):
    """A hub row of 10 000 ranges (an edge to every fourth vertex; the
    products' hub rows keep 10 000 ranges) does not fit a side partition of
    the slot form's input conversion: the build moves to range CRS (or the
    exact build when the slot form is forced), and repeated builds stay
    exact."""
    n = 40000
    rows = [[[v, v] for v in range(4, n, 4)]] + [[[i + 1, i + 1]] for i in range(1, n - 1)] + [[]]
    en = np.array([x for r in rows for pair in r for x in pair], np.uint32)
    rm = np.concatenate([[0], np.cumsum([2 * len(r) for r in rows])]).astype(np.uint64)
    a = dv.RangeCRS(n, en, rm)
    ctx.set_tuning(layout=layout)
    want = orc.dindex(a, 1, 4)
    for build in range(3):
        print(f"hub row build {build} layout {layout}", flush=True)
        d, st = ctx.create_distance_index(a, 1, 4, stats=True)
        assert_identical(d, want, f"hub build {build}")
        if layout == 0 and build > 0:
            assert st["layout"] == 1 and st["retried"] == 0


def far_noncanonical(n, rng, far=3000):
    """Rows of two or three ranges in random order, one of them far from the
    row (outside the reference accumulator's 2048-bit band around it)."""
    rows = []
    for i in range(n):
        r = [[min(i + 1, n - 1), min(i + 2, n - 1)]]
        if rng.random() < 0.3:
            c = int((i + rng.choice([-far, far])) % n)
            r.append([c, min(c + int(rng.integers(0, 3)), n - 1)])
        if rng.random() < 0.2:
            c = int(rng.integers(0, n))
            r.append([c, c])
        rng.shuffle(r)
        rows.append(r)
    en = np.array([x for r in rows for pair in r for x in pair], np.uint32)
    rm = np.concatenate([[0], np.cumsum([2 * len(r) for r in rows])]).astype(np.uint64)
    return dv.RangeCRS(n, en, rm)


def test_spgemm_noncanonical_far_ranges(ctx):
    """A . B with non-canonical rows of B reaching outside the reference
    accumulator's band: the reference keeps only what falls in its growing
    window (range_sparse.hpp:1192-1242), and so does the product here."""
    rng = np.random.default_rng(17)
    b = far_noncanonical(9000, rng)
    a = far_noncanonical(9000, rng)
    want = orc.spgemm(a, b)
    assert_identical(ctx.range_spgemm(a, b), want, "A.B non-canonical")
    # the window matters here: the product with B's rows as sets differs
    union = orc.spgemm(a, from_dense(to_dense(b)))
    assert not (len(union.entries) == len(want.entries)
                and np.array_equal(union.entries, want.entries))
    for k in (1, 2, 5):
        assert_identical(ctx.range_power(b, k), orc.power(b, k), f"B^{k} non-canonical")


@pytest.mark.parametrize("layout", [0, 1, 2])
def test_distance_index_noncanonical_far_ranges(ctx, layout):
    """Distance indexes of an input whose non-canonical rows reach outside the
    accumulator band, in every layout, first and repeated builds."""
    rng = np.random.default_rng(23)
    a = far_noncanonical(7000, rng)
    ctx.set_tuning(layout=layout)
    for dlo, dup in [(3, 9), (0, 6)]:
        want = orc.dindex(a, dlo, dup)
        print(f"non-canonical far ranges [{dlo},{dup}] layout {layout}", flush=True)
        for build in range(2):
            assert_identical(ctx.create_distance_index(a, dlo, dup), want,
                             f"D[{dlo},{dup}] build {build}")


def test_batched_with_noncanonical_component(ctx):
    """A batched call with one non-canonical component: that component is
    built on its own, the others in the assembly; all equal the oracle."""
    rng = np.random.default_rng(29)
    comps = [dv.generate_graph(1, backbone_bp=20000, seed=3), far_noncanonical(5000, rng),
             dv.generate_graph(2, backbone_bp=30000, seed=4)]
    outs = ctx.create_distance_index_batched(comps, 4, 12)
    for c, (a, d) in enumerate(zip(comps, outs)):
        assert_identical(d, orc.dindex(a, 4, 12), f"component {c}")
