"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact: rowmap, entries and nnz (check_identical, test_dindex.cpp:34-53).
"""
import numpy as np
import pytest

import diverg_amd as dv
from oracle import oracle as orc
from tests.rcrs_util import (assert_identical, bool_matmul, from_dense, random_dense,
                             to_dense)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph_1e5():
    return dv.generate_graph(1)


@pytest.fixture(autouse=True)
def default_tuning(ctx):
    t = ctx.tuning()
    yield
    ctx.set_tuning(**{k: v for k, v in t.items() if k != "autotune"}, autotune=t["autotune"])


@pytest.mark.parametrize("size_bytes", [4, 8])
@pytest.mark.parametrize("shape,density,run", [
    ((1, 1, 1), 1.0, 1),
    ((40, 70, 33), 0.2, 3),
    ((300, 500, 700), 0.05, 8),
    ((257, 257, 4100), 0.02, 64),
    ((64, 3000, 129), 0.3, 5),
])
def test_spgemm_random(ctx, size_bytes, shape, density, run):
    m, k, n = shape
    rng = np.random.default_rng(hash((shape, size_bytes)) % (1 << 32))
    da = random_dense(rng, m, k, density, run)
    db = random_dense(rng, k, n, density, run)
    a, b = from_dense(da, size_bytes), from_dense(db, size_bytes)
    print(f"spgemm random {shape} S={size_bytes}", flush=True)
    c = ctx.range_spgemm(a, b)
    assert np.array_equal(to_dense(c), bool_matmul(da, db))
    assert_identical(c, orc.spgemm(a, b), "spgemm vs oracle")


def test_spgemm_empty_rows_and_matrix(ctx):
    a = from_dense(np.zeros((5, 7), bool))
    b = from_dense(np.ones((7, 9), bool))
    c = ctx.range_spgemm(a, b)
    assert c.nrows == 5 and len(c.entries) == 0 and c.nnz == 0
    assert np.all(c.rowmap == 0)


TIERS = [
    # lds_words, stage, lanes_small, small_slice_max, lane_slice_max
    (512, 8, 16, 64, 4096),   # defaults: lane tier + 16-lane + 64-lane groups
    (512, 8, 16, 64, 16),     # lane tier up to 16 slice ranges
    (1, 8, 16, 64, 16),       # one-word windows: every group row is chunked
    (2, 2, 16, 64, 0),        # lane tier off, two-entry stage (most rows recomputed)
    (3, 4, 4, 64, 0),         # 4-lane groups, 3-word windows
    (64, 2, 8, 100000, 0),    # every row to 8-lane groups
    (512, 8, 32, 100000, 0),  # every row to 32-lane groups
    (512, 8, 16, 0, 0),       # every row to 64-lane groups
    (8, 6, 64, 64, 3),        # small lane tier, 64-lane medium groups
]


@pytest.mark.parametrize("tier", TIERS, ids=[str(t) for t in TIERS])
def test_spgemm_tiers_chunks_and_stage(ctx, graph_1e5, tier):
    """Every tier, tiny LDS windows (chunked path) and tiny stages (direct
    recompute) give the same product."""
    lds, stage, lanes, mid_max, lane_max = tier
    ctx.set_tuning(lds_words=lds, stage_entries=stage, lanes_small=lanes,
                   small_slice_max=mid_max, lane_slice_max=lane_max)
    ai = orc.spadd(orc.identity(graph_1e5.nrows), graph_1e5)
    ai = dv.RangeCRS(ai.ncols, ai.entries, ai.rowmap, ai.nnz)
    print(f"spgemm tiers {tier}", flush=True)
    for k in (1, 4, 37):
        p = orc.power(ai, k)
        p = dv.RangeCRS(p.ncols, p.entries, p.rowmap, p.nnz)
        assert_identical(ctx.range_spgemm(p, p), orc.spgemm(p, p), f"P{k}.P{k} {tier}")
    rng = np.random.default_rng(3)
    da, db = random_dense(rng, 120, 900, 0.05, 30), random_dense(rng, 900, 2000, 0.02, 90)
    a, b = from_dense(da), from_dense(db)
    assert_identical(ctx.range_spgemm(a, b), orc.spgemm(a, b), f"random {tier}")


@pytest.mark.parametrize("size_bytes", [4, 8])
def test_spadd_identity_plus_a(ctx, graph_1e5, size_bytes):
    a = graph_1e5.as_size(size_bytes)
    i = ctx.identity(a.nrows, size_bytes)
    assert_identical(i, orc.identity(a.nrows), "identity")
    c = ctx.range_spadd(i, a)
    assert_identical(c, orc.spadd(i, a), "I + A")


def test_spadd_random_and_noncanonical(ctx):
    rng = np.random.default_rng(7)
    da, db = random_dense(rng, 200, 300, 0.1, 6), random_dense(rng, 200, 300, 0.1, 6)
    a, b = from_dense(da), from_dense(db)
    c = ctx.range_spadd(a, b)
    assert np.array_equal(to_dense(c), da | db)
    assert_identical(c, orc.spadd(a, b), "spadd random")
    # a non-canonical row: the tail after the first non-touching range is
    # copied verbatim (range_sparse.hpp:328-355)
    a = dv.RangeCRS(40, np.array([0, 1, 5, 6], np.uint32), np.array([0, 4], np.uint64))
    b = dv.RangeCRS(40, np.array([2, 3, 20, 30, 8, 9], np.uint32), np.array([0, 6], np.uint64))
    c = ctx.range_spadd(a, b)
    assert c.entries.tolist() == [0, 3, 5, 6, 20, 30, 8, 9]
    assert c.nnz == 19
    assert_identical(c, orc.spadd(a, b), "spadd non-canonical")


@pytest.mark.parametrize("k", [0, 1, 2, 5, 16, 150])
@pytest.mark.parametrize("plan_mode", [0, 1, 2])
def test_power(ctx, graph_1e5, k, plan_mode):
    ctx.set_tuning(plan_mode=plan_mode)
    print(f"power k={k} plan_mode={plan_mode}", flush=True)
    assert_identical(ctx.range_power(graph_1e5, k), orc.power(graph_1e5, k), f"A^{k}")


def test_power_noncanonical_input(ctx):
    """The first product of a chain is I.A when A is the caller's matrix, so
    a non-canonical input comes back canonical, as in the reference."""
    a = dv.RangeCRS(6, np.array([0, 1, 1, 3, 5, 5, 4, 4], np.uint32),
                    np.array([0, 4, 4, 6, 8, 8, 8], np.uint64))
    for k in (1, 3):
        assert_identical(ctx.range_power(a, k), orc.power(a, k), f"A^{k} non-canonical")


def test_distance_index_autotuned(ctx, graph_1e5):
    """The autotuner tries many launch shapes; every one must give the same
    index, and the cached winner is reused for the same shape."""
    want = orc.dindex(graph_1e5, 150, 450)
    ctx.set_tuning(autotune=1)
    print("dindex autotune 1e5-bp graph [150,450]", flush=True)
    for _ in range(2):
        assert_identical(ctx.create_distance_index(graph_1e5, 150, 450), want, "autotuned D")
    t = ctx.last_tuning()
    assert t["autotune"] == 1
    assert t["lanes_small"] in (8, 16, 32) and t["stage_entries"] in (4, 8, 16)


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("summary_span", [0, 2, 16, 64, 200])
def test_distance_index_block_summaries(ctx, summary_span, layout):
    """Block summaries on and off, on a graph whose bubbles make some 32-row
    blocks fail the single-interval test (they fall back to their ranges), in
    range-CRS and slot-form builds."""
    ctx.set_tuning(summary_span=summary_span, layout=layout)
    a = dv.generate_graph(3, backbone_bp=60000, seed=8, snp_rate=0.05, indel_rate=0.01,
                          sv_rate=2e-3, sv_min=10, sv_max=300)
    print(f"dindex summaries span={summary_span}", flush=True)
    for dlo, dup in [(0, 300), (40, 700)]:
        assert_identical(ctx.create_distance_index(a, dlo, dup), orc.dindex(a, dlo, dup),
                         f"D[{dlo},{dup}] summary_span={summary_span}")


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("dlo,dup", [(0, 2100), (900, 2100)])
def test_distance_index_wide_summary_blocks(ctx, dlo, dup, layout):
    """Windows wide enough that the last products summarise in blocks of 64,
    128 and 256 rows (a quarter to an eighth of the ranges' span)."""
    ctx.set_tuning(layout=layout)
    a = dv.generate_graph(3, backbone_bp=12000, seed=9, snp_rate=0.03, indel_rate=0.005,
                          sv_rate=1e-3, sv_min=10, sv_max=200)
    print(f"dindex wide summary blocks [{dlo},{dup}] {a.nrows} loci", flush=True)
    assert_identical(ctx.create_distance_index(a, dlo, dup), orc.dindex(a, dlo, dup),
                     f"D[{dlo},{dup}]")


def test_distance_index_many_scan_tiles(ctx):
    """A graph of over 8.4 million loci: the row-pointer scans have more than
    4096 tiles of 2048 rows and take the spine path."""
    a = dv.generate_graph(2, backbone_bp=9_000_000, seed=12, size_bytes=8)
    assert a.nrows > 4096 * 2048
    print(f"dindex many tiles: {a.nrows} loci", flush=True)
    assert_identical(ctx.create_distance_index(a, 5, 24), orc.dindex(a, 5, 24), "D[5,24] 9e6 bp")


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("groups_only", [False, True])
@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("dlo,dup", [(300, 600), (40, 300), (200, 210), (1, 8), (3, 4)])
def test_distance_index_schedules(ctx, schedule, dlo, dup, groups_only, layout):
    """The paired schedule (A (A + I))^m . rest, whose first product adds the
    identity inside the SpGEMM, and the reference's order A^d1 (A + I)^(d2 - d1)
    give the same index, in the lane tier and in the group tiers, in range-CRS
    and slot-form builds."""
    a = dv.generate_graph(2, backbone_bp=40000, seed=33, size_bytes=4)
    ctx.set_tuning(schedule=schedule, lane_slice_max=0 if groups_only else 4096, layout=layout)
    print(f"dindex schedule={schedule} [{dlo},{dup}] groups_only={groups_only}", flush=True)
    d, st = ctx.create_distance_index(a, dlo, dup, stats=True)
    assert_identical(d, orc.dindex(a, dlo, dup), f"D[{dlo},{dup}] schedule {schedule}")
    assert st["layout"] == (0 if layout == 2 else 1)
    if schedule == 0 and (dlo, dup) == (300, 600):
        # M = A (A + I), then M^300 along an addition chain no longer than
        # the binary method's 11 products
        assert st["n_spgemm"] <= 12 and st["n_spadd"] == 0


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("size_bytes", [4, 8])
@pytest.mark.parametrize("dlo,dup", [(150, 450), (0, 250), (7, 7), (1, 8), (5, 5), (0, 1)])
@pytest.mark.parametrize("plan_mode", [0, 1, 2])
def test_distance_index_graph_1e5(ctx, graph_1e5, size_bytes, dlo, dup, plan_mode, layout):
    """Planned builds (one readback), exact builds (a readback per op) and
    planned builds forced through the overflow fallback agree with the oracle,
    with range-CRS and slot-form products."""
    ctx.set_tuning(plan_mode=plan_mode, layout=layout)
    a = graph_1e5.as_size(size_bytes)
    print(f"dindex 1e5-bp graph [{dlo},{dup}] S={size_bytes} plan_mode={plan_mode}", flush=True)
    d, st = ctx.create_distance_index(a, dlo, dup, stats=True)
    assert_identical(d, orc.dindex(a, dlo, dup), f"D[{dlo},{dup}]")
    assert st["n_spadd"] <= 1 and st["n_spgemm"] > 0
    assert st["planned"] == (plan_mode == 0)
    assert st["retried"] == (plan_mode == 2)


def test_distance_index_layout_switch(ctx):
    """Automatic layout: a graph whose products have many rows of several
    ranges is built in slot form once, then in range CRS; both are exact."""
    a = dv.generate_graph(3, backbone_bp=30000, seed=5, snp_rate=0.08, indel_rate=0.02,
                          sv_rate=3e-3, sv_min=10, sv_max=200)
    ctx.set_tuning(layout=0)
    want = orc.dindex(a, 40, 160)
    d1, st1 = ctx.create_distance_index(a, 40, 160, stats=True)
    d2, st2 = ctx.create_distance_index(a, 40, 160, stats=True)
    assert_identical(d1, want, "first build")
    assert_identical(d2, want, "second build")
    assert st1["layout"] == 0 and st2["layout"] == 1


def test_distance_index_slot_side_overflow(ctx):
    """Slot-form products whose side arrays run out (a tiny prediction, as in
    plan_mode 2) fall back to the exact build."""
    ctx.set_tuning(layout=2, plan_mode=2)
    a = dv.generate_graph(1, backbone_bp=20000, seed=3)
    d, st = ctx.create_distance_index(a, 10, 60, stats=True)
    assert_identical(d, orc.dindex(a, 10, 60), "D[10,60]")
    assert st["retried"] == 1


def test_distance_index_bad_bounds(ctx, graph_1e5):
    for dlo, dup in [(5, 4), (0, 0)]:
        with pytest.raises(dv.DiVerGError) as e:
            ctx.create_distance_index(graph_1e5, dlo, dup)
        assert e.value.status == 1


def test_device_resident_chain(ctx, graph_1e5):
    da = ctx.to_device(graph_1e5)
    sq = ctx.range_spgemm(da, da, location="device")
    p4 = ctx.range_spgemm(sq, sq, location="device")
    assert_identical(p4.to_host(), orc.power(graph_1e5, 4), "A^4 on device")
    d = ctx.create_distance_index(da, 3, 9, location="device")
    assert_identical(d.to_host(), orc.dindex(graph_1e5, 3, 9), "D on device")


@pytest.mark.parametrize("layout", [0, 2])
def test_distance_index_graph_replays(ctx, graph_1e5, layout):
    """Slot-form builds with graphs=1: the products are captured into a
    hipGraph once per shape and replayed.  Several replays of one shape on a
    resident input, with eager builds of another shape, host-input builds
    (which capture again) and other device work between them; every result
    against the oracle."""
    ctx.set_tuning(graphs=1, layout=layout)
    print(f"dindex graph replays layout={layout}", flush=True)
    want = orc.dindex(graph_1e5, 150, 450)
    small = dv.generate_graph(3, backbone_bp=20000, seed=3, snp_rate=0.02, indel_rate=0.005)
    want_small = orc.dindex(small, 20, 90)
    da = ctx.to_device(graph_1e5)
    for rep in range(3):
        # the first build of the shape captures, the next ones replay
        for k in range(3):
            d = ctx.create_distance_index(da, 150, 450, location="device")
            assert_identical(d.to_host(), want, f"round {rep} build {k}")
            d.free()
        # between rounds: eager products and a build of another shape from a
        # host input (its own capture)
        sq = ctx.range_spgemm(da, da, location="device")
        sq.free()
        assert_identical(ctx.create_distance_index(small, 20, 90), want_small, f"other {rep}")
    ctx.set_tuning(graphs=0)
    assert_identical(ctx.create_distance_index(da, 150, 450), want, "eager after replays")


def test_query_batch(ctx, graph_1e5):
    d = orc.dindex(graph_1e5, 150, 450)
    d = dv.RangeCRS(d.ncols, d.entries, d.rowmap, d.nnz)
    rng = np.random.default_rng(11)
    n = 200000
    qi = rng.integers(0, d.nrows, n, dtype=np.uint32)
    near = rng.integers(0, 600, n, dtype=np.uint32)
    qj = np.where(rng.random(n) < 0.5, np.minimum(qi + near, d.ncols - 1),
                  rng.integers(0, d.ncols, n)).astype(np.uint32)
    got = ctx.query(d, qi, qj)
    exp, rc = orc.query(d, qi, qj)
    assert rc == 0
    assert np.array_equal(got, exp)
    assert 0 < got.sum() < n
    dd = ctx.to_device(d)
    assert np.array_equal(ctx.query(dd, qi, qj), exp)
    with pytest.raises(dv.DiVerGError) as e:
        ctx.query(d, np.array([d.nrows], np.uint32), np.array([0], np.uint32))
    assert e.value.status == 6


def test_malformed_inputs_rejected(ctx):
    bad = dv.RangeCRS(10, np.array([3, 12], np.uint32), np.array([0, 2], np.uint64))
    with pytest.raises(dv.DiVerGError):
        ctx.range_spgemm(bad, from_dense(np.ones((10, 10), bool)))
