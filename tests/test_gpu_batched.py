"""Batched builds (one planned build over a block-diagonal assembly of
components) and hipGraph replay of planned builds, on the GPU."""
import numpy as np
import pytest

import diverg_amd as dv
from oracle import oracle as orc
from tests.rcrs_util import assert_identical

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def default_tuning(ctx):
    t = ctx.tuning()
    yield
    ctx.set_tuning(**t)


def components():
    comps = [
        dv.generate_graph(1, backbone_bp=30000),
        dv.generate_graph(3, backbone_bp=20000, seed=77, sv_rate=2e-3, sv_min=20, sv_max=400),
        dv.generate_graph(2, backbone_bp=5000, size_bytes=4),
        dv.RangeCRS(1, np.zeros(0, np.uint32), np.array([0, 0], np.uint64)),  # one locus, no edge
        dv.generate_graph(1, backbone_bp=2, seed=5),
        dv.generate_graph(2, backbone_bp=60000, seed=9),
    ]
    return comps


@pytest.mark.parametrize("graphs", [0, 1])
@pytest.mark.parametrize("dlo,dup", [(30, 90), (0, 40)])
def test_batched_equals_plain_and_oracle(ctx, graphs, dlo, dup):
    ctx.set_tuning(graphs=graphs)
    comps = components()
    print(f"batched {len(comps)} components [{dlo},{dup}] graphs={graphs}", flush=True)
    outs, st = ctx.create_distance_index_batched(comps, dlo, dup, stats=True)
    assert len(outs) == len(comps)
    for c, (a, d) in enumerate(zip(comps, outs)):
        assert d.size_bytes == a.size_bytes
        assert_identical(d, ctx.create_distance_index(a, dlo, dup), f"component {c} vs plain")
        assert_identical(d, orc.dindex(a, dlo, dup), f"component {c} vs oracle")
    assert st["n_spgemm"] > 0


def test_batched_device_outputs(ctx):
    comps = components()[:3]
    dcomps = [ctx.to_device(a) for a in comps]
    outs = ctx.create_distance_index_batched(dcomps, 3, 11, location="device")
    for a, d in zip(comps, outs):
        assert_identical(d.to_host(), orc.dindex(a, 3, 11), "device-resident batched")


def test_batched_bad_bounds(ctx):
    with pytest.raises(dv.DiVerGError):
        ctx.create_distance_index_batched(components()[:2], 9, 3)


def test_graph_replays_match_eager(ctx):
    """Planned builds captured in a hipGraph: every replay equals the eager
    build, with eager work and synchronisations between the replays."""
    a = dv.generate_graph(1)
    da = ctx.to_device(a)
    ctx.set_tuning(graphs=0)
    eager = ctx.create_distance_index(da, 150, 450)
    expected = orc.dindex(a, 150, 450)
    assert_identical(eager, expected, "eager")
    ctx.set_tuning(graphs=1)
    for r in range(4):
        d = ctx.create_distance_index(da, 150, 450)
        assert_identical(d, expected, f"replay {r}")
        # unrelated eager work between replays
        ctx.set_tuning(graphs=0)
        assert_identical(ctx.create_distance_index(da, 2, 5), orc.dindex(a, 2, 5), "eager between")
        ctx.set_tuning(graphs=1)


def test_batched_size_classes_autotuned(ctx):
    """With the autotuner on, components are built per size class (below 2^20
    rows, below 2^24, larger), each class one block-diagonal build with its
    own tuned launch shape; every component still equals the oracle."""
    comps = [dv.generate_graph(2, backbone_bp=30000, seed=41),
             dv.generate_graph(2, backbone_bp=1_200_000, seed=42),
             dv.generate_graph(1, backbone_bp=8000, seed=43),
             dv.generate_graph(2, backbone_bp=1_100_000, seed=44, size_bytes=4)]
    assert sum(a.nrows >= 1 << 20 for a in comps) == 2
    ctx.set_tuning(autotune=1)
    print(f"batched size classes {[a.nrows for a in comps]}", flush=True)
    for rep in range(2):
        outs, st = ctx.create_distance_index_batched(comps, 5, 20, stats=True)
        assert st["builds"] == 2
        for c, (a, d) in enumerate(zip(comps, outs)):
            assert_identical(d, orc.dindex(a, 5, 20), f"component {c} rep {rep}")
