"""GPU distance indexes of hand-made micro graphs (chains, a SNP bubble, a
deletion, self-loops, cycles, two components) against the committed
walk-length fixtures and the chain closed form, through every intermediate
layout and plan mode: the smallest shapes reach the kernels' edge cases
(one-block grids, single-row slices, empty rows, narrow products)."""
import json
import os

import numpy as np
import pytest

import diverg_amd as dv
from oracle import oracle as orc
from tests.rcrs_util import assert_identical, from_dense, to_dense, walk_index_dense

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def micro_graphs():
    with open(os.path.join(GOLD, "micro_graphs.json")) as f:
        return json.load(f)


def rcrs(j, size_bytes):
    m = dv.RangeCRS(j["ncols"], np.array(j["entries"], np.uint32),
                    np.array(j["rowmap"], np.uint64), j.get("nnz"))
    return m.as_size(size_bytes)


@pytest.fixture(autouse=True)
def default_tuning(ctx):
    t = ctx.tuning()
    yield
    ctx.set_tuning(**{k: v for k, v in t.items() if k != "autotune"}, autotune=t["autotune"])


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("plan_mode", [0, 1])
@pytest.mark.parametrize("size_bytes", [4, 8])
@pytest.mark.parametrize("name", ["chain12", "snp", "deletion", "self_loop",
                                  "two_components", "cycle"])
def test_micro_graph_indexes(ctx, name, size_bytes, plan_mode, layout):
    g = micro_graphs()[name]
    a = rcrs(g["A"], size_bytes)
    ctx.set_tuning(plan_mode=plan_mode, layout=layout)
    for case in g["cases"]:
        print(f"micro {name} [{case['dlo']},{case['dup']}] S={size_bytes} "
              f"plan_mode={plan_mode} layout={layout}", flush=True)
        got = ctx.create_distance_index(a, case["dlo"], case["dup"])
        assert_identical(got, rcrs(case["D"], size_bytes),
                         f"{name} [{case['dlo']},{case['dup']}]")


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("L,d1,d2", [(1, 0, 1), (2, 1, 1), (500, 13, 77), (700, 0, 300),
                                     (300, 299, 600)])
def test_chain_closed_form(ctx, L, d1, d2, layout):
    """On a chain of L loci, row i of D is [i + d1, min(i + d2, L - 1)] (empty
    when i + d1 > L - 1)."""
    a = from_dense(np.eye(L, k=1, dtype=bool))
    ctx.set_tuning(layout=layout)
    print(f"chain L={L} [{d1},{d2}] layout={layout}", flush=True)
    d = ctx.create_distance_index(a, d1, d2)
    for i in range(L):
        exp = [] if i + d1 > L - 1 else [i + d1, min(i + d2, L - 1)]
        assert d.entries[d.rowmap[i]:d.rowmap[i + 1]].tolist() == exp, f"row {i}"
    assert_identical(d, orc.dindex(a, d1, d2), f"chain {L} [{d1},{d2}]")


@pytest.mark.parametrize("layout", [1, 2])
def test_dense_walks_random_small(ctx, layout):
    """Random sparse digraphs of 1-40 vertices (self-loops and back edges
    allowed): D against the dense walk-length layering."""
    rng = np.random.default_rng(7)
    ctx.set_tuning(layout=layout)
    for trial in range(12):
        n = int(rng.integers(1, 41))
        dense = rng.random((n, n)) < min(1.0, 2.5 / n)
        a = from_dense(dense)
        dlo = int(rng.integers(0, 6))
        dup = dlo + int(rng.integers(0 if dlo else 1, 12))
        print(f"random n={n} [{dlo},{dup}] layout={layout}", flush=True)
        d = ctx.create_distance_index(a, dlo, dup)
        assert np.array_equal(to_dense(d), walk_index_dense(dense, dlo, dup)), (n, dlo, dup)
        assert_identical(d, orc.dindex(a, dlo, dup), f"random {trial}")
