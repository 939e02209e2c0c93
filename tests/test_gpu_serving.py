"""Serving loop on the GPU: continuous batching of queries over resident
indexes, builds admitted by HBM budget, answers bit-exact against the CPU
lookup, and every result the same alone and in the mix."""
import numpy as np
import pytest

import diverg_amd as dv
from oracle import oracle as orc
from tests.rcrs_util import assert_identical

pytestmark = pytest.mark.gpu


def as_rcrs(o):
    return dv.RangeCRS(o.ncols, o.entries, o.rowmap, o.nnz)


def pairs(rng, d, n, near=600):
    qi = rng.integers(0, d.nrows, n, dtype=np.uint32)
    qj = np.where(rng.random(n) < 0.5, np.minimum(qi.astype(np.int64) + rng.integers(0, near, n),
                                                  d.ncols - 1),
                  rng.integers(0, d.ncols, n)).astype(np.uint32)
    return qi, qj


@pytest.fixture(scope="module")
def indexes():
    a1 = dv.generate_graph(1)
    a2 = dv.generate_graph(3, backbone_bp=40000, seed=31, sv_rate=1e-3, sv_min=20, sv_max=300)
    return [as_rcrs(orc.dindex(a1, 150, 450)), as_rcrs(orc.dindex(a2, 0, 50))]


def test_server_requests(indexes):
    rng = np.random.default_rng(1)
    with dv.Server(hbm_budget=8 << 30) as srv:
        ids = [srv.add_index(d) for d in indexes]
        tickets = []
        for r in range(20):
            k = r % 2
            qi, qj = pairs(rng, indexes[k], 5000)
            tickets.append((srv.submit_queries(ids[k], qi, qj), k, qi, qj))
        for t, k, qi, qj in tickets:
            info = srv.wait(t)
            assert info["status"] == 0 and info["n"] == len(qi)
            assert info["done_us"] >= info["submit_us"]
            got = srv.results(t, len(qi))
            exp, rc = orc.query(indexes[k], qi, qj)
            assert np.array_equal(got, exp)
        with pytest.raises(dv.DiVerGError) as e:
            srv.submit_queries(99, qi, qj)
        assert e.value.status == 8
        # an out-of-range pair answers 2
        t = srv.submit_queries(ids[0], np.array([indexes[0].nrows], np.uint32),
                               np.array([0], np.uint32))
        srv.wait(t)
        assert srv.results(t, 1)[0] == 2


def test_server_replay_mix_equals_alone(ctx, indexes):
    rng = np.random.default_rng(2)
    n = 200000
    rate = 1.0e6  # offered queries per second
    arrival = np.cumsum(rng.exponential(1e6 / rate, n))
    qidx = rng.integers(0, 2, n, dtype=np.uint32)
    qi = np.zeros(n, np.uint32)
    qj = np.zeros(n, np.uint32)
    for k in range(2):
        sel = qidx == k
        qi[sel], qj[sel] = pairs(rng, indexes[k], int(sel.sum()))
    comps = [dv.generate_graph(1, backbone_bp=50000, seed=4),
             dv.generate_graph(3, backbone_bp=30000, seed=6, sv_rate=1e-3, sv_min=20, sv_max=200)]
    builds = [(comps[0], 15, 45, 20000.0), (comps[1], 0, 30, 60000.0)]
    with dv.Server(hbm_budget=16 << 30) as srv:
        ids = [srv.add_index(d) for d in indexes]
        assert ids == [0, 1]
        print(f"replay {n} queries at {rate:.0f}/s with {len(builds)} builds", flush=True)
        res = srv.replay(qidx, qi, qj, arrival, builds)
        # answers in the mix == CPU lookup
        for k in range(2):
            sel = qidx == k
            exp, rc = orc.query(indexes[k], qi[sel], qj[sel])
            assert np.array_equal(res["answers"][sel], exp)
        lat = res["latency_us"]
        assert np.all(lat > 0) and np.isfinite(lat).all()
        print(f"p50 {np.percentile(lat, 50):.1f} us  p99 {np.percentile(lat, 99):.1f} us  "
              f"elapsed {res['elapsed_us'] / 1e3:.1f} ms  builds {res['build_done_us']}", flush=True)
        # builds in the mix == builds alone == oracle
        for b, (a, dlo, dup, _) in enumerate(builds):
            assert res["build_index"][b] >= 2
            got = srv.get_index(int(res["build_index"][b]))
            assert_identical(got, orc.dindex(a, dlo, dup), f"build {b} in the mix")
            assert_identical(got, ctx.create_distance_index(a, dlo, dup), f"build {b} alone")
        # the same queries alone (one request per index) == in the mix
        for k in range(2):
            sel = qidx == k
            t = srv.submit_queries(ids[k], qi[sel], qj[sel])
            srv.wait(t)
            assert np.array_equal(srv.results(t, int(sel.sum())), res["answers"][sel])
        # the new indexes answer queries too
        nid = int(res["build_index"][0])
        d0 = orc.dindex(comps[0], 15, 45)
        q1, q2 = pairs(rng, d0, 10000)
        t = srv.submit_queries(nid, q1, q2)
        srv.wait(t)
        assert np.array_equal(srv.results(t, 10000), orc.query(d0, q1, q2)[0])


def test_server_budget_admission():
    a = dv.generate_graph(1)
    with dv.Server(hbm_budget=1 << 20) as srv:
        with pytest.raises(dv.DiVerGError) as e:
            srv.submit_build(a, 5, 10)
        assert e.value.status == 5
        small = dv.generate_graph(1, backbone_bp=500)
        t = srv.submit_build(small, 2, 6)
        info = srv.wait(t)
        assert info["status"] == 0
        assert_identical(srv.get_index(info["index_id"]), orc.dindex(small, 2, 6), "small build")
