"""Benchmark: distance-index build on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (synthetic, seeded): the 5e6 bp variation graph (SNP 0.01/bp, indel
0.001/bp, geometric length mean 4 capped at 50), uint32 range-CRS, distance
index for [d1, d2] = [300, 600].  A step is one full index build
D = A^300 . (A + I)^300 with A already resident in HBM; the output stays in
HBM.  `value` is loci indexed per second (rows of D per second of build).

Further legs (fields of the same JSON line; --legs selects them):
  chromosome    the 5e7 bp graph with structural variants (uint64 rowmap,
                [200, 1000]): build time, peak HBM, rows checked against
                the oracle's independent walk-length BFS
  queries       batched locus-pair queries on the HBM-resident chromosome
                index (pairs drawn as in the serving workload, batches of
                65536 resident in HBM): queries/s, per-batch p50/p99 latency,
                the CPU lookup's rate on the same pairs, answers bit-exact
  serving       the chromosome index (and the 5e6 bp one) resident in HBM, a
                Poisson stream of locus-pair queries (mean 2e6/s offered;
                sources uniform, half the targets 0-1000 bp downstream on the
                backbone) with component builds of 1e5-2e7 bp interleaved:
                queries/s, p50/p99 latency, build completion times, answers
                checked against the CPU lookup, builds in the mix against
                plain builds; then the same queries all at once (saturation)
                through the server and through plain dv_query_batch calls
  wide          an index of about 2.1 ranges per row (the figure reported for
                real graphs; the plain workload's index has 1.003): 1.6e6 bp
                backbone, SNP 0.1/bp, indel 0.05/bp, SV 3e-3/bp (half
                inversions), [300, 600]; build time in the default layout
                (slots), in range CRS and autotuned, each against the oracle
  batched       24 components (log-normal lengths, median 5e6 bp) built in
                one planned build over their block-diagonal assembly vs 24
                plain builds; bit-exact per component

The JSON line also carries:
  roofline      the SpGEMM lane tier, the dominant work of a build: per
                product, its lane-per-row launch (slot form: k_slot_fast for
                products read through chains or of narrow slices, else
                k_slot_lane; range CRS: k_spgemm_lane), timed with HIP events
                on the library stream;
                algorithmic bytes bytes(X) + bytes(Y) + bytes(C) of the
                product over that time, against 8 TB/s
  cpu_baseline  the CPU oracle (the reference's OpenMP algorithm restated in
                C) building the same index on the host: one thread per
                physical core of the affinity mask within the job's CPU share,
                OMP_PROC_BIND=close, OMP_PLACES=cores, 1 warm-up and the median
                of 3, plus one 1-thread build
  bit_exact     the GPU index equals the oracle's (rowmap, entries, nnz)
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = ("index build s + achieved HBM GB/s; batched queries/s with p50/p99, "
          "bit-exact")
HBM_PEAK_GBS = 8000.0


# The process's CPU mask at start.  With OMP_PROC_BIND set, the OpenMP runtime
# binds the thread that loads it to the first place (one core) when the
# oracle library loads; threads started later from that thread (the serving
# loop, its build thread) would inherit that one core.
START_CPUS = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None


def unbind_main_thread():
    """Give the calling thread the process's starting CPU mask again."""
    if START_CPUS is not None:
        os.sched_setaffinity(0, START_CPUS)


def hip_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def measured_traffic(args, kernel):
    """HBM bytes per lane-tier kernel launch from the committed rocprofv3
    counter runs (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy's readings),
    for the workload and kernel they were taken on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    w = t.get("workload", {})
    if (w.get("preset"), w.get("d1"), w.get("d2")) != (args.preset, args.dlo, args.dup):
        return None
    if t.get("kernel") != kernel:
        return None
    return t.get("traffic_bytes_per_launch")


def host_cores():
    """The host's cores as the CPU baseline counts them: logical CPUs in this
    process's affinity mask (nproc), the physical cores they belong to
    (sysfs topology), lscpu's summary, and the CPU share the machine grants
    the job (OMP_NUM_THREADS when the environment sets it)."""
    import subprocess
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology"
        try:
            with open(base + "/physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "/core_id") as f:
                core = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("cpu", c))
    lscpu = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)",
                             "Model name"):
                lscpu[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    share = os.environ.get("OMP_NUM_THREADS", "")
    return {"nproc": len(cpus), "physical_cores_in_mask": len(cores), "lscpu": lscpu,
            "cpu_share": int(share) if share.isdigit() and int(share) > 0 else None}


def cpu_threads(cores=None):
    """Threads of the CPU baseline: one per physical core in the affinity
    mask, within the CPU share granted to the job."""
    c = cores or host_cores()
    n = c["physical_cores_in_mask"]
    if c["cpu_share"]:
        n = min(n, c["cpu_share"])
    return max(1, n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--dlo", type=int, default=300)
    ap.add_argument("--dup", type=int, default=600)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--legs", default="plain,chromosome,queries,serving,wide,batched",
                    help="comma-separated legs to run (plain is always run)")
    ap.add_argument("--full-check", action="store_true",
                    help="compare the whole chromosome index and every batched component with "
                         "the oracle (minutes of CPU time); by default rows of the chromosome "
                         "index are checked against the oracle's walk-length BFS and the "
                         "batched components below 1e6 loci against the oracle")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-rate", type=float, default=2.0e6)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy kernel three times "
                         "(the reference for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        x = torch.arange(1 << 28, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": 4 << 28, "dispatches": 3}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(components shard across GPUs with no exchange; see the design notes). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline's OpenMP threads: one per core, packed (read by the
    # OpenMP runtime when the oracle library loads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")

    import diverg_amd as dv

    t0 = time.time()
    a = dv.generate_graph(args.preset, size_bytes=4)
    gen_s = time.time() - t0
    print(f"# graph: {a.nrows} loci, {len(a.entries) // 2} ranges, generated in {gen_s:.2f}s",
          file=sys.stderr, flush=True)

    ctx = dv.Context(0)
    da = ctx.to_device(a)  # resident before timing
    for w in range(args.warmup):
        # the last warmup build is profiled, so that the timing events exist
        # before the timed region
        ctx.profile(w + 1 == args.warmup)
        d = ctx.create_distance_index(da, args.dlo, args.dup, location="device")
        d.free()
        print(f"# warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    ctx.profile(False)

    # Kernel timing (two HIP events per product, on the build's stream) holds
    # the queue for a few microseconds per event, so only the last fifth of
    # the timed builds (at least one) carry it; the others run as a caller's
    # would.  Every timed build does the same work.
    nprof = max(1, args.steps // 5)
    stats = []
    hip_sync()
    t_start = time.perf_counter()
    for s in range(args.steps):
        if s == args.steps - nprof:
            ctx.profile(True)
        d, st = ctx.create_distance_index(da, args.dlo, args.dup, location="device",
                                          stats=True)
        stats.append(st)
        if s + 1 < args.steps:
            d.free()
    hip_sync()
    elapsed = time.perf_counter() - t_start
    prof = ctx.profile_read()
    ctx.profile(False)
    print(f"# timed {args.steps} builds in {elapsed:.3f}s", file=sys.stderr, flush=True)

    ms_per_step = 1e3 * elapsed / args.steps
    build_ms = float(np.median([s["build_ms"] for s in stats]))
    value = a.nrows * args.steps / elapsed

    # one "launch" is one product's lane tier (a kernel that depends on the
    # product's form)
    lane_kernel = "slot_lane_tier" if stats[-1]["layout"] == 0 else "k_spgemm_lane"
    cnt = prof["spgemm_lane"]
    mean_ms = cnt["total_ms"] / max(1, cnt["launches"])
    alg_per_launch = cnt["alg_bytes"] / max(1, cnt["launches"])
    achieved = alg_per_launch / (mean_ms * 1e-3) / 1e9 if mean_ms > 0 else 0.0
    slice_gbs = (cnt["slice_bytes"] / max(1, cnt["launches"])) / (mean_ms * 1e-3) / 1e9 \
        if mean_ms > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": lane_kernel,
        "kernels": (["k_slot_fast", "k_slot_lane"]
                    if lane_kernel == "slot_lane_tier" else ["k_spgemm_lane"]),
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": measured_traffic(args, lane_kernel),
        "launches": cnt["launches"],
        "mean_launch_ms": round(mean_ms, 4),
        "alg_bytes_per_launch": int(alg_per_launch),
        "timed_builds_profiled": nprof,
        "gathered_slice_GBps": round(slice_gbs, 1),
    }

    result_host = d.to_host()
    d.free()

    # A^d1 on its own (the reference's intermediate rA^dlo, dindex.hpp:151-157)
    pw = ctx.range_power(da, args.dlo, location="device")
    power_host = pw.to_host()
    pw.free()

    cpu = None
    bit_exact = None
    power_exact = None
    if not args.no_cpu:
        from oracle import oracle as orc
        cores = host_cores()
        th = cpu_threads(cores)
        orc.set_threads(th)
        # one warm-up build, then the median of three, all on `th` threads
        runs = []
        e = None
        for r in range(4):
            t = time.perf_counter()
            e = orc.dindex(a, args.dlo, args.dup)
            if r:
                runs.append(time.perf_counter() - t)
            print(f"# cpu oracle build {r + 1}/4 ({'warm-up' if r == 0 else 'timed'}) on "
                  f"{th} threads", file=sys.stderr, flush=True)
        cpu_s = float(np.median(runs))
        bit_exact = bool(result_host.nnz == e.nnz % (1 << 32)
                         and np.array_equal(result_host.rowmap.astype(np.uint64), e.rowmap)
                         and np.array_equal(result_host.entries, e.entries))
        del e
        ep = orc.power(a, args.dlo)
        power_exact = bool(power_host.nnz == ep.nnz % (1 << 32)
                           and np.array_equal(power_host.rowmap.astype(np.uint64), ep.rowmap)
                           and np.array_equal(power_host.entries, ep.entries))
        del ep
        # one thread: a single timed build
        orc.set_threads(1)
        t = time.perf_counter()
        orc.dindex(a, args.dlo, args.dup)
        one_s = time.perf_counter() - t
        orc.set_threads(th)
        cpu = {
            "value": round(a.nrows / cpu_s, 1),
            "unit": "loci/s",
            "cores": th,
            "kind": "port",
            "sample": (f"the full workload: builds of the same {a.nrows}-locus index "
                       f"[{args.dlo},{args.dup}]; 1 warm-up, median of 3 on {th} threads "
                       f"(OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                       f"OMP_PLACES={os.environ.get('OMP_PLACES')}), and one 1-thread build"),
            "build_s": round(cpu_s, 3),
            "runs_s": [round(x, 3) for x in runs],
            "threads": th,
            "one_thread_s": round(one_s, 3),
            "one_thread_loci_per_s": round(a.nrows / one_s, 1),
            "host": cores,
        }
        print(f"# cpu oracle: median {cpu_s:.2f}s on {th} threads, 1 thread {one_s:.2f}s, "
              f"bit_exact={bit_exact}, A^{args.dlo} bit_exact={power_exact}",
              file=sys.stderr, flush=True)
        unbind_main_thread()

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "loci/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"distance-index build, synthetic variation graph preset {args.preset}",
            "backbone_bp": int(dv.graph_preset(args.preset).backbone_bp),
            "loci": a.nrows,
            "ranges_A": len(a.entries) // 2,
            "d1": args.dlo,
            "d2": args.dup,
            "size_type": "uint32",
            "index_ranges": len(result_host.entries) // 2,
            "index_nnz": int(result_host.nnz),
        },
        "build_s": round(build_ms / 1e3, 5),
        "alg_GBps_build": round(stats[-1]["alg_bytes"] / (build_ms * 1e-3) / 1e9, 2),
        "spgemms_per_build": stats[-1]["n_spgemm"],
        "peak_device_bytes": stats[-1]["peak_device_bytes"],
        "bit_exact": bit_exact,
        "a_d1_bit_exact": power_exact,
        "a_d1_ranges": len(power_host.entries) // 2,
        "roofline": roofline,
        "cpu_baseline": cpu,
        # per build, over the profiled builds
        "kernels_ms": {k: round(v["total_ms"] / nprof, 4) for k, v in prof.items()},
    }
    legs = set(args.legs.split(","))
    if "chromosome" in legs or "serving" in legs or "queries" in legs:
        out["chromosome"], chrom = leg_chromosome(ctx, args.full_check)
        if "queries" in legs:
            out["queries"] = leg_queries(ctx, chrom)
        if "serving" in legs:
            out["serving"] = leg_serving(chrom, result_host, args)
        del chrom
    if "wide" in legs:
        out["wide"] = leg_wide(ctx, args.no_cpu)
    if "batched" in legs:
        out["batched"] = leg_batched(ctx, args.full_check)
    ctx.close()
    print(json.dumps(out), flush=True)


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def oracle_dindex_logged(a, dlo, dup, what):
    """orc.dindex(a, dlo, dup) composed from the oracle's own operations in
    the same order (dindex.hpp:46-189: rI + rA, rA^dlo and (rA + rI)^(dup -
    dlo) by range_power_inplace, range_sparse.hpp:2376-2431, then one
    product), with a line after each product, for builds that run minutes."""
    from oracle import oracle as orc
    t0 = time.time()
    done = [0]

    def mul(x, y):
        z = orc.spgemm(x, y)
        done[0] += 1
        log(f"{what}: oracle product {done[0]} done ({time.time() - t0:.0f}s)")
        return z

    def power(p, k):
        c = orc.identity(p.ncols)
        while True:
            if k & 1:
                c = mul(c, p)
            k >>= 1
            if k == 0:
                return c
            p = mul(p, p)

    rai = orc.spadd(orc.identity(a.nrows), a)
    if dlo == 0:
        return power(rai, dup)
    rad = power(a, dlo)
    raid = power(rai, dup - dlo)
    del rai
    return mul(rad, raid)


def leg_chromosome(ctx, full_check=False):
    """5e7 bp graph with SVs, uint64 rowmap, [200, 1000]."""
    import diverg_amd as dv
    from oracle import oracle as orc
    t = time.time()
    a, vpos = dv.generate_graph(3, size_bytes=8, with_vpos=True)
    log(f"chromosome graph: {a.nrows} loci, {len(a.entries) // 2} ranges ({time.time() - t:.1f}s)")
    da = ctx.to_device(a)
    d = ctx.create_distance_index(da, 200, 1000, location="device")  # warm
    d.free()
    d, st = ctx.create_distance_index(da, 200, 1000, location="device", stats=True)
    dh = d.to_host()
    d.free()
    del da
    # sampled rows against the oracle's independent walk-length BFS; with
    # full_check the whole index against the oracle
    orc.set_threads(cpu_threads())
    full_ok, oracle_s = None, None
    if full_check:
        t = time.perf_counter()
        e = oracle_dindex_logged(a, 200, 1000, "chromosome")
        oracle_s = time.perf_counter() - t
        full_ok = bool(dh.nnz == e.nnz and np.array_equal(dh.rowmap, e.rowmap)
                       and np.array_equal(dh.entries, e.entries))
        del e
    rng = np.random.default_rng(0xD1E5)
    rows = np.sort(rng.choice(a.nrows, 1500, replace=False)).astype(np.uint32)
    w = orc.walk_rows(a, rows, 200, 1000)
    ok = all(np.array_equal(dh.entries[dh.rowmap[r]:dh.rowmap[r + 1]],
                            w.entries[w.rowmap[k]:w.rowmap[k + 1]]) for k, r in enumerate(rows))
    res = {
        "workload": "5e7 bp backbone, SNP 0.01, indel 0.001, SV 1e-5 (30% inversions), "
                    "uint64 rowmap, [200,1000]",
        "loci": a.nrows,
        "build_s": round(st["build_ms"] / 1e3, 5),
        "loci_per_s": round(a.nrows / (st["build_ms"] / 1e3), 1),
        "spgemms": st["n_spgemm"],
        "alg_GBps": round(st["alg_bytes"] / (st["build_ms"] * 1e-3) / 1e9, 2),
        "peak_hbm_bytes": st["peak_device_bytes"],
        "index_ranges": len(dh.entries) // 2,
        "index_nnz": dh.nnz,
        "bit_exact_vs_oracle": full_ok,
        "oracle_s": round(oracle_s, 2) if oracle_s else None,
        "rows_checked_vs_walk_bfs": len(rows),
        "rows_bit_exact": bool(ok),
    }
    log(f"chromosome build {res['build_s']}s, whole index vs oracle {full_ok}, "
        f"BFS rows bit-exact {ok}")
    return res, (a, vpos, dh)


class DeviceBuffer:
    """Raw device memory through the HIP runtime (for query arrays)."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        if self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(1, nbytes))) != 0:
            raise RuntimeError("hipMalloc failed")

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        if self.hip.hipMemcpy(self.ptr, arr.ctypes.data_as(ctypes.c_void_p),
                              ctypes.c_size_t(arr.nbytes), 1) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def download(self, arr):
        if self.hip.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), self.ptr,
                              ctypes.c_size_t(arr.nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy D2H failed")

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def leg_queries(ctx, chrom, batch=65536, nbatch=256):
    """Batched locus-pair queries on the HBM-resident chromosome index: pairs
    drawn as the serving workload draws them (sources uniform; half the
    targets 0-1000 bp downstream on the backbone, half uniform), in batches
    already resident in HBM; per-batch latency p50/p99 and throughput;
    every answer checked against the CPU lookup (crs_matrix.hpp:149-169)."""
    from oracle import oracle as orc
    a, vpos, dh = chrom
    rng = np.random.default_rng(0xD1E5A)
    n = batch * nbatch
    src = rng.integers(0, a.nrows, n, dtype=np.int64)
    down = np.minimum(vpos[src].astype(np.int64) + rng.integers(0, 1001, n), int(vpos[-1]))
    near_t = np.searchsorted(vpos, down, side="left").astype(np.int64)
    qi = src.astype(np.uint32)
    qj = np.where(rng.random(n) < 0.5, near_t, rng.integers(0, a.nrows, n)).astype(np.uint32)
    d = ctx.to_device(dh)
    bi, bj, bo = DeviceBuffer(4 * n), DeviceBuffer(4 * n), DeviceBuffer(n)
    bi.upload(qi)
    bj.upload(qj)
    ctx.query_device(d, bi.ptr.value, bj.ptr.value, bo.ptr.value, batch)  # warm
    lat = np.zeros(nbatch)
    hip_sync()
    t0 = time.perf_counter()
    for b in range(nbatch):
        t = time.perf_counter()
        ctx.query_device(d, bi.ptr.value + 4 * b * batch, bj.ptr.value + 4 * b * batch,
                         bo.ptr.value + b * batch, batch)
        lat[b] = time.perf_counter() - t
    span = time.perf_counter() - t0
    # one launch over every pair: the device's own query rate
    hip_sync()
    t = time.perf_counter()
    ctx.query_device(d, bi.ptr.value, bj.ptr.value, bo.ptr.value, n)
    big_s = time.perf_counter() - t
    got = np.zeros(n, np.uint8)
    bo.download(got)
    for x in (bi, bj, bo):
        x.free()
    d.free()
    orc.set_threads(cpu_threads())
    t = time.perf_counter()
    want, rc = orc.query(dh, qi, qj)
    cpu_s = time.perf_counter() - t
    res = {
        "workload": "chromosome index resident in HBM; pairs as the serving workload (sources "
                    "uniform, half the targets 0-1000 bp downstream on the backbone, half "
                    f"uniform); {nbatch} batches of {batch} pairs resident in HBM",
        "queries": n,
        "batched_queries_per_s": round(n / span, 1),
        "batch_latency_p50_us": round(float(np.percentile(lat, 50)) * 1e6, 2),
        "batch_latency_p99_us": round(float(np.percentile(lat, 99)) * 1e6, 2),
        "one_launch_queries_per_s": round(n / big_s, 1),
        "cpu_queries_per_s": round(n / cpu_s, 1),
        "cpu_threads": cpu_threads(),
        "hits": int(got.sum()),
        "bit_exact": bool(rc == 0 and np.array_equal(got, want)),
    }
    log(f"queries: {res['batched_queries_per_s']:.3g}/s in batches, p50 "
        f"{res['batch_latency_p50_us']} us p99 {res['batch_latency_p99_us']} us, "
        f"one launch {res['one_launch_queries_per_s']:.3g}/s, exact {res['bit_exact']}")
    return res


def leg_serving(chrom, idx2, args):
    """Continuous batching of a Poisson query stream with interleaved builds."""
    import diverg_amd as dv
    from oracle import oracle as orc
    a, vpos, dh = chrom
    indexes = [dh, idx2]
    rng = np.random.default_rng(0xD1E55)
    n = int(args.serve_rate * args.serve_seconds)
    arrival = np.cumsum(rng.exponential(1e6 / args.serve_rate, n))
    qidx = (rng.random(n) < 0.1).astype(np.uint32)  # 90% chromosome index, 10% the 5e6 bp one
    qi = np.zeros(n, np.uint32)
    qj = np.zeros(n, np.uint32)
    # chromosome index: targets 0-1000 bp downstream on the backbone for half
    sel = qidx == 0
    m = int(sel.sum())
    src = rng.integers(0, a.nrows, m, dtype=np.int64)
    down = vpos[src].astype(np.int64) + rng.integers(0, 1001, m)
    down = np.minimum(down, int(vpos[-1]))
    near_t = np.searchsorted(vpos, down, side="left").astype(np.int64)
    uni_t = rng.integers(0, a.nrows, m, dtype=np.int64)
    qi[sel] = src
    qj[sel] = np.where(rng.random(m) < 0.5, near_t, uni_t)
    sel2 = ~sel
    m2 = int(sel2.sum())
    s2 = rng.integers(0, idx2.nrows, m2, dtype=np.int64)
    qi[sel2] = s2
    qj[sel2] = np.where(rng.random(m2) < 0.5, np.minimum(s2 + rng.integers(0, 1001, m2),
                                                         idx2.ncols - 1),
                        rng.integers(0, idx2.ncols, m2))
    # builds of components of 1e5 - 2e7 bp, log-uniform, spread over the stream
    nb = 4
    lens = np.exp(rng.uniform(np.log(1e5), np.log(2e7), nb)).astype(np.int64)
    builds = []
    for b in range(nb):
        comp = dv.generate_graph(2, backbone_bp=int(lens[b]), seed=0xD1E55000 + b, size_bytes=8)
        builds.append((comp, 200, 1000, float((b + 0.5) * args.serve_seconds * 1e6 / nb)))
    log(f"serving: {n} queries over {args.serve_seconds}s, builds of {lens.tolist()} bp")
    # the server's threads start from this one: all of the process's CPUs
    unbind_main_thread()
    host_cpus = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    with dv.Server() as srv:
        ids = [srv.add_index(x) for x in indexes]
        res = srv.replay(qidx, qi, qj, arrival, builds)
        built = [srv.get_index(int(i)) for i in res["build_index"]]
    ok = True
    exps = []
    for k in range(2):
        s = qidx == k
        exp, rc = orc.query(indexes[k], qi[s], qj[s])
        exps.append(exp)
        ok = ok and rc == 0 and np.array_equal(res["answers"][s], exp)
    lat = res["latency_us"]
    done = arrival + lat
    span = (done.max() - arrival.min()) / 1e6
    # saturation: every query at once, through the server (continuous
    # batching, batches overlapping on the device) and through plain
    # dv_query_batch calls of the same batch size, one index at a time
    with dv.Server() as srv:
        for x in indexes:
            srv.add_index(x)
        sat = srv.replay(qidx, qi, qj, np.zeros(n), ())
    sat_ok = all(np.array_equal(sat["answers"][qidx == k], exps[k]) for k in range(2))
    sat_qps = n / (sat["elapsed_us"] / 1e6)
    with dv.Context(0) as c3:
        dix = [c3.to_device(x) for x in indexes]
        plain_ans = np.zeros(n, np.uint8)
        batch = 65536
        order = [np.flatnonzero(qidx == k) for k in range(2)]
        hip_sync()
        t = time.perf_counter()
        for k in range(2):
            sel = order[k]
            for b0 in range(0, len(sel), batch):
                part = sel[b0:b0 + batch]
                plain_ans[part] = c3.query(dix[k], qi[part], qj[part])
        plain_s = time.perf_counter() - t
        for x in dix:
            x.free()
    plain_ok = all(np.array_equal(plain_ans[qidx == k], exps[k]) for k in range(2))
    # builds in the mix against plain builds of the same components
    bok = True
    with dv.Context(0) as c2:
        for (comp, dlo, dup, _), got in zip(builds, built):
            ref = c2.create_distance_index(comp, dlo, dup)
            bok = bok and dv.identical(got, ref)
    out = {
        "workload": "indexes: chromosome graph [200,1000] + 5e6 bp graph [300,600]; Poisson "
                    f"stream {args.serve_rate:.0f} req/s offered for {args.serve_seconds}s; "
                    "builds of 1e5-2e7 bp components interleaved",
        "queries": n,
        "offered_qps": args.serve_rate,
        "queries_per_s": round(n / span, 1),
        "p50_us": round(float(np.percentile(lat, 50)), 2),
        "p99_us": round(float(np.percentile(lat, 99)), 2),
        "max_us": round(float(lat.max()), 2),
        "host_cpus_of_server_threads": host_cpus,
        "builds_bp": lens.tolist(),
        "build_completion_s": [round(x / 1e6, 4) for x in res["build_done_us"].tolist()],
        "answers_bit_exact_vs_cpu_lookup": bool(ok),
        "builds_in_mix_equal_alone": bool(bok),
        "saturation": {
            "workload": f"the same {n} queries all arriving at once, host arrays in and out",
            "served_queries_per_s": round(sat_qps, 1),
            "served_p50_us": round(float(np.percentile(sat["latency_us"], 50)), 2),
            "served_p99_us": round(float(np.percentile(sat["latency_us"], 99)), 2),
            "plain_batch_calls_queries_per_s": round(n / plain_s, 1),
            "served_over_plain": round(sat_qps * plain_s / n, 3),
            "bit_exact": bool(sat_ok and plain_ok),
        },
    }
    log(f"serving: {out['queries_per_s']:.0f} q/s, p50 {out['p50_us']} us, "
        f"p99 {out['p99_us']} us, exact {ok}, builds {out['build_completion_s']}; "
        f"saturated {sat_qps:.3g} q/s served vs {n / plain_s:.3g} q/s plain calls")
    return out


def leg_wide(ctx, no_cpu=False, steps=5):
    """The index shape of real graphs (about 2.1 ranges per row): build time
    in the slot layout (the default), in range CRS, and autotuned, each the
    median of `steps` timed builds after two warm-ups (the autotuned one's
    first warm-up tunes), every output against the oracle."""
    import diverg_amd as dv
    from oracle import oracle as orc
    params = dict(backbone_bp=1_600_000, snp_rate=0.1, indel_rate=0.05, sv_rate=3e-3,
                  sv_inversion_frac=0.5, sv_deletion_frac=0.3)
    a = dv.generate_graph(2, size_bytes=4, **params)
    log(f"wide: {a.nrows} loci, {len(a.entries) // 2} ranges")
    da = ctx.to_device(a)
    user = ctx.tuning()["layout"]
    res = {"workload": "1.6e6 bp backbone, SNP 0.1, indel 0.05, SV 3e-3 (50% inversions, "
                       "30% deletions), uint32 rowmap, [300,600]",
           "loci": a.nrows, "ranges_A": len(a.entries) // 2}
    outs = {}
    user_autotune = ctx.tuning()["autotune"]
    for name, layout, tune in (("slots", 0, 0), ("range_crs", 1, 0), ("autotuned", 0, 1)):
        # autotuned: the first build of the shape times candidate launch
        # shapes (untimed here, in the warm-up), later builds reuse the winner
        ctx.set_tuning(layout=layout, autotune=tune)
        first_ms = None
        for _ in range(2):  # warm-up: the first builds learn capacity predictions
            d, st = ctx.create_distance_index(da, 300, 600, location="device", stats=True)
            first_ms = st["build_ms"] if first_ms is None else first_ms
            d.free()
        ms, st, retried = [], None, 0
        for _ in range(steps):
            d, st = ctx.create_distance_index(da, 300, 600, location="device", stats=True)
            ms.append(st["build_ms"])
            retried += st["retried"]
            if len(ms) < steps:
                d.free()
        outs[name] = d.to_host()
        d.free()
        build_ms = float(np.median(ms))
        res[name] = {"build_s": round(build_ms / 1e3, 5),
                     # the shape's first build (autotuned: the tuning builds included)
                     "first_build_s": round(first_ms / 1e3, 5),
                     "loci_per_s": round(a.nrows / (build_ms / 1e3), 1),
                     "layout_used": st["layout"], "timed_builds_retried": retried,
                     "spgemms": st["n_spgemm"],
                     "alg_GBps": round(st["alg_bytes"] / (build_ms * 1e-3) / 1e9, 2)}
        if tune:
            lt = ctx.last_tuning()
            res[name]["tuning"] = {k: lt[k] for k in ("lds_words", "stage_entries",
                                                      "rows_per_group", "lanes_small",
                                                      "small_slice_max", "lane_slice_max",
                                                      "summary_span", "chain_span")}
    ctx.set_tuning(layout=user, autotune=user_autotune)
    del da
    dh = outs["slots"]
    res["index_ranges"] = len(dh.entries) // 2
    res["ranges_per_row"] = round(len(dh.entries) / 2 / max(1, dh.nrows), 3)
    res["index_nnz"] = dh.nnz
    ok = {n: None for n in outs}
    if not no_cpu:
        orc.set_threads(cpu_threads())
        t = time.perf_counter()
        e = orc.dindex(a, 300, 600)
        res["oracle_s"] = round(time.perf_counter() - t, 2)
        res["oracle_threads"] = cpu_threads()
        for n, o in outs.items():
            ok[n] = bool(o.nnz == e.nnz and np.array_equal(o.rowmap, e.rowmap)
                         and np.array_equal(o.entries, e.entries))
    res["bit_exact"] = ok
    log(f"wide: {res['ranges_per_row']} ranges per row; slots {res['slots']['build_s']}s "
        f"(layout {res['slots']['layout_used']}, retried {res['slots']['timed_builds_retried']}), "
        f"range CRS {res['range_crs']['build_s']}s, autotuned {res['autotuned']['build_s']}s; "
        f"bit-exact {ok}")
    return res


def leg_batched(ctx, full_check=False):
    """24 components in one planned build vs 24 plain builds."""
    import diverg_amd as dv
    t = time.time()
    comps = [dv.generate_graph(4, component=c, size_bytes=8) for c in range(24)]
    total = sum(c.nrows for c in comps)
    log(f"batched: 24 components, {total} loci ({time.time() - t:.1f}s to generate)")
    dcomps = [ctx.to_device(c) for c in comps]
    ctx.set_tuning(graphs=1)
    outs = ctx.create_distance_index_batched(dcomps, 300, 600, location="device")  # warm
    for o in outs:
        o.free()
    hip_sync()
    t = time.perf_counter()
    outs, st = ctx.create_distance_index_batched(dcomps, 300, 600, location="device", stats=True)
    hip_sync()
    batched_s = time.perf_counter() - t
    # per size class, each class with its autotuned launch shape (the first
    # call tunes, untimed)
    ctx.set_tuning(autotune=1)
    t = time.perf_counter()
    for o in ctx.create_distance_index_batched(dcomps, 300, 600, location="device"):
        o.free()
    hip_sync()
    tune_s = time.perf_counter() - t
    log(f"batched: per-class autotuning took {tune_s:.1f}s")
    t = time.perf_counter()
    touts, tst = ctx.create_distance_index_batched(dcomps, 300, 600, location="device",
                                                   stats=True)
    hip_sync()
    tuned_s = time.perf_counter() - t
    ctx.set_tuning(autotune=0)
    ctx.set_tuning(graphs=0)
    hip_sync()
    t = time.perf_counter()
    plain = []
    for dc in dcomps:
        plain.append(ctx.create_distance_index(dc, 300, 600, location="device"))
    hip_sync()
    plain_s = time.perf_counter() - t
    from oracle import oracle as orc
    orc.set_threads(cpu_threads())
    exact = oracle_exact = True
    checked = 0
    t = time.perf_counter()
    for c, (comp, o, to, p) in enumerate(zip(comps, outs, touts, plain)):
        got = o.to_host()
        exact = exact and dv.identical(got, p.to_host()) and dv.identical(got, to.to_host())
        if full_check or comp.nrows < 1_000_000:
            oracle_exact = oracle_exact and dv.identical(got, orc.dindex(comp, 300, 600))
            checked += 1
            log(f"batched: component {c} ({comp.nrows} loci) checked against the oracle")
        o.free()
        to.free()
        p.free()
    oracle_s = time.perf_counter() - t
    res = {
        "workload": "24 components, log-normal lengths (median 5e6 bp, sigma 1, clipped "
                    "1e5-1.2e8), SNP 0.01, indel 0.001, uint64 rowmap, [300,600]",
        "components": 24,
        "loci": total,
        "batched_s": round(batched_s, 4),
        "size_class_tuned_s": round(tuned_s, 4),
        "size_class_builds": tst["builds"],
        "size_class_tuning_s": round(tune_s, 2),
        "sequential_plain_s": round(plain_s, 4),
        "speedup": round(plain_s / batched_s, 3),
        "batched_loci_per_s": round(total / batched_s, 1),
        "spgemms": st["n_spgemm"],
        "peak_hbm_bytes": st["peak_device_bytes"],
        "bit_exact_per_component": bool(exact),
        "bit_exact_vs_oracle_per_component": bool(oracle_exact),
        "components_checked_vs_oracle": checked,
        "oracle_s": round(oracle_s, 2),
    }
    log(f"batched {batched_s:.3f}s (per size class, tuned: {tuned_s:.3f}s in "
        f"{tst['builds']} builds) vs sequential {plain_s:.3f}s, exact {exact}, "
        f"oracle exact {oracle_exact}")
    return res


if __name__ == "__main__":
    main()
