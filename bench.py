"""Benchmark: distance-index build on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (synthetic, seeded): the 5e6 bp variation graph (SNP 0.01/bp, indel
0.001/bp, geometric length mean 4 capped at 50), uint32 range-CRS, distance
index for [d1, d2] = [300, 600].  A step is one full index build
D = A^300 . (A + I)^300 with A already resident in HBM; the output stays in
HBM.  `value` is loci indexed per second (rows of D per second of build).

The JSON line also carries:
  roofline      the SpGEMM lane-per-row kernel k_spgemm_lane (the dominant
                kernel): algorithmic bytes bytes(X) + bytes(Y) + bytes(C) of its
                product per launch over its mean launch time (HIP events on the
                library stream), against 8 TB/s
  cpu_baseline  the CPU oracle (the reference's OpenMP algorithm restated in
                C) building the same index on the host, timed once
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


def hip_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--dlo", type=int, default=300)
    ap.add_argument("--dup", type=int, default=600)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
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

    import diverg_amd as dv

    t0 = time.time()
    a = dv.generate_graph(args.preset, size_bytes=4)
    gen_s = time.time() - t0
    print(f"# graph: {a.nrows} loci, {len(a.entries) // 2} ranges, generated in {gen_s:.2f}s",
          file=sys.stderr, flush=True)

    ctx = dv.Context(0)
    da = ctx.to_device(a)  # resident before timing
    for w in range(args.warmup):
        d = ctx.create_distance_index(da, args.dlo, args.dup, location="device")
        d.free()
        print(f"# warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)

    ctx.profile(True)
    stats = []
    hip_sync()
    t_start = time.perf_counter()
    for s in range(args.steps):
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

    cnt = prof["spgemm_lane"]
    mean_ms = cnt["total_ms"] / max(1, cnt["launches"])
    alg_per_launch = cnt["alg_bytes"] / max(1, cnt["launches"])
    achieved = alg_per_launch / (mean_ms * 1e-3) / 1e9 if mean_ms > 0 else 0.0
    slice_gbs = (cnt["slice_bytes"] / max(1, cnt["launches"])) / (mean_ms * 1e-3) / 1e9 \
        if mean_ms > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": "k_spgemm_lane",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None,
        "launches": cnt["launches"],
        "mean_launch_ms": round(mean_ms, 4),
        "alg_bytes_per_launch": int(alg_per_launch),
        "gathered_slice_GBps": round(slice_gbs, 1),
    }

    result_host = d.to_host()
    d.free()

    cpu = None
    bit_exact = None
    if not args.no_cpu:
        from oracle import oracle as orc
        th = cpu_threads()
        orc.set_threads(th)
        t = time.perf_counter()
        e = orc.dindex(a, args.dlo, args.dup)
        cpu_s = time.perf_counter() - t
        bit_exact = bool(result_host.nnz == e.nnz % (1 << 32)
                         and np.array_equal(result_host.rowmap.astype(np.uint64), e.rowmap)
                         and np.array_equal(result_host.entries, e.entries))
        cpu = {
            "value": round(a.nrows / cpu_s, 1),
            "unit": "loci/s",
            "cores": th,
            "kind": "port",
            "sample": (f"the full workload: one build of the same {a.nrows}-locus index "
                       f"[{args.dlo},{args.dup}], {cpu_s:.2f} s"),
            "build_s": round(cpu_s, 3),
        }
        print(f"# cpu oracle: {cpu_s:.2f}s on {th} threads, bit_exact={bit_exact}",
              file=sys.stderr, flush=True)

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
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_ms": {k: round(v["total_ms"] / args.steps, 4) for k, v in prof.items()},
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
