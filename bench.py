"""Benchmark of the multi-source search path on one MI355X.

Default workload (the headline): batched multi-source BFS (distances +
uint64 path counts) on R-MAT scale 20, edge factor 16, 64 sources per batch.
A step is one pass of the search over one batch of 64 non-isolated sources
with every input resident in HBM and both outputs ([64 x n] int32 distances,
[64 x n] uint64 path counts) written to HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--variant auto]

Other workloads (one JSON line each, same fields):
    --workload grid   2048x2048 grid, 10% edges deleted, largest component:
                      256-source BFS + path counts per step, and diameter
                      sampling over 64 sources
    --workload bc     sampled Brandes BC on R-MAT scale 22: 1024 sources per
                      step in batches of 64/128/256 (best reported)
    --workload lbs    load-balanced search over 2^26 items, Zipf(1.1) counts
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

METRIC = "GTEPS multi-source BFS/BC on R-MAT s20; served queries/s with p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def model_bytes(n, k, st, sigma=True):
    """Algorithmic HBM bytes of one batch: the per-source compulsory-traffic
    model summed over its k sources (n_s = reached vertices, m_s = edges
    scanned from them, m'_s = shortest-path DAG edges, all counted by the
    engine):
      BFS    4n + 20 n_s + 8 m_s
      sigma  8n +  8 n_s + 16 m'_s"""
    b = 4 * n * k + 20 * st.reached + 8 * st.teps_units
    if sigma:
        b += 8 * n * k + 8 * st.reached + 16 * st.dag_edges
    return int(b)


def profile_traffic(name):
    """HBM bytes per step measured by the rocprofv3 counter runs summarised in
    profiles/<name> (line 'traffic_bytes_per_step: B'), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_step:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def copy_gbs(gu, nbytes=1 << 30, reps=5):
    """Measured device-to-device copy bandwidth (read + write bytes / time),
    the box's STREAM-like ceiling reported beside the 8 TB/s peak."""
    a, b = gu.DeviceBuffer(nbytes), gu.DeviceBuffer(nbytes)
    try:
        a.zero()
        b.copy_from(a)
        gu.synchronize()
        with gu.Timer() as t:
            for _ in range(reps):
                b.copy_from(a)
        return round(2 * nbytes * reps / (t.ms / 1e3) / 1e9, 1)
    finally:
        a.close()
        b.close()


def cpu_threads():
    try:
        c = len(os.sched_getaffinity(0))
    except AttributeError:
        c = os.cpu_count() or 1
    return max(1, min(16, c))


def _line(metric, value, unit, args, ms, dtype, config, extra):
    out = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded generator, no dataset)", "config": config,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


def bench_grid(args, gu, props):
    """2048x2048 grid minus 10% of edges, largest component: a step is one
    256-source BFS with path counts (one 256-lane batch); then diameter
    sampling over 64 sources.  Thousands of levels: measures per-level cost."""
    t0 = time.time()
    hg = gu.HostGraph.grid(2048, 0.10, seed=30)
    gen_s = time.time() - t0
    n, k = hg.n, 256
    batches = [hg.sample_sources(k, seed=3000 + i) for i in range(args.warmup + args.steps)]
    dist = gu.DeviceBuffer(4 * k * n)
    sig = gu.DeviceBuffer(8 * k * n)
    cfg = gu.make_config(args.variant, direction=args.direction, batch_width=args.width or 256,
                         alpha=args.alpha, beta=args.beta, engine=args.engine)
    with gu.DeviceGraph(hg) as dg:
        for i in range(args.warmup):
            dg.multi_search_device(batches[i], dist, sig, cfg)
        gu.synchronize()
        stats = []
        t1 = time.perf_counter()
        for i in range(args.warmup, args.warmup + args.steps):
            stats.append(dg.multi_search_device(batches[i], dist, sig, cfg))
        gu.synchronize()
        wall = time.perf_counter() - t1
        dsrc = hg.sample_sources(64, seed=3999)
        dcfg = gu.make_config(args.variant, direction=args.direction, alpha=args.alpha, beta=args.beta,
                              engine=args.engine)
        dg.diameter(dsrc, dcfg)
        t2 = time.perf_counter()
        diam, dst = dg.diameter(dsrc, dcfg)
        dwall = time.perf_counter() - t2
    units = sum(s.teps_units for s in stats)
    dev_ms = sum(s.device_ms for s in stats)
    levels = stats[-1].max_level + 1
    alg = sum(model_bytes(n, k, st) for st in stats)
    _line("GTEPS 256-source BFS + path counts, 2048^2 grid -10% LCC", round(units / wall / 1e9, 3),
          "GTEPS", args, wall / args.steps * 1e3, "int32+u64",
          {"workload": "2048x2048 4-neighbour grid, 10% edges deleted (seeded), largest component; "
                       "256 sources per step", "n": n, "m_directed": hg.m,
           "batch_width": args.width or 256, "variant": args.variant, "direction": args.direction,
           "engine": stats[-1].engine_used,
           "generator_seconds": round(gen_s, 2), "device": props["name"], "arch": props["arch"]},
          {"levels_per_step": levels, "us_per_level": round(dev_ms / args.steps / levels * 1e3, 2),
           "roofline": {"bound": "hbm", "achieved": round(alg / (dev_ms / 1e3) / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(alg / (dev_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                        "traffic": profile_traffic("grid2048_256src_traffic.txt")},
           "diameter_sampling": {"sources": 64, "diameter": diam, "ms": round(dwall * 1e3, 3),
                                 "levels": dst.max_level + 1}})


def bench_bc(args, gu, props):
    """Sampled Brandes BC on R-MAT scale 22 (ef16): 1024 sources per step,
    dependencies in FP32 with compensated sums, BC accumulated in FP64; the
    batch width (64/128/256) that runs fastest is reported."""
    t0 = time.time()
    hg = gu.HostGraph.rmat(22, 16, 0.57, 0.19, 0.19, seed=22)
    gen_s = time.time() - t0
    n, k = hg.n, 1024
    src = hg.sample_sources(k, seed=2200)
    bc = gu.DeviceBuffer(8 * n)
    res = {}
    with gu.DeviceGraph(hg) as dg:
        for width in (64, 128, 256):
            cfg = gu.make_config(args.variant, batch_width=width)
            bc.zero()
            dg.betweenness_device(src[:width], None, bc, cfg)  # warm-up
            gu.synchronize()
            t1 = time.perf_counter()
            units = 0
            for _ in range(max(args.steps, 1)):
                bc.zero()
                st = dg.betweenness_device(src, None, bc, cfg)
                units += st.teps_units
            gu.synchronize()
            wall = (time.perf_counter() - t1) / max(args.steps, 1)
            res[width] = (wall, units / max(args.steps, 1), st)
        best = min(res, key=lambda w: res[w][0])
        check = {}
        if args.bc_check:
            from oracle import oracle as orc

            ks = src[:args.bc_check]
            bc.zero()
            dg.betweenness_device(ks, None, bc, gu.make_config(args.variant, batch_width=best))
            got = bc.to_host(np.float64)
            print("bc-check: FP64 oracle over %d sources ..." % len(ks), file=sys.stderr, flush=True)
            t2 = time.time()
            ref = orc.bc_f64_mt(hg.R, hg.C, ks, min(16, os.cpu_count() or 1))
            nz = ref > 0
            rel = np.abs(got[nz] - ref[nz]) / ref[nz]
            check = {"sources": int(len(ks)), "max_rel_err": float(rel.max(initial=0.0)),
                     "zero_where_oracle_zero": bool(np.all(got[~nz] == 0)),
                     "oracle_seconds": round(time.time() - t2, 1)}
    wall, units, st = res[best]
    _line("GTEPS sampled BC (1024 sources), R-MAT s22", round(units / wall / 1e9, 3), "GTEPS", args,
          wall * 1e3, "f32 delta / f64 BC / u64 sigma",
          {"workload": "R-MAT s22 ef16, 1024 sampled non-isolated sources per step", "n": n,
           "m_directed": hg.m, "generator_seconds": round(gen_s, 2), "device": props["name"],
           "arch": props["arch"]},
          {"batch_width": best,
           "ms_by_width": {str(w): round(res[w][0] * 1e3, 3) for w in res},
           "levels_last_batch": st.max_level + 1, "bc_check": check})


def bench_lbs(args, gu, props):
    """Device-wide load-balanced search: 2^26 work items, Zipf(1.1) segment
    counts capped at 2^20 (int32); a step is scan + search."""
    W = 1 << 26
    c = gu.zipf_counts(1.1, 1 << 20, W, seed=26)
    S = len(c)
    cb = gu.DeviceBuffer.from_array(c)
    sb = gu.DeviceBuffer(4 * S)
    rb = gu.DeviceBuffer(4 * W)
    for _ in range(args.warmup):
        gu.lbs_device(cb, S, sb, rb, W)
    gu.synchronize()
    dev = 0.0
    t1 = time.perf_counter()
    for _ in range(args.steps):
        _, st = gu.lbs_device(cb, S, sb, rb, W)
        dev += st.device_ms
    gu.synchronize()
    wall = time.perf_counter() - t1
    alg = 8 * S + 4 * W
    extra = {"roofline": {"bound": "hbm", "achieved": round(alg / (dev / args.steps / 1e3) / 1e9, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(alg / (dev / args.steps / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                          "traffic": profile_traffic("lbs_zipf_2p26_traffic.txt")},
             "device_ms_per_step": round(dev / args.steps, 4), "segments": S}
    if not args.no_cpu:
        from oracle import oracle as orc

        _, _, secs = orc.lbs_timed(c)
        extra["cpu_baseline"] = {"value": round(W / secs / 1e9, 4), "unit": "G items/s", "cores": 1,
                                 "kind": "port", "sample": "the full 2^26-item input, "
                                 "scan + merge (lbs.cu:18-31), %.2f s" % secs}
    _line("G items/s load-balanced search, 2^26 items, Zipf(1.1)", round(W / (dev / args.steps / 1e3) / 1e9, 3),
          "G items/s", args, wall / args.steps * 1e3, "int32",
          {"workload": "2^26 work items in %d Zipf(1.1) segments (cap 2^20)" % S,
           "device": props["name"], "arch": props["arch"]}, extra)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="bfs", choices=["bfs", "grid", "bc", "lbs"])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variant", default="auto")
    ap.add_argument("--alpha", type=int, default=0)
    ap.add_argument("--direction", type=int, default=0, help="0 auto, 1 top-down, 2 bottom-up")
    ap.add_argument("--width", type=int, default=0, help="sources per batch (64/128/256; 0 = default)")
    ap.add_argument("--engine", type=int, default=0, help="0 auto, 1 batched lanes, 2 one workgroup per source")
    ap.add_argument("--beta", type=int, default=0)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-variants", action="store_true", help="skip the per-variant sweep")
    ap.add_argument("--bc-check", type=int, default=0,
                    help="bc workload: compare BC over this many sources with the FP64 CPU oracle")
    ap.add_argument("--calib-bytes", type=int, default=0,
                    help="first run one device-to-device copy of this many bytes (a kernel of known "
                         "traffic, to scale HBM counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: multi-GPU source sharding is not part of this "
              "build (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    gu = __graft_entry__.load_package()
    props = gu.choose_device(0)
    if args.calib_bytes:
        a = gu.DeviceBuffer(args.calib_bytes)
        b = gu.DeviceBuffer(args.calib_bytes)
        a.zero()
        b.copy_from(a)
        gu.synchronize()
        a.close()
        b.close()
    if args.workload != "bfs":
        {"grid": bench_grid, "bc": bench_bc, "lbs": bench_lbs}[args.workload](args, gu, props)
        return
    t0 = time.time()
    hg = gu.HostGraph.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=20)
    gen_s = time.time() - t0
    n = hg.n
    k = args.batch
    nb = args.warmup + args.steps
    batches = [hg.sample_sources(k, seed=1000 + i) for i in range(nb)]
    dist = gu.DeviceBuffer(4 * k * n)
    sig = gu.DeviceBuffer(8 * k * n)
    cfg = gu.make_config(args.variant, alpha=args.alpha, beta=args.beta)
    with gu.DeviceGraph(hg) as dg:
        for i in range(args.warmup):
            dg.multi_search_device(batches[i], dist, sig, cfg)
        gu.synchronize()
        units = 0
        dev_ms = 0.0
        stats = []
        t1 = time.perf_counter()
        for i in range(args.warmup, nb):
            st = dg.multi_search_device(batches[i], dist, sig, cfg)
            units += st.teps_units
            dev_ms += st.device_ms
            stats.append(st)
        gu.synchronize()
        wall = time.perf_counter() - t1
        # algorithmic bytes over the timed steps
        alg_bytes = sum(model_bytes(n, k, st) for st in stats)

        variants = {}
        if not args.no_variants:
            for name in ["shuffle", "warp", "cta_warp", "scan", "edge_parallel", "linear"]:
                vc = gu.make_config(name)
                dg.multi_search_device(batches[0], dist, sig, vc)
                u, ms = 0, 0.0
                for i in range(args.warmup, min(nb, args.warmup + 3)):
                    st = dg.multi_search_device(batches[i], dist, sig, vc)
                    u += st.teps_units
                    ms += st.device_ms
                variants[name] = round(u / (ms / 1e3) / 1e9, 2)

    gteps = units / wall / 1e9
    achieved = alg_bytes / (dev_ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+u64",
        "data": "synthetic (seeded R-MAT generator, no dataset)",
        "config": {
            "workload": "R-MAT s%d ef16 (a,b,c,d)=(.57,.19,.19,.05) symmetrised; %d-source "
                        "multi-source BFS, int32 distances + uint64 path counts" % (args.scale, k),
            "n": n, "m_directed": hg.m, "sources_per_step": k, "variant": args.variant,
            "generator_seconds": round(gen_s, 2), "device": props["name"], "arch": props["arch"],
        },
        "device_ms_per_step": round(dev_ms / args.steps, 4),
        "levels_last_step": stats[-1].max_level + 1,
        "topdown_levels_last": stats[-1].topdown_levels,
        "bottomup_levels_last": stats[-1].bottomup_levels,
        "variant_used": stats[-1].variant_used,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profile_traffic("bfs_rmat20_64src_traffic.txt"),
            "scope": "one batched search pass (all kernels of the launch sequence); algorithmic "
                     "bytes = per-source model summed over the batch",
            "algorithmic_bytes_per_step": int(alg_bytes / args.steps),
            "pass_ms": round(dev_ms / args.steps, 4),
            "measured_copy_gbs": copy_gbs(gu),
        },
        "variants_gteps": variants,
    }
    if not args.no_cpu:
        from oracle import oracle as orc

        sample = np.concatenate(batches[args.warmup:args.warmup + 2])
        orc.sigma_batch_timed(hg.R, hg.C, sample[:1], 1)  # warm-up
        u, secs = orc.sigma_batch_timed(hg.R, hg.C, sample, 1)
        out["cpu_baseline"] = {
            "value": round(u / secs / 1e9, 5),
            "unit": "GTEPS",
            "cores": 1,
            "kind": "port",
            "sample": "%d sources (the first two timed batches), sequential BFS + path counts, "
                      "one source after another (oracle restating shortest_paths.cu:11-39), "
                      "%.1f s" % (len(sample), secs),
        }
        nt = cpu_threads()
        sample = np.concatenate(batches[args.warmup:args.warmup + min(4, args.steps)])
        u, secs = orc.sigma_batch_timed(hg.R, hg.C, sample, nt)
        out["cpu_baseline_threads"] = {
            "value": round(u / secs / 1e9, 5), "unit": "GTEPS", "cores": nt, "kind": "port",
            "sample": "%d sources, static partition over %d threads, %.1f s" % (len(sample), nt, secs),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
