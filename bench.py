"""Benchmark: batched multi-source BFS (distances + uint64 path counts) on
R-MAT scale 20, edge factor 16, 64 sources per batch -- one MI355X.

A step is one pass of the search over one batch of 64 non-isolated sources
with every input resident in HBM and both outputs ([64 x n] int32 distances,
[64 x n] uint64 path counts) written to HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--variant auto]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variant", default="auto")
    ap.add_argument("--alpha", type=int, default=0)
    ap.add_argument("--beta", type=int, default=0)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-variants", action="store_true", help="skip the per-variant sweep")
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
            "traffic": None,
            "scope": "one batched search pass (all kernels of the launch sequence); algorithmic "
                     "bytes = per-source model summed over the batch",
            "algorithmic_bytes_per_step": int(alg_bytes / args.steps),
            "pass_ms": round(dev_ms / args.steps, 4),
        },
        "variants_gteps": variants,
    }
    if not args.no_cpu:
        from oracle import oracle as orc

        sample = batches[args.warmup][:8]
        orc.sigma_batch_timed(hg.R, hg.C, sample[:1], 1)  # warm-up
        u, secs = orc.sigma_batch_timed(hg.R, hg.C, sample, 1)
        out["cpu_baseline"] = {
            "value": round(u / secs / 1e9, 5),
            "unit": "GTEPS",
            "cores": 1,
            "kind": "port",
            "sample": "8 of the 64 sources of one batch, sequential BFS + path counts "
                      "(oracle restating shortest_paths.cu:11-39), %.1f s" % secs,
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
