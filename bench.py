"""Benchmark of the multi-source search path on one MI355X.

Default workload (the headline): batched multi-source BFS (distances +
uint64 path counts) on R-MAT scale 20, edge factor 16, 64 sources per batch.
A step is one pass of the search over one batch of 64 non-isolated sources
with every input resident in HBM and both outputs ([64 x n] int32 distances,
[64 x n] uint64 path counts) written to HBM.  Prints one JSON line; its
"serving" object holds the served queries/s and p50/p99 latency of the mixed
query stream (the serve workload below, --no-serve skips it).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--variant auto]

Other workloads (one JSON line each, same fields):
    --workload grid   2048x2048 grid, 10% edges deleted, largest component:
                      256-source BFS + path counts per step, and diameter
                      sampling over 64 sources
    --workload bc     sampled Brandes BC on R-MAT scale 22: 1024 sources per
                      step in batches of 64/128/256 (best reported)
    --workload lbs    load-balanced search over 2^26 items, Zipf(1.1) counts
                      (--lbs-segments S: over S Zipf-weighted segments, e.g.
                      the reference example's 100,000)
    --workload serve  query server over R-MAT s18 / s20 and a 1024^2 grid:
                      served vs plain queries/s and p50/p99 latency
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


def model_bytes_bc(n, k, batches, st):
    """Algorithmic bytes of a sampled Brandes call over k sources in
    `batches` launches: the BFS + sigma forward model, plus the backward
    20 n_s + 4 n_s (stack) + 8 m_s + 12 m'_s per source, plus the reduce
    (4n read of delta per source, 16n FP64 read-modify-write per batch)."""
    b = model_bytes(n, k, st)
    b += 24 * st.reached + 8 * st.teps_units + 12 * st.dag_edges
    b += 4 * n * k + 16 * n * batches
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


def power_leg(gu, step, seconds=2.0, period_ms=10):
    """Board power while the line's own passes repeat for about `seconds`
    (power_measurer of util_device.cu:79-130, through AMD SMI): the mean of
    the socket-power samples, the energy counter's average over the same
    interval, and GTEPS per watt.  None where the host cannot read the
    device's power."""
    pm = gu.PowerMeasurer(period_ms)
    if not pm.start():
        return {"available": False, "reason": "AMD SMI reports no power for this device (status %d)"
                % pm.status}
    units, i, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        units += step(i)
        i += 1
        if i % 16 == 0:
            gu.synchronize()
    avg = pm.end()
    wall = time.perf_counter() - t0
    r = pm.result
    gteps = units / wall / 1e9
    watts = r.energy_avg_w if r.energy_avg_w > 0 else avg
    return {"available": True, "avg_w": round(avg, 1), "max_w": round(r.max_w, 1),
            "energy_avg_w": round(r.energy_avg_w, 1) if r.energy_avg_w > 0 else None,
            "samples": r.samples, "seconds": round(r.seconds, 2), "passes": i,
            "gteps": round(gteps, 2), "gteps_per_w": round(gteps / watts, 3) if watts > 0 else None,
            "sampler": "AMD SMI socket power every %d ms on a host thread; energy counter read at both "
                       "ends" % period_ms}


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


def _leg_result(metric, value, unit, steps, warmup, ms, dtype, config, extra):
    """One workload's result: a line of its own (--workload) or a sub-object
    of the default line."""
    out = {"metric": metric, "value": value, "unit": unit, "steps": steps, "warmup": warmup,
           "ms_per_step": round(ms, 4), "dtype": dtype, "config": config}
    out.update(extra)
    return out


def _print_leg(args, r):
    extra = {k: v for k, v in r.items()
             if k not in ("metric", "value", "unit", "steps", "warmup", "ms_per_step", "dtype", "config")}
    _line(r["metric"], r["value"], r["unit"], args, r["ms_per_step"], r["dtype"], r["config"], extra)


def _roofline(alg_bytes, pass_ms, traffic, **more):
    """HBM roofline of one pass: model (algorithmic) bytes / pass time against
    the 8 TB/s peak, and the same for the bytes the profiler counted
    (traffic, from the committed counter runs) when known."""
    gbs = alg_bytes / (pass_ms / 1e3) / 1e9
    r = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
         "frac_counter": (round(traffic / (pass_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None),
         "traffic_over_model": round(traffic / alg_bytes, 3) if traffic else None,
         "algorithmic_bytes_per_step": int(alg_bytes), "pass_ms": round(pass_ms, 4)}
    r.update(more)
    return r


def bench_grid(args, gu, props, steps=None, warmup=None, cpu_sources=96):
    """2048x2048 grid minus 10% of edges, largest component: a step is one
    256-source BFS with path counts (one 256-lane batch); then diameter
    sampling over 64 sources.  Thousands of levels: measures per-level cost.
    Returns the result dict."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    t0 = time.time()
    hg = gu.HostGraph.grid(2048, 0.10, seed=30)
    gen_s = time.time() - t0
    n, k = hg.n, 256
    batches = [hg.sample_sources(k, seed=3000 + i) for i in range(warmup + steps)]
    dist = gu.DeviceBuffer(4 * k * n)
    sig = gu.DeviceBuffer(8 * k * n)
    cfg = gu.make_config(args.variant, direction=args.direction, batch_width=args.width or 256,
                         alpha=args.alpha, beta=args.beta, engine=args.engine)
    with gu.DeviceGraph(hg) as dg:
        tuned = None
        if not args.no_tune and args.variant == "auto" and args.engine == 0:
            tuned = dg.autotune(hg.sample_sources(k, seed=3998), gu.QUERY_BFS)
        for i in range(warmup):
            dg.multi_search_device(batches[i], dist, sig, cfg)
        gu.synchronize()
        stats = []
        t1 = time.perf_counter()
        for i in range(warmup, warmup + steps):
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
        dist.close()
        sig.close()
        queue = queue_leg(gu, hg, dg, args.engine, 0 if args.no_queue else args.queue_sources)
    if queue:
        # uneven depths: sources over one 2048^2 grid and 48 grids of 256^2
        hu = grid_union(gu)
        with gu.DeviceGraph(hu) as du:
            queue = {"grid2048": queue,
                     "grid2048_plus_48_grid256": dict(queue_leg(gu, hu, du, args.engine, args.queue_sources),
                                                      n=hu.n)}
    units = sum(s.teps_units for s in stats)
    dev_ms = sum(s.device_ms for s in stats)
    levels = stats[-1].max_level + 1
    alg = sum(model_bytes(n, k, st) for st in stats)
    extra = {}
    if not args.no_cpu:
        from oracle import oracle as orc

        sample = batches[warmup][:cpu_sources]
        u, secs = orc.sigma_batch_timed(hg.R, hg.C, sample, 1)
        extra["cpu_baseline"] = {
            "value": round(u / secs / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": "the first %d sources of the first timed batch, sequential BFS + path counts "
                      "(oracle restating shortest_paths.cu:11-39), %.1f s" % (len(sample), secs)}
    return _leg_result(
        "GTEPS 256-source BFS + path counts, 2048^2 grid -10% LCC", round(units / wall / 1e9, 3),
        "GTEPS", steps, warmup, wall / steps * 1e3, "int32+u64",
        {"workload": "2048x2048 4-neighbour grid, 10% edges deleted (seeded), largest component; "
                     "256 sources per step", "n": n, "m_directed": hg.m,
         "batch_width": args.width or 256, "variant": args.variant, "direction": args.direction,
         "engine": stats[-1].engine_used,
         "generator_seconds": round(gen_s, 2), "device": props["name"], "arch": props["arch"]},
        {"levels_per_step": levels, "us_per_level": round(dev_ms / steps / levels * 1e3, 2),
         "autotuned": tuned,
         "roofline": _roofline(alg / steps, dev_ms / steps, profile_traffic("grid2048_256src_traffic.txt"),
                               scope="one 256-source per-source-engine pass (traversal and the "
                                     "row write-back kernels)"),
         "diameter_sampling": {"sources": 64, "diameter": diam, "ms": round(dwall * 1e3, 3),
                               "levels": dst.max_level + 1},
         "source_queue": queue, **extra})


def grid_union(gu, big=2048, small=256, copies=48, drop=0.10, seed=31):
    """One big grid and `copies` small ones as one graph (block-diagonal
    CSR, each grid with `drop` of its edges deleted, largest component):
    sources drawn uniformly over its vertices have very uneven depths
    (about 2k-4k levels in the big grid, a few hundred in a small one)."""
    parts = [gu.HostGraph.grid(big, drop, seed=seed)]
    parts += [gu.HostGraph.grid(small, drop, seed=seed + 1 + i) for i in range(copies)]
    Rs, Cs, off, base = [np.zeros(1, dtype=np.int64)], [], 0, 0
    for p in parts:
        Rs.append(p.R[1:].astype(np.int64) + base)
        Cs.append(p.C.astype(np.int64) + off)
        off += p.n
        base += p.m
    R = np.concatenate(Rs).astype(np.int32)
    return gu.HostGraph(R, np.concatenate(Cs).astype(np.int32))


def queue_leg(gu, hg, dg, engine, k, seed=3997):
    """One call of k sources (several launches' worth) with path counts, as
    launches of 256 sources and as the queue-fed grid (workgroups take
    sources from a device counter): device time, GTEPS and the deepest and
    mean source depth, which bound the two forms (each launch waits for its
    deepest source; the queue only for the last ones)."""
    if not k:
        return None
    n = hg.n
    src = hg.sample_sources(k, seed=seed)
    db, sb = gu.DeviceBuffer(4 * k * n), gu.DeviceBuffer(8 * k * n)
    out = {"sources": k}
    try:
        for name, mode in (("launches_of_256", 2), ("queue", 1)):
            cfg = gu.make_config(engine=engine, source_queue=mode)
            dg.multi_search_device(src, db, sb, cfg)  # warm-up (workspace)
            st = dg.multi_search_device(src, db, sb, cfg)
            out[name] = {"ms": round(st.device_ms, 3), "gteps": round(st.teps_units / st.device_ms / 1e6, 2),
                         "launches": st.batches, "max_level": st.max_level}
        ecc, _ = dg.multi_search_rows(src, None, None, want_ecc=True,
                                      config=gu.make_config(engine=engine, source_queue=1))
        out["depth_max"], out["depth_mean"] = int(ecc.max()), round(float(ecc.mean()), 1)
        # launches of 256 wait for each launch's deepest source
        out["depth_sum_of_launch_maxima"] = int(sum(ecc[i:i + 256].max() for i in range(0, k, 256)))
        out["speedup"] = round(out["launches_of_256"]["ms"] / out["queue"]["ms"], 3)
    finally:
        db.close()
        sb.close()
    return out


def bench_bc(args, gu, props, steps=None, cpu_sources=8):
    """Sampled Brandes BC on R-MAT scale 22 (ef16): 1024 sources per step,
    dependencies in FP32 with compensated sums, BC accumulated in FP64; the
    batch width (64/128/256) that runs fastest is reported.  Returns the
    result dict."""
    steps = max(args.steps if steps is None else steps, 1)
    t0 = time.time()
    hg = gu.HostGraph.rmat(22, 16, 0.57, 0.19, 0.19, seed=22)
    gen_s = time.time() - t0
    n, k = hg.n, 1024
    src = hg.sample_sources(k, seed=2200)
    bc = gu.DeviceBuffer(8 * n)
    res = {}
    with gu.DeviceGraph(hg) as dg:
        tuned = None
        if not args.no_tune and args.variant == "auto":
            tuned = dg.autotune(src[:256], gu.QUERY_BC)
        for width in ((args.width,) if args.width else (64, 128, 256)):
            cfg = gu.make_config(args.variant, batch_width=width)
            bc.zero()
            dg.betweenness_device(src[:width], None, bc, cfg)  # warm-up
            gu.synchronize()
            t1 = time.perf_counter()
            units = 0
            for _ in range(steps):
                bc.zero()
                st = dg.betweenness_device(src, None, bc, cfg)
                units += st.teps_units
            gu.synchronize()
            wall = (time.perf_counter() - t1) / steps
            res[width] = (wall, units / steps, st)
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
    bc.close()
    wall, units, st = res[best]
    alg = model_bytes_bc(n, k, (k + best - 1) // best, st)
    extra = {"roofline": _roofline(alg, st.device_ms, profile_traffic("bc_rmat22_1024src_traffic.txt"),
                                   scope="one 1024-source Brandes call at the best width (forward, "
                                         "backward and FP64 reduce kernels); traffic measured at width 256")}
    if not args.no_cpu:
        from oracle import oracle as orc

        cs = src[:cpu_sources]
        u, secs = orc.brandes_timed(hg.R, hg.C, cs)
        extra["cpu_baseline"] = {
            "value": round(u / secs / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": "the first %d of the 1024 sources, sequential FP32 Brandes (oracle restating "
                      "bc.cu:11-57), %.1f s; the rate applies linearly to all 1024" % (len(cs), secs)}
    return _leg_result(
        "GTEPS sampled BC (1024 sources), R-MAT s22", round(units / wall / 1e9, 3), "GTEPS", steps, 1,
        wall * 1e3, "f32 delta / f64 BC / u64 sigma",
        {"workload": "R-MAT s22 ef16, 1024 sampled non-isolated sources per step", "n": n,
         "m_directed": hg.m, "generator_seconds": round(gen_s, 2), "device": props["name"],
         "arch": props["arch"]},
        {"batch_width": best, "autotuned": tuned,
         "ms_by_width": {str(w): round(res[w][0] * 1e3, 3) for w in res},
         "levels_last_batch": st.max_level + 1, "bc_check": check, **extra})


def bench_lbs(args, gu, props, steps=None, warmup=None):
    """Device-wide load-balanced search: 2^26 work items, Zipf(1.1) segment
    counts capped at 2^20 (int32); a step is scan + search.  Returns the
    result dict."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    W = 1 << 26
    if args.lbs_segments:
        # a fixed segment count (the reference's lbs.cu example uses 100,000):
        # 2^26 items spread by Zipf(1.1) weights over S segments, shuffled
        S = args.lbs_segments
        w = 1.0 / np.arange(1, S + 1) ** 1.1
        c = np.floor(w / w.sum() * W).astype(np.int64)
        c[0] += W - c.sum()
        np.random.default_rng(26).shuffle(c)
        c = c.astype(np.int32)
    else:
        c = gu.zipf_counts(1.1, 1 << 20, W, seed=26)
    S = len(c)
    cb = gu.DeviceBuffer.from_array(c)
    sb = gu.DeviceBuffer(4 * S)
    rb = gu.DeviceBuffer(4 * W)
    for _ in range(warmup):
        gu.lbs_device(cb, S, sb, rb, W)
    gu.synchronize()
    dev = 0.0
    t1 = time.perf_counter()
    for _ in range(steps):
        _, st = gu.lbs_device(cb, S, sb, rb, W)
        dev += st.device_ms
    gu.synchronize()
    wall = time.perf_counter() - t1
    for b in (cb, sb, rb):
        b.close()
    alg = 8 * S + 4 * W
    extra = {"roofline": _roofline(alg, dev / steps,
                                   None if args.lbs_segments else profile_traffic("lbs_zipf_2p26_traffic.txt"),
                                   scope="one scan + search call"),
             "device_ms_per_step": round(dev / steps, 4), "segments": S}
    if not args.no_cpu:
        from oracle import oracle as orc

        _, _, secs = orc.lbs_timed(c)
        extra["cpu_baseline"] = {"value": round(W / secs / 1e9, 4), "unit": "G items/s", "cores": 1,
                                 "kind": "port", "sample": "the full 2^26-item input, "
                                 "scan + merge (lbs.cu:18-31), %.2f s" % secs}
    return _leg_result(
        "G items/s load-balanced search, 2^26 items, Zipf(1.1)", round(W / (dev / steps / 1e3) / 1e9, 3),
        "G items/s", steps, warmup, wall / steps * 1e3, "int32",
        {"workload": ("2^26 work items in %d Zipf(1.1)-weighted segments" % S if args.lbs_segments
                      else "2^26 work items in %d Zipf(1.1) segments (cap 2^20)" % S),
         "device": props["name"], "arch": props["arch"]}, extra)


def _serve_queries(gu, hgs, nq, seed):
    """Seeded query stream: graph uniform over the resident graphs, kind
    70% BFS (distances + path counts) / 20% sampled BC / 10% diameter
    sampling, sources per query log-uniform integers in [1, 512]."""
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(nq):
        g = int(rng.integers(len(hgs)))
        u = rng.random()
        kind = gu.QUERY_BFS if u < 0.7 else (gu.QUERY_BC if u < 0.9 else gu.QUERY_DIAMETER)
        k = int(min(512, max(1, np.floor(np.exp(rng.random() * np.log(513))))))
        src = hgs[g].sample_sources(k, seed=int(rng.integers(1 << 31)))
        qs.append((g, kind, src, gu.QF_SIGMA if kind == gu.QUERY_BFS else 0))
    return qs


def _pct(x, p):
    return round(float(np.percentile(np.asarray(x), p)), 2) if len(x) else None


def _warm(gu, srv, ids, dgs, slots, host):
    """Every slot sees every kind at full width on every graph, with results
    where the stream will ask for them."""
    warm = []
    for g, dg in enumerate(dgs):
        src = np.arange(256, dtype=np.int32) % dg.n
        for kind, fl in ((gu.QUERY_BFS, gu.QF_SIGMA), (gu.QUERY_BC, 0), (gu.QUERY_DIAMETER, 0)):
            for _ in range(slots):
                warm.append(srv.submit(ids[g], kind, src, fl | (gu.QF_HOST if host else 0)))
    for t in warm:
        srv.wait(t)
        srv.release(t)


_LEG_COUNTERS = ("submitted", "completed", "rejected", "launches", "sources")


def _served(gu, srv, ids, qs, arrivals_ms, host):
    """Runs the stream through the server; returns (latencies ms, makespan ms,
    stats).  With host every query asks for its results in host memory
    (GU_QF_HOST: the server copies them to its pinned host arena before
    reporting the query done), else they stay in HBM.  A collector thread
    takes finished queries as they come and releases them; latency = done
    (results where they were asked for) - scheduled arrival, on the server
    clock.  The stats' counters cover this stream only; the peaks cover the
    server's life so far."""
    import threading

    st0 = srv.stats()
    done = {}
    nq = len(qs)

    def collect():
        for _ in range(nq):
            t, r = srv.next_done(600000)
            done[t] = (r.done_ms, r.status, r.launches)
            srv.release(t)

    th = threading.Thread(target=collect)
    th.start()
    tick = [0] * nq
    start = srv.now_ms() + 2.0
    for i, (g, kind, src, fl) in enumerate(qs):
        wait = start + arrivals_ms[i] - srv.now_ms()
        if wait > 0.2:
            time.sleep(wait / 1e3)
        tick[i] = srv.submit(ids[g], kind, src, fl | (gu.QF_HOST if host else 0))
    th.join()
    st = srv.stats()
    for k in _LEG_COUNTERS:
        st[k] -= st0[k]
    lat = [done[tick[i]][0] - (start + arrivals_ms[i]) for i in range(nq)]
    bad = sum(1 for i in range(nq) if done[tick[i]][1] != 0)
    makespan = max(done[t][0] for t in tick) - start
    st["failed"] = bad
    st["launches_per_query"] = round(st["launches"] / max(1, nq), 3)
    return lat, makespan, st


def _plain(gu, dgs, qs, arrivals_ms, to_host):
    """The plain path: one query per call on the graph's own handle, serial,
    in arrival order; outputs allocated per query, copied to host memory when
    to_host (the drop-in *_setup pattern: alloc, launch, D2H, free)."""
    lat = []
    t0 = time.perf_counter() * 1e3 + 2.0
    for i, (g, kind, src, fl) in enumerate(qs):
        wait = t0 + arrivals_ms[i] - time.perf_counter() * 1e3
        if wait > 0.2:
            time.sleep(wait / 1e3)
        dg = dgs[g]
        k, n = len(src), dg.n
        if kind == gu.QUERY_BFS:
            db, sb = gu.DeviceBuffer(4 * k * n), gu.DeviceBuffer(8 * k * n)
            dg.multi_search_device(src, db, sb)
            if to_host:
                db.to_host(np.int32)
                sb.to_host(np.uint64)
            db.close()
            sb.close()
        elif kind == gu.QUERY_BC:
            bb = gu.DeviceBuffer(8 * n)
            bb.zero()
            dg.betweenness_device(src, None, bb)
            if to_host:
                bb.to_host(np.float64)
            bb.close()
        else:
            dg.diameter(src)
        lat.append(time.perf_counter() * 1e3 - (t0 + arrivals_ms[i]))
    makespan = time.perf_counter() * 1e3 - t0
    return lat, makespan


def _output_bytes(gu, hgs, qs):
    """Bytes of results the stream asks for: distance + path-count rows of
    BFS queries, BC vectors, diameters."""
    b = 0
    for g, kind, src, fl in qs:
        n = hgs[g].n
        b += len(src) * n * 12 if kind == gu.QUERY_BFS else (8 * n if kind == gu.QUERY_BC else 4)
    return b


def _leg(lat, mk, nq):
    return {"qps": round(nq / (mk / 1e3), 2), "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99)}


def serve_stream(args, gu, hgs, tune, plain_host=True):
    """Runs the seeded serving stream over the resident graphs `hgs`, for two
    result destinations:
      host: results in host memory (the survey's definition of a served
        query: arrival -> host-visible result).  Served = the server with
        GU_QF_HOST; plain = one query per call through the drop-in pattern
        (allocate outputs, launch, copy to pageable host memory, free),
        serial, in arrival order.
      device: results left in HBM on both sides.
    For each destination: the server's saturated rate (every query queued at
    t=0), then Poisson arrivals at 90% of that rate through the server and,
    identical, through the plain path; the device side also at 70%.
    Returns the result dict and the host-results served run's makespan."""
    nq = args.queries
    qs = _serve_queries(gu, hgs, nq, seed=4242)
    budget = args.budget_gb << 30
    host_budget = args.host_budget_gb << 30
    names = ("rmat18", "rmat20", "grid1024")
    dgs = [gu.DeviceGraph(h) for h in hgs]
    unit, tuned = {}, {}
    unit_gaps = np.random.default_rng(99).exponential(1.0, nq)
    arr_of = lambda rate: (np.cumsum(unit_gaps) - unit_gaps[0]) * (1e3 / rate)
    zero = np.zeros(nq)
    try:
        if tune:
            for name, h, dg in zip(names, hgs, dgs):
                smp = h.sample_sources(256, seed=17)
                tuned[name] = {"bfs": dg.autotune(smp, gu.QUERY_BFS), "bc": dg.autotune(smp, gu.QUERY_BC)}
            print("serve: tuned %s" % json.dumps(tuned), file=sys.stderr, flush=True)
        # device time of one launch per graph and kind, 1 and 256 sources
        for name, h, dg in zip(names, hgs, dgs):
            src = h.sample_sources(256, seed=7)
            db, sb, bb = gu.DeviceBuffer(4 * 256 * h.n), gu.DeviceBuffer(8 * 256 * h.n), gu.DeviceBuffer(8 * h.n)
            row = {}
            for k in (1, 256):
                dg.multi_search_device(src[:k], db, sb)
                row["bfs_sigma_%d" % k] = round(dg.multi_search_device(src[:k], db, sb).device_ms, 3)
                dg.betweenness_device(src[:k], None, bb)
                row["bc_%d" % k] = round(dg.betweenness_device(src[:k], None, bb).device_ms, 3)
                row["diameter_%d" % k] = round(dg.diameter(src[:k])[1].device_ms, 3)
            for b in (db, sb, bb):
                b.close()
            unit[name] = row
        print("serve: unit costs %s" % json.dumps(unit), file=sys.stderr, flush=True)
        # One server per destination, configured as it would be deployed
        # (a host arena only for host results), its graphs registered (each
        # slot's search workspace reserved) and warmed once; every leg of
        # that destination, plain one included, runs while it is alive.
        # Freeing a server just before a plain leg slowed the plain path's
        # per-query allocations by seconds at a time, and the host arena's
        # delivery stream is a fifth stream on four hardware queues.
        with gu.Server(hbm_budget=budget, slots=args.slots, max_pack=256, host_budget=host_budget) as srv:
            ids = [srv.add_graph(dg) for dg in dgs]
            _warm(gu, srv, ids, dgs, args.slots, True)
            lat_hs, mk_hs, st_hs = _served(gu, srv, ids, qs, zero, True)
            sat_h = nq / (mk_hs / 1e3)
            qps_h = args.qps if args.qps > 0 else 0.9 * sat_h
            arr_h = arr_of(qps_h)
            lat_h, mk_h, st_h = _served(gu, srv, ids, qs, arr_h, True)
            print("serve: host results, saturated %.1f q/s, served at %.1f q/s offered" % (sat_h, qps_h),
                  file=sys.stderr, flush=True)
            if plain_host:
                lat_ph, mk_ph = _plain(gu, dgs, qs, arr_h, True)
                print("serve: plain (host results) done", file=sys.stderr, flush=True)
        with gu.Server(hbm_budget=budget, slots=args.slots, max_pack=256) as srv:
            ids = [srv.add_graph(dg) for dg in dgs]
            _warm(gu, srv, ids, dgs, args.slots, False)
            lat_s, mk_s, st_s = _served(gu, srv, ids, qs, zero, False)
            sat_d = nq / (mk_s / 1e3)
            qps_d = args.qps if args.qps > 0 else 0.9 * sat_d
            arr_d, arr70 = arr_of(qps_d), arr_of(0.7 * sat_d)
            lat_70, mk_70, st_70 = _served(gu, srv, ids, qs, arr70, False)
            lat_d, mk_d, st_d = _served(gu, srv, ids, qs, arr_d, False)
            lat_pd, mk_pd = _plain(gu, dgs, qs, arr_d, False)
            print("serve: device results, saturated %.1f q/s; plain done" % sat_d, file=sys.stderr, flush=True)
    finally:
        for dg in dgs:
            dg.close()
    out_b = _output_bytes(gu, hgs, qs)
    hostr = {"saturated_qps": round(sat_h, 2), "offered_qps": round(qps_h, 2),
             "served": dict(_leg(lat_h, mk_h, nq), server=st_h),
             "results_gb": round(out_b / 1e9, 2),
             "served_results_gbs": round(out_b / (mk_h / 1e3) / 1e9, 2)}
    if plain_host:
        hostr["plain"] = _leg(lat_ph, mk_ph, nq)
        hostr["like_for_like_ratio"] = round(hostr["served"]["qps"] / hostr["plain"]["qps"], 2)
        hostr["p99_ratio"] = round(hostr["plain"]["p99_ms"] / hostr["served"]["p99_ms"], 2)
    devr = {"saturated_qps": round(sat_d, 2), "offered_qps": round(qps_d, 2),
            "served": dict(_leg(lat_d, mk_d, nq), server=st_d),
            "served_at_70pct": dict(_leg(lat_70, mk_70, nq), offered_qps=round(0.7 * sat_d, 2)),
            "plain": _leg(lat_pd, mk_pd, nq)}
    devr["ratio"] = round(devr["served"]["qps"] / devr["plain"]["qps"], 2)
    devr["p99_ratio"] = round(devr["plain"]["p99_ms"] / devr["served"]["p99_ms"], 2)
    res = {"queries": nq, "slots": args.slots,
           # hardware queues of the process: HIP's default is 4, and a server
           # with a host arena has one stream more than its slots
           "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "default (4)"),
           "autotuned": tuned, "launch_ms": unit,
           "qps": hostr["served"]["qps"], "p50_ms": hostr["served"]["p50_ms"],
           "p99_ms": hostr["served"]["p99_ms"],
           "like_for_like_ratio": hostr.get("like_for_like_ratio"),
           "host_results": hostr, "device_results": devr}
    return res, mk_h


def serve_graphs(gu, rmat20=None):
    """The three resident graphs of the serving stream: R-MAT s18 ef16,
    R-MAT s20 ef16 (the same seeded graph as the BFS workload) and a 1024^2
    grid with 10% of its edges deleted, largest component."""
    return [gu.HostGraph.rmat(18, 16, 0.57, 0.19, 0.19, seed=18),
            rmat20 if rmat20 is not None else gu.HostGraph.rmat(20, 16, 0.57, 0.19, 0.19, seed=20),
            gu.HostGraph.grid(1024, 0.10, seed=10)]


SERVE_WORKLOAD = ("R-MAT s18 ef16 + R-MAT s20 ef16 + 1024^2 grid -10%% LCC resident; %d queries, "
                  "70%% BFS (d + sigma) / 20%% BC / 10%% diameter, 1-512 sources log-uniform, "
                  "Poisson arrivals")


def bench_serve(args, gu, props):
    """Serving: R-MAT s18 ef16, R-MAT s20 ef16 and a 1024^2 grid (-10% edges,
    largest component) resident in HBM; a seeded stream of BFS / BC /
    diameter queries of 1-512 sources, reporting queries/s and p50/p99
    latency (scheduled arrival -> result ready) served and plain."""
    t0 = time.time()
    hgs = serve_graphs(gu)
    gen_s = time.time() - t0
    res, mk_l = serve_stream(args, gu, hgs, not args.no_tune)
    nq = args.queries
    _line("served queries/s, mixed BFS/BC/diameter stream over 3 resident graphs", res["qps"],
          "queries/s", args, mk_l / nq, "int32+u64 / f32+f64 BC",
          {"workload": SERVE_WORKLOAD % nq,
           "hbm_budget_gb": args.budget_gb, "slots": args.slots, "max_pack": 256,
           "generator_seconds": round(gen_s, 2), "device": props["name"], "arch": props["arch"],
           "graphs": [{"n": h.n, "m_directed": h.m} for h in hgs]},
          res)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="bfs", choices=["bfs", "grid", "bc", "lbs", "serve"])
    ap.add_argument("--queries", type=int, default=200, help="serve workload: queries in the stream")
    ap.add_argument("--qps", type=float, default=0.0, help="serve workload: offered load of the "
                    "comparison run (0 = 90%% of the measured saturated rate)")
    ap.add_argument("--slots", type=int, default=4, help="serve workload: launches in flight")
    ap.add_argument("--budget-gb", type=int, default=64, help="serve workload: HBM budget for outputs")
    ap.add_argument("--host-budget-gb", type=int, default=64,
                    help="serve workload: pinned host memory for results delivered to the host")
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
    ap.add_argument("--lbs-segments", type=int, default=0,
                    help="--workload lbs: spread the 2^26 items over this many Zipf-weighted segments "
                         "(default: Zipf(1.1) counts capped at 2^20, 1805 segments)")
    ap.add_argument("--no-variants", action="store_true", help="skip the per-variant sweep")
    ap.add_argument("--queue-sources", type=int, default=1024,
                    help="grid workload: sources of the call timed as launches of 256 and as the "
                         "queue-fed grid")
    ap.add_argument("--no-queue", action="store_true", help="grid workload: skip that comparison")
    ap.add_argument("--no-power", action="store_true", help="bfs workload: skip the power leg")
    ap.add_argument("--no-serve", action="store_true",
                    help="bfs workload: skip the serving stream reported beside the GTEPS figure")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (degree-profile defaults)")
    ap.add_argument("--no-legs", action="store_true",
                    help="bfs workload: skip the grid, BC and LBS legs reported beside the headline")
    ap.add_argument("--level-graphs", type=int, default=0,
                    help="batched engine: 1 replays captured chunks of levels as HIP graphs, 0 launches "
                         "the kernels directly")
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
    if args.workload == "serve":
        bench_serve(args, gu, props)
        return
    if args.workload != "bfs":
        _print_leg(args, {"grid": bench_grid, "bc": bench_bc, "lbs": bench_lbs}[args.workload](args, gu, props))
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
    cfg = gu.make_config(args.variant, alpha=args.alpha, beta=args.beta, level_graphs=args.level_graphs,
                         batch_width=args.width)
    with gu.DeviceGraph(hg) as dg:
        tuned = None
        if not args.no_tune and args.variant == "auto":
            tuned = dg.autotune(hg.sample_sources(k, seed=999), gu.QUERY_BFS)
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

        # the same passes with both outputs copied to (pageable) host memory,
        # as the drop-in *_setup wrappers return them
        e2e_units, t2 = 0, time.perf_counter()
        for i in range(args.warmup, min(nb, args.warmup + 3)):
            st = dg.multi_search_device(batches[i], dist, sig, cfg)
            dist.to_host(np.int32)
            sig.to_host(np.uint64)
            e2e_units += st.teps_units
        e2e_gteps = e2e_units / (time.perf_counter() - t2) / 1e9

        power = None if args.no_power else power_leg(
            gu, lambda i: dg.multi_search_device(batches[args.warmup + i % args.steps], dist, sig, cfg).teps_units)

        variants, variants_td, td_levels = {}, {}, {}
        if not args.no_variants:
            for name in ["shuffle", "warp", "cta_warp", "scan", "edge_parallel", "linear"]:
                vc = gu.make_config(name)
                dg.multi_search_device(batches[0], dist, sig, vc)
                u, ms = 0, 0.0
                for i in range(args.warmup, nb):  # the timed batches of the line itself
                    st = dg.multi_search_device(batches[i], dist, sig, vc)
                    u += st.teps_units
                    ms += st.device_ms
                variants[name] = round(u / (ms / 1e3) / 1e9, 2)
            # every level top-down (the reference kernels' direction), so
            # the expanders themselves are compared; the first 5 timed batches
            for name in ["shuffle", "warp", "cta_warp", "scan", "edge_parallel", "linear"]:
                vc = gu.make_config(name, direction=gu.DIR_TOPDOWN)
                dg.multi_search_device(batches[0], dist, sig, vc)
                u, ms = 0, 0.0
                for i in range(args.warmup, min(nb, args.warmup + 5)):
                    st = dg.multi_search_device(batches[i], dist, sig, vc)
                    u += st.teps_units
                    ms += st.device_ms
                variants_td[name] = round(u / (ms / 1e3) / 1e9, 2)
                td_levels[name] = st.topdown_levels

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
        "autotuned": tuned,
        "roofline": _roofline(alg_bytes / args.steps, dev_ms / args.steps,
                              profile_traffic("bfs_rmat20_64src_traffic.txt"),
                              scope="one batched search pass (all kernels of the launch sequence); "
                                    "algorithmic bytes = per-source model summed over the batch",
                              measured_copy_gbs=copy_gbs(gu)),
        "variants_gteps": variants,
        "variants_topdown_gteps": variants_td,
        "pcie_inclusive_gteps": round(e2e_gteps, 2),
        "power": power,
    }
    dist.close()
    sig.close()
    if not args.no_legs:
        # the other benchmark configurations, each with its roofline and CPU
        # baseline, on fewer steps than their own --workload lines
        t0 = time.time()
        out["grid"] = bench_grid(args, gu, props, steps=3, warmup=1, cpu_sources=32)
        out["bc"] = bench_bc(args, gu, props, steps=1, cpu_sources=4)
        out["lbs"] = bench_lbs(args, gu, props, steps=20, warmup=3)
        out["legs_seconds"] = round(time.time() - t0, 1)
    if not args.no_serve:
        # the metric's second half: the mixed query stream over three resident
        # graphs (the R-MAT s20 above is one of them), served and plain
        t0 = time.time()
        hgs = serve_graphs(gu, hg)
        res, _ = serve_stream(args, gu, hgs, not args.no_tune)
        res["workload"] = SERVE_WORKLOAD % args.queries
        res["hbm_budget_gb"], res["slots"], res["max_pack"] = args.budget_gb, args.slots, 256
        res["seconds"] = round(time.time() - t0, 1)
        out["serving"] = res
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
