"""Aggregate GTEPS of concurrent multi-source BFS on one MI355X, plain vs
batched vs served.

Plain (the headline `value`): R-MAT scale 22, edge factor 16 (a,b,c =
.57,.19,.19, seeded, vertex ids permuted, cleaned and symmetrised), 1,024
seeded sources of degree > 0, GroupBy-ordered, run as 8 groups of 128 lanes
one after another with the direction-optimising policy; one step = the 8
groups with all 1,024 depth arrays written to device memory.  GTEPS = sum
over sources of deg(v) over reached v (directed CSR entries) / wall time.

Batched: the same 1,024 sources through the batched launcher (wider lane
words, several traversal contexts in flight, device-driven levels).
Served: Poisson-arriving queries of 1-256 sources (log-uniform) against the
R-MAT s22 and a Chung-Lu graph, both resident; p50/p99 latency per rate.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import ibfs_amd  # noqa: E402

METRIC = "aggregate GTEPS (128 srcs, R-MAT s22) + p50/p99 query latency, plain vs served"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "s22": dict(scale=22, sources=1024),
    "s20": dict(scale=20, sources=1024),
    "s16": dict(scale=16, sources=128),
}

# Plain-leg engine settings per workload: inspector workgroup size and the
# degree classes' thresholds (the reference's gpu_ibfs takes sml_shed and
# lrg_shed as constructor arguments, gpu_ibfs.cu:36-45).  For R-MAT s22 a sweep
# of the three on 8 x 128 lanes took 15.12 ms per 1,024 sources with these
# against 15.43-15.47 ms with the library defaults (512, 64, 512); other
# workloads keep the defaults.
PLAIN_SETTINGS = {"s22": dict(wg_size=1024, sml_shed=128, lrg_shed=2048)}


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def alg_bytes_per_kernel(log_levels, V, Vt, W, k, hub_frac=0.0):
    """Algorithmic bytes of each kernel launch of one traversal, from the
    level log (B=8 offset, A=4 adjacency, Q=4 queue entry, W status bytes,
    F=1 changed flag; status words exist for the Vt vertices with an edge).
    A hybrid level reads each candidate's word and its entries to the pulled
    hubs, counted as hub_frac (the pulled hubs' share of all adjacency
    entries) of the candidates' edges, an upper bound since a candidate stops
    once it covers the active lanes; the other frontier vertices push."""
    B, A, Q, F = 8, 4, 4, 1
    out = {"expand_td": 0, "expand_bu": 0, "expand_hy": 0, "inspect": 0, "emit": 0}
    for i, li in enumerate(log_levels):
        L = li["level"]
        q = sum(li["q_td"][:2]) * Q + li["q_td"][2] * 8 + sum(li["q_bu"]) * Q
        # flag + degree + S_L read + S_{L+1} written for every vertex; S_{L-1}
        # read and the flag cleared for the changed ones
        out["inspect"] += Vt * (F + 4 + 2 * W) + li["n_changed"] * ((W if L > 0 else 0) + F) + q
        if L == 0:
            continue
        prev = log_levels[i - 1]
        if li["direction"] == 0:
            out["expand_td"] += (prev["n_frontier"] * (2 * B + W) + prev["m_frontier"] * (A + W)
                                 + (W + F) * li["n_changed"])
        elif li["direction"] == 2:
            out["expand_hy"] += (prev["n_unvisited"] * (2 * B + W)
                                 + int(hub_frac * prev["m_unvisited"]) * (A + W)
                                 + prev["m_frontier_nonhub"] * (A + W) + W * li["n_changed"])
        else:
            out["expand_bu"] += (li["bu_work"] * (2 * B + W) + li["bu_prefix"] * (A + W)
                                 + (W + F) * li["n_changed"])
    nsnap = len(log_levels) - 1  # snapshots 0..last
    out["emit"] = nsnap * W * Vt + 4 * V + 4 * k * V  # snapshots, vertex map, depths
    return out


def plain_leg(g, src, K, steps, warmup, tune, emit_mode, out):
    """Two timed regions of `steps` steps each on the same engine settings:
    the first without kernel events (the headline rate), the second with a
    HIP event pair around every kernel on the engine's stream (per-kernel
    durations for the roofline; the events add gaps between kernels)."""
    res = {}
    for timing in (0, 1):
        eng = ibfs_amd.Engine(g, lane_bits=K, kernel_timing=timing, emit_mode=emit_mode, **tune)
        for i in range(warmup):
            eng.run_groups(src, out=out)
        kern = {"expand_td": 0.0, "expand_bu": 0.0, "expand_hy": 0.0, "inspect": 0.0, "emit": 0.0}
        launches = dict.fromkeys(kern, 0)
        edges = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.run_groups(src, out=out)
            st = eng.last_stats
            edges += st.traversed_edges
            kern["expand_td"] += st.expand_td_ms
            kern["expand_bu"] += st.expand_bu_ms
            kern["expand_hy"] += st.expand_hy_ms
            kern["inspect"] += st.inspect_ms
            kern["emit"] += st.emit_ms
            launches["expand_td"] += st.n_td_launches
            launches["expand_bu"] += st.n_bu_launches
            launches["expand_hy"] += st.n_hy_launches
            launches["inspect"] += st.n_inspect_launches
            launches["emit"] += st.n_emit_launches
        wall = time.perf_counter() - t0
        res[timing] = dict(gteps=edges / wall / 1e9, ms_per_step=wall * 1e3 / steps,
                           edges_per_step=edges // steps,
                           kern_ms={k: v / steps for k, v in kern.items()},
                           launches={k: v // steps for k, v in launches.items()}, cfg=eng.cfg)
        eng.close()
    out0 = res[0]
    out0["kern_ms"] = res[1]["kern_ms"]
    out0["launches"] = res[1]["launches"]
    out0["events_ms_per_step"] = res[1]["ms_per_step"]
    return out0


def roofline_leg(g, src, K, V, tune, kern_ms, launches, steps):
    Vt = g.traversal_vertices
    model = ibfs_amd.Engine(g, lane_bits=K, work_model=1, **tune)
    scratch = ibfs_amd.DeviceBuffer(4 * K * V)
    alg = {"expand_td": 0, "expand_bu": 0, "expand_hy": 0, "inspect": 0, "emit": 0}
    hub_frac = g.hub_entries(model.cfg.hybrid_hubs) / max(g.edge_count, 1)
    levels = []
    for g0 in range(0, len(src), K):
        grp = src[g0:g0 + K]
        model.traversal(grp, out=scratch)
        lv = model.level_log()
        levels.append([(x["level"], x["direction"]) for x in lv])
        for kname, b in alg_bytes_per_kernel(lv, V, Vt, K // 8, len(grp), hub_frac).items():
            alg[kname] += b
    dom = max(kern_ms, key=kern_ms.get)
    achieved = alg[dom] / (kern_ms[dom] * 1e-3) / 1e9 if kern_ms[dom] > 0 else 0.0
    kernels = {k: {"ms_per_step": round(kern_ms[k], 4), "launches_per_step": launches[k],
                   "alg_GB_per_step": round(alg[k] / 1e9, 4),
                   "alg_GBps": round(alg[k] / (kern_ms[k] * 1e-3) / 1e9, 1) if kern_ms[k] > 0 else None}
               for k in kern_ms}
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}
    return roof, kernels, levels


def batch_alg_bytes(g, beg, adj, src, opts):
    """Algorithmic bytes per step of the batched form's kernels: the sources
    grouped as the batched launcher packs them (descending two-hop size,
    stable), each group traversed once by a work-model engine with the same
    settings (the device plan picks the host rule's directions); expansion
    = top-down + bottom-up + hybrid levels (one k_expand_dd launch each)."""
    K = int(opts.get("lane_bits", 256))
    deg = np.diff(beg)
    key = np.array([int(deg[adj[beg[s]:beg[s + 1]]].sum()) for s in src], dtype=np.int64)
    order = np.argsort(-key, kind="stable") if opts.get("pack_sources", 1) else np.arange(len(src))
    psrc = src[order]
    cfg = {k: v for k, v in opts.items() if k not in ("contexts", "burst")}
    model = ibfs_amd.Engine(g, work_model=1, **cfg)
    V, Vt = g.vert_count, g.traversal_vertices
    scratch = ibfs_amd.DeviceBuffer(4 * K * V)
    hub_frac = g.hub_entries(model.cfg.hybrid_hubs) / max(g.edge_count, 1)
    alg = {"expand": 0, "inspect": 0, "emit": 0}
    launches = {"expand": 0, "inspect": 0, "emit": 0}
    for g0 in range(0, len(psrc), K):
        grp = psrc[g0:g0 + K]
        model.traversal(grp, out=scratch)
        lv = model.level_log()
        words = 1  # status words of this group: the narrowest width that holds it
        while words * 64 < len(grp) and words * 64 < K:
            words *= 2
        b = alg_bytes_per_kernel(lv, V, Vt, 8 * words, len(grp), hub_frac)
        alg["expand"] += b["expand_td"] + b["expand_bu"] + b["expand_hy"]
        alg["inspect"] += b["inspect"]
        alg["emit"] += b["emit"]
        launches["expand"] += len(lv) - 1
        launches["inspect"] += len(lv)
        launches["emit"] += 1
    model.close()
    scratch.free()
    return {k: {"alg_GB_per_step": round(alg[k] / 1e9, 4), "launches_per_step": launches[k]}
            for k in alg}


KERNEL_NAMES = {"expand_bu": "k_expand_bu", "expand_td": "k_expand_td", "expand_hy": "k_expand_hy",
                "inspect": "k_inspect", "emit": "k_emit"}


def measured_traffic(workload, K, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (scaled by the 1 GiB copy yardstick
    recorded there), or None when no such summary exists."""
    path = ROOT / "profiles" / f"hbm_traffic_rmat_{workload}_plain{K}.md"
    if not path.exists():
        return None, None
    want = f"{KERNEL_NAMES[kernel]}<{K // 64}"
    col = None
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if "HBM bytes/launch" in cells:
            col = cells.index("HBM bytes/launch")
        elif col is not None and len(cells) > col and cells[0].startswith(want):
            try:
                return float(cells[col]), path.name
            except ValueError:
                return None, None
    return None, None


# Batched-leg settings: the "best" record of the batched autotuner's report
# for this workload (ibfs_amd.autotune_batch on the R-MAT s22 graph with the
# 1,024 GroupBy-ordered sources, saved with save_to=); the library defaults
# with 256 lanes, 2 contexts and bursts of 4 when no report is present.
TUNED_DIR = ROOT / "ibfs_amd" / "tuned"


def batch_settings(workload):
    return ibfs_amd.tuned_batch_settings(f"rmat_{workload}")


def batch_leg(g, src, steps, warmup, tune, emit_mode, out, workload, contexts=None, burst=None):
    opts, c0, b0, source = batch_settings(workload)
    opts = dict(opts, **tune)
    if emit_mode is not None:
        opts["emit_mode"] = emit_mode
    contexts = contexts or c0
    burst = burst or b0
    b = ibfs_amd.Batch(g, contexts=contexts, burst=burst, **opts)
    for _ in range(warmup):
        b.run(src, out=out)
    edges = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(src, out=out)
        edges += b.last_stats.traversed_edges
    wall = time.perf_counter() - t0
    st = b.last_stats
    b.close()
    return dict(gteps=edges / wall / 1e9, ms_per_step=wall * 1e3 / steps,
                lane_bits=opts.get("lane_bits", 256), contexts=contexts, burst=burst,
                config=opts, tuned_from=source,
                levels_per_direction={"td": st.n_td_launches, "bu": st.n_bu_launches,
                                      "hybrid": st.n_hy_launches, "tail": st.n_tail_levels},
                tail_reruns=st.n_tail_reruns)


def edges_of(g, sources, chunk=256):
    """Traversed edges of each source set (untimed, through the batched form)."""
    b = ibfs_amd.Batch(g, contexts=1, burst=8, lane_bits=256)
    buf = ibfs_amd.DeviceBuffer(4 * chunk * g.vert_count)
    total = 0
    for i in range(0, len(sources), chunk):
        b.run(sources[i:i + chunk], out=buf)
        total += b.last_stats.traversed_edges
    b.close()
    buf.free()
    return total


def served_leg(graphs, rates, n_queries, seed, K, admit_window=None, check=True, threads=16,
               server_opts=None):
    """Poisson arrivals at each rate; every query names one graph and carries
    a log-uniform 1..256 source set.  Outputs come from a device row pool per
    graph, recycled on release (one allocation before the arrivals).  Latency
    = completion - scheduled arrival, so any wait of the arrival thread for
    pool space counts.  Per rate, 2 rows of each of the first 3 queries on
    each graph are copied out at completion and compared with the oracle."""
    rng = np.random.default_rng(seed)
    opts = dict(server_opts or {})
    servers = [ibfs_amd.Server(gr["g"], lane_bits=K, status_bytes=gr["status_bytes"],
                               admit_window=admit_window, **opts) for gr in graphs]
    pools = [ibfs_amd.RowPool(gr["g"].vert_count, gr["pool_rows"]) for gr in graphs]
    results = []
    for lam in rates:
        gi = rng.integers(0, len(graphs), size=n_queries)
        sizes = np.floor(np.exp(rng.uniform(0, np.log(256), size=n_queries))).astype(int)
        qs = [rng.choice(graphs[i]["pool"], size=int(n)).astype(np.int32) for i, n in zip(gi, sizes)]
        gaps = rng.exponential(1.0 / lam, size=n_queries)
        ids = [None] * n_queries
        views = [None] * n_queries
        t_sched = [0.0] * n_queries
        lat = [None] * n_queries
        samples = {}  # query -> (graph, rows of its sources, host copies)
        want_sample = {k: 0 for k in range(len(graphs))}
        origin = time.perf_counter()

        def producer():
            t_next = time.perf_counter()
            for i in range(n_queries):
                t_next += gaps[i]
                t_sched[i] = t_next
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                views[i] = pools[gi[i]].alloc(len(qs[i]))
                ids[i] = servers[gi[i]].submit(qs[i], views[i])

        def consumer():
            for i in range(n_queries):
                while ids[i] is None:
                    time.sleep(0.0002)
                info = servers[gi[i]].wait(ids[i], timeout_ms=120000)
                t_done = time.perf_counter()
                lat[i] = (info, t_done)
                k = gi[i]
                if check and want_sample[k] < 3:
                    want_sample[k] += 1
                    V = graphs[k]["g"].vert_count
                    rows = list(range(min(2, len(qs[i]))))
                    samples[i] = (k, rows, [views[i].to_host(count=V, offset=r * V) for r in rows])
                servers[k].release(ids[i])
                pools[k].free(views[i])
                views[i] = None

        waits0 = [p.waits for p in pools]
        tp, tc = threading.Thread(target=producer), threading.Thread(target=consumer)
        tp.start()
        tc.start()
        tp.join()
        tc.join()
        waits = sum(p.waits for p in pools) - sum(waits0)
        lats = np.array([(lat[i][1] - t_sched[i]) * 1e3 for i in range(n_queries)])
        t_first = min(t_sched) - origin
        t_last = max(x[1] for x in lat) - origin
        span_s = t_last - t_first
        verified = 0
        if check:  # untimed: the sampled rows against the oracle
            import oracle
            for k, gr in enumerate(graphs):
                sel = [(i, r) for i, (kk, rows, _) in samples.items() if kk == k for r in rows]
                if not sel:
                    continue
                want = oracle.bfs_many(gr["beg"], gr["adj"],
                                       np.array([qs[i][r] for i, r in sel], dtype=np.int32),
                                       threads=threads)
                for (i, r), w in zip(sel, want):
                    got = samples[i][2][samples[i][1].index(r)]
                    if not np.array_equal(got, w):
                        raise RuntimeError(f"served query {i} row {r} ({gr['name']}) differs "
                                           f"from the oracle")
                    verified += 1
        edges = 0
        for k, gr in enumerate(graphs):
            sel = [q for q, i in zip(qs, gi) if i == k]
            if sel:
                edges += edges_of(gr["g"], np.concatenate(sel))
        results.append({"lambda_qps": lam, "queries": n_queries,
                        "sources": int(sum(len(q) for q in qs)),
                        "p50_ms": round(float(np.percentile(lats, 50)), 3),
                        "p99_ms": round(float(np.percentile(lats, 99)), 3),
                        "mean_ms": round(float(lats.mean()), 3),
                        "gteps": round(edges / span_s / 1e9, 3), "span_ms": round(span_s * 1e3, 2),
                        "pool_waits": waits, "verified_rows": verified})
        log(f"served lambda={lam}/s: p50 {results[-1]['p50_ms']} ms p99 {results[-1]['p99_ms']} ms "
            f"{results[-1]['gteps']} GTEPS over {results[-1]['span_ms']} ms "
            f"({verified} rows checked, {waits} pool waits)")
    info = {"graphs": [gr["name"] for gr in graphs], "lane_bits": K,
            "cohort_lanes": (f"narrowest of 64..{K} holding the waiting sources"
                             if not admit_window else K),
            "admit_window": 0 if admit_window is None else admit_window,
            "server_options": opts,
            "contexts": [s.contexts for s in servers],
            "admitted_in_flight": [s.counters()["admitted_in_flight"] for s in servers],
            "output_pool_rows": [gr["pool_rows"] for gr in graphs],
            "latency": "completion - scheduled arrival", "rates": results}
    for s_ in servers:
        s_.close()
    for p_ in pools:
        p_.close()
    return info


def config_leg(name, h, src0, threads, K=128, batch_opts=None, contexts=2, burst=4, steps=5,
               cpu_n=64):
    """One configuration of the workload list: plain groups of K lanes and the
    batched launcher on the same GroupBy-ordered sources (each output checked
    against the oracle on up to cpu_n spread rows), the bottom-up kernel's
    algorithmic HBM fraction, and the oracle's GTEPS on those rows."""
    import oracle  # checker and CPU baseline only
    beg, adj = h.arrays()
    V = h.vert_count
    g = ibfs_amd.DeviceGraph(h)
    src = g.groupby(src0)
    n = len(src)
    out = ibfs_amd.DeviceBuffer(4 * n * V)
    p = plain_leg(g, src, K, steps, 1, {}, 0, out)
    rows = np.arange(0, n, max(1, n // cpu_n))[:cpu_n]
    t0 = time.perf_counter()
    want = oracle.bfs_many(beg, adj, src[rows], threads=threads)
    cpu_s = time.perf_counter() - t0
    cpu_edges = sum(oracle.traversed_edges(beg, d) for d in want)
    ok = verify_sample(out, rows, want, V, name + " plain")
    roof, kernels, _ = roofline_leg(g, src, K, V, {}, p["kern_ms"], p["launches"], steps)
    b = ibfs_amd.Batch(g, contexts=contexts, burst=burst, **(batch_opts or {}))
    b.run(src, out=out)
    edges = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(src, out=out)
        edges += b.last_stats.traversed_edges
    bt = (time.perf_counter() - t0) / steps
    ok = verify_sample(out, rows, want, V, name + " batched") and ok
    b.close()
    out.free()
    bu = kernels["expand_bu"]
    rec = {"sources": n, "plain_gteps": round(p["gteps"], 1),
           "plain_ms_per_step": round(p["ms_per_step"], 3),
           "batched_gteps": round(edges / steps / bt / 1e9, 1),
           "batched_ms_per_step": round(bt * 1e3, 3),
           "bu_alg_fraction": round((bu["alg_GBps"] or 0.0) / HBM_PEAK_GBS, 3),
           "dominant_kernel": roof["kernel"], "dominant_fraction": roof["frac"],
           "cpu_gteps": round(cpu_edges / cpu_s / 1e9, 3), "cpu_threads": threads,
           "verified_rows": len(rows), "verified": bool(ok)}
    log(f"{name}: {json.dumps(rec)}")
    return g, rec


def all_configs_leg(threads):
    """The workload list beyond the headline: R-MAT s16 with 128 sources,
    Chung-Lu 3.07M with 512 sources, and R-MAT s18 with every vertex as a
    source for closeness (plain GTEPS on 4,096 of them; the batched
    closeness sums and reach over all, checked against the oracle on 64)."""
    import oracle  # checker and CPU baseline only
    recs = {}
    h = ibfs_amd.HostCsr.rmat(16, 16, 0.57, 0.19, 0.19, seed=1, permute=True, threads=threads)
    g, recs["rmat_s16_128src"] = config_leg("rmat_s16_128src", h, h.gen_sources(128, seed=2),
                                            threads, batch_opts={"lane_bits": 128}, contexts=1,
                                            steps=20, cpu_n=128)
    g.close()
    h = ibfs_amd.HostCsr.chung_lu(3_070_000, 2.2, 76.0, 33000.0, seed=1, threads=threads)
    g, recs["chung_lu_3.07m_512src"] = config_leg("chung_lu_3.07m_512src", h,
                                                  h.gen_sources(512, seed=2), threads,
                                                  batch_opts={"lane_bits": 256})
    g.close()
    h = ibfs_amd.HostCsr.rmat(18, 16, 0.57, 0.19, 0.19, seed=1, permute=True, threads=threads)
    beg, adj = h.arrays()
    V = h.vert_count
    g = ibfs_amd.DeviceGraph(h)
    sample = g.groupby(h.gen_sources(4096, seed=2))
    out = ibfs_amd.DeviceBuffer(4 * len(sample) * V)
    p = plain_leg(g, sample, 128, 3, 1, {}, 0, out)
    _, kernels, _ = roofline_leg(g, sample, 128, V, {}, p["kern_ms"], p["launches"], 3)
    out.free()
    allv = np.arange(V, dtype=np.int32)
    b = ibfs_amd.Batch(g, contexts=2, burst=4, lane_bits=256)
    b.closeness(allv[:4096])
    t0 = time.perf_counter()
    dsum, reach = b.closeness(allv)
    ct = time.perf_counter() - t0
    edges = b.last_stats.traversed_edges
    b.close()
    chk = np.arange(0, V, V // 64)[:64].astype(np.int32)
    t0 = time.perf_counter()
    want = oracle.bfs_many(beg, adj, chk, threads=threads)
    cpu_s = time.perf_counter() - t0
    cpu_edges = sum(oracle.traversed_edges(beg, d) for d in want)
    ok = (np.array_equal(dsum[chk], np.where(want >= 0, want, 0).sum(axis=1))
          and np.array_equal(reach[chk], (want >= 0).sum(axis=1)))
    recs["rmat_s18_closeness_all_sources"] = {
        "sources": V, "plain_gteps": round(p["gteps"], 1), "plain_sample_sources": len(sample),
        "batched_closeness_gteps": round(edges / ct / 1e9, 1),
        "batched_closeness_s": round(ct, 3),
        "bu_alg_fraction": round((kernels["expand_bu"]["alg_GBps"] or 0.0) / HBM_PEAK_GBS, 3),
        "cpu_gteps": round(cpu_edges / cpu_s / 1e9, 3), "cpu_threads": threads,
        "verified_rows": len(chk), "verified": bool(ok)}
    log(f"rmat_s18_closeness_all_sources: {json.dumps(recs['rmat_s18_closeness_all_sources'])}")
    g.close()
    return recs


def host_cpu_info():
    """Host cores: those this process may run on, the machine's count, the
    cgroup CPU quota if one is set, and the lscpu model name."""
    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "nproc": os.cpu_count()}
    try:
        q = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if q and q[0] != "max":
            info["cgroup_cpu_quota"] = round(int(q[0]) / int(q[1]), 2)
    except (OSError, ValueError, IndexError):
        pass
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("Socket(s):"):
                info["sockets"] = int(line.split(":", 1)[1])
    except Exception:
        pass
    return info


def verify_sample(out, rows, want, V, what):
    """Rows `rows` of a device depth buffer against the oracle's arrays."""
    for r, w in zip(rows, want):
        got = out.to_host(count=V, offset=int(r) * V)
        if not np.array_equal(got, w):
            log(f"PARITY FAILURE: {what} row {r} differs from the CPU oracle")
            return False
    return True


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 10 timed steps after 3 warm-up steps: ≈0.16 s of plain and ≈0.1 s of
    # batched timed work; 3 steps after 1 warm-up read the batched step
    # ≈1.5% slower and noisier
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="s22", choices=sorted(WORKLOADS))
    ap.add_argument("--lane-bits", type=int, default=128)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--emit-mode", type=int, default=None)
    ap.add_argument("--tune", action="store_true",
                    help="run the batched autotuner on this workload, save its report under "
                         "ibfs_amd/tuned/ (the batched leg's settings) and exit")
    ap.add_argument("--tune-out", default=None, help="report path for --tune")
    ap.add_argument("--hubs", type=int, default=None)
    ap.add_argument("--alpha", type=int, default=None)
    ap.add_argument("--beta", type=int, default=None)
    ap.add_argument("--inspect-vpt", type=int, default=None)
    ap.add_argument("--phase-offsets", type=int, default=None,
                    help="batched leg: lanes start at levels aligned by two-hop size (1) or all "
                         "at level 0 (0); default: the tuned setting")
    ap.add_argument("--tail-div", type=int, default=None,
                    help="batched leg: sparse tail levels once the frontier has at most "
                         "V/tail_div vertices and E/tail_div edges (0: never); default: the "
                         "tuned setting, else the library default")
    ap.add_argument("--batch-contexts", type=int, default=None)
    ap.add_argument("--burst", type=int, default=None)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched and served legs only (profiling runs)")
    ap.add_argument("--serve-queries", type=int, default=1000)
    ap.add_argument("--serve-rates", default="200,800,2000,4000")
    ap.add_argument("--serve-s22-rates", default="2000,4000",
                    help="rates of the R-MAT-only served line (empty: none)")
    ap.add_argument("--admit-window", type=int, default=None)
    ap.add_argument("--serve-contexts", type=int, default=1,
                    help="traversal contexts per served graph (status budget)")
    ap.add_argument("--serve-s22-contexts", type=int, default=3,
                    help="traversal contexts of the R-MAT-only served line")
    ap.add_argument("--serve-continuous-rates", default="800,2000,4000",
                    help="rates of the served line with in-flight admission (sources join "
                         "running cohorts within 2 levels of their start; empty: none)")
    ap.add_argument("--serve-phase-offsets", type=int, default=None,
                    help="phase offsets of served cohorts (default: the batched leg's tuned value)")
    ap.add_argument("--all-configs", action="store_true",
                    help="measure the other workloads instead (R-MAT s16 x 128, Chung-Lu "
                         "3.07M x 512, R-MAT s18 all-source closeness) and print their JSON")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device 4 times (HBM counter yardstick)")
    args = ap.parse_args()
    tune = {}
    if args.alpha is not None:
        tune["alpha"] = args.alpha
    if args.beta is not None:
        tune["beta"] = args.beta
    if args.inspect_vpt is not None:
        tune["inspect_vpt"] = args.inspect_vpt
    if args.phase_offsets is not None:
        tune["phase_offsets"] = args.phase_offsets
    if args.tail_div is not None:
        tune["tail_div"] = args.tail_div
    if args.gpus != 1:
        print(f"bench.py measures one GPU: libibfs has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); source groups would shard across GPUs with no exchange.",
              file=sys.stderr)
        return 2
    if ibfs_amd.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 3
    if args.all_configs:
        threads = max(1, min(16, host_cpu_info()["affinity_cpus"]))
        print(json.dumps({"configs": all_configs_leg(threads), "host": host_cpu_info()}))
        return 0
    copy_cal = None
    if args.copy_calibration > 0:
        # plain device copy of known size: the yardstick for HBM byte counters
        nb = args.copy_calibration << 20
        a, b = ibfs_amd.DeviceBuffer(nb), ibfs_amd.DeviceBuffer(nb)
        b.copy_from(a)
        t0 = time.perf_counter()
        for _ in range(3):
            b.copy_from(a)
        dt = (time.perf_counter() - t0) / 3
        copy_cal = {"bytes": nb, "copies": 4, "ms": round(dt * 1e3, 3),
                    "GBps_read_plus_write": round(2 * nb / dt / 1e9, 1)}
        log(f"copy calibration: {nb >> 20} MiB in {dt * 1e3:.3f} ms")
        a.free()
        b.free()
    cpu = host_cpu_info()
    # host threads: the cores this process may use, at most the 16-core
    # share of one GPU on these machines
    threads = max(1, min(16, cpu["affinity_cpus"]))
    wl = WORKLOADS[args.workload]
    log(f"generating R-MAT scale {wl['scale']} ({threads} threads)")
    t0 = time.perf_counter()
    h = ibfs_amd.HostCsr.rmat(wl["scale"], 16, 0.57, 0.19, 0.19, seed=1, permute=True,
                              threads=threads)
    V, E = h.vert_count, h.edge_count
    log(f"graph V={V} E={E} in {time.perf_counter() - t0:.1f}s")
    src0 = h.gen_sources(wl["sources"], seed=2)
    g = ibfs_amd.DeviceGraph(h, device=0, hub_count=args.hubs)
    g.groupby(src0)
    t0 = time.perf_counter()
    src = g.groupby(src0)
    groupby_ms = (time.perf_counter() - t0) * 1e3
    K = args.lane_bits
    n = len(src)
    if args.tune:
        path = Path(args.tune_out) if args.tune_out else TUNED_DIR / f"batch_rmat_{args.workload}.json"
        path.parent.mkdir(parents=True, exist_ok=True)
        log(f"batched autotuner on the {n} sources")
        t0 = time.perf_counter()
        best, c, b, rep = ibfs_amd.autotune_batch(g, src, contexts=2, burst=4, passes=2, reps=2)
        rep["workload"] = f"rmat_s{wl['scale']}_ef16_{n}src_groupby"
        rep["seconds"] = round(time.perf_counter() - t0, 1)
        path.write_text(json.dumps(rep, indent=1) + "\n")
        log(f"best {rep['best']} at {rep['best_gteps']} GTEPS; report in {path}")
        return 0
    emit_mode = 0 if args.emit_mode is None else args.emit_mode
    out = ibfs_amd.DeviceBuffer(4 * n * V)
    # the sample of sources checked against the oracle (and timed on the
    # CPU): spread over every group
    stride = max(1, n // max(1, min(args.cpu_sample, n)))
    sample_rows = np.arange(0, n, stride)[:args.cpu_sample]
    sample = src[sample_rows]
    beg = adj = want = None
    cpu_rec = None
    if not args.no_cpu:
        import oracle  # CPU baseline and checker only
        beg, adj = h.arrays()
        cth = args.cpu_threads or threads
        log(f"cpu baseline: {len(sample)} sources on {cth} threads")
        t0 = time.perf_counter()
        want = oracle.bfs_many(beg, adj, sample, threads=cth)
        cpu_s = time.perf_counter() - t0
        cpu_edges = sum(oracle.traversed_edges(beg, d) for d in want)
        cpu_rec = {"value": round(cpu_edges / cpu_s / 1e9, 4), "unit": "GTEPS", "cores": cth,
                   "kind": "port",
                   "sample": f"textbook BFS of {len(sample)} of the {n} GroupBy-ordered sources "
                             f"(every {stride}th) on the same graph, {cpu_s:.1f} s",
                   "host": cpu}
        one = sample[:8]
        t0 = time.perf_counter()
        want1 = oracle.bfs_many(beg, adj, one, threads=1)
        one_s = time.perf_counter() - t0
        one_edges = sum(oracle.traversed_edges(beg, d) for d in want1)
        cpu_rec["single_thread"] = {"value": round(one_edges / one_s / 1e9, 4), "unit": "GTEPS",
                                    "cores": 1, "sample": f"{len(one)} sources, {one_s:.1f} s"}
        log(f"cpu {cpu_rec['value']} GTEPS on {cth} threads, {cpu_rec['single_thread']['value']} "
            f"on one")

    plain = roof = kernels = levels = None
    verified = {}
    if not args.no_plain:
        plain_tune = dict(PLAIN_SETTINGS.get(args.workload, {}), **tune)
        plain = plain_leg(g, src, K, args.steps, args.warmup, plain_tune, emit_mode, out)
        log(f"plain: {plain['ms_per_step']:.2f} ms/step, {plain['gteps']:.1f} GTEPS "
            f"({plain['events_ms_per_step']:.2f} ms/step with kernel events)")
        if want is not None:  # untimed: the plain leg's own output
            if not verify_sample(out, sample_rows, want, V, "plain"):
                return 4
            verified["plain"] = len(sample_rows)
        roof, kernels, levels = roofline_leg(g, src, K, V, plain_tune, plain["kern_ms"],
                                             plain["launches"], args.steps)
        roof["traffic"], roof["traffic_source"] = measured_traffic(args.workload, K, roof["kernel"])
        if roof["traffic"]:
            roof["alg_bytes_per_launch"] = round(kernels[roof["kernel"]]["alg_GB_per_step"] * 1e9
                                                 / max(1, kernels[roof["kernel"]]["launches_per_step"]))
        log(f"dominant kernel {roof['kernel']}: {roof['achieved']:.0f} GB/s algorithmic")

    batched = None
    if not args.no_batch:
        out_b = ibfs_amd.DeviceBuffer(4 * n * V)
        batched = batch_leg(g, src, args.steps, args.warmup, tune, args.emit_mode, out_b,
                            args.workload, args.batch_contexts, args.burst)
        batched["gteps"] = round(batched["gteps"], 3)
        if plain:
            batched["vs_plain"] = round(batched["gteps"] / plain["gteps"], 3)
        ratio = f" ({batched['vs_plain']}x plain)" if "vs_plain" in batched else ""
        log(f"batched: {batched['ms_per_step']:.2f} ms/step, {batched['gteps']} GTEPS{ratio}")
        if want is not None:
            if not verify_sample(out_b, sample_rows, want, V, "batched"):
                return 4
            verified["batched"] = len(sample_rows)
        out_b.free()
        if beg is None:
            beg, adj = h.arrays()
        batched["kernels_alg"] = batch_alg_bytes(g, beg, adj, src, batched["config"])
    if verified:
        log(f"device depth arrays verified bit-exact against the oracle: {verified}")
    out.free()

    served = served_s22 = served_cont = None
    if not args.no_serve:
        log("generating Chung-Lu graph (3.07M vertices, gamma 2.2, mean 76, cap 33000)")
        t0 = time.perf_counter()
        hc = ibfs_amd.HostCsr.chung_lu(3_070_000, 2.2, 76.0, 33000.0, seed=1, threads=threads)
        gc = ibfs_amd.DeviceGraph(hc, device=0)
        log(f"Chung-Lu V={hc.vert_count} E={hc.edge_count} in {time.perf_counter() - t0:.1f}s")

        # status budget of one traversal context per graph (32 snapshots of
        # 256-lane words per vertex with an edge): with admission at cohort
        # starts, one context per graph gave the largest cohorts and the best
        # latency and throughput (3 + 2 contexts: p50 21 ms and 9.7K GTEPS at
        # 2,000 queries/s; 1 + 1: 11 ms and 10.3K)
        def one_context(gr):
            return args.serve_contexts * 32 * gr.traversal_vertices * 32 + 4096
        check = not args.no_cpu
        if check and beg is None:
            beg, adj = h.arrays()
        cbeg, cadj = hc.arrays() if check else (None, None)
        # output rows per graph: room for the queries in flight at the highest
        # rate (about 60 GB of depth arrays each)
        graphs = [
            {"name": f"rmat_s{wl['scale']}", "g": g, "pool": h.gen_sources(4096, seed=3),
             "status_bytes": one_context(g), "beg": beg, "adj": adj,
             "pool_rows": int(60e9 // (4 * V))},
            {"name": "chung_lu_3.07M", "g": gc, "pool": hc.gen_sources(4096, seed=3),
             "status_bytes": one_context(gc), "beg": cbeg, "adj": cadj,
             "pool_rows": int(60e9 // (4 * hc.vert_count))},
        ]
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        po = args.serve_phase_offsets
        if po is None:
            po = int(batch_settings(args.workload)[0].get("phase_offsets", 0))
        sopts = {"phase_offsets": po}
        served = served_leg(graphs, rates, args.serve_queries, seed=3, K=256,
                            admit_window=args.admit_window, check=check, threads=threads,
                            server_opts=sopts)
        rates_c = [float(x) for x in args.serve_continuous_rates.split(",") if x]
        if rates_c:
            log("served, in-flight admission (window of 2 levels)")
            served_cont = served_leg(graphs, rates_c, args.serve_queries, seed=5, K=256,
                                     admit_window=2, check=check, threads=threads,
                                     server_opts=sopts)
        rates1 = [float(x) for x in args.serve_s22_rates.split(",") if x]
        if rates1:
            # one graph resident: three contexts overlap cohorts (the two-graph
            # mix already overlaps its two graphs' cohorts); measured at 4,000
            # queries/s: 1 context 12.0K GTEPS (p50 154 ms), 2: 12.8K, 3: 13.3K
            # (p50 127 ms)
            log(f"served, R-MAT s22 alone ({args.serve_s22_contexts} contexts)")
            alone = [dict(graphs[0], status_bytes=args.serve_s22_contexts * 32
                          * g.traversal_vertices * 32 + 4096)]
            served_s22 = served_leg(alone, rates1, args.serve_queries, seed=4, K=256,
                                    admit_window=args.admit_window, check=check, threads=threads,
                                    server_opts=sopts)

    res = {
        "metric": METRIC,
        "value": round(plain["gteps"], 3) if plain else None,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(plain["ms_per_step"], 3) if plain else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"rmat_s{wl['scale']}_ef16_{n}src_{K}lane_groups_groupby",
                   "scale": wl["scale"], "edge_factor": 16, "V": V, "E_directed": E,
                   "V_with_edges": g.traversal_vertices,
                   "sources": n, "lane_bits": K, "groups": (n + K - 1) // K,
                   "graph_seed": 1, "source_seed": 2, "permuted_ids": True,
                   "policy": (f"alpha_beta({plain['cfg'].alpha},{plain['cfg'].beta})"
                              if plain else None),
                   "engine_settings": ({"wg_size": plain["cfg"].wg_size,
                                        "sml_shed": plain["cfg"].sml_shed,
                                        "lrg_shed": plain["cfg"].lrg_shed} if plain else None),
                   "depth_out": "device int32 [sources][V]"},
        "roofline": roof,
        "cpu_baseline": cpu_rec,
        "graph500_gteps": round(plain["gteps"] / 2, 3) if plain else None,
        "groupby_ms": round(groupby_ms, 3),
        "plain_with_kernel_events_ms_per_step": (round(plain["events_ms_per_step"], 3)
                                                  if plain else None),
        "kernels": kernels,
        "batched": batched,
        "served": served,
        "served_rmat_only": served_s22,
        "served_continuous": served_cont,
        "levels_per_group": levels,
        "verified_sources": verified,
    }
    if copy_cal:
        res["copy_calibration"] = copy_cal
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
