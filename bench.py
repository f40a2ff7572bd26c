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
    eng = ibfs_amd.Engine(g, lane_bits=K, kernel_timing=1, emit_mode=emit_mode, **tune)
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
    return dict(gteps=edges / wall / 1e9, ms_per_step=wall * 1e3 / steps, edges_per_step=edges // steps,
                kern_ms={k: v / steps for k, v in kern.items()},
                launches={k: v // steps for k, v in launches.items()}, cfg=eng.cfg)


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
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0].startswith(want):
            try:
                return float(cells[4]), path.name
            except ValueError:
                return None, None
    return None, None


# batched-leg settings found by ibfs_autotune_batch on this workload (R-MAT s22,
# the 1,024 GroupBy-ordered sources): 256 lanes, 2 contexts, bursts of 4,
# lrg_shed 4096, beta 96, top-down hub chunks of 8,192 edges
BATCH_TUNED = {"lrg_shed": 4096, "beta": 96, "td_chunk": 8192}


def batch_leg(g, src, K, contexts, burst, steps, warmup, tune, emit_mode, out):
    opts = dict(BATCH_TUNED, **tune)
    b = ibfs_amd.Batch(g, contexts=contexts, burst=burst, lane_bits=K, emit_mode=emit_mode, **opts)
    for _ in range(warmup):
        b.run(src, out=out)
    edges = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(src, out=out)
        edges += b.last_stats.traversed_edges
    wall = time.perf_counter() - t0
    return dict(gteps=edges / wall / 1e9, ms_per_step=wall * 1e3 / steps, lane_bits=K,
                contexts=contexts, burst=burst, config=opts)


def edges_of(g, sources, chunk=256):
    """Traversed edges of each source set (untimed, through the batched form)."""
    b = ibfs_amd.Batch(g, contexts=1, burst=8, lane_bits=256)
    buf = ibfs_amd.DeviceBuffer(4 * chunk * g.vert_count)
    total = 0
    for i in range(0, len(sources), chunk):
        b.run(sources[i:i + chunk], out=buf)
        total += b.last_stats.traversed_edges
    return total


def served_leg(graphs, rates, n_queries, seed, K, admit_window=None):
    """Poisson arrivals at each rate; every query names one graph and carries
    a log-uniform 1..256 source set.  Latency = done - submit."""
    rng = np.random.default_rng(seed)
    servers = [ibfs_amd.Server(gr["g"], lane_bits=K, status_bytes=gr["status_bytes"],
                               admit_window=admit_window) for gr in graphs]
    checkers = [ibfs_amd.Batch(gr["g"], contexts=1, burst=8, lane_bits=256) for gr in graphs]
    results = []
    for lam in rates:
        gi = rng.integers(0, len(graphs), size=n_queries)
        sizes = np.floor(np.exp(rng.uniform(0, np.log(256), size=n_queries))).astype(int)
        qs = [rng.choice(graphs[i]["pool"], size=int(n)).astype(np.int32) for i, n in zip(gi, sizes)]
        gaps = rng.exponential(1.0 / lam, size=n_queries)
        ids = [None] * n_queries
        # outputs allocated before the arrivals start: no allocation or free
        # inside the measured window
        bufs = [ibfs_amd.DeviceBuffer(4 * len(q) * graphs[i]["g"].vert_count) for q, i in zip(qs, gi)]
        lat = [None] * n_queries

        def producer():
            t_next = time.perf_counter()
            for i in range(n_queries):
                t_next += gaps[i]
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[i] = servers[gi[i]].submit(qs[i], bufs[i])

        def consumer():
            for i in range(n_queries):
                while ids[i] is None:
                    time.sleep(0.0002)
                info = servers[gi[i]].wait(ids[i], timeout_ms=120000)
                lat[i] = info
                servers[gi[i]].release(ids[i])

        tp, tc = threading.Thread(target=producer), threading.Thread(target=consumer)
        tp.start()
        tc.start()
        tp.join()
        tc.join()
        # untimed check: served depth arrays equal the batched form's for the
        # first queries of the rate
        verified = 0
        for i in range(min(3, n_queries)):
            gr = graphs[gi[i]]
            want = checkers[gi[i]].run(qs[i])
            got = bufs[i].to_host(count=len(qs[i]) * gr["g"].vert_count).reshape(want.shape)
            if not np.array_equal(got, want):
                raise RuntimeError(f"served query {i} differs from the batched form")
            verified += 1
        for b in bufs:
            b.free()
        t_first = min(x["submit_ms"] for x in lat)
        t_last = max(x["done_ms"] for x in lat)
        lats = np.array([x["latency_ms"] for x in lat])
        edges = 0
        for k, gr in enumerate(graphs):
            sel = [q for q, i in zip(qs, gi) if i == k]
            if sel:
                edges += edges_of(gr["g"], np.concatenate(sel))
        span_s = (t_last - t_first) / 1e3
        results.append({"lambda_qps": lam, "queries": n_queries,
                        "sources": int(sum(len(q) for q in qs)),
                        "p50_ms": round(float(np.percentile(lats, 50)), 3),
                        "p99_ms": round(float(np.percentile(lats, 99)), 3),
                        "mean_ms": round(float(lats.mean()), 3),
                        "gteps": round(edges / span_s / 1e9, 3), "span_ms": round(span_s * 1e3, 2),
                        "verified_queries": verified})
        log(f"served lambda={lam}/s: p50 {results[-1]['p50_ms']} ms p99 {results[-1]['p99_ms']} ms "
            f"{results[-1]['gteps']} GTEPS over {results[-1]['span_ms']} ms")
    info = {"graphs": [gr["name"] for gr in graphs], "lane_bits": K,
            "cohort_lanes": (f"narrowest of 64..{K} holding the waiting sources"
                             if not admit_window else K),
            "admit_window": 0 if admit_window is None else admit_window,
            "contexts": [s.contexts for s in servers], "rates": results}
    for s in servers:
        s.close()
    return info


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="s22", choices=sorted(WORKLOADS))
    ap.add_argument("--lane-bits", type=int, default=128)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--emit-mode", type=int, default=0)
    ap.add_argument("--hubs", type=int, default=None)
    ap.add_argument("--alpha", type=int, default=None)
    ap.add_argument("--beta", type=int, default=None)
    ap.add_argument("--inspect-vpt", type=int, default=None)
    ap.add_argument("--batch-contexts", type=int, default=2)
    ap.add_argument("--batch-lanes", type=int, default=256)
    ap.add_argument("--burst", type=int, default=4)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-queries", type=int, default=120)
    ap.add_argument("--serve-rates", default="200,800,2000,4000")
    ap.add_argument("--admit-window", type=int, default=None)
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
    if args.gpus != 1:
        print(f"bench.py measures one GPU: libibfs has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); source groups would shard across GPUs with no exchange.",
              file=sys.stderr)
        return 2
    if ibfs_amd.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 3
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
    wl = WORKLOADS[args.workload]
    threads = min(16, os.cpu_count() or 1)
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
    out = ibfs_amd.DeviceBuffer(4 * n * V)

    plain = plain_leg(g, src, K, args.steps, args.warmup, tune, args.emit_mode, out)
    log(f"plain: {plain['ms_per_step']:.2f} ms/step, {plain['gteps']:.1f} GTEPS")
    roof, kernels, levels = roofline_leg(g, src, K, V, tune, plain["kern_ms"], plain["launches"],
                                         args.steps)
    roof["traffic"], roof["traffic_source"] = measured_traffic(args.workload, K, roof["kernel"])
    if roof["traffic"]:
        roof["alg_bytes_per_launch"] = round(kernels[roof["kernel"]]["alg_GB_per_step"] * 1e9
                                             / max(1, kernels[roof["kernel"]]["launches_per_step"]))
    log(f"dominant kernel {roof['kernel']}: {roof['achieved']:.0f} GB/s algorithmic")

    batched = None
    if not args.no_batch:
        batched = batch_leg(g, src, args.batch_lanes, args.batch_contexts, args.burst, args.steps,
                            args.warmup, tune, args.emit_mode, out)
        batched["gteps"] = round(batched["gteps"], 3)
        batched["vs_plain"] = round(batched["gteps"] / plain["gteps"], 3)
        log(f"batched: {batched['ms_per_step']:.2f} ms/step, {batched['gteps']} GTEPS "
            f"({batched['vs_plain']}x plain)")

    cpu = None
    verified = 0
    if not args.no_cpu:
        import oracle  # CPU baseline leg only
        beg, adj = h.arrays()
        sample = src[:min(args.cpu_sample, n)]
        cth = args.cpu_threads or threads
        log(f"cpu baseline: {len(sample)} sources on {cth} threads")
        t0 = time.perf_counter()
        want = oracle.bfs_many(beg, adj, sample, threads=cth)
        cpu_s = time.perf_counter() - t0
        cpu_edges = sum(oracle.traversed_edges(beg, d) for d in want)
        cpu = {"value": round(cpu_edges / cpu_s / 1e9, 4), "unit": "GTEPS", "cores": cth,
               "kind": "port",
               "sample": f"textbook BFS of the first {len(sample)} of the {n} GroupBy-ordered "
                         f"sources on the same graph, {cpu_s:.1f} s"}
        # the same BFS on one thread, over the first two sources of the sample
        one = sample[:2]
        t0 = time.perf_counter()
        want1 = oracle.bfs_many(beg, adj, one, threads=1)
        one_s = time.perf_counter() - t0
        one_edges = sum(oracle.traversed_edges(beg, d) for d in want1)
        cpu["single_thread"] = {"value": round(one_edges / one_s / 1e9, 4), "unit": "GTEPS",
                                "cores": 1, "sample": f"first {len(one)} sources, {one_s:.1f} s"}
        # untimed check: the device depth arrays of the sample equal the CPU ones
        got = out.to_host(count=len(sample) * V).reshape(len(sample), V)
        if not np.array_equal(got, want):
            log("PARITY FAILURE: device depths differ from the CPU oracle")
            return 4
        verified = len(sample)
        log(f"cpu {cpu['value']} GTEPS; {verified} device depth arrays verified bit-exact")
    del out

    served = None
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
            return 32 * gr.traversal_vertices * 32 + 4096
        graphs = [
            {"name": f"rmat_s{wl['scale']}", "g": g, "pool": h.gen_sources(4096, seed=3),
             "status_bytes": one_context(g)},
            {"name": "chung_lu_3.07M", "g": gc, "pool": hc.gen_sources(4096, seed=3),
             "status_bytes": one_context(gc)},
        ]
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        served = served_leg(graphs, rates, args.serve_queries, seed=3, K=256,
                            admit_window=args.admit_window)

    res = {
        "metric": METRIC,
        "value": round(plain["gteps"], 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(plain["ms_per_step"], 3),
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
                   "policy": f"alpha_beta({plain['cfg'].alpha},{plain['cfg'].beta})",
                   "depth_out": "device int32 [sources][V]"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "graph500_gteps": round(plain["gteps"] / 2, 3),
        "groupby_ms": round(groupby_ms, 3),
        "kernels": kernels,
        "batched": batched,
        "served": served,
        "levels_per_group": levels,
        "verified_sources": verified,
    }
    if copy_cal:
        res["copy_calibration"] = copy_cal
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
