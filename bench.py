"""Aggregate GTEPS of concurrent multi-source BFS on one MI355X.

Default workload: R-MAT scale 22, edge factor 16 (a,b,c = .57,.19,.19,
seeded, vertex ids permuted, cleaned and symmetrised), 1,024 seeded sources
of degree > 0, GroupBy-ordered, run as 8 groups of 128 lanes with the
direction-optimising policy.  One step = the 8 groups (all 1,024 depth arrays
written to device memory).  GTEPS = sum over sources of deg(v) over reached v
(directed CSR entries) / wall time of the step.

Prints one JSON line (see the README of the result fields in DESIGN.md).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
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


def alg_bytes_per_kernel(log_levels, V, W, k):
    """Algorithmic bytes of each kernel launch of one traversal, from the
    level log (B=8 offset, A=4 adjacency, Q=4 queue entry, W status bytes)."""
    B, A, Q = 8, 4, 4
    out = {"expand_td": 0, "expand_bu": 0, "inspect": 0, "emit": 0}
    for i, li in enumerate(log_levels):
        L = li["level"]
        q = sum(li["q_td"][:2]) * Q + li["q_td"][2] * 8 + sum(li["q_bu"]) * Q
        reads = 2 if L > 0 else 1
        out["inspect"] += V * (4 + (reads + 1) * W) + q
        if L == 0:
            continue
        prev = log_levels[i - 1]
        if li["direction"] == 0:
            out["expand_td"] += (prev["n_frontier"] * (2 * B + W) + prev["m_frontier"] * (A + W)
                                 + W * li["n_changed"])
        else:
            out["expand_bu"] += (li["bu_work"] * (2 * B + W) + li["bu_prefix"] * (A + W)
                                 + W * li["n_changed"])
    nsnap = len(log_levels) - 1  # snapshots 0..last
    out["emit"] = nsnap * W * V + 4 * k * V
    return out


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
    ap.add_argument("--batch", type=int, default=0,
                    help="contexts of the batched launcher (0: plain engine)")
    ap.add_argument("--burst", type=int, default=8)
    ap.add_argument("--dd-grid", type=int, default=0)
    args = ap.parse_args()
    tune = {}
    if args.alpha is not None:
        tune["alpha"] = args.alpha
    if args.beta is not None:
        tune["beta"] = args.beta
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); source groups would shard across GPUs with no exchange.",
              file=sys.stderr)
        return 2
    if ibfs_amd.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 3
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
    # GroupBy on the device (timed separately, outside the step)
    g.groupby(src0)
    t0 = time.perf_counter()
    src = g.groupby(src0)
    groupby_ms = (time.perf_counter() - t0) * 1e3
    K = args.lane_bits
    n = len(src)
    if args.batch:
        eng = ibfs_amd.Batch(g, contexts=args.batch, burst=args.burst, lane_bits=K,
                             emit_mode=args.emit_mode, dd_grid=args.dd_grid, **tune)
        eng.run_groups = eng.run
    else:
        eng = ibfs_amd.Engine(g, lane_bits=K, kernel_timing=1, emit_mode=args.emit_mode, **tune)
    out = ibfs_amd.DeviceBuffer(4 * n * V)
    log(f"engine ready: lanes={K} groups={(n + K - 1) // K} depth buffer {4 * n * V / 2**30:.1f} GiB")
    for i in range(args.warmup):
        eng.run_groups(src, out=out)
        log(f"warmup {i + 1}/{args.warmup}: {eng.last_stats.total_ms:.2f} ms device")
    kern = {"expand_td": 0.0, "expand_bu": 0.0, "inspect": 0.0, "emit": 0.0}
    launches = {"expand_td": 0, "expand_bu": 0, "inspect": 0, "emit": 0}
    edges = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        eng.run_groups(src, out=out)
        st = eng.last_stats
        edges += st.traversed_edges
        kern["expand_td"] += st.expand_td_ms
        kern["expand_bu"] += st.expand_bu_ms
        kern["inspect"] += st.inspect_ms
        kern["emit"] += st.emit_ms
        launches["expand_td"] += st.n_td_launches
        launches["expand_bu"] += st.n_bu_launches
        launches["inspect"] += st.n_inspect_launches
        launches["emit"] += st.n_emit_launches
    wall = time.perf_counter() - t0
    gteps = edges / wall / 1e9
    ms_per_step = wall * 1e3 / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, {gteps:.1f} GTEPS")

    # algorithmic bytes per kernel from one untimed pass with the work model
    model = ibfs_amd.Engine(g, lane_bits=K, work_model=1, **tune)
    scratch = ibfs_amd.DeviceBuffer(4 * K * V)  # keeps the timed outputs intact
    alg = {"expand_td": 0, "expand_bu": 0, "inspect": 0, "emit": 0}
    levels = []
    for g0 in range(0, n, K):
        grp = src[g0:g0 + K]
        model.traversal(grp, out=scratch)
        lv = model.level_log()
        levels.append([(x["level"], x["direction"]) for x in lv])
        for kname, b in alg_bytes_per_kernel(lv, V, K // 8, len(grp)).items():
            alg[kname] += b
    per_step_ms = {k: v / args.steps for k, v in kern.items()}
    dom = max(per_step_ms, key=per_step_ms.get)
    achieved = alg[dom] / (per_step_ms[dom] * 1e-3) / 1e9 if per_step_ms[dom] > 0 else 0.0
    kernels = {k: {"ms_per_step": round(per_step_ms[k], 4),
                   "launches_per_step": launches[k] // args.steps,
                   "alg_GB_per_step": round(alg[k] / 1e9, 4),
                   "alg_GBps": round(alg[k] / (per_step_ms[k] * 1e-3) / 1e9, 1)
                   if per_step_ms[k] > 0 else None}
               for k in kern}
    log(f"dominant kernel {dom}: {achieved:.0f} GB/s algorithmic")

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
        # untimed check: the device depth arrays of the sample equal the CPU ones
        got = out.to_host(count=len(sample) * V).reshape(len(sample), V)
        if not np.array_equal(got, want):
            log("PARITY FAILURE: device depths differ from the CPU oracle")
            return 4
        verified = len(sample)
        log(f"cpu {cpu['value']} GTEPS; {verified} device depth arrays verified bit-exact")

    res = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"rmat_s{wl['scale']}_ef16_{n}src_{K}lane_groups_groupby",
                   "scale": wl["scale"], "edge_factor": 16, "V": V, "E_directed": E,
                   "sources": n, "lane_bits": K, "groups": (n + K - 1) // K,
                   "graph_seed": 1, "source_seed": 2, "permuted_ids": True,
                   "policy": f"alpha_beta({eng.cfg.alpha},{eng.cfg.beta})", "depth_out": "device int32 [sources][V]"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "graph500_gteps": round(gteps / 2, 3),
        "groupby_ms": round(groupby_ms, 3),
        "kernels": kernels,
        "levels_per_group": levels,
        "verified_sources": verified,
    }
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
