"""Parity at the benchmarked configurations: R-MAT scale 22 with 1,024
GroupBy-ordered sources (plain 8 x 128 lanes and the batched launcher), and
the serving workload with the R-MAT s22 and the 3.07M-vertex Chung-Lu graph
resident together under Poisson-arriving queries of 1-256 sources.  Every
depth array is compared with the CPU oracle's textbook BFS.  Needs an MI355X."""
import os
import sys
import threading
import time

import numpy as np
import pytest

import ibfs_amd
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    # the host cores this process may run on (at most 16: the box's share)
    return max(1, min(16, len(os.sched_getaffinity(0))))


# settings of the batched leg of the measurement: the autotuner's saved
# report for this workload (ibfs_amd/tuned/batch_rmat_s22.json)
BATCH_OPTS, BATCH_CONTEXTS, BATCH_BURST, _ = ibfs_amd.tuned_batch_settings("rmat_s22")


def test_rmat_s22_1024_sources_plain_and_batched():
    """The headline workload (the reference's group loop, main.cu:76-90, at
    concurrent-count 128, readme.md:27): R-MAT s22 EF16 (seed 1, permuted
    ids), 1,024 sources (seed 2) in GroupBy order, as 8 groups of 128 lanes
    on the engine and through the batched launcher (the tuned settings, then
    512 lanes with the tuned scheduling, then 64 lanes unpacked with no tail).
    Every one of the 1,024 depth arrays of every form equals the oracle's,
    and the traversed-edge counts equal the oracle's sum."""
    th = cpu_threads()
    say(f"[gpu] headline: generating R-MAT s22 ({th} threads)")
    h = ibfs_amd.HostCsr.rmat(22, 16, 0.57, 0.19, 0.19, seed=1, permute=True, threads=th)
    beg, adj = h.arrays()
    V = h.vert_count
    g = ibfs_amd.DeviceGraph(h)
    src = g.groupby(h.gen_sources(1024, seed=2))
    n = len(src)
    out_p = ibfs_amd.DeviceBuffer(4 * n * V)
    out_b = ibfs_amd.DeviceBuffer(4 * n * V)
    say("[gpu] headline: plain 8 x 128 lanes")
    eng = ibfs_amd.Engine(g, lane_bits=128)
    eng.run_groups(src, out=out_p)
    edges_p = eng.last_stats.traversed_edges
    say(f"[gpu] headline: batched {BATCH_OPTS} x {BATCH_CONTEXTS} contexts")
    b = ibfs_amd.Batch(g, BATCH_CONTEXTS, BATCH_BURST, **BATCH_OPTS)
    b.run(src, out=out_b)
    edges_b = b.last_stats.traversed_edges
    assert edges_b == edges_p
    b.close()
    out_w = ibfs_amd.DeviceBuffer(4 * n * V)
    out_n = ibfs_amd.DeviceBuffer(4 * n * V)
    say("[gpu] headline: batched 512 lanes; 64 lanes unpacked, no tail")
    b = ibfs_amd.Batch(g, 2, BATCH_BURST, **dict(BATCH_OPTS, lane_bits=512))
    b.run(src, out=out_w)
    assert b.last_stats.traversed_edges == edges_p
    b.close()
    b = ibfs_amd.Batch(g, 3, 8, lane_bits=64, pack_sources=0, tail_div=0)
    b.run(src, out=out_n)
    assert b.last_stats.traversed_edges == edges_p
    b.close()
    chunk = 64
    host = np.zeros(chunk * V, dtype=np.int32)
    want_edges = 0
    for i in range(0, n, chunk):
        say(f"[gpu] headline: oracle rows {i}..{i + chunk - 1}")
        w = oracle.bfs_many(beg, adj, src[i:i + chunk], threads=th)
        want_edges += sum(oracle.traversed_edges(beg, d) for d in w)
        for form, buf in (("plain", out_p), ("batched", out_b), ("batched 512", out_w),
                          ("batched 64", out_n)):
            got = buf.to_host(count=len(w) * V, offset=i * V, out=host).reshape(w.shape)
            if not np.array_equal(got, w):
                bad = np.argwhere(got != w)
                r, v = bad[0]
                raise AssertionError(
                    f"{form}: {len(bad)} entries differ in source rows {i}..{i + len(w) - 1}; "
                    f"first: row {i + r} (source {src[i + r]}) vertex {v}: {got[r, v]} != {w[r, v]}")
        del w
    assert want_edges == edges_p


def _one_context(g, lanes=256, ring=32):
    return ring * g.traversal_vertices * (lanes // 8) + 4096


def _serve_stream(graphs, nq, rate, seed, label):
    """Submits nq queries as a Poisson stream at `rate` per second from a
    producer thread, each naming one graph and a log-uniform set of 1-256
    distinct sources of its pool, into recycled device rows; checks every row
    of every query against the oracle as it completes.  Returns three of the
    queries (index -> depth arrays) for the alone-vs-mix check."""
    rng = np.random.default_rng(seed)
    gi = rng.integers(0, 2, size=nq)
    sizes = np.floor(np.exp(rng.uniform(0, np.log(256), size=nq))).astype(int)
    idx = [rng.choice(256, size=int(k), replace=False) for k in sizes]
    gaps = rng.exponential(1.0 / rate, size=nq)
    ids, views = [None] * nq, [None] * nq
    err = []
    stop = threading.Event()  # set when the checking side gives up

    def producer():
        try:
            t = time.perf_counter()
            for i in range(nq):
                t += gaps[i]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                gr = graphs[gi[i]]
                while views[i] is None:
                    if stop.is_set():
                        return
                    views[i] = gr["rows"].alloc(len(idx[i]), timeout=1.0)
                ids[i] = gr["srv"].submit(gr["pool"][idx[i]], views[i])
        except BaseException as e:  # reported by the main thread
            err.append(e)

    # one host buffer for the largest query, touched once
    host = np.zeros(256 * max(gr["V"] for gr in graphs), dtype=np.int32)
    tp = threading.Thread(target=producer)
    tp.start()
    checked_rows = 0
    keep = {}  # three queries kept for the alone-vs-mix check
    try:
        for i in range(nq):
            while ids[i] is None:
                if err or not tp.is_alive() and ids[i] is None:
                    raise AssertionError(f"producer stopped: {err}")
                time.sleep(0.0005)
            gr = graphs[gi[i]]
            gr["srv"].wait(ids[i], timeout_ms=300000)
            k = len(idx[i])
            got = views[i].to_host(count=k * gr["V"], out=host).reshape(k, gr["V"])
            for r, j in enumerate(idx[i]):
                if not np.array_equal(got[r], gr["want"][j]):
                    v = int(np.argmax(got[r] != gr["want"][j]))
                    raise AssertionError(
                        f"{label}: query {i} ({gr['name']}, {k} sources) row {r} source "
                        f"{gr['pool'][j]} vertex {v}: {got[r, v]} != {gr['want'][j][v]}")
            checked_rows += k
            if len(keep) < 3 and k >= 2:
                keep[i] = (gi[i], idx[i], got.copy())
            gr["srv"].release(ids[i])
            gr["rows"].free(views[i])
            views[i] = None
            if i % 25 == 0:
                say(f"[gpu] serving ({label}): {i + 1}/{nq} queries checked ({checked_rows} rows)")
    finally:
        stop.set()
        tp.join()
    assert not err, err
    say(f"[gpu] serving ({label}): {nq} queries, {checked_rows} rows equal the oracle")
    return keep


def test_serving_two_graphs_poisson_queries():
    """The serving workload: R-MAT s22 and Chung-Lu 3.07M resident at once,
    one server each (256-lane contexts, the measured phase offsets of cohort
    lanes), queries arriving as a Poisson stream (1,500 per second), each
    naming one graph and a log-uniform set of 1-256 distinct sources drawn
    from a seeded pool of 256 sources per graph; outputs come from recycled
    device row pools.  First 200 queries with admission at cohort starts (the
    measured mode), then 120 with in-flight admission (sources join a running
    context within 2 levels of its cohort's start).  Every row of every query
    equals the oracle's BFS of its source.  Then three of the first queries
    run alone and give the same arrays as in the mix."""
    th = cpu_threads()
    say("[gpu] serving: generating R-MAT s22 and Chung-Lu 3.07M")
    hs = ibfs_amd.HostCsr.rmat(22, 16, 0.57, 0.19, 0.19, seed=1, permute=True, threads=th)
    hc = ibfs_amd.HostCsr.chung_lu(3_070_000, 2.2, 76.0, 33000.0, seed=1, threads=th)
    graphs = []
    for name, h in (("rmat_s22", hs), ("chung_lu", hc)):
        beg, adj = h.arrays()
        pool = h.gen_sources(256, seed=3)
        say(f"[gpu] serving: oracle for the {name} pool")
        want = oracle.bfs_many(beg, adj, pool, threads=th)
        g = ibfs_amd.DeviceGraph(h)
        del beg, adj
        graphs.append(dict(name=name, g=g, V=h.vert_count, pool=pool, want=want,
                           rows=ibfs_amd.RowPool(h.vert_count, 1536),
                           srv=ibfs_amd.Server(g, lane_bits=256, status_bytes=_one_context(g),
                                               phase_offsets=BATCH_OPTS.get("phase_offsets", 0))))
    keep = _serve_stream(graphs, 200, 1500.0, 11, "cohort admission")
    for gr in graphs:
        gr["srv"].set_admission(2)
    _serve_stream(graphs, 120, 1500.0, 12, "in-flight admission")
    inflight = sum(gr["srv"].counters()["admitted_in_flight"] for gr in graphs)
    say(f"[gpu] serving: sources admitted in flight {inflight}")
    for gr in graphs:
        gr["srv"].set_admission(0)
    say("[gpu] serving: alone runs")
    for i, (g_i, q_idx, mixed) in keep.items():
        gr = graphs[g_i]
        view = gr["rows"].alloc(len(q_idx))
        q = gr["srv"].submit(gr["pool"][q_idx], view)
        gr["srv"].wait(q, timeout_ms=300000)
        alone = view.to_host(count=len(q_idx) * gr["V"]).reshape(mixed.shape)
        gr["srv"].release(q)
        gr["rows"].free(view)
        np.testing.assert_array_equal(alone, mixed, err_msg=f"query {i} alone vs in the mix")
    for gr in graphs:
        gr["srv"].close()
        gr["rows"].close()
