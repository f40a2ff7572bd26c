"""Bit-exact parity of the HIP traversal (through the C ABI) against the CPU
oracle and the golden fixtures.  Needs an MI355X."""
import sys

import numpy as np
import pytest

import ibfs_amd
import oracle
from conftest import GOLDEN, golden_bfs_names, load_golden, symmetric_csr_from_edges

pytestmark = pytest.mark.gpu

POLICIES = {
    "ab": dict(policy=ibfs_amd.POLICY_ALPHA_BETA),
    "td": dict(policy=ibfs_amd.POLICY_TD_ONLY),
    "bu": dict(policy=ibfs_amd.POLICY_BU_ONLY),
    "sw1": dict(policy=ibfs_amd.POLICY_FIXED, sw_level=1),
    "sw2": dict(policy=ibfs_amd.POLICY_FIXED, sw_level=2),
    "hy": dict(policy=ibfs_amd.POLICY_HYBRID),
    # alpha/beta with hybrid levels whenever pulling is priced at 1/64 of a push
    "abh": dict(policy=ibfs_amd.POLICY_ALPHA_BETA, hybrid_gamma=1),
    # alpha/beta without hybrid levels
    "ab0": dict(policy=ibfs_amd.POLICY_ALPHA_BETA, hybrid_gamma=0),
}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def s16():
    h = ibfs_amd.HostCsr.rmat(16, 16, seed=1)
    beg, adj = h.arrays()
    src = h.gen_sources(128, seed=2)
    g = ibfs_amd.DeviceGraph(h)
    want = oracle.bfs_many(beg, adj, src, threads=8)
    return dict(h=h, beg=beg, adj=adj, src=src, g=g, want=want)


@pytest.mark.parametrize("name", golden_bfs_names())
@pytest.mark.parametrize("pol", list(POLICIES))
def test_golden_all_sources(name, pol):
    say(f"[gpu] golden {name} policy {pol}")
    z = load_golden(name)
    h = ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"])
    g = ibfs_amd.DeviceGraph(h)
    eng = ibfs_amd.Engine(g, **POLICIES[pol])
    V = len(z["beg_pos"]) - 1
    src = np.arange(V, dtype=np.int32)
    got = eng.run_groups(src)
    np.testing.assert_array_equal(got, z["depth"])


@pytest.mark.parametrize("lane_bits", [64, 128, 256, 512])
@pytest.mark.parametrize("pol", ["ab", "bu", "td", "hy", "abh"])
def test_s16_128_sources(s16, lane_bits, pol):
    say(f"[gpu] s16 lane_bits={lane_bits} policy={pol}")
    eng = ibfs_amd.Engine(s16["g"], lane_bits=lane_bits, **POLICIES[pol])
    got = eng.run_groups(s16["src"])
    np.testing.assert_array_equal(got, s16["want"])
    st = eng.last_stats
    want_edges = sum(oracle.traversed_edges(s16["beg"], d) for d in s16["want"])
    assert st.traversed_edges == want_edges


@pytest.mark.parametrize("sml,lrg,wg", [(1, 1, 256), (32, 512, 512), (128, 2048, 1024),
                                         (16384, 65536, 1024)])
def test_s16_degree_class_thresholds(s16, sml, lrg, wg):
    """The degree-class thresholds (thread / wave / workgroup classes of the
    top-down frontier and the bottom-up candidates, inspector.cu:130-134) and
    the inspector's workgroup size only move vertices between queues (1/1:
    every vertex of degree above 1 in the workgroup class; 16384/65536: every
    vertex in the thread class, the largest degree being 9,752),
    and the depths equal the oracle's under each policy."""
    for pol in ("ab", "td", "bu", "hy"):
        say(f"[gpu] s16 thresholds sml={sml} lrg={lrg} wg={wg} policy={pol}")
        eng = ibfs_amd.Engine(s16["g"], sml_shed=sml, lrg_shed=lrg, wg_size=wg, **POLICIES[pol])
        np.testing.assert_array_equal(eng.run_groups(s16["src"]), s16["want"])


def test_s16_fixed_switch_every_level(s16):
    src = s16["src"][:64]
    for sw in range(0, 8):
        say(f"[gpu] s16 fixed sw={sw}")
        eng = ibfs_amd.Engine(s16["g"], policy=ibfs_amd.POLICY_FIXED, sw_level=sw)
        np.testing.assert_array_equal(eng.traversal(src), s16["want"][:64])


def test_partial_group_duplicates_and_isolated_sources():
    say("[gpu] duplicates / isolated")
    z = load_golden("bfs_twocomp.npz")
    h = ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"])
    g = ibfs_amd.DeviceGraph(h)
    src = np.array([7, 0, 0, 3, 7, 6, 3], dtype=np.int32)  # 7 is isolated
    for pol in POLICIES:
        eng = ibfs_amd.Engine(g, **POLICIES[pol])
        got = eng.traversal(src)
        np.testing.assert_array_equal(got, z["depth"][src])


def _directed_with_sinks(V=400, m=1500, seed=4):
    """Random directed CSR where 1/4 of the vertices have no out-edges (some
    of them reachable sinks) and vertices V-10..V-1 have no edges at all."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, V - 10, size=m)
    b = rng.integers(0, V - 10, size=m)
    keep = (a % 4 != 0) & (a != b)
    a, b = a[keep], b[keep]
    key = np.unique(a.astype(np.int64) * V + b)
    a, b = (key // V).astype(np.int64), (key % V).astype(np.int32)
    beg = np.zeros(V + 1, dtype=np.int64)
    np.add.at(beg, a + 1, 1)
    return np.cumsum(beg), b


def test_directed_sinks_and_isolated_top_down():
    """Top-down only (the direction that is defined on a directed CSR):
    sinks (out-degree 0, reachable) and isolated vertices keep exact depths
    in every runner."""
    say("[gpu] directed sinks / isolated")
    beg, adj = _directed_with_sinks()
    V = len(beg) - 1
    deg = np.diff(beg)
    indeg = np.bincount(adj, minlength=V)
    assert ((deg == 0) & (indeg > 0)).any() and ((deg == 0) & (indeg == 0)).any()
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(beg, adj))
    assert g.traversal_vertices == int(((deg > 0) | (indeg > 0)).sum())
    src = np.arange(V, dtype=np.int32)
    want = oracle.bfs_many(beg, adj, src, threads=8)
    td = POLICIES["td"]
    np.testing.assert_array_equal(ibfs_amd.Engine(g, **td).run_groups(src), want)
    b = ibfs_amd.Batch(g, contexts=2, burst=3, ring_levels=5, **td)
    np.testing.assert_array_equal(b.run(src), want)
    dsum, reach = b.closeness(src)
    want_sum, want_reach = _close_ref(want)
    np.testing.assert_array_equal(reach, want_reach)
    np.testing.assert_array_equal(dsum, want_sum)
    srv = ibfs_amd.Server(g, lane_bits=64, ring_levels=8, **td)
    queries = [src[i:i + n] for i, n in ((390, 10), (0, 64), (100, 3), (395, 5), (200, 150))]
    outs, _ = _serve(srv, queries, V, gap_s=0.001)
    for q, o in zip(queries, outs):
        np.testing.assert_array_equal(o, want[q])
    srv.close()


def test_server_isolated_sources():
    say("[gpu] server isolated sources")
    z = load_golden("bfs_twocomp.npz")
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
    V = len(z["beg_pos"]) - 1
    assert g.traversal_vertices == V - 1  # vertex 7 has no edge
    queries = [np.array(q, dtype=np.int32) for q in ([7], [7, 7, 0], [3, 7, 6, 1], [7] * 70)]
    srv = ibfs_amd.Server(g, lane_bits=64, ring_levels=8)
    outs, _ = _serve(srv, queries, V, gap_s=0.001)
    for q, o in zip(queries, outs):
        np.testing.assert_array_equal(o, z["depth"][q])
    srv.close()


def _degenerate(case):
    """(beg, adj) of the degenerate graphs: one vertex, no edges at all (no
    vertex has a status word), a single undirected edge among isolated
    vertices, and a star whose centre is the only hub."""
    if case == "one_vertex":
        return np.zeros(2, dtype=np.int64), np.zeros(0, dtype=np.int32)
    if case == "no_edges":
        return np.zeros(6, dtype=np.int64), np.zeros(0, dtype=np.int32)
    if case == "one_edge":
        return symmetric_csr_from_edges(5, [(1, 3)])
    return symmetric_csr_from_edges(300, [(0, i) for i in range(1, 300)])


@pytest.mark.parametrize("case", ["one_vertex", "no_edges", "one_edge", "star"])
def test_degenerate_graphs_every_runner(case):
    """Every runner on graphs with no or almost no edges: engine (host and
    device output, the group loop), GroupBy, batched runs and closeness, a
    server, and the allocation-free entry point.  Depths equal the oracle's;
    a source's reach counts the source itself."""
    say(f"[gpu] degenerate graph {case}")
    beg, adj = _degenerate(case)
    V = len(beg) - 1
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(beg, adj))
    deg = np.diff(beg)
    indeg = np.bincount(adj, minlength=V) if len(adj) else np.zeros(V, dtype=np.int64)
    assert g.traversal_vertices == int(((deg > 0) | (indeg > 0)).sum())
    src = np.array([0, V - 1, 0, V // 2, min(1, V - 1), min(3, V - 1)] * 2, dtype=np.int32)
    want = oracle.bfs_many(beg, adj, src, threads=4)
    eng = ibfs_amd.Engine(g, lane_bits=64)
    np.testing.assert_array_equal(eng.traversal(src), want)
    buf = ibfs_amd.DeviceBuffer(4 * len(src) * V)
    eng.traversal(src, out=buf)
    np.testing.assert_array_equal(buf.to_host().reshape(len(src), V), want)
    np.testing.assert_array_equal(eng.run_groups(src), want)
    assert eng.last_stats.traversed_edges == sum(oracle.traversed_edges(beg, d) for d in want)
    np.testing.assert_array_equal(g.groupby(src), oracle.groupby(beg, adj, src))
    b = ibfs_amd.Batch(g, contexts=2, burst=4, lane_bits=64)
    np.testing.assert_array_equal(b.run(src), want)
    dsum, reach = b.closeness(src)
    d64 = want.astype(np.int64)
    np.testing.assert_array_equal(reach, (d64 >= 0).sum(axis=1))
    np.testing.assert_array_equal(dsum, np.where(d64 >= 0, d64, 0).sum(axis=1))
    b.close()
    srv = ibfs_amd.Server(g, lane_bits=64, ring_levels=8)
    queries = [src[:1], src, src[3:7]]
    outs, _ = _serve(srv, queries, V)
    for q, o in zip(queries, outs):
        np.testing.assert_array_equal(o, oracle.bfs_many(beg, adj, q, threads=4))
    srv.close()
    g.traversal_noalloc(src, buf)
    np.testing.assert_array_equal(buf.to_host().reshape(len(src), V), want)
    eng.close()


def test_lane_position_invariance():
    say("[gpu] lane position")
    z = load_golden("bfs_grid6x7.npz")
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
    eng = ibfs_amd.Engine(g, lane_bits=512)
    for lane in (0, 63, 64, 127, 128, 255, 256, 383, 448, 511):
        src = np.full(lane + 1, 5, dtype=np.int32)
        src[lane] = 17
        got = eng.traversal(src)
        np.testing.assert_array_equal(got[lane], z["depth"][17])


@pytest.mark.parametrize("ring", [4, 5, 8, 255])
def test_ring_wrap_long_path(ring):
    say(f"[gpu] ring {ring}")
    z = load_golden("bfs_classes1500.npz")
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
    src = np.array([0, 700, 801, 1499, 1200, 5, 5], dtype=np.int32)
    eng = ibfs_amd.Engine(g, ring_levels=ring)
    got = eng.traversal(src)
    np.testing.assert_array_equal(got, z["depth"][src])
    if ring < 255:
        assert eng.last_stats.emit_windows > 1


@pytest.mark.parametrize("hubs", [0, 3, 1000, None])
@pytest.mark.parametrize("emit_mode", [0, 1, 2, 3])
def test_hub_cache_and_emission_variants(hubs, emit_mode):
    say(f"[gpu] hubs={hubs} emit_mode={emit_mode}")
    for name in ("bfs_classes1500.npz", "bfs_er300.npz", "bfs_path10.npz"):
        z = load_golden(name)
        h = ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"])
        g = ibfs_amd.DeviceGraph(h, hub_count=hubs)
        V = len(z["beg_pos"]) - 1
        for pol in ("ab", "bu"):
            for ring in (4, 32):
                eng = ibfs_amd.Engine(g, emit_mode=emit_mode, ring_levels=ring, **POLICIES[pol])
                src = np.arange(0, V, max(1, V // 200), dtype=np.int32)
                np.testing.assert_array_equal(eng.run_groups(src), z["depth"][src],
                                              err_msg=f"{name} {pol} ring={ring}")


@pytest.mark.parametrize("lane_bits", [64, 128, 256, 512])
def test_wide_emission_widths_and_row_alignment(s16, lane_bits):
    """Emission mode 3 (a wave per status word of 256 vertices, 16-byte
    stores): every width, rows 16-byte aligned and not (a device output at a
    4-byte offset takes the 4-byte fallback), vertex counts not divisible by
    4 (golden graphs of 9, 42 and 1,500 vertices), ring wrap (continuation
    windows skip entries), the engine, the batched launcher and the server."""
    say(f"[gpu] wide emission lanes={lane_bits}")
    V = s16["h"].vert_count
    n = len(s16["src"])
    buf = ibfs_amd.DeviceBuffer(4 * n * V + 16)
    for off in (0, 4):
        view = ibfs_amd.DeviceView(buf, off, 4 * n * V)
        eng = ibfs_amd.Engine(s16["g"], lane_bits=lane_bits, emit_mode=3)
        eng.run_groups(s16["src"], out=view)
        np.testing.assert_array_equal(view.to_host().reshape(n, V), s16["want"], err_msg=f"off {off}")
        b = ibfs_amd.Batch(s16["g"], contexts=2, burst=4, lane_bits=lane_bits, emit_mode=3)
        b.run(s16["src"], out=view)
        np.testing.assert_array_equal(view.to_host().reshape(n, V), s16["want"], err_msg=f"off {off}")
    for name in ("bfs_cycle9.npz", "bfs_grid6x7.npz", "bfs_classes1500.npz"):
        z = load_golden(name)
        g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
        Vg = len(z["beg_pos"]) - 1
        src = np.arange(Vg, dtype=np.int32)
        for ring in (4, 32):
            eng = ibfs_amd.Engine(g, lane_bits=lane_bits, emit_mode=3, ring_levels=ring)
            np.testing.assert_array_equal(eng.run_groups(src), z["depth"], err_msg=f"{name} {ring}")
            b = ibfs_amd.Batch(g, contexts=2, burst=3, lane_bits=lane_bits, emit_mode=3,
                               ring_levels=ring)
            np.testing.assert_array_equal(b.run(src), z["depth"], err_msg=f"{name} {ring}")
        srv = ibfs_amd.Server(g, lane_bits=lane_bits, emit_mode=3, ring_levels=8, admit_window=2)
        queries = [src[i::7] for i in range(7)]
        outs, _ = _serve(srv, queries, Vg, gap_s=0.0005)
        for q, o in zip(queries, outs):
            np.testing.assert_array_equal(o, z["depth"][q], err_msg=name)
        srv.close()


@pytest.mark.parametrize("vpt", [1, 4, 8, 16])
def test_inspector_chunking(s16, vpt):
    say(f"[gpu] inspect_vpt={vpt}")
    for lb in (64, 256):
        eng = ibfs_amd.Engine(s16["g"], lane_bits=lb, inspect_vpt=vpt)
        np.testing.assert_array_equal(eng.run_groups(s16["src"]), s16["want"])
    z = load_golden("bfs_classes1500.npz")
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
    V = len(z["beg_pos"]) - 1
    src = np.arange(V, dtype=np.int32)
    b = ibfs_amd.Batch(g, contexts=2, burst=4, inspect_vpt=vpt)
    np.testing.assert_array_equal(b.run(src), z["depth"])


@pytest.mark.parametrize("hubs", [0, 64, None])
def test_s16_hub_cache(s16, hubs):
    say(f"[gpu] s16 hubs={hubs}")
    g = ibfs_amd.DeviceGraph(s16["h"], hub_count=hubs)
    for lb in (128, 256):
        eng = ibfs_amd.Engine(g, lane_bits=lb, policy=ibfs_amd.POLICY_BU_ONLY)
        np.testing.assert_array_equal(eng.run_groups(s16["src"]), s16["want"])


@pytest.mark.parametrize("lane_bits", [64, 128, 256, 512])
@pytest.mark.parametrize("contexts,burst", [(1, 1), (2, 3), (4, 8)])
def test_batch_s16(s16, lane_bits, contexts, burst):
    say(f"[gpu] batch s16 lanes={lane_bits} ctx={contexts} burst={burst}")
    b = ibfs_amd.Batch(s16["g"], contexts=contexts, burst=burst, lane_bits=lane_bits)
    src = np.concatenate([s16["src"], s16["src"][:37]])  # uneven: 165 sources
    want = np.concatenate([s16["want"], s16["want"][:37]])
    np.testing.assert_array_equal(b.run(src), want)
    st = b.last_stats
    assert st.traversed_edges == sum(oracle.traversed_edges(s16["beg"], d) for d in want)
    V = s16["h"].vert_count
    buf = ibfs_amd.DeviceBuffer(4 * len(src) * V)
    b.run(src, out=buf)
    np.testing.assert_array_equal(buf.to_host().reshape(len(src), V), want)


@pytest.mark.parametrize("pol", list(POLICIES))
def test_batch_golden_policies_and_ring_wrap(pol):
    say(f"[gpu] batch golden {pol}")
    for name in ("bfs_classes1500.npz", "bfs_twocomp.npz", "bfs_grid6x7.npz"):
        z = load_golden(name)
        g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
        V = len(z["beg_pos"]) - 1
        src = np.arange(V, dtype=np.int32)
        for ring in (4, 7, 32):
            b = ibfs_amd.Batch(g, contexts=3, burst=5, ring_levels=ring, **POLICIES[pol])
            np.testing.assert_array_equal(b.run(src), z["depth"], err_msg=f"{name} ring={ring}")


def _close_ref(depth):
    d = depth.astype(np.int64)
    reached = d >= 0
    return np.where(reached, d, 0).sum(axis=1), reached.sum(axis=1)


@pytest.mark.parametrize("lane_bits", [64, 128, 256, 512])
def test_closeness_matches_depth_arrays_s16(s16, lane_bits):
    say(f"[gpu] closeness s16 lanes={lane_bits}")
    b = ibfs_amd.Batch(s16["g"], contexts=2, burst=4, lane_bits=lane_bits)
    src = np.concatenate([s16["src"], s16["src"][:21]])
    dsum, reach = b.closeness(src)
    want_sum, want_reach = _close_ref(np.concatenate([s16["want"], s16["want"][:21]]))
    np.testing.assert_array_equal(reach, want_reach)
    np.testing.assert_array_equal(dsum, want_sum)


@pytest.mark.parametrize("ring", [4, 6, 32])
def test_closeness_golden_all_sources(ring):
    say(f"[gpu] closeness golden ring={ring}")
    for name in ("bfs_classes1500.npz", "bfs_twocomp.npz", "bfs_er300.npz"):
        z = load_golden(name)
        g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
        V = len(z["beg_pos"]) - 1
        b = ibfs_amd.Batch(g, contexts=2, burst=3, ring_levels=ring)
        dsum, reach = b.closeness(np.arange(V, dtype=np.int32))
        want_sum, want_reach = _close_ref(z["depth"])
        np.testing.assert_array_equal(reach, want_reach, err_msg=name)
        np.testing.assert_array_equal(dsum, want_sum, err_msg=name)


def _serve(server, queries, V, gap_s=0.0):
    import threading
    import time
    bufs = [ibfs_amd.DeviceBuffer(4 * len(q) * V) for q in queries]
    ids = [None] * len(queries)

    def producer():
        for i, q in enumerate(queries):
            ids[i] = server.submit(q, bufs[i])
            if gap_s:
                time.sleep(gap_s)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    infos = [server.wait(i, timeout_ms=60000) for i in ids]
    outs = [b.to_host().reshape(len(q), V) for b, q in zip(bufs, queries)]
    return outs, infos


@pytest.mark.parametrize("lane_bits", [64, 256, 512])
def test_server_mixed_equals_alone_s16(s16, lane_bits):
    say(f"[gpu] server s16 lanes={lane_bits}")
    rng = np.random.default_rng(3)
    V = s16["h"].vert_count
    pool = s16["h"].gen_sources(2000, seed=11)
    sizes = np.floor(np.exp(rng.uniform(0, np.log(256), size=24))).astype(int)
    queries = [rng.choice(pool, size=int(n), replace=True).astype(np.int32) for n in sizes]
    srv = ibfs_amd.Server(s16["g"], status_bytes=0, lane_bits=lane_bits)
    outs, infos = _serve(srv, queries, V, gap_s=0.002)
    for q, o, inf in zip(queries, outs, infos):
        assert inf["done"] and inf["latency_ms"] >= 0
        np.testing.assert_array_equal(o, oracle.bfs_many(s16["beg"], s16["adj"], q, 8))
    # one query alone on a fresh server gives the same arrays as in the mix
    alone = ibfs_amd.Server(s16["g"], lane_bits=lane_bits)
    o2, _ = _serve(alone, [queries[5]], V)
    np.testing.assert_array_equal(o2[0], outs[5])
    srv.close()
    alone.close()


@pytest.mark.parametrize("host_levels", [True, False])
@pytest.mark.parametrize("ring,window", [(8, 1), (9, 0), (32, 1), (8, 1 << 20), (32, 3)])
def test_server_ring_flush_and_recycling(ring, window, host_levels):
    say(f"[gpu] server ring={ring} admit_window={window} host_levels={host_levels}")
    z = load_golden("bfs_classes1500.npz")
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
    V = len(z["beg_pos"]) - 1
    rng = np.random.default_rng(ring)
    queries = [rng.integers(0, V, size=int(n)).astype(np.int32) for n in (1, 7, 64, 3, 100, 2, 40)]
    # 64 lanes < 217 sources
    srv = ibfs_amd.Server(g, lane_bits=64, ring_levels=ring, admit_window=window,
                          host_levels=host_levels)
    outs, infos = _serve(srv, queries, V, gap_s=0.001)
    for q, o in zip(queries, outs):
        np.testing.assert_array_equal(o, z["depth"][q])
    srv.close()


def test_autotune_returns_a_valid_faster_or_equal_config(s16, tmp_path):
    say("[gpu] autotune")
    cfg, rep = ibfs_amd.autotune(s16["g"], s16["src"][:128], passes=1, reps=1,
                                 save_to=str(tmp_path / "tune.json"))
    assert rep["best_gteps"] >= rep["trials"][0]["gteps"]
    assert cfg.lane_bits in (64, 128, 256, 512)
    eng = ibfs_amd.Engine(s16["g"], cfg)
    np.testing.assert_array_equal(eng.run_groups(s16["src"]), s16["want"])
    assert (tmp_path / "tune.json").exists()


def test_autotune_batch_returns_a_valid_config(s16, tmp_path):
    say("[gpu] autotune batch")
    src = s16["src"][:64]
    cfg, ctx, burst, rep = ibfs_amd.autotune_batch(s16["g"], src, passes=1, reps=1,
                                                   save_to=str(tmp_path / "tune.json"))
    assert rep["mode"] == "batch" and rep["best_gteps"] >= rep["trials"][0]["gteps"]
    assert 1 <= ctx <= 4 and burst in (4, 8, 16) and cfg.wg_size in (256, 512, 1024)
    b = ibfs_amd.Batch(s16["g"], contexts=ctx, burst=burst, cfg=cfg)
    np.testing.assert_array_equal(b.run(src), s16["want"][:64])
    back = ibfs_amd.config_from_dict(rep["best"])
    assert back.dd_grid == cfg.dd_grid and back.lane_bits == cfg.lane_bits


@pytest.mark.parametrize("wg", [256, 1024])
def test_inspector_workgroup_sizes(s16, wg):
    say(f"[gpu] wg_size={wg}")
    for lb in (64, 128, 256, 512):
        eng = ibfs_amd.Engine(s16["g"], lane_bits=lb, wg_size=wg)
        np.testing.assert_array_equal(eng.run_groups(s16["src"]), s16["want"])
    b = ibfs_amd.Batch(s16["g"], contexts=2, burst=4, lane_bits=256, wg_size=wg)
    np.testing.assert_array_equal(b.run(s16["src"]), s16["want"])


def test_device_output_and_determinism(s16):
    say("[gpu] device output")
    eng = ibfs_amd.Engine(s16["g"])
    V = s16["h"].vert_count
    buf = ibfs_amd.DeviceBuffer(4 * 128 * V)
    eng.traversal(s16["src"], out=buf)
    got = buf.to_host().reshape(128, V)
    np.testing.assert_array_equal(got, s16["want"])
    again = eng.traversal(s16["src"])
    np.testing.assert_array_equal(again, got)


def test_level_log_matches_oracle_model(s16):
    say("[gpu] level log")
    src = s16["src"][:100]
    eng = ibfs_amd.Engine(s16["g"], policy=ibfs_amd.POLICY_BU_ONLY)
    eng.traversal(src)
    log = eng.level_log()
    m = oracle.level_model(s16["beg"], s16["adj"], src, dirs=[1] * 64)
    assert len(log) == m["levels"] + 1
    for L, li in enumerate(log):
        assert li["n_changed"] == m["n_changed"][L], L
        assert li["n_frontier"] == m["n_front"][L], L
        assert li["m_frontier"] == m["m_front"][L], L
        assert li["n_unvisited"] == m["n_cand"][L], L
        assert li["m_unvisited"] == m["m_cand"][L], L
        assert sum(li["q_bu"]) == m["n_cand"][L]


def test_groupby_known_answer_gpu():
    say("[gpu] groupby known")
    z = load_golden("groupby_known.npz")
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
    np.testing.assert_array_equal(g.groupby(z["src"]), z["expected"])


def test_groupby_matches_oracle_s16(s16):
    say("[gpu] groupby s16")
    src = s16["h"].gen_sources(3000, seed=9)
    got = s16["g"].groupby(src)
    np.testing.assert_array_equal(got, oracle.groupby(s16["beg"], s16["adj"], src))


@pytest.mark.parametrize("n", [1, 2, 255, 1_000_000])
def test_groupby_large_and_tiny_s16(s16, n):
    """The placement is a stable radix sort: a million sources drawn with
    repeats (every vertex ~15 times, isolated ones included) keep input order
    among equal hub keys, as do one and two sources."""
    say(f"[gpu] groupby s16 n={n}")
    src = np.random.default_rng(4).integers(0, s16["h"].vert_count, size=n).astype(np.int32)
    got = s16["g"].groupby(src)
    np.testing.assert_array_equal(got, oracle.groupby(s16["beg"], s16["adj"], src))


def test_reference_mirror_classes(tmp_path):
    say("[gpu] mirror classes")
    h = ibfs_amd.HostCsr.rmat(12, 16, seed=4)
    h.save(str(tmp_path / "g_beg_pos.bin"), str(tmp_path / "g_csr.bin"))
    gr = ibfs_amd.graph(str(tmp_path / "g_beg_pos.bin"), str(tmp_path / "g_csr.bin"), 256)
    gr.gen_src(seed=3)
    before = gr.src_list.copy()
    gr.groupby()
    assert sorted(before.tolist()) == sorted(gr.src_list.tolist())
    np.testing.assert_array_equal(gr.src_list, oracle.groupby(gr.beg_pos, gr.csr, before))
    inst = ibfs_amd.gpu_ibfs(gr, 128, sw_level=2)
    for grp in range(2):
        d = inst.traversal(grp * 128)
        want = oracle.bfs_many(gr.beg_pos, gr.csr, gr.src_list[grp * 128:(grp + 1) * 128], 8)
        np.testing.assert_array_equal(d, want)


def test_chung_lu_small():
    say("[gpu] chung-lu")
    h = ibfs_amd.HostCsr.chung_lu(30000, 2.2, 30, 3000, seed=5)
    beg, adj = h.arrays()
    src = h.gen_sources(256, seed=6)
    g = ibfs_amd.DeviceGraph(h)
    for lb in (128, 256):
        eng = ibfs_amd.Engine(g, lane_bits=lb)
        np.testing.assert_array_equal(eng.run_groups(src), oracle.bfs_many(beg, adj, src, 8))


def test_engine_argument_errors(s16):
    with pytest.raises(ibfs_amd.IbfsError):
        ibfs_amd.Engine(s16["g"], lane_bits=96)
    eng = ibfs_amd.Engine(s16["g"])
    with pytest.raises(ibfs_amd.IbfsError):
        eng.traversal(np.arange(129, dtype=np.int32))
    with pytest.raises(ibfs_amd.IbfsError):
        eng.traversal(np.array([70000], dtype=np.int32))
    # the engine still works after rejected calls
    np.testing.assert_array_equal(eng.traversal(s16["src"][:3]), s16["want"][:3])


def test_cli_runs_reference_arguments(tmp_path):
    say("[gpu] cli")
    import subprocess
    from conftest import ROOT
    h = ibfs_amd.HostCsr.rmat(12, 16, seed=4)
    h.save(str(tmp_path / "g_beg_pos.bin"), str(tmp_path / "g_csr.bin"))
    r = subprocess.run([str(ROOT / "ibfs_amd" / "gpu-ibfs"), str(tmp_path / "g_beg_pos.bin"),
                        str(tmp_path / "g_csr.bin"), "128", "2", "256"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Traversal-time:") == 2 and "Total time:" in r.stdout


@pytest.mark.parametrize("cut", [0, 100])
@pytest.mark.parametrize("hubs", [0, 64, 4096, None])
@pytest.mark.parametrize("pol", ["hy", "abh"])
def test_hybrid_levels_every_hub_split(s16, hubs, pol, cut):
    """Hybrid levels (candidates pull through their prefix of pulled hubs,
    the other frontier vertices push) with no hubs (all push), few, many and
    the default count (every vertex a hub on s16: all pull), pulling through
    every cached hub (cut 0) or the top 100, through the engine, the batched
    launcher and both serving paths."""
    say(f"[gpu] hybrid hubs={hubs} policy={pol} cut={cut}")
    g = ibfs_amd.DeviceGraph(s16["h"], hub_count=hubs) if hubs is not None else s16["g"]
    src, want = s16["src"], s16["want"]
    pol = dict(POLICIES[pol], hybrid_hubs=cut)
    for lb in (64, 256):
        eng = ibfs_amd.Engine(g, lane_bits=lb, kernel_timing=1, **pol)
        np.testing.assert_array_equal(eng.run_groups(src), want)
        dirs = [x["direction"] for x in eng.level_log()]
        if pol["policy"] == ibfs_amd.POLICY_HYBRID:
            assert set(dirs[1:]) == {2} and eng.last_stats.n_hy_launches > 0
        b = ibfs_amd.Batch(g, contexts=2, burst=3, lane_bits=lb, **pol)
        np.testing.assert_array_equal(b.run(src), want)
    V = s16["h"].vert_count
    for window in (0, 2):
        srv = ibfs_amd.Server(g, lane_bits=64, ring_levels=8, admit_window=window, **pol)
        queries = [src[0:40], src[40:41], src[41:128]]
        outs, _ = _serve(srv, queries, V, gap_s=0.0005)
        for q, o in zip(queries, outs):
            np.testing.assert_array_equal(o, oracle.bfs_many(s16["beg"], s16["adj"], q, threads=8))
        srv.close()


def test_hybrid_is_chosen_on_hub_heavy_levels(s16):
    """With pulling priced low, an alpha/beta traversal of R-MAT runs its
    hub-heavy top-down levels hybrid, and the non-hub frontier edges it
    reports never exceed the frontier edges."""
    g = ibfs_amd.DeviceGraph(s16["h"], hub_count=2048)
    eng = ibfs_amd.Engine(g, lane_bits=128, **POLICIES["abh"])
    np.testing.assert_array_equal(eng.traversal(s16["src"]), s16["want"])
    log = eng.level_log()
    assert 2 in [x["direction"] for x in log]
    assert all(0 <= x["m_frontier_nonhub"] <= x["m_frontier"] for x in log)
    # no hubs: every frontier edge is a non-hub edge
    g0 = ibfs_amd.DeviceGraph(s16["h"], hub_count=0)
    eng0 = ibfs_amd.Engine(g0, lane_bits=128)
    eng0.traversal(s16["src"])
    assert all(x["m_frontier_nonhub"] == x["m_frontier"] for x in eng0.level_log())


def test_direction_config_checks(s16):
    for bad in (dict(policy=5), dict(hybrid_gamma=-1), dict(hybrid_gamma=5000),
                dict(hybrid_hubs=-1), dict(split_cov=-1), dict(split_cov=101)):
        with pytest.raises(ibfs_amd.IbfsError):
            ibfs_amd.Engine(s16["g"], **bad)
        with pytest.raises(ibfs_amd.IbfsError):
            ibfs_amd.Batch(s16["g"], **bad)
        with pytest.raises(ibfs_amd.IbfsError):
            ibfs_amd.Server(s16["g"], **bad)


def test_hub_entries_match_the_adjacency(s16):
    """ibfs_graph_hub_entries counts the adjacency entries pointing to the
    top hubs by degree (ties by id), for every cut."""
    beg, adj = s16["beg"], s16["adj"]
    deg = np.diff(beg)
    nz = np.flatnonzero((deg > 0) | (np.bincount(adj, minlength=len(deg)) > 0))
    order = nz[np.argsort(-deg[nz], kind="stable")]
    g = ibfs_amd.DeviceGraph(s16["h"], hub_count=5000)
    for top in (1, 100, 4096, 5000, 0, 10 ** 9):
        r = 5000 if top <= 0 or top > 5000 else top
        hub = np.zeros(len(deg), bool)
        hub[order[:r]] = True
        assert g.hub_entries(top) == int(hub[adj].sum()), top


def test_server_cohort_widths_and_window_changes(s16):
    """One server whose admission window changes between batches of queries:
    device-driven cohorts at the narrowest width that holds their sources
    (64/128/256 lanes: query sizes 1, 100, 200) alternate with host-driven
    full-width levels on the same contexts; every depth array equals the
    oracle's."""
    say("[gpu] server cohort widths / window changes")
    V = s16["h"].vert_count
    pool = s16["h"].gen_sources(600, seed=21)
    srv = ibfs_amd.Server(s16["g"], lane_bits=256, ring_levels=8)
    off = 0
    for window, host in ((0, False), (3, True), (0, False), (1, True), (2, False), (0, True)):
        srv.set_admission(window)
        srv.set_levels(host)
        sizes = (1, 100, 200) if window == 0 else (5, 40)
        queries = []
        for n in sizes:
            queries.append(pool[off:off + n])
            off = (off + n) % (len(pool) - 200)
        outs, _ = _serve(srv, queries, V, gap_s=0.002)
        for q, o in zip(queries, outs):
            np.testing.assert_array_equal(o, oracle.bfs_many(s16["beg"], s16["adj"], q, threads=8))
    srv.close()


@pytest.mark.parametrize("lane_bits", [64, 256])
def test_server_sources_join_running_device_cohorts(s16, lane_bits):
    """Device-driven cohorts with an admission window: queries that arrive
    while a cohort runs its first levels join it at a level boundary (the
    device-side admission list between that level's expansion and
    inspection); their depths are relative to their own admission level and
    equal the oracle's, and the same queries alone give the same arrays."""
    say(f"[gpu] server in-flight admission lanes={lane_bits}")
    V = s16["h"].vert_count
    pool = s16["h"].gen_sources(3000, seed=31)
    rng = np.random.default_rng(lane_bits)
    queries = [rng.choice(pool, size=int(n), replace=False).astype(np.int32)
               for n in (1, 3, 2, 9, 1, 17, 4, 2, 30, 5, 1, 12)]
    srv = ibfs_amd.Server(s16["g"], lane_bits=lane_bits, admit_window=3, ring_levels=16)
    outs, infos = _serve(srv, queries, V, gap_s=0.0002)
    for q, o in zip(queries, outs):
        np.testing.assert_array_equal(o, oracle.bfs_many(s16["beg"], s16["adj"], q, threads=8))
    cnt = srv.counters()
    assert cnt["admitted"] == sum(len(q) for q in queries)
    srv.close()
    alone = ibfs_amd.Server(s16["g"], lane_bits=lane_bits, admit_window=3, ring_levels=16)
    o2, _ = _serve(alone, [queries[5]], V)
    np.testing.assert_array_equal(o2[0], outs[5])
    alone.close()


@pytest.mark.parametrize("p,window", [(2, 0), (3, 0), (2, 2), (1, 3)])
def test_server_phase_offsets(s16, p, window):
    """Served cohorts whose first lanes start at levels set by their sources'
    two-hop sizes (phase offsets), alone and with sources joining running
    cohorts (admission window): isolated, degree-1 and hub sources mixed in
    queries of 1-200 sources; every depth array equals the oracle's."""
    say(f"[gpu] server phase offsets p={p} window={window}")
    beg, adj = s16["beg"], s16["adj"]
    deg = np.diff(beg)
    V = s16["h"].vert_count
    rng = np.random.default_rng(p * 10 + window)
    pool = np.concatenate([s16["h"].gen_sources(400, seed=41), np.flatnonzero(deg == 1)[:50],
                           np.argsort(-deg)[:20], np.flatnonzero(deg == 0)[:5]]).astype(np.int32)
    queries = [rng.choice(pool, size=int(n), replace=False).astype(np.int32)
               for n in (1, 200, 7, 60, 3, 120, 1, 33)]
    srv = ibfs_amd.Server(s16["g"], lane_bits=256, ring_levels=8, phase_offsets=p,
                          admit_window=window)
    outs, _ = _serve(srv, queries, V, gap_s=0.0003)
    for q, o in zip(queries, outs):
        np.testing.assert_array_equal(o, oracle.bfs_many(beg, adj, q, threads=8))
    srv.close()


def test_server_in_flight_admission_on_a_long_path():
    """A path of 4,000 vertices: a cohort started by a source at one end runs
    for thousands of levels, so the queries submitted after it has started
    all join it in flight (unbounded admission window) at different levels,
    across many ring flushes (16 snapshots).  Depths are |i - s| exactly."""
    say("[gpu] server in-flight admission, path graph")
    V = 4000
    edges = [(i, i + 1) for i in range(V - 1)]
    beg, adj = symmetric_csr_from_edges(V, edges)
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(beg, adj))
    srv = ibfs_amd.Server(g, lane_bits=64, admit_window=1 << 20, ring_levels=16)
    first = np.array([0], dtype=np.int32)
    b0 = ibfs_amd.DeviceBuffer(4 * V)
    q0 = srv.submit(first, b0)
    import time
    t0 = time.perf_counter()
    while srv.counters()["admitted"] < 1:
        assert time.perf_counter() - t0 < 30, "first query never admitted"
        time.sleep(0.0001)
    rng = np.random.default_rng(9)
    queries = [rng.integers(0, V, size=int(n)).astype(np.int32) for n in (1, 5, 2, 9, 3)]
    outs, _ = _serve(srv, queries, V, gap_s=0.001)
    srv.wait(q0, timeout_ms=60000)
    idx = np.arange(V)
    np.testing.assert_array_equal(b0.to_host(), idx)
    for q, o in zip(queries, outs):
        np.testing.assert_array_equal(o, np.abs(idx[None, :] - q[:, None]))
    cnt = srv.counters()
    assert cnt["admitted_in_flight"] == sum(len(q) for q in queries), cnt
    srv.close()


@pytest.mark.parametrize("lane_bits", [64, 256, 512])
def test_batch_phase_offsets(s16, lane_bits):
    """Batched runs whose lanes start at different levels (phase offsets:
    sources with small two-hop sizes start up to two levels earlier): depth
    arrays and closeness sums equal the oracle's, with the ring wrapping
    (5 snapshots) and without, packed and in caller order, and with a
    mix of isolated, low-degree and hub sources."""
    say(f"[gpu] batch phase offsets lanes={lane_bits}")
    h, beg, adj = s16["h"], s16["beg"], s16["adj"]
    deg = np.diff(beg)
    rng = np.random.default_rng(lane_bits)
    lowdeg = np.flatnonzero(deg == 1)[:40].astype(np.int32)
    hubs = np.argsort(-deg)[:10].astype(np.int32)
    iso = np.flatnonzero(deg == 0)[:3].astype(np.int32)
    src = np.concatenate([s16["src"], lowdeg, hubs, iso]).astype(np.int32)
    rng.shuffle(src)
    want = oracle.bfs_many(beg, adj, src, threads=8)
    wsum = np.where(want >= 0, want, 0).sum(axis=1)
    wreach = (want >= 0).sum(axis=1)
    for ring, pack, p in ((5, 0, 1), (5, 1, 2), (32, 0, 3), (32, 1, 2), (32, 1, 1)):
        b = ibfs_amd.Batch(s16["g"], contexts=2, burst=3, lane_bits=lane_bits, ring_levels=ring,
                           phase_offsets=p, pack_sources=pack)
        np.testing.assert_array_equal(b.run(src), want, err_msg=f"ring {ring} pack {pack} p {p}")
        dsum, reach = b.closeness(src)
        np.testing.assert_array_equal(reach, wreach)
        np.testing.assert_array_equal(dsum, wsum)
    z = load_golden("bfs_classes1500.npz")
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
    V = len(z["beg_pos"]) - 1
    gsrc = np.arange(V, dtype=np.int32)
    b = ibfs_amd.Batch(g, contexts=2, burst=3, lane_bits=lane_bits, ring_levels=4, phase_offsets=1)
    np.testing.assert_array_equal(b.run(gsrc), z["depth"])


def test_batch_packing_keeps_caller_order(s16):
    """Batched runs group sources by two-hop size (default) or in the
    caller's order; depth arrays (host and device output) and closeness come
    back in the caller's order either way, equal to the oracle's."""
    say("[gpu] batch packing")
    V = s16["h"].vert_count
    pool = s16["h"].gen_sources(700, seed=31)
    src = np.concatenate([pool, pool[:50]])  # duplicates across groups
    want = oracle.bfs_many(s16["beg"], s16["adj"], src, threads=8)
    want_sum, want_reach = _close_ref(want)
    for pack in (1, 0):
        b = ibfs_amd.Batch(s16["g"], contexts=3, burst=4, lane_bits=128, pack_sources=pack)
        np.testing.assert_array_equal(b.run(src), want)
        buf = ibfs_amd.DeviceBuffer(4 * len(src) * V)
        b.run(src, out=buf)
        np.testing.assert_array_equal(buf.to_host().reshape(len(src), V), want)
        dsum, reach = b.closeness(src)
        np.testing.assert_array_equal(dsum, want_sum)
        np.testing.assert_array_equal(reach, want_reach)
        b.close()


def test_split_levels_every_coverage(s16):
    """Split levels (dense lanes pull over whole lists with atomic commits,
    sparse lanes push the bits they gained one level before) at coverage
    thresholds 1, 50 and 99 %, in caller order so that slow and fast sources
    share groups, with rings of 4 and 32 levels, batched and served: every
    depth array and closeness equals the oracle's, and split levels occur."""
    V = s16["h"].vert_count
    src = s16["h"].gen_sources(512, seed=41)
    want = oracle.bfs_many(s16["beg"], s16["adj"], src, threads=8)
    want_sum, want_reach = _close_ref(want)
    splits = 0
    for cov in (1, 50, 99):
        for ring in (4, 32):
            say(f"[gpu] split levels cov={cov} ring={ring}")
            b = ibfs_amd.Batch(s16["g"], contexts=2, burst=4, lane_bits=256, pack_sources=0,
                               split_cov=cov, ring_levels=ring)
            np.testing.assert_array_equal(b.run(src), want, err_msg=f"cov={cov} ring={ring}")
            splits += b.last_stats.n_split_levels
            dsum, reach = b.closeness(src)
            np.testing.assert_array_equal(dsum, want_sum)
            np.testing.assert_array_equal(reach, want_reach)
            b.close()
        srv = ibfs_amd.Server(s16["g"], lane_bits=256, split_cov=cov)
        queries = [src[:200], src[200:201], src[201:512]]
        outs, _ = _serve(srv, queries, V, gap_s=0.002)
        for q, o in zip(queries, outs):
            np.testing.assert_array_equal(o, oracle.bfs_many(s16["beg"], s16["adj"], q, threads=8))
        srv.close()
    assert splits > 0


@pytest.mark.parametrize("opts", [dict(), dict(policy=ibfs_amd.POLICY_FIXED, sw_level=2),
                                  dict(hybrid_gamma=0), dict(hybrid_gamma=64, hybrid_hubs=64)])
def test_device_plan_chooses_the_host_directions(s16, opts):
    """One group through the host-driven engine and through the batched
    launcher's device-driven plan: both apply the same direction rule to the
    same counters, so the levels and the number of top-down, bottom-up and
    hybrid levels agree (and no split level occurs with split levels off)."""
    say(f"[gpu] device plan directions {opts}")
    src = s16["src"][:128]
    eng = ibfs_amd.Engine(s16["g"], lane_bits=128, **opts)
    eng.traversal(src)
    levels = eng.last_stats.levels
    dirs = [li["direction"] for li in eng.level_log() if 1 <= li["level"] < levels]
    b = ibfs_amd.Batch(s16["g"], contexts=1, burst=3, lane_bits=128, pack_sources=0, **opts)
    b.run(src)
    st = b.last_stats
    assert st.levels == levels
    assert (st.n_td_launches, st.n_bu_launches, st.n_hy_launches, st.n_split_levels) == \
        (dirs.count(0), dirs.count(1), dirs.count(2), 0)
    b.close()


def test_engines_on_threads_share_a_graph(s16):
    """One engine per host thread on a shared graph (the supported threading
    model): three engines of 64, 128 and 256 lanes with kernel timing run
    their groups concurrently, twice each; every depth array equals the
    oracle's and each engine's traversed edges agree across its runs."""
    import threading
    say("[gpu] engines on threads")
    pool = s16["h"].gen_sources(900, seed=51)
    jobs = [(64, pool[:150]), (128, pool[150:450]), (256, pool[450:900])]
    want = [oracle.bfs_many(s16["beg"], s16["adj"], src, threads=8) for _, src in jobs]
    engines = [ibfs_amd.Engine(s16["g"], lane_bits=lb, kernel_timing=1) for lb, _ in jobs]
    got = [[None, None] for _ in jobs]
    edges = [[0, 0] for _ in jobs]
    errors = []

    def work(i):
        try:
            for rep in range(2):
                got[i][rep] = engines[i].run_groups(jobs[i][1])
                edges[i][rep] = engines[i].last_stats.traversed_edges
        except Exception as exc:  # reported on the main thread
            errors.append(exc)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(len(jobs)):
        for rep in range(2):
            np.testing.assert_array_equal(got[i][rep], want[i], err_msg=f"engine {i} run {rep}")
        assert edges[i][0] == edges[i][1] > 0


def test_wait_limit_ends_a_wait_and_the_engine_recovers():
    """A host wait that exceeds the library's wait limit returns
    IBFS_ERR_STATE instead of spinning; every kernel still runs to its end.
    With the limit restored, the same engine and the same batch (whose next
    run first drains the failed run's work) and a new engine give the
    oracle's depths.  The condition is forced from the host with a tiny limit
    on runs of several milliseconds."""
    say("[gpu] wait limit")
    h = ibfs_amd.HostCsr.rmat(20, 16, seed=3)
    beg, adj = h.arrays()
    g = ibfs_amd.DeviceGraph(h)
    src = h.gen_sources(512, seed=4)
    want = oracle.bfs_many(beg, adj, src[:16], threads=8)
    old = ibfs_amd.get_wait_timeout()
    eng = ibfs_amd.Engine(g, lane_bits=512)
    bat = ibfs_amd.Batch(g, lane_bits=256, contexts=2, burst=4)
    buf = ibfs_amd.DeviceBuffer(4 * len(src) * h.vert_count)
    try:
        try:
            ibfs_amd.set_wait_timeout(50e-6)
            with pytest.raises(ibfs_amd.IbfsError) as ei:
                eng.traversal(src)
            assert ei.value.code == ibfs_amd._lib.IBFS_ERR_STATE
            assert "wait limit" in str(ei.value)
            ibfs_amd.set_wait_timeout(1e-9)
            with pytest.raises(ibfs_amd.IbfsError) as ei:
                bat.run(src, out=buf)
            assert ei.value.code == ibfs_amd._lib.IBFS_ERR_STATE
            assert "wait limit" in str(ei.value)
        finally:
            ibfs_amd.set_wait_timeout(old)
        np.testing.assert_array_equal(eng.traversal(src[:16]), want)
        np.testing.assert_array_equal(bat.run(src[:16]), want)
    finally:
        eng.close()
        bat.close()
        buf.free()
    eng = ibfs_amd.Engine(g, lane_bits=512)
    np.testing.assert_array_equal(eng.traversal(src[:16]), want)


def test_host_copies_every_size_class():
    """Device-to-host copies (ibfs_memcpy_d2h, which DeviceBuffer.to_host and
    host-placed depth arrays use): tiny copies, just under and at the size
    where copies start streaming through the pinned chunks (4 MiB), a 64-MiB
    chunk plus a few bytes, three chunks plus an odd tail; device sources
    at odd offsets, host destinations at odd addresses, nothing written past
    the destination; then two threads copying at once, and a copy whose wait
    passes the limit part-way."""
    import threading
    say("[gpu] host copies")
    L = ibfs_amd._lib.lib()
    rng = np.random.default_rng(5)
    total = 3 * (64 << 20) + 4096 + 7
    pattern = rng.integers(0, 256, size=total, dtype=np.uint8)
    buf = ibfs_amd.DeviceBuffer(total)
    assert L.ibfs_memcpy_h2d(buf.ptr, pattern.ctypes.data, total) == 0
    cases = [(0, 1), (3, 4095), (1, (4 << 20) - 1), (0, 4 << 20), (5, (64 << 20) + 3),
             (0, 3 * (64 << 20) + 7), (11, total - 11)]
    for off, n in cases:
        host = np.full(n + 16, 0xAB, dtype=np.uint8)
        dst = host[1:1 + n]
        assert L.ibfs_memcpy_d2h(dst.ctypes.data, buf.ptr + off, n) == 0, (off, n)
        np.testing.assert_array_equal(dst, pattern[off:off + n], err_msg=f"{off}+{n}")
        assert host[0] == 0xAB and (host[1 + n:] == 0xAB).all()
    np.testing.assert_array_equal(buf.to_host(dtype=np.uint8), pattern)
    got = [None, None]

    def worker(i):
        got[i] = buf.to_host(dtype=np.uint8, count=total - 64 * i, offset=64 * i)

    th = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in range(2):
        np.testing.assert_array_equal(got[i], pattern[64 * i:])
    # a wait past the limit part-way through a pipelined copy (the limit set
    # from the host, the DMA itself ends normally): the call fails with
    # IBFS_ERR_STATE unless the copy won the race, and the next copy (on new
    # pinned chunks) is exact
    old = ibfs_amd.get_wait_timeout()
    host = np.zeros(total, dtype=np.uint8)
    ibfs_amd.set_wait_timeout(1e-9)
    try:
        rc = L.ibfs_memcpy_d2h(host.ctypes.data, buf.ptr, total)
    finally:
        ibfs_amd.set_wait_timeout(old)
    assert rc in (0, ibfs_amd._lib.IBFS_ERR_STATE), rc
    if rc == 0:
        np.testing.assert_array_equal(host, pattern)
    np.testing.assert_array_equal(buf.to_host(dtype=np.uint8), pattern)
    buf.free()
