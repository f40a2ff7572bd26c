"""The allocation-free entry point ibfs_traversal_noalloc against the
engine's ibfs_traversal (the entry point it is a variant of), through the
C ABI: two calls in a row, the second with a smaller graph and fewer
sources, each bit-exact against the engine and the oracle; then the limits.
Needs an MI355X."""
import sys

import numpy as np
import pytest

import ibfs_amd
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_noalloc_matches_traversal_on_two_calls_in_a_row():
    cases = [(16, 128, 2, 1), (13, 37, 5, 4)]  # (scale, sources, graph seed, source seed)
    for scale, k, gs, ss in cases:
        say(f"[gpu] no-alloc: R-MAT s{scale}, {k} sources")
        h = ibfs_amd.HostCsr.rmat(scale, 16, seed=gs)
        beg, adj = h.arrays()
        V = h.vert_count
        g = ibfs_amd.DeviceGraph(h)
        src = h.gen_sources(k, seed=ss)
        eng = ibfs_amd.Engine(g, lane_bits=128)
        want = eng.traversal(src)
        want_edges = eng.last_stats.traversed_edges
        out = ibfs_amd.DeviceBuffer(4 * k * V)
        edges = g.traversal_noalloc(src, out)
        got = out.to_host().reshape(k, V)
        np.testing.assert_array_equal(got, want)
        assert edges == want_edges
        np.testing.assert_array_equal(got, oracle.bfs_many(beg, adj, src, threads=8))


def test_noalloc_policies_ring_flushes_and_isolated_sources():
    """Fixed switch levels, top-down and bottom-up only; a 1,500-vertex
    golden graph whose BFS is longer than the 8-snapshot ring (windows are
    flushed); isolated and duplicate sources; 64 and 128 lanes."""
    from conftest import load_golden
    z = load_golden("bfs_classes1500.npz")
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
    V = len(z["beg_pos"]) - 1
    src = np.array([0, 700, 801, 1499, 1200, 5, 5], dtype=np.int32)
    out = ibfs_amd.DeviceBuffer(4 * 128 * V)
    for pol in (dict(policy=ibfs_amd.POLICY_FIXED, sw_level=2), dict(policy=ibfs_amd.POLICY_TD_ONLY),
                dict(policy=ibfs_amd.POLICY_BU_ONLY), dict()):
        cfg = ibfs_amd.default_config(**pol)
        g.traversal_noalloc(src, out, cfg)
        np.testing.assert_array_equal(out.to_host(count=len(src) * V).reshape(len(src), V),
                                      z["depth"][src], err_msg=str(pol))
    all_src = np.arange(0, V, 13, dtype=np.int32)[:128]
    g.traversal_noalloc(all_src, out)
    np.testing.assert_array_equal(out.to_host(count=len(all_src) * V).reshape(len(all_src), V),
                                  z["depth"][all_src])
    z2 = load_golden("bfs_twocomp.npz")
    g2 = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z2["beg_pos"], z2["adj"]))
    V2 = len(z2["beg_pos"]) - 1
    s2 = np.arange(V2, dtype=np.int32)
    out2 = ibfs_amd.DeviceBuffer(4 * V2 * V2)
    g2.traversal_noalloc(s2, out2)
    np.testing.assert_array_equal(out2.to_host().reshape(V2, V2), z2["depth"])


def test_noalloc_limits_are_rejected():
    h = ibfs_amd.HostCsr.rmat(19, 16, seed=1)
    g = ibfs_amd.DeviceGraph(h)
    out = ibfs_amd.DeviceBuffer(4 * h.vert_count)
    with pytest.raises(ibfs_amd.IbfsError) as ei:
        g.traversal_noalloc(h.gen_sources(1, seed=1), out)
    assert ei.value.code == ibfs_amd._lib.IBFS_ERR_ARG and "scratch limits" in str(ei.value)
    hs = ibfs_amd.HostCsr.rmat(12, 16, seed=1)
    gs = ibfs_amd.DeviceGraph(hs)
    big = ibfs_amd.DeviceBuffer(4 * 129 * hs.vert_count)
    with pytest.raises(ibfs_amd.IbfsError):
        gs.traversal_noalloc(hs.gen_sources(129, seed=1), big)
    with pytest.raises(ibfs_amd.IbfsError):
        gs.traversal_noalloc(hs.gen_sources(4, seed=1), big, ibfs_amd.default_config(td_chunk=1024))
