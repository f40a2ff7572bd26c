"""Sparse tail levels of batched runs (ibfs_config::tail_div): once the
frontier is small, the remaining levels push only newly gained bits into one
cumulative status slot and log every gain; the depths of those levels are
written from the log after the emission.  Depth arrays, closeness sums and the
TEPS numerator must equal the oracle's bit for bit, including runs whose log
overflows (the group is rerun without the tail).  Needs an MI355X."""
import sys

import numpy as np
import pytest

import ibfs_amd
import oracle
from conftest import load_golden, symmetric_csr_from_edges

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def s16():
    h = ibfs_amd.HostCsr.rmat(16, 16, seed=1)
    beg, adj = h.arrays()
    g = ibfs_amd.DeviceGraph(h)
    deg = np.diff(beg)
    # ordinary sources, degree-1 sources (long tails), hubs and isolated ones
    src = np.concatenate([h.gen_sources(200, seed=5), np.flatnonzero(deg == 1)[:40],
                          np.argsort(-deg)[:8], np.flatnonzero(deg == 0)[:3]]).astype(np.int32)
    np.random.default_rng(3).shuffle(src)
    want = oracle.bfs_many(beg, adj, src, threads=8)
    return dict(h=h, beg=beg, adj=adj, g=g, src=src, want=want)


def _close_ref(depth):
    d = depth.astype(np.int64)
    return np.where(d >= 0, d, 0).sum(axis=1), (d >= 0).sum(axis=1)


@pytest.mark.parametrize("lane_bits", [64, 256, 512])
@pytest.mark.parametrize("tail_div", [0, 1, 16, 256, 4096])
def test_tail_s16_depths_edges_closeness(s16, lane_bits, tail_div):
    """tail_div 1 enters the tail as soon as the frontier stops growing (dense
    tail levels, many records, overflowing logs at 512 lanes); larger values
    only for the last levels; 0 never.  Host and device outputs, phase
    offsets on and off, ring wrapping (5 snapshots) and not."""
    say(f"[gpu] tail s16 lanes={lane_bits} tail_div={tail_div}")
    src, want = s16["src"], s16["want"]
    V = s16["h"].vert_count
    edges = sum(oracle.traversed_edges(s16["beg"], d) for d in want)
    wsum, wreach = _close_ref(want)
    tails = reruns = 0
    for ring, p in ((32, 0), (5, 0), (32, 3), (5, 2)):
        b = ibfs_amd.Batch(s16["g"], contexts=2, burst=3, lane_bits=lane_bits, ring_levels=ring,
                           phase_offsets=p, tail_div=tail_div)
        np.testing.assert_array_equal(b.run(src), want, err_msg=f"ring {ring} p {p}")
        st = b.last_stats
        assert st.traversed_edges == edges
        tails += st.n_tail_levels
        reruns += st.n_tail_reruns
        buf = ibfs_amd.DeviceBuffer(4 * len(src) * V)
        b.run(src, out=buf)
        np.testing.assert_array_equal(buf.to_host().reshape(len(src), V), want)
        dsum, reach = b.closeness(src)
        np.testing.assert_array_equal(reach, wreach)
        np.testing.assert_array_equal(dsum, wsum)
        b.close()
    say(f"      tail levels {tails} reruns {reruns}")
    if tail_div == 0:
        assert tails == 0 and reruns == 0
    elif tail_div <= 256:
        assert tails > 0


@pytest.mark.parametrize("pol", ["ab", "td", "bu", "sw2"])
def test_tail_golden_all_sources(pol):
    """Every vertex of the golden graphs as a source (scipy-written depths),
    each direction policy before the tail, rings of 4 and 32 snapshots."""
    pols = dict(ab=dict(policy=ibfs_amd.POLICY_ALPHA_BETA),
                td=dict(policy=ibfs_amd.POLICY_TD_ONLY),
                bu=dict(policy=ibfs_amd.POLICY_BU_ONLY),
                sw2=dict(policy=ibfs_amd.POLICY_FIXED, sw_level=2))
    say(f"[gpu] tail golden policy {pol}")
    for name in ("bfs_classes1500.npz", "bfs_twocomp.npz", "bfs_grid6x7.npz", "bfs_er300.npz"):
        z = load_golden(name)
        g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(z["beg_pos"], z["adj"]))
        V = len(z["beg_pos"]) - 1
        src = np.arange(V, dtype=np.int32)
        want_sum, want_reach = _close_ref(z["depth"])
        for ring in (4, 32):
            for tail_div in (1, 64):
                b = ibfs_amd.Batch(g, contexts=3, burst=5, ring_levels=ring, tail_div=tail_div,
                                   **pols[pol])
                msg = f"{name} ring={ring} tail_div={tail_div}"
                np.testing.assert_array_equal(b.run(src), z["depth"], err_msg=msg)
                dsum, reach = b.closeness(src)
                np.testing.assert_array_equal(reach, want_reach, err_msg=msg)
                np.testing.assert_array_equal(dsum, want_sum, err_msg=msg)
                b.close()


def test_tail_long_path_and_log_overflow():
    """A path of 4,000 vertices: from level 2 on every level is a tail level.
    Eight sources fit the log (one record per lane and vertex); 64 sources
    append about 64 x 4,000 records, more than the log holds (traversal
    vertices + 65,536), so the group is rerun without the tail.  Depths are
    |i - s| in both cases."""
    say("[gpu] tail long path")
    V = 4000
    beg, adj = symmetric_csr_from_edges(V, [(i, i + 1) for i in range(V - 1)])
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(beg, adj))
    idx = np.arange(V)
    rng = np.random.default_rng(11)
    for n, overflow in ((8, False), (64, True)):
        src = rng.integers(0, V, size=n).astype(np.int32)
        b = ibfs_amd.Batch(g, contexts=1, burst=16, lane_bits=64, ring_levels=16, tail_div=1)
        np.testing.assert_array_equal(b.run(src), np.abs(idx[None, :] - src[:, None]))
        st = b.last_stats
        say(f"      n={n} tail levels {st.n_tail_levels} reruns {st.n_tail_reruns}")
        assert st.n_tail_reruns == (1 if overflow else 0)
        if not overflow:
            assert st.n_tail_levels > 1000
        dsum, reach = b.closeness(src)
        np.testing.assert_array_equal(reach, np.full(n, V))
        np.testing.assert_array_equal(dsum, np.abs(idx[None, :] - src[:, None]).sum(axis=1))
        b.close()

