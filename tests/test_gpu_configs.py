"""Parity at the configured sizes of the workloads other than the headline
one: all-source closeness on R-MAT scale 18 and 512 sources on the 3.07M-vertex
Chung-Lu graph.  Needs an MI355X."""
import sys

import numpy as np
import pytest

import ibfs_amd
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _components(beg, adj):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    V = len(beg) - 1
    m = csr_matrix((np.ones(len(adj), dtype=np.int8), adj, beg), shape=(V, V))
    n, lab = connected_components(m, directed=False)
    return lab, np.bincount(lab)


def test_closeness_every_vertex_rmat_s18():
    """R-MAT s18, every vertex a source (262,144 BFS) through the batched
    launcher: reach equals the source's component size for every source
    (closed form), and sums of depths equal the plain 128-lane form's depth
    arrays and the oracle on samples."""
    say("[gpu] closeness s18 all sources")
    h = ibfs_amd.HostCsr.rmat(18, 16, seed=1)
    beg, adj = h.arrays()
    V = h.vert_count
    g = ibfs_amd.DeviceGraph(h)
    src = np.arange(V, dtype=np.int32)
    b = ibfs_amd.Batch(g, contexts=2, burst=4, lane_bits=256)
    dsum, reach = b.closeness(src)
    lab, size = _components(beg, adj)
    np.testing.assert_array_equal(reach, size[lab])
    assert np.all(dsum >= reach - 1)
    rng = np.random.default_rng(18)
    sample = rng.choice(V, size=256, replace=False).astype(np.int32)
    d = ibfs_amd.Engine(g, lane_bits=128).run_groups(sample).astype(np.int64)
    np.testing.assert_array_equal(dsum[sample], np.where(d >= 0, d, 0).sum(axis=1))
    np.testing.assert_array_equal(reach[sample], (d >= 0).sum(axis=1))
    want = oracle.bfs_many(beg, adj, sample[:16], threads=16).astype(np.int64)
    np.testing.assert_array_equal(dsum[sample[:16]], np.where(want >= 0, want, 0).sum(axis=1))


def test_chung_lu_3m_512_sources():
    """Chung-Lu 3.07M vertices (gamma 2.2, mean 76, cap 33,000), 512 sources
    in GroupBy-ordered groups of 128 and through the batched launcher: every
    depth array of both forms equals the oracle's, and the traversed edges of
    both forms agree."""
    say("[gpu] chung-lu 3.07M x 512 sources")
    h = ibfs_amd.HostCsr.chung_lu(3_070_000, 2.2, 76.0, 33000.0, seed=1)
    beg, adj = h.arrays()
    V = h.vert_count
    g = ibfs_amd.DeviceGraph(h)
    src = g.groupby(h.gen_sources(512, seed=2))
    out = ibfs_amd.DeviceBuffer(4 * len(src) * V)
    eng = ibfs_amd.Engine(g, lane_bits=128)
    eng.run_groups(src, out=out)
    edges = eng.last_stats.traversed_edges
    chunk = 64
    want = [oracle.bfs_many(beg, adj, src[i:i + chunk], threads=16) for i in range(0, len(src), chunk)]

    def check(form):
        for i, w in zip(range(0, len(src), chunk), want):
            got = out.to_host(count=len(w) * V, offset=i * V).reshape(w.shape)
            np.testing.assert_array_equal(got, w, err_msg=f"{form}, source rows {i}..{i + len(w) - 1}")

    check("plain")
    b = ibfs_amd.Batch(g, contexts=2, burst=4, lane_bits=256)
    b.run(src, out=out)
    assert b.last_stats.traversed_edges == edges
    check("batched")


def test_hybrid_levels_with_pushes_into_workgroup_candidates():
    """Hybrid levels on a Chung-Lu graph whose high-degree vertices are
    bottom-up candidates of the workgroup class while the frontier vertices
    outside the pulled hubs push into them: every workgroup must branch on one
    read of its candidate's word (regression: the waves used to read it
    separately, disagree on covering the active lanes after a concurrent push,
    and meet at mismatched barriers).  Every level is hybrid, with 64 and
    4,096 pulled hubs, three runs each, against the oracle."""
    h = ibfs_amd.HostCsr.chung_lu(400_000, 2.2, 76.0, 20000.0, seed=5, threads=8)
    beg, adj = h.arrays()
    g = ibfs_amd.DeviceGraph(h)
    src = g.groupby(h.gen_sources(512, seed=6))
    want = oracle.bfs_many(beg, adj, src, threads=8)
    for cut in (64, 4096):
        opts = dict(policy=ibfs_amd.POLICY_HYBRID, hybrid_hubs=cut, lane_bits=256)
        eng = ibfs_amd.Engine(g, **opts)
        b = ibfs_amd.Batch(g, contexts=2, burst=4, **opts)
        for rep in range(3):
            say(f"[gpu] chung-lu 400K hybrid cut={cut} run {rep}")
            np.testing.assert_array_equal(eng.run_groups(src), want)
            np.testing.assert_array_equal(b.run(src), want)


def test_split_levels_chung_lu_400k():
    """Split levels on a Chung-Lu graph (heavy-tailed degrees: workgroup-class
    candidates pull the dense lanes while the sparse lanes' frontier pushes
    into the same words), 512 sources in caller order, coverage 50 %, three
    runs against the oracle."""
    h = ibfs_amd.HostCsr.chung_lu(400_000, 2.2, 76.0, 20000.0, seed=5, threads=8)
    beg, adj = h.arrays()
    g = ibfs_amd.DeviceGraph(h)
    src = h.gen_sources(512, seed=7)
    want = oracle.bfs_many(beg, adj, src, threads=8)
    b = ibfs_amd.Batch(g, contexts=2, burst=4, lane_bits=256, pack_sources=0, split_cov=50)
    for rep in range(3):
        say(f"[gpu] chung-lu 400K split run {rep}")
        np.testing.assert_array_equal(b.run(src), want)
    assert b.last_stats.n_split_levels > 0


def _grid_csr(n):
    """Symmetric CSR of the n x n grid graph (4-neighbourhood), sorted lists."""
    V = n * n
    idx = np.arange(V).reshape(n, n)
    a = np.concatenate([idx[:, :-1].ravel(), idx[:, 1:].ravel(), idx[:-1, :].ravel(),
                        idx[1:, :].ravel()])
    b = np.concatenate([idx[:, 1:].ravel(), idx[:, :-1].ravel(), idx[1:, :].ravel(),
                        idx[:-1, :].ravel()])
    o = np.lexsort((b, a))
    a, b = a[o], b[o]
    beg = np.zeros(V + 1, dtype=np.int64)
    np.add.at(beg, a + 1, 1)
    return np.cumsum(beg), b.astype(np.int32)


def test_grid_1024_long_traversals():
    """A 1,024 x 1,024 grid (1M vertices, diameter 2,046): 256 random sources
    take up to about 2,000 levels, so the 32-snapshot ring flushes dozens of
    times per group.  While the frontier is small most vertices still miss
    some lane, and a tail log would overflow (every vertex gains each lane
    at its own level): the default tail setting must not start a tail that
    ends in a dense rerun, while tail_div 1 does and reruns.  Depth arrays
    of the plain engine (two 128-lane groups) and of the batched launcher
    (256 lanes, tail off, default and 1), TEPS numerators and closeness
    equal the oracle's."""
    say("[gpu] grid 1024 x 1024, 256 sources")
    beg, adj = _grid_csr(1024)
    V = len(beg) - 1
    g = ibfs_amd.DeviceGraph(ibfs_amd.HostCsr.from_arrays(beg, adj))
    src = np.random.default_rng(7).integers(0, V, size=256).astype(np.int32)
    want = oracle.bfs_many(beg, adj, src, threads=8)
    edges = sum(oracle.traversed_edges(beg, d) for d in want)
    d64 = want.astype(np.int64)
    want_sum, want_reach = np.where(d64 >= 0, d64, 0).sum(axis=1), (d64 >= 0).sum(axis=1)
    out = ibfs_amd.DeviceBuffer(4 * len(src) * V)
    eng = ibfs_amd.Engine(g, lane_bits=128)
    eng.run_groups(src, out=out)
    np.testing.assert_array_equal(out.to_host().reshape(len(src), V), want)
    assert eng.last_stats.traversed_edges == edges
    say(f"      plain: {eng.last_stats.levels} levels, {eng.last_stats.emit_windows} windows")
    eng.close()
    for tail_div in (0, 32, 1):
        b = ibfs_amd.Batch(g, contexts=2, burst=8, lane_bits=256, tail_div=tail_div)
        b.run(src, out=out)
        np.testing.assert_array_equal(out.to_host().reshape(len(src), V), want,
                                      err_msg=f"tail_div {tail_div}")
        st = b.last_stats
        assert st.traversed_edges == edges
        say(f"      batched tail_div={tail_div}: tail levels {st.n_tail_levels}, "
            f"reruns {st.n_tail_reruns}")
        if tail_div != 1:
            assert st.n_tail_reruns == 0
        dsum, reach = b.closeness(src)
        np.testing.assert_array_equal(reach, want_reach)
        np.testing.assert_array_equal(dsum, want_sum)
        b.close()
