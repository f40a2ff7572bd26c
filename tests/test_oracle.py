"""Pins every function of the CPU oracle (oracle/ibfs_oracle.c) against
library routines, brute force, closed forms, hand-derived known answers and
the reference's own compiled converter."""
import itertools

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden_bfs_names, load_golden, symmetric_csr_from_edges


def floyd_warshall(beg, adj):
    V = len(beg) - 1
    INF = 10 ** 9
    d = np.full((V, V), INF, dtype=np.int64)
    for v in range(V):
        d[v, v] = 0
        for j in range(beg[v], beg[v + 1]):
            d[v, adj[j]] = min(d[v, adj[j]], 1)
    for k in range(V):
        d = np.minimum(d, d[:, k:k + 1] + d[k:k + 1, :])
    return np.where(d >= INF, -1, d).astype(np.int32)


@pytest.mark.parametrize("name", golden_bfs_names())
def test_bfs_matches_scipy_fixture(name):
    g = load_golden(name)
    V = len(g["beg_pos"]) - 1
    srcs = range(V) if V <= 300 else range(0, V, 7)
    for s in srcs:
        d = oracle.bfs(g["beg_pos"], g["adj"], s)
        np.testing.assert_array_equal(d, g["depth"][s], err_msg=f"{name} src {s}")


@pytest.mark.parametrize("name", ["bfs_path10.npz", "bfs_star10.npz", "bfs_k6.npz",
                                  "bfs_twocomp.npz", "bfs_cycle9.npz", "bfs_er60.npz",
                                  "bfs_grid6x7.npz"])
def test_bfs_matches_floyd_warshall(name):
    g = load_golden(name)
    fw = floyd_warshall(g["beg_pos"], g["adj"])
    V = len(g["beg_pos"]) - 1
    got = oracle.bfs_many(g["beg_pos"], g["adj"], np.arange(V), threads=3)
    np.testing.assert_array_equal(got, fw)


def test_bfs_closed_forms():
    # path: d = |i - s|
    V = 50
    beg, adj = symmetric_csr_from_edges(V, [(i, i + 1) for i in range(V - 1)])
    for s in (0, 17, 49):
        np.testing.assert_array_equal(oracle.bfs(beg, adj, s), np.abs(np.arange(V) - s))
    # star with a leaf source: hub 1, other leaves 2
    beg, adj = symmetric_csr_from_edges(8, [(0, i) for i in range(1, 8)])
    d = oracle.bfs(beg, adj, 3)
    assert d[3] == 0 and d[0] == 1 and all(d[i] == 2 for i in (1, 2, 4, 5, 6, 7))
    # isolated source: only itself
    beg, adj = symmetric_csr_from_edges(5, [(0, 1), (1, 2)])
    d = oracle.bfs(beg, adj, 4)
    np.testing.assert_array_equal(d, [-1, -1, -1, -1, 0])


def test_bfs_invariants_and_checker_catches_mistakes():
    g = load_golden("bfs_er300.npz")
    beg, adj = g["beg_pos"], g["adj"]
    for s in range(0, 300, 13):
        d = oracle.bfs(beg, adj, s)
        assert oracle.check_depth(beg, adj, s, d) == 0
        # planted mistakes must be caught
        reached = np.nonzero(d > 0)[0]
        if len(reached):
            bad = d.copy()
            bad[reached[0]] += 1
            assert oracle.check_depth(beg, adj, s, bad) != 0
            bad = d.copy()
            bad[reached[-1]] = -1
            assert oracle.check_depth(beg, adj, s, bad) != 0
        bad = d.copy()
        bad[s] = 1
        assert oracle.check_depth(beg, adj, s, bad) != 0


@pytest.mark.parametrize("name", golden_bfs_names())
def test_bitwise_restatement_equals_textbook_under_every_schedule(name):
    g = load_golden(name)
    beg, adj = g["beg_pos"], g["adj"]
    V = len(beg) - 1
    rng = np.random.default_rng(len(name))
    src = rng.integers(0, V, size=min(V, 128)).astype(np.int32)
    src[-1] = src[0]  # a duplicate source in two lanes
    want = g["depth"][src]
    for sw in (0, 1, 2, 3, 5, 1 << 30):
        got, levels = oracle.ibfs_bitwise(beg, adj, src, sw)
        np.testing.assert_array_equal(got, want, err_msg=f"{name} sw={sw}")
        assert levels == want.max() + 1


def test_bitwise_lane_position_does_not_matter():
    g = load_golden("bfs_grid6x7.npz")
    beg, adj = g["beg_pos"], g["adj"]
    for lane in (0, 1, 63, 64, 127, 200, 255):
        src = np.zeros(lane + 1, dtype=np.int32)
        src[lane] = 17
        got, _ = oracle.ibfs_bitwise(beg, adj, src, 1)
        np.testing.assert_array_equal(got[lane], g["depth"][17])


def test_groupby_known_answer():
    g = load_golden("groupby_known.npz")
    got = oracle.groupby(g["beg_pos"], g["adj"], g["src"])
    np.testing.assert_array_equal(got, g["expected"])


def test_groupby_is_a_stable_key_sort():
    # independent statement: stable sort by (first hub id or +inf)
    g = load_golden("bfs_classes1500.npz")
    beg, adj = g["beg_pos"], g["adj"]
    deg = np.diff(beg)
    rng = np.random.default_rng(7)
    src = rng.choice(len(deg), size=300, replace=False).astype(np.int32)

    def key(s):
        if deg[s] > 128:
            return 1 << 40
        for j in range(beg[s], beg[s + 1]):
            if deg[adj[j]] > 128:
                return int(adj[j])
        return 1 << 40

    expected = sorted(src.tolist(), key=key)  # Python sort is stable
    np.testing.assert_array_equal(oracle.groupby(beg, adj, src), expected)


def test_degree_class_thresholds():
    assert oracle.degree_class(0) == 0
    assert oracle.degree_class(1) == 1
    assert oracle.degree_class(31) == 1
    assert oracle.degree_class(32) == 2
    assert oracle.degree_class(512) == 2
    assert oracle.degree_class(513) == 3
    assert oracle.degree_class(5, sml=4, lrg=4) == 3


def test_traversed_edges_closed_form():
    g = load_golden("bfs_twocomp.npz")
    beg = g["beg_pos"]
    # source 0 reaches the triangle: 3 vertices of degree 2
    assert oracle.traversed_edges(beg, g["depth"][0]) == 6
    # source 3 reaches the path 3-4-5-6: degrees 1+2+2+1
    assert oracle.traversed_edges(beg, g["depth"][3]) == 6
    assert oracle.traversed_edges(beg, g["depth"][7]) == 0


def test_read_csr_byte_fixture(tmp_path):
    beg, adj = oracle.read_csr(GOLDEN / "tri_beg_pos.bin", GOLDEN / "tri_csr.bin")
    np.testing.assert_array_equal(beg, [0, 2, 4, 6])
    np.testing.assert_array_equal(adj, [1, 2, 0, 2, 0, 1])
    # truncated csr: size not a multiple of 8 -> format error
    (tmp_path / "c.bin").write_bytes((GOLDEN / "tri_csr.bin").read_bytes()[:-3])
    assert oracle.read_csr(GOLDEN / "tri_beg_pos.bin", tmp_path / "c.bin") == -3
    assert oracle.read_csr(tmp_path / "missing.bin", GOLDEN / "tri_csr.bin") == -2
    # an id out of range
    bad = (GOLDEN / "tri_csr.bin").read_bytes()[:-8] + (7).to_bytes(8, "little")
    (tmp_path / "c2.bin").write_bytes(bad)
    assert oracle.read_csr(GOLDEN / "tri_beg_pos.bin", tmp_path / "c2.bin") == -3


@pytest.mark.parametrize("name", ["conv_plain", "conv_header", "conv_directed", "conv_tabs"])
def test_text_converter_matches_reference_vectors(name):
    z = load_golden(f"{name}.npz")
    beg, adj = oracle.text_to_csr(GOLDEN / f"{name}.txt", bool(z["reverse"]), int(z["skip"]))
    np.testing.assert_array_equal(beg, z["beg_pos"])
    np.testing.assert_array_equal(adj, z["adj"])


def test_level_model_small_cases():
    # path 0-1-2-3-4, sources {0}: one new vertex per level, lam_pop 1
    beg, adj = symmetric_csr_from_edges(5, [(i, i + 1) for i in range(4)])
    m = oracle.level_model(beg, adj, [0], dirs=[0, 1, 1, 1, 1, 1])
    assert m["levels"] == 5
    np.testing.assert_array_equal(m["n_changed"][:6], [1, 1, 1, 1, 1, 0])
    np.testing.assert_array_equal(m["lam_pop"][:5], [1, 1, 1, 1, 1])
    # bottom-up at level L reads, for every unreached vertex of degree > 0,
    # its whole adjacency unless the vertex at depth L finds 0's lane at its
    # first neighbour (L-1): vertex L reads 1 entry; vertices > L read deg.
    # level 1: v1 reads 1; v2,v3 read 2; v4 reads 1 -> 6
    assert m["bu_prefix"][1] == 6
    # level 2: v2 reads 1 (neighbour 1 first), v3 reads 2, v4 reads 1 -> 4
    assert m["bu_prefix"][2] == 4
