"""Writes the committed golden fixtures of tests/golden/.

Run from the repository root:  python tests/golden/make_fixtures.py

* bfs_<name>.npz  - small undirected graphs (beg_pos int64, adj int32) with
  the all-pairs hop distances (depth int32 [V][V], -1 unreachable) computed
  by scipy.sparse.csgraph.shortest_path(unweighted=True), a library routine
  independent of this project.
* groupby_known.npz - a hand-built graph, a source list and the GroupBy
  order derived by hand from the rules of groupby.cu:36-149 (see
  groupby_case below).
* tri_beg_pos.bin / tri_csr.bin - the reference's on-disk CSR format
  (uint64 little-endian, graph.cu:71-112) of a triangle, written byte by
  byte with struct.pack.
* conv_*.txt / conv_*.npz - text edge lists and the CSR written for them by
  the reference's own converter (tuple_text_to_bin.cpp, compiled from the
  reference sources into oracle/_ref/ by `make -C oracle ref`); regenerated
  only when that binary is present.
"""
from __future__ import annotations

import shutil
import struct
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent


def csr_from_edges(V: int, edges) -> tuple[np.ndarray, np.ndarray]:
    """Symmetric, loop-free, duplicate-free CSR with ascending adjacency."""
    adj = [set() for _ in range(V)]
    for u, v in edges:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    beg = np.zeros(V + 1, dtype=np.int64)
    for v in range(V):
        beg[v + 1] = beg[v] + len(adj[v])
    a = np.array([w for v in range(V) for w in sorted(adj[v])], dtype=np.int32)
    return beg, a


def apsp(beg, adj) -> np.ndarray:
    V = len(beg) - 1
    m = sp.csr_matrix((np.ones(len(adj)), adj, beg), shape=(V, V))
    d = shortest_path(m, method="D", directed=False, unweighted=True)
    out = np.where(np.isinf(d), -1, d).astype(np.int32)
    return out


def graphs():
    rng = np.random.default_rng(20161)
    g = {}
    g["path10"] = csr_from_edges(10, [(i, i + 1) for i in range(9)])
    g["star10"] = csr_from_edges(10, [(0, i) for i in range(1, 10)])
    g["cycle9"] = csr_from_edges(9, [(i, (i + 1) % 9) for i in range(9)])
    g["k6"] = csr_from_edges(6, [(i, j) for i in range(6) for j in range(i + 1, 6)])
    # two components (triangle, path of 4) plus an isolated vertex 7
    g["twocomp"] = csr_from_edges(8, [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 6)])
    grid = []
    R, Cc = 6, 7
    for r in range(R):
        for c in range(Cc):
            v = r * Cc + c
            if c + 1 < Cc:
                grid.append((v, v + 1))
            if r + 1 < R:
                grid.append((v, v + Cc))
    g["grid6x7"] = csr_from_edges(R * Cc, grid)
    # sparse random graphs with several components and isolated vertices
    for name, V, m in (("er60", 60, 70), ("er300", 300, 450)):
        e = rng.integers(0, V, size=(m, 2))
        g[name] = csr_from_edges(V, [tuple(x) for x in e])
    # a hub-and-spoke graph with degrees across the 32/512 class thresholds
    V = 1500
    e = [(0, i) for i in range(1, 700)]            # hub of degree 699 (> 512)
    e += [(700, i) for i in range(701, 801)]       # mid vertex of degree 100
    e += [(i, i + 1) for i in range(801, 1499)]    # long path (deep levels)
    e += [(1, 700), (800, 801)]
    g["classes1500"] = csr_from_edges(V, e)
    return g


def groupby_case():
    """Hub H1=10 (degree 203), hub H2=5 (degree 151), vertex 20 of degree
    exactly 128 (not a hub: the rule is > 128), vertex 30 of degree exactly
    128 (still eligible: the rule is <= 128) whose first neighbour is H1.

    Sources in input order and their key by the rules of groupby.cu:
      150 -> first hub 10 | 10  -> degree 203 > 128, remainder
      310 -> first hub 5  | 480 -> no hub neighbour, remainder
      460 -> adjacency [7, 10]: first hub 10 (7 is not a hub)
      120 -> hub 10       | 470 -> adjacency [5, 10]: first hub 5
      305 -> hub 5        | 5   -> degree 151 > 128, remainder
      200 -> hub 10       | 510 -> only neighbour 20 (degree 128): remainder
      30  -> degree 128, first neighbour 10: hub 10
    Hubs are laid out in ascending id (prefix sum over all vertices), each
    group stable in input order, then the remainder in input order."""
    e = [(10, i) for i in range(100, 300)] + [(5, i) for i in range(300, 450)]
    e += [(460, 7), (460, 10), (470, 5), (470, 10), (480, 481)]
    e += [(20, i) for i in range(500, 628)]
    e += [(30, 10)] + [(30, i) for i in range(631, 758)]
    beg, adj = csr_from_edges(800, e)
    src = np.array([150, 10, 310, 480, 460, 120, 470, 305, 5, 200, 510, 30], dtype=np.int32)
    expected = np.array([310, 470, 305, 150, 460, 120, 200, 30, 10, 480, 5, 510], dtype=np.int32)
    deg = np.diff(beg)
    assert deg[10] == 203 and deg[5] == 151 and deg[20] == 128 and deg[30] == 128
    return beg, adj, src, expected


def write_tri_bin():
    beg = [0, 2, 4, 6]
    csr = [1, 2, 0, 2, 0, 1]
    (HERE / "tri_beg_pos.bin").write_bytes(b"".join(struct.pack("<Q", x) for x in beg))
    (HERE / "tri_csr.bin").write_bytes(b"".join(struct.pack("<Q", x) for x in csr))


CONV_CASES = {
    # name: (text, reverse, skip_lines)
    "conv_plain": ("1 2\n2 3\n3 1\n4 2\n", 1, 0),
    "conv_header": ("# header line\n% another\n10 12\n11 10\n12 13\n13 10\n", 1, 2),
    "conv_directed": ("0 3\n0 1\n2 0\n1 2\n3 3\n", 0, 0),
    "conv_tabs": ("5\t7\n7 \t 6\n6 5\n\n", 1, 0),
}


def write_converter_cases():
    ref = ROOT / "oracle" / "_ref" / "text_to_bin.bin"
    for name, (text, rev, skip) in CONV_CASES.items():
        (HERE / f"{name}.txt").write_text(text)
    if not ref.exists():
        print("reference converter not built; converter vectors left unchanged")
        return
    for name, (text, rev, skip) in CONV_CASES.items():
        with tempfile.TemporaryDirectory() as td:
            p = Path(td) / "g.txt"
            p.write_text(text)
            subprocess.run([str(ref), str(p), str(rev), str(skip)], check=True,
                           stdout=subprocess.DEVNULL, cwd=td)
            beg = np.frombuffer((Path(td) / "g.txt_beg_pos.bin").read_bytes(), dtype="<i8")
            adj = np.frombuffer((Path(td) / "g.txt_csr.bin").read_bytes(), dtype="<i8")
            np.savez(HERE / f"{name}.npz", beg_pos=beg.astype(np.int64), adj=adj.astype(np.int64),
                     reverse=np.int64(rev), skip=np.int64(skip))


def main():
    for name, (beg, adj) in graphs().items():
        np.savez_compressed(HERE / f"bfs_{name}.npz", beg_pos=beg, adj=adj, depth=apsp(beg, adj))
    beg, adj, src, exp = groupby_case()
    np.savez_compressed(HERE / "groupby_known.npz", beg_pos=beg, adj=adj, src=src, expected=exp)
    write_tri_bin()
    write_converter_cases()
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
