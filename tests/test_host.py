"""CPU tests of libibfs: exported C ABI, host CSR I/O, generators, converter,
source selection and error paths (no GPU needed)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import ibfs_amd
import oracle
from ibfs_amd import _lib
from conftest import GOLDEN, ROOT, load_golden


def header_symbols():
    text = (ROOT / "include" / "ibfs.h").read_text()
    return sorted(set(re.findall(r"^int\s+(ibfs_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(ibfs_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    l = _lib.lib()
    for s in syms:
        assert hasattr(l, s)
    # the Python binding declares every entry point of the header
    assert set(syms) <= set(_lib.SIGNATURES)


def test_abi_version_and_defaults():
    assert _lib.lib().ibfs_abi_version() == 1
    cfg = ibfs_amd.default_config()
    assert cfg.lane_bits == 128 and cfg.sml_shed == 64 and cfg.lrg_shed == 512
    assert cfg.policy == _lib.POLICY_ALPHA_BETA
    assert cfg.hybrid_gamma == 16 and cfg.hybrid_hubs == 4096 and cfg.split_cov == 0
    assert cfg.ring_levels >= 4


def test_hostcsr_roundtrip_reference_format(tmp_path):
    h = ibfs_amd.HostCsr.load(str(GOLDEN / "tri_beg_pos.bin"), str(GOLDEN / "tri_csr.bin"))
    beg, adj = h.arrays()
    np.testing.assert_array_equal(beg, [0, 2, 4, 6])
    np.testing.assert_array_equal(adj, [1, 2, 0, 2, 0, 1])
    g = load_golden("bfs_er300.npz")
    h = ibfs_amd.HostCsr.from_arrays(g["beg_pos"], g["adj"])
    h.save(str(tmp_path / "b.bin"), str(tmp_path / "c.bin"))
    # byte layout: uint64 little endian, checked with numpy and the oracle reader
    raw = np.frombuffer((tmp_path / "b.bin").read_bytes(), dtype="<u8")
    np.testing.assert_array_equal(raw, g["beg_pos"])
    b2, a2 = oracle.read_csr(tmp_path / "b.bin", tmp_path / "c.bin")
    np.testing.assert_array_equal(a2, g["adj"])


def test_hostcsr_load_errors(tmp_path):
    with pytest.raises(ibfs_amd.IbfsError) as ei:
        ibfs_amd.HostCsr.load(str(tmp_path / "nope.bin"), str(GOLDEN / "tri_csr.bin"))
    assert ei.value.code == _lib.IBFS_ERR_IO
    (tmp_path / "c.bin").write_bytes((GOLDEN / "tri_csr.bin").read_bytes()[:-8])
    with pytest.raises(ibfs_amd.IbfsError) as ei:  # beg_pos[V] != E
        ibfs_amd.HostCsr.load(str(GOLDEN / "tri_beg_pos.bin"), str(tmp_path / "c.bin"))
    assert ei.value.code == _lib.IBFS_ERR_FORMAT
    (tmp_path / "c3.bin").write_bytes(b"\x01" * 12)
    with pytest.raises(ibfs_amd.IbfsError) as ei:
        ibfs_amd.HostCsr.load(str(GOLDEN / "tri_beg_pos.bin"), str(tmp_path / "c3.bin"))
    assert ei.value.code == _lib.IBFS_ERR_FORMAT
    with pytest.raises(ibfs_amd.IbfsError):
        ibfs_amd.HostCsr.from_arrays([0, 1, 2], [1, 5])  # id out of range
    with pytest.raises(ibfs_amd.IbfsError):
        ibfs_amd.HostCsr.from_arrays([0, 2, 1], [1, 0])  # decreasing beg_pos


def _check_clean_symmetric(beg, adj):
    V = len(beg) - 1
    deg = np.diff(beg)
    src = np.repeat(np.arange(V), deg)
    assert np.all(src != adj), "self loop"
    key = src.astype(np.int64) * V + adj
    assert np.all(np.diff(key) > 0), "adjacency not strictly ascending (duplicate or unsorted)"
    rkey = adj.astype(np.int64) * V + src
    assert np.array_equal(np.sort(rkey), key), "not symmetric"


def test_rmat_generator_properties_and_determinism():
    h1 = ibfs_amd.HostCsr.rmat(10, 16, seed=1, threads=3)
    h2 = ibfs_amd.HostCsr.rmat(10, 16, seed=1, threads=1)
    b1, a1 = h1.arrays()
    b2, a2 = h2.arrays()
    np.testing.assert_array_equal(b1, b2)
    np.testing.assert_array_equal(a1, a2)
    _check_clean_symmetric(b1, a1)
    assert h1.vert_count == 1024
    assert 10000 < h1.edge_count < 2 * 16 * 1024
    b3, a3 = ibfs_amd.HostCsr.rmat(10, 16, seed=2).arrays()
    assert not np.array_equal(a1, a3)
    # skew: R-MAT has a heavy tail
    assert np.diff(b1).max() > 8 * np.diff(b1).mean()


def test_chung_lu_generator_properties():
    h = ibfs_amd.HostCsr.chung_lu(20000, gamma=2.2, mean_degree=20, max_degree=500, seed=3,
                                  threads=4)
    beg, adj = h.arrays()
    _check_clean_symmetric(beg, adj)
    deg = np.diff(beg)
    assert 10 < deg.mean() < 21            # dedup removes some of the expected mass
    assert deg.max() <= 3 * 500
    b2, a2 = ibfs_amd.HostCsr.chung_lu(20000, 2.2, 20, 500, seed=3, threads=1).arrays()
    np.testing.assert_array_equal(adj, a2)


def test_gen_sources_seeded_distinct_nonisolated():
    g = load_golden("bfs_twocomp.npz")
    h = ibfs_amd.HostCsr.from_arrays(g["beg_pos"], g["adj"])
    s = h.gen_sources(7, seed=5)
    assert len(set(s.tolist())) == 7 and 7 not in s  # vertex 7 is isolated
    np.testing.assert_array_equal(s, h.gen_sources(7, seed=5))
    with pytest.raises(ibfs_amd.IbfsError):
        h.gen_sources(8, seed=5)


@pytest.mark.parametrize("name", ["conv_plain", "conv_header", "conv_directed", "conv_tabs"])
def test_converter_matches_reference_vectors(name):
    z = load_golden(f"{name}.npz")
    h = ibfs_amd.HostCsr.from_text(str(GOLDEN / f"{name}.txt"), bool(z["reverse"]),
                                   int(z["skip"]))
    beg, adj = h.arrays()
    np.testing.assert_array_equal(beg, z["beg_pos"])
    np.testing.assert_array_equal(adj, z["adj"])


def test_converter_against_compiled_reference(tmp_path):
    ref = ROOT / "oracle" / "_ref" / "text_to_bin.bin"
    if not ref.exists():
        pytest.skip("reference converter not built here")
    rng = np.random.default_rng(11)
    e = rng.integers(3, 400, size=(2000, 2))
    lines = "\n".join(f"{a} {b}" for a, b in e) + "\n"
    for rev in (0, 1):
        # fresh directory per run: the reference opens its outputs without
        # truncation, so a reversed self-loop's unwritten slot would show an
        # earlier run's bytes instead of 0
        d = tmp_path / f"rev{rev}"
        d.mkdir()
        p = d / "g.txt"
        p.write_text("skip me\n" + lines + "7 7\n")
        subprocess.run([str(ref), str(p), str(rev), "1"], check=True, cwd=d,
                       stdout=subprocess.DEVNULL)
        want_b = np.frombuffer((d / "g.txt_beg_pos.bin").read_bytes(), dtype="<i8")
        want_a = np.frombuffer((d / "g.txt_csr.bin").read_bytes(), dtype="<i8")
        beg, adj = ibfs_amd.HostCsr.from_text(str(p), bool(rev), 1).arrays()
        np.testing.assert_array_equal(beg, want_b)
        np.testing.assert_array_equal(adj, want_a)


SUFFIXES = ("_beg_pos.bin", "_csr.bin", "_head.bin", "_weight.bin", "_deg.bin")


def test_text_to_bin_writes_the_reference_converters_files(tmp_path):
    """The text-to-bin command (arguments of tuple_text_to_bin.cpp:61-67)
    writes the same five files, byte for byte, as the reference converter
    compiled from its own source: with and without reversed edges, skipped
    header lines, a self-loop, tabs, and an odd trailing token."""
    ref = ROOT / "oracle" / "_ref" / "text_to_bin.bin"
    cli = ROOT / "ibfs_amd" / "text-to-bin"
    if not ref.exists() or not cli.exists():
        pytest.skip("reference converter or text-to-bin not built here")
    rng = np.random.default_rng(12)
    e = rng.integers(5, 300, size=(1500, 2))
    body = "\n".join(f"{a}\t{b}" if i % 7 == 0 else f"{a} {b}" for i, (a, b) in enumerate(e))
    texts = {"plain": "# header one\n# header two\n" + body + "\n9 9\n",
             "odd": body + "\n9 9\n12\n"}
    for name, text in texts.items():
        for rev in (0, 1):
            outs = {}
            for who, exe in (("ref", ref), ("ours", cli)):
                d = tmp_path / f"{name}{rev}{who}"  # fresh files for both
                d.mkdir()
                (d / "g.txt").write_text(text)
                skip = 2 if name == "plain" else 0
                r = subprocess.run([str(exe), str(d / "g.txt"), str(rev), str(skip)], cwd=d,
                                   capture_output=True, text=True)
                assert r.returncode == 0, r.stdout
                outs[who] = {sfx: (d / ("g.txt" + sfx)).read_bytes() for sfx in SUFFIXES}
            for sfx in SUFFIXES:
                assert outs["ours"][sfx] == outs["ref"][sfx], (name, rev, sfx)
    r = subprocess.run([str(cli)], capture_output=True, text=True)
    assert r.returncode != 0 and "Wrong input" in r.stdout


def test_traversed_edges_host():
    g = load_golden("bfs_twocomp.npz")
    h = ibfs_amd.HostCsr.from_arrays(g["beg_pos"], g["adj"])
    for s in range(8):
        assert h.traversed_edges(g["depth"][s]) == oracle.traversed_edges(g["beg_pos"],
                                                                          g["depth"][s])


def test_device_entry_points_fail_cleanly_without_gpu(gpu_available):
    if gpu_available:
        pytest.skip("a GPU is present")
    g = load_golden("bfs_twocomp.npz")
    h = ibfs_amd.HostCsr.from_arrays(g["beg_pos"], g["adj"])
    with pytest.raises(ibfs_amd.IbfsError) as ei:
        ibfs_amd.DeviceGraph(h)
    assert ei.value.code == _lib.IBFS_ERR_NODEVICE
    with pytest.raises(ibfs_amd.IbfsError):
        ibfs_amd.DeviceBuffer(16)


def test_c_abi_argument_checks():
    l = _lib.lib()
    cfg = _lib.Config()
    assert l.ibfs_config_default(None) == _lib.IBFS_ERR_ARG
    assert l.ibfs_engine_create(None, C.byref(cfg), None) == _lib.IBFS_ERR_ARG
    assert l.ibfs_traversal(None, None, 0, None, 0, None) == _lib.IBFS_ERR_ARG
    assert l.ibfs_groupby(None, None, 0, None) == _lib.IBFS_ERR_ARG
    assert l.ibfs_graph_info(None, None, None) == _lib.IBFS_ERR_ARG
    assert l.ibfs_graph_layout(None, None, None) == _lib.IBFS_ERR_ARG
    assert l.ibfs_graph_hub_entries(None, 0, None) == _lib.IBFS_ERR_ARG
    assert l.ibfs_server_set_admission(None, 1) == _lib.IBFS_ERR_ARG
    assert l.ibfs_autotune_batch(None, None, 0, None, 2, 8, 1, 1, None, None, None, None, 0) \
        == _lib.IBFS_ERR_ARG
    assert "null" in _lib.last_error()
    assert l.ibfs_graph_destroy(None) == 0 and l.ibfs_engine_destroy(None) == 0


def test_cli_usage_and_errors(tmp_path):
    cli = ROOT / "ibfs_amd" / "gpu-ibfs"
    if not cli.exists():
        pytest.skip("CLI not built")
    r = subprocess.run([str(cli)], capture_output=True, text=True)
    assert r.returncode != 0 and "Wrong input" in r.stdout
    r = subprocess.run([str(cli), str(tmp_path / "nope"), str(tmp_path / "nope2"), "128", "2",
                        "128"], capture_output=True, text=True)
    assert r.returncode != 0 and "cannot open" in r.stdout


def test_wait_timeout_setting():
    """The process-wide limit of the library's host waits on device work."""
    old = ibfs_amd.get_wait_timeout()
    assert old == 120.0
    try:
        ibfs_amd.set_wait_timeout(2.5)
        assert ibfs_amd.get_wait_timeout() == 2.5
        for bad in (0.0, -1.0, float("nan"), 1e9):
            with pytest.raises(ibfs_amd.IbfsError) as ei:
                ibfs_amd.set_wait_timeout(bad)
            assert ei.value.code == _lib.IBFS_ERR_ARG
        assert ibfs_amd.get_wait_timeout() == 2.5
        assert _lib.lib().ibfs_get_wait_timeout(None) == _lib.IBFS_ERR_ARG
    finally:
        ibfs_amd.set_wait_timeout(old)


def test_row_ranges_first_fit_and_coalescing():
    """Bookkeeping of the recycled output rows of served queries."""
    r = ibfs_amd.RowRanges(16)
    a, b, c = r.take(4), r.take(8), r.take(4)
    assert (a, b, c) == (0, 4, 12) and r.take(1) is None and not r.fits(1)
    r.give(b, 8)
    assert r.free_ranges == [(4, 8)] and r.take(9) is None
    d = r.take(3)
    assert d == 4 and r.free_ranges == [(7, 5)]
    r.give(a, 4)
    r.give(c, 4)
    assert r.free_ranges == [(0, 4), (7, 9)]
    r.give(d, 3)  # closes both gaps
    assert r.free_ranges == [(0, 16)]
    with pytest.raises(ValueError):
        r.give(2, 3)  # already free
    with pytest.raises(ValueError):
        r.give(15, 2)  # outside
    with pytest.raises(ValueError):
        r.take(17)
    # random alloc/free sequences never hand out overlapping rows
    rng = np.random.default_rng(5)
    r = ibfs_amd.RowRanges(64)
    held = []
    for _ in range(2000):
        if held and (rng.random() < 0.5 or not r.fits(1)):
            r0, n = held.pop(rng.integers(len(held)))
            r.give(r0, n)
        else:
            n = int(rng.integers(1, 9))
            r0 = r.take(n)
            if r0 is not None:
                assert all(r0 + n <= x or x + m <= r0 for x, m in held)
                held.append((r0, n))
        assert sum(m for _, m in held) + sum(c for _, c in r.free_ranges) == 64
