"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the ibfs_amd product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_CONVERTER = HERE / "_ref" / "text_to_bin.bin"

_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)
_lib = None


def build(reference: bool = False) -> None:
    """Compiles liboracle.so (and, if asked and present, the reference's
    converter into _ref/)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    if reference and Path("/root/reference/tuple_text_to_binary_csr/tuple_text_to_bin.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        l.oracle_bfs.argtypes = [C.c_int64, _I64P, _I32P, C.c_int32, _I32P]
        l.oracle_bfs_many.argtypes = [C.c_int64, _I64P, _I32P, _I32P, C.c_int64, _I32P, C.c_int]
        l.oracle_ibfs_bitwise.argtypes = [C.c_int64, _I64P, _I32P, _I32P, C.c_int, C.c_int, _I32P]
        l.oracle_groupby.argtypes = [C.c_int64, _I64P, _I32P, _I32P, C.c_int64, _I32P]
        l.oracle_degree_class.argtypes = [C.c_int64, C.c_int64, C.c_int64]
        l.oracle_traversed_edges.argtypes = [C.c_int64, _I64P, _I32P]
        l.oracle_traversed_edges.restype = C.c_int64
        l.oracle_check_depth.argtypes = [C.c_int64, _I64P, _I32P, C.c_int32, _I32P]
        l.oracle_read_csr.argtypes = [C.c_char_p, C.c_char_p, _I64P, _I64P,
                                      C.POINTER(_I64P), C.POINTER(_I32P)]
        l.oracle_text_to_csr.argtypes = [C.c_char_p, C.c_int, C.c_int64, _I64P, _I64P,
                                         C.POINTER(_I64P), C.POINTER(_I32P)]
        l.oracle_free.argtypes = [C.c_void_p]
        l.oracle_level_model.argtypes = [C.c_int64, _I64P, _I32P, _I32P, C.c_int,
                                         C.POINTER(C.c_int8), C.c_int] + [_I64P] * 6 + [_I32P]
        _lib = l
    return _lib


def _p64(a):
    return a.ctypes.data_as(_I64P)


def _p32(a):
    return a.ctypes.data_as(_I32P)


def _csr(beg, adj):
    return (np.ascontiguousarray(beg, dtype=np.int64), np.ascontiguousarray(adj, dtype=np.int32))


def bfs(beg, adj, src: int) -> np.ndarray:
    beg, adj = _csr(beg, adj)
    V = len(beg) - 1
    d = np.empty(V, dtype=np.int32)
    rc = lib().oracle_bfs(V, _p64(beg), _p32(adj), int(src), _p32(d))
    assert rc == 0, rc
    return d


def bfs_many(beg, adj, sources, threads: int = 1) -> np.ndarray:
    beg, adj = _csr(beg, adj)
    s = np.ascontiguousarray(sources, dtype=np.int32)
    V = len(beg) - 1
    d = np.empty((len(s), V), dtype=np.int32)
    rc = lib().oracle_bfs_many(V, _p64(beg), _p32(adj), _p32(s), len(s), _p32(d), threads)
    assert rc == 0, rc
    return d


def ibfs_bitwise(beg, adj, sources, sw_level: int) -> tuple[np.ndarray, int]:
    beg, adj = _csr(beg, adj)
    s = np.ascontiguousarray(sources, dtype=np.int32)
    V = len(beg) - 1
    d = np.empty((len(s), V), dtype=np.int32)
    levels = lib().oracle_ibfs_bitwise(V, _p64(beg), _p32(adj), _p32(s), len(s), sw_level, _p32(d))
    assert levels >= 0, levels
    return d, levels


def groupby(beg, adj, sources) -> np.ndarray:
    beg, adj = _csr(beg, adj)
    s = np.ascontiguousarray(sources, dtype=np.int32)
    out = np.empty_like(s)
    rc = lib().oracle_groupby(len(beg) - 1, _p64(beg), _p32(adj), _p32(s), len(s), _p32(out))
    assert rc == 0, rc
    return out


def degree_class(deg: int, sml: int = 32, lrg: int = 512) -> int:
    return lib().oracle_degree_class(deg, sml, lrg)


def traversed_edges(beg, depth) -> int:
    beg = np.ascontiguousarray(beg, dtype=np.int64)
    d = np.ascontiguousarray(depth, dtype=np.int32)
    return lib().oracle_traversed_edges(len(beg) - 1, _p64(beg), _p32(d))


def check_depth(beg, adj, src: int, depth) -> int:
    beg, adj = _csr(beg, adj)
    d = np.ascontiguousarray(depth, dtype=np.int32)
    return lib().oracle_check_depth(len(beg) - 1, _p64(beg), _p32(adj), int(src), _p32(d))


def _take(V, E, bp, ap):
    beg = np.ctypeslib.as_array(bp, shape=(V + 1,)).copy()
    adj = np.ctypeslib.as_array(ap, shape=(max(E, 1),))[:E].copy()
    lib().oracle_free(C.cast(bp, C.c_void_p))
    lib().oracle_free(C.cast(ap, C.c_void_p))
    return beg, adj


def read_csr(beg_file: str, csr_file: str):
    V, E, bp, ap = C.c_int64(), C.c_int64(), _I64P(), _I32P()
    rc = lib().oracle_read_csr(str(beg_file).encode(), str(csr_file).encode(), C.byref(V),
                               C.byref(E), C.byref(bp), C.byref(ap))
    if rc:
        return rc
    return _take(V.value, E.value, bp, ap)


def text_to_csr(path: str, reverse: bool, skip_lines: int = 0):
    V, E, bp, ap = C.c_int64(), C.c_int64(), _I64P(), _I32P()
    rc = lib().oracle_text_to_csr(str(path).encode(), int(reverse), skip_lines, C.byref(V),
                                  C.byref(E), C.byref(bp), C.byref(ap))
    if rc:
        return rc
    return _take(V.value, E.value, bp, ap)


def level_model(beg, adj, sources, dirs, max_levels: int = 64) -> dict:
    beg, adj = _csr(beg, adj)
    s = np.ascontiguousarray(sources, dtype=np.int32)
    dv = np.zeros(max_levels, dtype=np.int8)
    dv[:len(dirs)] = dirs[:max_levels]
    outs = [np.zeros(max_levels, dtype=np.int64) for _ in range(6)]
    lam = np.zeros(max_levels, dtype=np.int32)
    n = lib().oracle_level_model(len(beg) - 1, _p64(beg), _p32(adj), _p32(s), len(s),
                                 dv.ctypes.data_as(C.POINTER(C.c_int8)), max_levels,
                                 *[_p64(o) for o in outs], _p32(lam))
    assert n >= 0, n
    names = ["n_changed", "n_front", "m_front", "n_cand", "m_cand", "bu_prefix"]
    res = {k: v[:n + 1] for k, v in zip(names, outs)}
    res["lam_pop"] = lam[:n + 1]
    res["levels"] = n
    return res
