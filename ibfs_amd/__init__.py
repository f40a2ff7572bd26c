"""ibfs_amd: MI355X-native concurrent multi-source BFS (iBFS operator surface).

Python host over the C ABI of libibfs.so (include/ibfs.h).  The classes
`graph` and `gpu_ibfs` mirror the reference's host interface (graph.h,
gpu_ibfs.cuh); `HostCsr`, `DeviceGraph` and `Engine` are the general API.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import _lib
from ._lib import (IbfsError, OUT_DEVICE, OUT_HOST, POLICY_ALPHA_BETA, POLICY_BU_ONLY,
                   POLICY_FIXED, POLICY_HYBRID, POLICY_TD_ONLY, Config, LevelInfo, Stats, check,
                   lib)

__all__ = [
    "IbfsError", "HostCsr", "DeviceGraph", "Engine", "Batch", "Server", "DeviceBuffer",
    "DeviceView", "RowPool", "RowRanges", "graph", "gpu_ibfs",
    "default_config", "device_count", "set_wait_timeout", "get_wait_timeout", "POLICY_FIXED",
    "POLICY_ALPHA_BETA", "POLICY_TD_ONLY",
    "POLICY_BU_ONLY", "POLICY_HYBRID",
]


def _i32p(a: np.ndarray):
    return a.ctypes.data_as(_lib.I32P)


def _i64p(a: np.ndarray):
    return a.ctypes.data_as(_lib.I64P)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().ibfs_device_count(C.byref(n)))
    return n.value


def set_wait_timeout(seconds: float) -> None:
    """Limit of every host wait of the library on device work (process-wide)."""
    check(lib().ibfs_set_wait_timeout(float(seconds)))


def get_wait_timeout() -> float:
    x = C.c_double()
    check(lib().ibfs_get_wait_timeout(C.byref(x)))
    return x.value


def default_config(**overrides) -> Config:
    cfg = Config()
    check(lib().ibfs_config_default(C.byref(cfg)))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(f"unknown config field {k}")
        setattr(cfg, k, v)
    return cfg


class HostCsr:
    """Host CSR: int64 beg_pos[V+1], int32 adjacency[E] (owned by libibfs)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def from_arrays(cls, beg_pos, adj) -> "HostCsr":
        beg = np.ascontiguousarray(beg_pos, dtype=np.int64)
        a = np.ascontiguousarray(adj, dtype=np.int32)
        h = C.c_void_p()
        check(lib().ibfs_hostcsr_create(_i64p(beg), len(beg) - 1, _i32p(a), len(a), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, beg_file: str, csr_file: str) -> "HostCsr":
        h = C.c_void_p()
        check(lib().ibfs_hostcsr_load(os.fsencode(beg_file), os.fsencode(csr_file), C.byref(h)))
        return cls(h.value)

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, seed: int = 1, permute: bool = True, threads: int = 0) -> "HostCsr":
        h = C.c_void_p()
        check(lib().ibfs_gen_rmat(scale, edge_factor, a, b, c, seed, int(permute), threads,
                                  C.byref(h)))
        return cls(h.value)

    @classmethod
    def chung_lu(cls, vert_count: int, gamma: float = 2.2, mean_degree: float = 76.0,
                 max_degree: float = 33000.0, seed: int = 1, threads: int = 0) -> "HostCsr":
        h = C.c_void_p()
        check(lib().ibfs_gen_chung_lu(vert_count, gamma, mean_degree, max_degree, seed, threads,
                                      C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_text(cls, path: str, reverse: bool = True, skip_lines: int = 0) -> "HostCsr":
        h = C.c_void_p()
        check(lib().ibfs_convert_text(os.fsencode(path), int(reverse), skip_lines, C.byref(h)))
        return cls(h.value)

    def save(self, beg_file: str, csr_file: str) -> None:
        check(lib().ibfs_hostcsr_save(self._h, os.fsencode(beg_file), os.fsencode(csr_file)))

    @property
    def vert_count(self) -> int:
        v, e = C.c_int64(), C.c_int64()
        check(lib().ibfs_hostcsr_info(self._h, C.byref(v), C.byref(e)))
        return v.value

    @property
    def edge_count(self) -> int:
        v, e = C.c_int64(), C.c_int64()
        check(lib().ibfs_hostcsr_info(self._h, C.byref(v), C.byref(e)))
        return e.value

    def arrays(self) -> tuple[np.ndarray, np.ndarray]:
        """Copies of (beg_pos int64[V+1], adj int32[E])."""
        V, E = self.vert_count, self.edge_count
        bp, ap = _lib.I64P(), _lib.I32P()
        check(lib().ibfs_hostcsr_data(self._h, C.byref(bp), C.byref(ap)))
        beg = np.ctypeslib.as_array(bp, shape=(V + 1,)).copy()
        adj = np.ctypeslib.as_array(ap, shape=(max(E, 1),))[:E].copy() if E else np.zeros(0, np.int32)
        return beg, adj

    def gen_sources(self, count: int, seed: int = 2) -> np.ndarray:
        out = np.zeros(count, dtype=np.int32)
        check(lib().ibfs_gen_sources(self._h, count, seed, _i32p(out)))
        return out

    def traversed_edges(self, depth: np.ndarray) -> int:
        d = np.ascontiguousarray(depth, dtype=np.int32)
        e = C.c_int64()
        check(lib().ibfs_traversed_edges(self._h, _i32p(d), C.byref(e)))
        return e.value

    def close(self) -> None:
        if self._h:
            lib().ibfs_hostcsr_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _copy_out(ptr, nbytes, dtype, count, offset, out):
    item = np.dtype(dtype).itemsize
    n = (nbytes // item - offset) if count is None else count
    if offset < 0 or n < 0 or (offset + n) * item > nbytes:
        raise ValueError("range outside the device memory")
    if out is None:
        out = np.empty(n, dtype=dtype)
    else:
        if out.dtype != np.dtype(dtype) or not out.flags.c_contiguous or out.size < n:
            raise ValueError("out must be C-contiguous, of the dtype and large enough")
        out = out.reshape(-1)[:n]
    check(lib().ibfs_memcpy_d2h(out.ctypes.data, C.c_void_p(ptr + offset * item), n * item))
    return out


class DeviceBuffer:
    """Raw device allocation (for device-resident depth outputs)."""

    def __init__(self, nbytes: int, device: int = 0):
        p = C.c_void_p()
        check(lib().ibfs_device_alloc(device, nbytes, C.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    def to_host(self, dtype=np.int32, count: Optional[int] = None, offset: int = 0,
                out: Optional[np.ndarray] = None) -> np.ndarray:
        """Copies `count` elements starting at element `offset` to the host,
        into `out` (a C-contiguous array of at least `count` elements of
        dtype; a view of its first `count` is returned) or a new array."""
        return _copy_out(self.ptr, self.nbytes, dtype, count, offset, out)

    def copy_from(self, other: "DeviceBuffer", nbytes: Optional[int] = None) -> None:
        """Device-to-device copy of nbytes (default: all of `other`)."""
        n = other.nbytes if nbytes is None else nbytes
        if n > self.nbytes or n > other.nbytes:
            raise ValueError("copy larger than a buffer")
        check(lib().ibfs_memcpy_d2d(C.c_void_p(self.ptr), C.c_void_p(other.ptr), n))

    def free(self) -> None:
        if self.ptr:
            lib().ibfs_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView:
    """A sub-range of a DeviceBuffer (same to_host interface).  It holds a
    reference to its buffer, so the memory stays allocated while the view
    is alive."""

    def __init__(self, buf: DeviceBuffer, offset_bytes: int, nbytes: int):
        if offset_bytes < 0 or nbytes < 0 or offset_bytes + nbytes > buf.nbytes:
            raise ValueError("view outside the buffer")
        self.buf = buf
        self.offset = offset_bytes
        self.ptr = buf.ptr + offset_bytes
        self.nbytes = nbytes

    def to_host(self, dtype=np.int32, count: Optional[int] = None, offset: int = 0,
                out: Optional[np.ndarray] = None) -> np.ndarray:
        return _copy_out(self.ptr, self.nbytes, dtype, count, offset, out)


class RowRanges:
    """First-fit allocator of consecutive rows in [0, rows), with coalescing
    of freed ranges (the bookkeeping of RowPool; no device memory)."""

    def __init__(self, rows: int):
        if rows < 1:
            raise ValueError("need rows >= 1")
        self.rows = rows
        self.free_ranges = [(0, rows)]  # sorted (first row, count), coalesced

    def take(self, n: int) -> Optional[int]:
        if n < 1 or n > self.rows:
            raise ValueError(f"cannot hold {n} rows in {self.rows}")
        for i, (r0, cnt) in enumerate(self.free_ranges):
            if cnt >= n:
                if cnt == n:
                    self.free_ranges.pop(i)
                else:
                    self.free_ranges[i] = (r0 + n, cnt - n)
                return r0
        return None

    def fits(self, n: int) -> bool:
        return any(cnt >= n for _, cnt in self.free_ranges)

    def give(self, r0: int, n: int) -> None:
        if r0 < 0 or n < 1 or r0 + n > self.rows:
            raise ValueError("range outside the pool")
        for a, c in self.free_ranges:
            if r0 < a + c and a < r0 + n:
                raise ValueError(f"rows {r0}..{r0 + n - 1} are already free")
        merged = []
        for a, c in sorted(self.free_ranges + [(r0, n)]):
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + c)
            else:
                merged.append((a, c))
        self.free_ranges = merged


class RowPool:
    """Device arena of depth rows (int32[V] each) for query outputs that are
    recycled: alloc(n) returns a DeviceView of n consecutive rows (first fit,
    waiting up to `timeout` seconds for space; None after a timeout), free()
    returns them.  Thread-safe.  One allocation at construction: nothing is
    allocated or freed per query."""

    def __init__(self, vert_count: int, rows: int, device: int = 0):
        import threading
        if rows < 1 or vert_count < 1:
            raise ValueError("need rows >= 1 and vert_count >= 1")
        self.row_bytes = 4 * vert_count
        self.rows = rows
        self.buf = DeviceBuffer(self.row_bytes * rows, device)
        self._ranges = RowRanges(rows)
        self._cv = threading.Condition()
        self.waits = 0            # alloc calls that had to wait for space

    def alloc(self, n: int, timeout: Optional[float] = None) -> Optional[DeviceView]:
        with self._cv:
            r0 = self._ranges.take(n)
            if r0 is None:
                self.waits += 1
                if not self._cv.wait_for(lambda: self._ranges.fits(n), timeout=timeout):
                    return None
                r0 = self._ranges.take(n)
        return DeviceView(self.buf, r0 * self.row_bytes, n * self.row_bytes)

    def free(self, view: DeviceView) -> None:
        if view.buf is not self.buf:
            raise ValueError("view of another pool")
        with self._cv:
            self._ranges.give(view.offset // self.row_bytes, view.nbytes // self.row_bytes)
            self._cv.notify_all()

    def close(self) -> None:
        self.buf.free()


class DeviceGraph:
    """A CSR resident in device memory (read-only, shareable by engines)."""

    def __init__(self, csr: HostCsr, device: int = 0, hub_count: Optional[int] = None):
        h = C.c_void_p()
        if hub_count is None:
            check(lib().ibfs_graph_from_host(csr._h, device, C.byref(h)))
        else:
            opts = _lib.GraphOpts()
            check(lib().ibfs_graph_opts_default(C.byref(opts)))
            opts.hub_count = hub_count
            beg, adj = csr.arrays()
            check(lib().ibfs_graph_create_ex(_i64p(beg), len(beg) - 1, _i32p(adj), len(adj),
                                             device, C.byref(opts), C.byref(h)))
        self._h = h
        self.device = device
        self.vert_count = csr.vert_count
        self.edge_count = csr.edge_count
        vt, hubs = C.c_int64(), C.c_int64()
        check(lib().ibfs_graph_layout(h, C.byref(vt), C.byref(hubs)))
        self.traversal_vertices = vt.value  # vertices with an edge (status words)
        self.hub_count = hubs.value

    def hub_entries(self, top: int = 0) -> int:
        """Adjacency entries that point to the top `top` hubs (0: all hubs)."""
        n = C.c_int64()
        check(lib().ibfs_graph_hub_entries(self._h, top, C.byref(n)))
        return n.value

    def traversal_noalloc(self, sources, out, cfg: Optional[Config] = None) -> int:
        """One group of <= 128 sources through ibfs_traversal_noalloc (no
        allocation inside the library; scratch in a file-scope device array):
        depth arrays into `out` (a device buffer or view of k*V int32);
        returns the traversed-edge count."""
        s = np.ascontiguousarray(sources, dtype=np.int32)
        if out.nbytes < 4 * len(s) * self.vert_count:
            raise ValueError("device buffer too small")
        e = C.c_int64()
        check(lib().ibfs_traversal_noalloc(self._h, C.byref(cfg) if cfg is not None else None,
                                           _i32p(s), len(s), C.c_void_p(out.ptr), C.byref(e)))
        return e.value

    @classmethod
    def load_files(cls, beg_file: str, csr_file: str, device: int = 0) -> "DeviceGraph":
        return cls(HostCsr.load(beg_file, csr_file), device)

    def groupby(self, src_list) -> np.ndarray:
        """GPU GroupBy: the permutation of src_list of groupby.cu:36-149."""
        s = np.ascontiguousarray(src_list, dtype=np.int32)
        out = np.zeros_like(s)
        check(lib().ibfs_groupby(self._h, _i32p(s), len(s), _i32p(out)))
        return out

    def close(self) -> None:
        if self._h:
            lib().ibfs_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One traversal engine (HIP stream, status ring, queues) over a DeviceGraph."""

    def __init__(self, g: DeviceGraph, cfg: Optional[Config] = None, **overrides):
        if cfg is None:
            cfg = default_config(device=g.device, **overrides)
        elif overrides:
            for k, v in overrides.items():
                setattr(cfg, k, v)
        h = C.c_void_p()
        check(lib().ibfs_engine_create(g._h, C.byref(cfg), C.byref(h)))
        self._h = h
        self.graph = g
        self.cfg = cfg
        self.last_stats = Stats()

    @property
    def lane_bits(self) -> int:
        return self.cfg.lane_bits

    def traversal(self, sources, out: Optional[DeviceBuffer] = None) -> Optional[np.ndarray]:
        """Runs one group (<= lane_bits sources).  Returns depth int32[k][V] on
        the host, or writes into `out` (device) and returns None."""
        s = np.ascontiguousarray(sources, dtype=np.int32)
        k, V = len(s), self.graph.vert_count
        st = Stats()
        if out is None:
            depth = np.empty((k, V), dtype=np.int32)
            check(lib().ibfs_traversal(self._h, _i32p(s), k, depth.ctypes.data, OUT_HOST,
                                       C.byref(st)))
            self.last_stats = st
            return depth
        if out.nbytes < 4 * k * V:
            raise ValueError("device buffer too small")
        check(lib().ibfs_traversal(self._h, _i32p(s), k, C.c_void_p(out.ptr), OUT_DEVICE,
                                   C.byref(st)))
        self.last_stats = st
        return None

    def run_groups(self, sources, out: Optional[DeviceBuffer] = None) -> Optional[np.ndarray]:
        s = np.ascontiguousarray(sources, dtype=np.int32)
        n, V = len(s), self.graph.vert_count
        st = Stats()
        if out is None:
            depth = np.empty((n, V), dtype=np.int32)
            check(lib().ibfs_run_groups(self._h, _i32p(s), n, depth.ctypes.data, OUT_HOST,
                                        C.byref(st)))
            self.last_stats = st
            return depth
        if out.nbytes < 4 * n * V:
            raise ValueError("device buffer too small")
        check(lib().ibfs_run_groups(self._h, _i32p(s), n, C.c_void_p(out.ptr), OUT_DEVICE,
                                    C.byref(st)))
        self.last_stats = st
        return None

    def level_log(self) -> list[dict]:
        n = C.c_int32()
        check(lib().ibfs_engine_level_log(self._h, None, 0, C.byref(n)))
        arr = (LevelInfo * max(n.value, 1))()
        check(lib().ibfs_engine_level_log(self._h, arr, n.value, C.byref(n)))
        out = []
        for i in range(n.value):
            li = arr[i]
            out.append({
                "level": li.level, "direction": li.direction, "n_changed": li.n_changed,
                "n_frontier": li.n_frontier, "m_frontier": li.m_frontier,
                "n_unvisited": li.n_unvisited, "m_unvisited": li.m_unvisited,
                "q_td": list(li.q_td), "q_bu": list(li.q_bu),
                "bu_work": li.bu_work, "bu_prefix": li.bu_prefix,
                "expand_ms": li.expand_ms, "inspect_ms": li.inspect_ms,
                "m_frontier_nonhub": li.m_frontier_nonhub,
            })
        return out

    def close(self) -> None:
        if self._h:
            lib().ibfs_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(g: DeviceGraph, sample_sources, passes: int = 2, reps: int = 2,
             start: Optional[Config] = None, save_to: Optional[str] = None) -> tuple[Config, dict]:
    """Tunes the engine configuration for graph g on the sample sources;
    returns (config, report).  save_to writes the report as JSON."""
    import json
    s = np.ascontiguousarray(sample_sources, dtype=np.int32)
    best = Config()
    buf = C.create_string_buffer(1 << 20)
    st = start if start is not None else default_config(device=g.device)
    check(lib().ibfs_autotune(g._h, _i32p(s), len(s), C.byref(st), passes, reps, C.byref(best),
                              buf, len(buf)))
    report = json.loads(buf.value.decode())
    if save_to:
        with open(save_to, "w") as f:
            json.dump(report, f, indent=1)
    return best, report


def autotune_batch(g: DeviceGraph, sample_sources, contexts: int = 2, burst: int = 8,
                   passes: int = 2, reps: int = 2, start: Optional[Config] = None,
                   save_to: Optional[str] = None) -> tuple[Config, int, int, dict]:
    """Tunes the batched launcher for graph g on the sample sources; returns
    (config, contexts, burst, report)."""
    import json
    s = np.ascontiguousarray(sample_sources, dtype=np.int32)
    best = Config()
    bc, bb = C.c_int32(), C.c_int32()
    buf = C.create_string_buffer(1 << 20)
    st = start if start is not None else default_config(device=g.device, lane_bits=256)
    check(lib().ibfs_autotune_batch(g._h, _i32p(s), len(s), C.byref(st), contexts, burst, passes,
                                    reps, C.byref(best), C.byref(bc), C.byref(bb), buf, len(buf)))
    report = json.loads(buf.value.decode())
    if save_to:
        with open(save_to, "w") as f:
            json.dump(report, f, indent=1)
    return best, bc.value, bb.value, report


TUNED_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned")


def tuned_batch_settings(name: str) -> tuple[dict, int, int, Optional[str]]:
    """Batched-launcher settings from the autotuner report saved as
    tuned/batch_<name>.json: (config fields, contexts, burst, report file
    name); defaults (256 lanes, 2 contexts, bursts of 4) when there is none."""
    import json
    path = os.path.join(TUNED_DIR, f"batch_{name}.json")
    if not os.path.exists(path):
        return {"lane_bits": 256}, 2, 4, None
    with open(path) as f:
        best = dict(json.load(f)["best"])
    contexts = int(best.pop("contexts", 2))
    burst = int(best.pop("burst", 4))
    return best, contexts, burst, os.path.basename(path)


def config_from_dict(d: dict, device: int = 0) -> Config:
    """Config from the "best" record of an autotune report (the batched
    launcher's contexts and burst, which are not config fields, are skipped)."""
    return default_config(device=device, **{k: int(v) for k, v in d.items()
                                            if k not in ("contexts", "burst")})


class Batch:
    """Batched launcher: any number of sources packed into lane groups,
    `contexts` traversals in flight, device-driven levels in captured
    hipGraphs (ibfs_batch_*)."""

    def __init__(self, g: DeviceGraph, contexts: int = 2, burst: int = 8,
                 cfg: Optional[Config] = None, **overrides):
        if cfg is None:
            cfg = default_config(device=g.device, **overrides)
        elif overrides:
            for k, v in overrides.items():
                setattr(cfg, k, v)
        h = C.c_void_p()
        check(lib().ibfs_batch_create(g._h, C.byref(cfg), contexts, burst, C.byref(h)))
        self._h = h
        self.graph = g
        self.cfg = cfg
        self.last_stats = Stats()

    def run(self, sources, out: Optional[DeviceBuffer] = None) -> Optional[np.ndarray]:
        s = np.ascontiguousarray(sources, dtype=np.int32)
        n, V = len(s), self.graph.vert_count
        st = Stats()
        if out is None:
            depth = np.empty((n, V), dtype=np.int32)
            check(lib().ibfs_batch_run(self._h, _i32p(s), n, depth.ctypes.data, OUT_HOST,
                                       C.byref(st)))
            self.last_stats = st
            return depth
        if out.nbytes < 4 * n * V:
            raise ValueError("device buffer too small")
        check(lib().ibfs_batch_run(self._h, _i32p(s), n, C.c_void_p(out.ptr), OUT_DEVICE,
                                   C.byref(st)))
        self.last_stats = st
        return None

    def closeness(self, sources) -> tuple[np.ndarray, np.ndarray]:
        """Per-source (sum of depths, reached vertices), int64 each."""
        s = np.ascontiguousarray(sources, dtype=np.int32)
        dsum = np.zeros(len(s), dtype=np.int64)
        reach = np.zeros(len(s), dtype=np.int64)
        st = Stats()
        check(lib().ibfs_batch_closeness(self._h, _i32p(s), len(s), _i64p(dsum), _i64p(reach),
                                         C.byref(st)))
        self.last_stats = st
        return dsum, reach

    def close(self) -> None:
        if self._h:
            lib().ibfs_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching BFS server over one DeviceGraph (ibfs_server_*).
    submit() is thread-safe and returns a query id; the query's depth arrays
    are written to the given device buffer ([n][V] int32)."""

    def __init__(self, g: DeviceGraph, status_bytes: int = 0, cfg: Optional[Config] = None,
                 admit_window: Optional[int] = None, host_levels: Optional[bool] = None,
                 **overrides):
        if cfg is None:
            overrides.setdefault("lane_bits", 256)
            cfg = default_config(device=g.device, **overrides)
        h = C.c_void_p()
        check(lib().ibfs_server_create(g._h, C.byref(cfg), status_bytes, C.byref(h)))
        self._h = h
        if host_levels is not None:
            check(lib().ibfs_server_set_levels(h, int(bool(host_levels))))
        if admit_window is not None:
            check(lib().ibfs_server_set_admission(h, admit_window))
        self.graph = g
        self.cfg = cfg
        c, l = C.c_int32(), C.c_int32()
        check(lib().ibfs_server_contexts(h, C.byref(c), C.byref(l)))
        self.contexts, self.lanes = c.value, l.value
        # output buffer of every query not yet released: the worker's kernels
        # write into it until the query is done, so it must stay allocated
        self._outs: dict[int, object] = {}

    def set_admission(self, window_levels: int) -> None:
        """Sources join a running context only within window_levels levels of
        its cohort's start (0: only at cohort starts)."""
        check(lib().ibfs_server_set_admission(self._h, window_levels))

    def set_levels(self, host_levels: bool) -> None:
        """False (default): device-driven cohort levels; True: host-driven
        levels with lanes freed and reused one by one."""
        check(lib().ibfs_server_set_levels(self._h, int(bool(host_levels))))

    def counters(self) -> dict:
        lv, ad, inf = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().ibfs_server_counters(self._h, C.byref(lv), C.byref(ad), C.byref(inf)))
        return {"levels": lv.value, "admitted": ad.value, "admitted_in_flight": inf.value}

    def submit(self, sources, out) -> int:
        """Queues a query; its depth arrays ([n][V] int32) go to `out` (a
        DeviceBuffer or DeviceView), which the server keeps referenced until
        release() or close()."""
        s = np.ascontiguousarray(sources, dtype=np.int32)
        if out.nbytes < 4 * len(s) * self.graph.vert_count:
            raise ValueError("device buffer too small")
        qid = C.c_int64()
        check(lib().ibfs_server_submit(self._h, _i32p(s), len(s), C.c_void_p(out.ptr),
                                       C.byref(qid)))
        self._outs[qid.value] = out
        return qid.value

    def wait(self, qid: int, timeout_ms: float = -1.0) -> dict:
        info = _lib.QueryInfo()
        check(lib().ibfs_server_wait(self._h, qid, timeout_ms, C.byref(info)))
        return {"id": info.id, "n_sources": info.n_sources, "done": bool(info.done),
                "submit_ms": info.submit_ms, "admit_ms": info.admit_ms,
                "done_ms": info.done_ms, "latency_ms": info.done_ms - info.submit_ms}

    def release(self, qid: int) -> None:
        check(lib().ibfs_server_release(self._h, qid))
        self._outs.pop(qid, None)

    def close(self) -> None:
        if self._h:
            # destroy joins the worker, so no kernel writes an output afterwards
            lib().ibfs_server_destroy(self._h)
            self._h = C.c_void_p()
        self._outs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Mirror of the reference's host classes (graph.h / graph.cu, gpu_ibfs.cuh).

class graph:
    """graph(beg_file, csr_file, src_count): loads the two-file CSR
    (graph.cu:63-119).  gen_src() draws src_count seeded distinct sources of
    degree > 0 (replacing graph.cu:121-148); groupby() reorders src_list with
    the GPU GroupBy kernel (groupby.cu:36-149)."""

    def __init__(self, beg_file: Optional[str] = None, csr_file: Optional[str] = None,
                 src_count: int = 0, csr: Optional[HostCsr] = None):
        self.host = csr if csr is not None else HostCsr.load(beg_file, csr_file)
        self.src_count = src_count
        self.vert_count = self.host.vert_count
        self.edge_count = self.host.edge_count
        self.beg_pos, self.csr = self.host.arrays()
        self.src_list: Optional[np.ndarray] = None
        self._dev: Optional[DeviceGraph] = None

    def device_graph(self, device: int = 0) -> DeviceGraph:
        if self._dev is None:
            self._dev = DeviceGraph(self.host, device)
        return self._dev

    def gen_src(self, seed: int = 2) -> np.ndarray:
        self.src_list = self.host.gen_sources(self.src_count, seed)
        return self.src_list

    def groupby(self, device: int = 0) -> np.ndarray:
        if self.src_list is None:
            raise IbfsError(_lib.IBFS_ERR_STATE, "groupby() before gen_src()")
        self.src_list = self.device_graph(device).groupby(self.src_list)
        return self.src_list


class gpu_ibfs:
    """gpu_ibfs(g, concurr_count, sw_level, gpu_count, sml_shed, lrg_shed)
    (gpu_ibfs.cu:36-45).  traversal(grp_beg) runs the group of concurr_count
    sources starting at g.src_list[grp_beg] (traversal.cu:78) and returns
    their depth arrays int32[concurr_count][V] (-1 unreached); the reference
    leaves them as device status words.  sw_level >= 0 selects the
    reference's fixed switch (top-down while level <= sw_level); pass
    sw_level=None for the alpha/beta policy."""

    def __init__(self, g: graph, concurr_count: int = 128, sw_level: Optional[int] = None,
                 gpu_count: int = 1, sml_shed: int = 32, lrg_shed: int = 512, device: int = 0,
                 **overrides):
        if gpu_count != 1:
            raise IbfsError(_lib.IBFS_ERR_ARG, "one GPU per engine; shard sources across engines")
        if concurr_count not in (64, 128, 256):
            raise IbfsError(_lib.IBFS_ERR_ARG, "concurr_count must be 64, 128 or 256")
        self.g = g
        self.concurr_count = concurr_count
        policy = POLICY_ALPHA_BETA if sw_level is None else POLICY_FIXED
        cfg = default_config(lane_bits=concurr_count, sml_shed=sml_shed, lrg_shed=lrg_shed,
                             policy=policy, sw_level=sw_level or 0, device=device, **overrides)
        self.engine = Engine(g.device_graph(device), cfg)

    def traversal(self, grp_beg: int) -> np.ndarray:
        if self.g.src_list is None:
            raise IbfsError(_lib.IBFS_ERR_STATE, "traversal() before gen_src()")
        src = self.g.src_list[grp_beg:grp_beg + self.concurr_count]
        return self.engine.traversal(src)
