"""ctypes binding of libibfs.so (the C ABI declared in include/ibfs.h).

The library is loaded from this package directory only; a missing library
raises ImportError-style errors instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libibfs.so"

IBFS_OK = 0
IBFS_ERR_ARG = -1
IBFS_ERR_IO = -2
IBFS_ERR_FORMAT = -3
IBFS_ERR_HIP = -4
IBFS_ERR_NOMEM = -5
IBFS_ERR_STATE = -6
IBFS_ERR_NODEVICE = -7

POLICY_FIXED = 0
POLICY_ALPHA_BETA = 1
POLICY_TD_ONLY = 2
POLICY_BU_ONLY = 3
POLICY_HYBRID = 4

OUT_HOST = 0
OUT_DEVICE = 1

_ERR_NAMES = {
    IBFS_ERR_ARG: "ERR_ARG", IBFS_ERR_IO: "ERR_IO", IBFS_ERR_FORMAT: "ERR_FORMAT",
    IBFS_ERR_HIP: "ERR_HIP", IBFS_ERR_NOMEM: "ERR_NOMEM", IBFS_ERR_STATE: "ERR_STATE",
    IBFS_ERR_NODEVICE: "ERR_NODEVICE",
}


class IbfsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("lane_bits", C.c_int32), ("sml_shed", C.c_int32), ("lrg_shed", C.c_int32),
        ("policy", C.c_int32), ("sw_level", C.c_int32), ("alpha", C.c_int32),
        ("beta", C.c_int32), ("wg_size", C.c_int32), ("td_chunk", C.c_int32),
        ("ring_levels", C.c_int32), ("kernel_timing", C.c_int32), ("device", C.c_int32),
        ("work_model", C.c_int32), ("emit_mode", C.c_int32), ("dd_grid", C.c_int32),
        ("inspect_vpt", C.c_int32), ("hybrid_gamma", C.c_int32), ("hybrid_hubs", C.c_int32),
        ("pack_sources", C.c_int32), ("split_cov", C.c_int32), ("phase_offsets", C.c_int32),
        ("tail_div", C.c_int32),
    ]


class LevelInfo(C.Structure):
    _fields_ = [
        ("level", C.c_int32), ("direction", C.c_int32), ("n_changed", C.c_int64),
        ("n_frontier", C.c_int64), ("m_frontier", C.c_int64), ("n_unvisited", C.c_int64),
        ("m_unvisited", C.c_int64), ("q_td", C.c_int64 * 3), ("q_bu", C.c_int64 * 3),
        ("bu_work", C.c_int64), ("bu_prefix", C.c_int64),
        ("expand_ms", C.c_float), ("inspect_ms", C.c_float), ("m_frontier_nonhub", C.c_int64),
    ]


class GraphOpts(C.Structure):
    _fields_ = [("hub_count", C.c_int32), ("nthreads", C.c_int32), ("reserved", C.c_int32 * 6)]


class QueryInfo(C.Structure):
    _fields_ = [("id", C.c_int64), ("n_sources", C.c_int32), ("done", C.c_int32),
                ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double)]


class Stats(C.Structure):
    _fields_ = [
        ("levels", C.c_int32), ("emit_windows", C.c_int32), ("traversed_edges", C.c_int64),
        ("total_ms", C.c_double), ("init_ms", C.c_double), ("expand_td_ms", C.c_double),
        ("expand_bu_ms", C.c_double), ("inspect_ms", C.c_double), ("emit_ms", C.c_double),
        ("n_td_launches", C.c_int32), ("n_bu_launches", C.c_int32),
        ("n_inspect_launches", C.c_int32), ("n_emit_launches", C.c_int32),
        ("expand_hy_ms", C.c_double), ("n_hy_launches", C.c_int32), ("n_split_levels", C.c_int32),
        ("n_tail_levels", C.c_int32), ("n_tail_reruns", C.c_int32),
    ]


P = C.c_void_p
I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every entry returns an int status
SIGNATURES = {
    "ibfs_abi_version": (C.c_int, []),
    "ibfs_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "ibfs_config_default": (C.c_int, [C.POINTER(Config)]),
    "ibfs_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ibfs_set_wait_timeout": (C.c_int, [C.c_double]),
    "ibfs_get_wait_timeout": (C.c_int, [C.POINTER(C.c_double)]),
    "ibfs_device_alloc": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(P)]),
    "ibfs_device_free": (C.c_int, [P]),
    "ibfs_memcpy_d2h": (C.c_int, [P, P, C.c_size_t]),
    "ibfs_memcpy_h2d": (C.c_int, [P, P, C.c_size_t]),
    "ibfs_memcpy_d2d": (C.c_int, [P, P, C.c_size_t]),
    "ibfs_hostcsr_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(P)]),
    "ibfs_hostcsr_save": (C.c_int, [P, C.c_char_p, C.c_char_p]),
    "ibfs_hostcsr_create": (C.c_int, [I64P, C.c_int64, I32P, C.c_int64, C.POINTER(P)]),
    "ibfs_hostcsr_info": (C.c_int, [P, I64P, I64P]),
    "ibfs_hostcsr_data": (C.c_int, [P, C.POINTER(I64P), C.POINTER(I32P)]),
    "ibfs_hostcsr_free": (C.c_int, [P]),
    "ibfs_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                C.c_uint64, C.c_int32, C.c_int32, C.POINTER(P)]),
    "ibfs_gen_chung_lu": (C.c_int, [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                    C.c_int32, C.POINTER(P)]),
    "ibfs_gen_sources": (C.c_int, [P, C.c_int64, C.c_uint64, I32P]),
    "ibfs_convert_text": (C.c_int, [C.c_char_p, C.c_int32, C.c_int64, C.POINTER(P)]),
    "ibfs_convert_text_files": (C.c_int, [C.c_char_p, C.c_int32, C.c_int64, C.c_char_p, I64P,
                                          I64P]),
    "ibfs_graph_create": (C.c_int, [I64P, C.c_int64, I32P, C.c_int64, C.c_int, C.POINTER(P)]),
    "ibfs_graph_opts_default": (C.c_int, [C.POINTER(GraphOpts)]),
    "ibfs_graph_create_ex": (C.c_int, [I64P, C.c_int64, I32P, C.c_int64, C.c_int,
                                       C.POINTER(GraphOpts), C.POINTER(P)]),
    "ibfs_graph_from_host": (C.c_int, [P, C.c_int, C.POINTER(P)]),
    "ibfs_graph_load_files": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(P)]),
    "ibfs_graph_info": (C.c_int, [P, I64P, I64P]),
    "ibfs_graph_layout": (C.c_int, [P, I64P, I64P]),
    "ibfs_graph_hub_entries": (C.c_int, [P, C.c_int64, I64P]),
    "ibfs_graph_destroy": (C.c_int, [P]),
    "ibfs_groupby": (C.c_int, [P, I32P, C.c_int64, I32P]),
    "ibfs_engine_create": (C.c_int, [P, C.POINTER(Config), C.POINTER(P)]),
    "ibfs_engine_destroy": (C.c_int, [P]),
    "ibfs_traversal": (C.c_int, [P, I32P, C.c_int32, P, C.c_int32, C.POINTER(Stats)]),
    "ibfs_engine_level_log": (C.c_int, [P, C.POINTER(LevelInfo), C.c_int32,
                                        C.POINTER(C.c_int32)]),
    "ibfs_run_groups": (C.c_int, [P, I32P, C.c_int64, P, C.c_int32, C.POINTER(Stats)]),
    "ibfs_traversal_noalloc": (C.c_int, [P, C.POINTER(Config), I32P, C.c_int32, P, I64P]),
    "ibfs_traversed_edges": (C.c_int, [P, I32P, I64P]),
    "ibfs_batch_create": (C.c_int, [P, C.POINTER(Config), C.c_int32, C.c_int32, C.POINTER(P)]),
    "ibfs_batch_run": (C.c_int, [P, I32P, C.c_int64, P, C.c_int32, C.POINTER(Stats)]),
    "ibfs_batch_destroy": (C.c_int, [P]),
    "ibfs_batch_closeness": (C.c_int, [P, I32P, C.c_int64, I64P, I64P, C.POINTER(Stats)]),
    "ibfs_server_create": (C.c_int, [P, C.POINTER(Config), C.c_int64, C.POINTER(P)]),
    "ibfs_server_contexts": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ibfs_server_set_admission": (C.c_int, [P, C.c_int32]),
    "ibfs_server_set_levels": (C.c_int, [P, C.c_int32]),
    "ibfs_server_counters": (C.c_int, [P, I64P, I64P, I64P]),
    "ibfs_server_submit": (C.c_int, [P, I32P, C.c_int32, P, C.POINTER(C.c_int64)]),
    "ibfs_server_wait": (C.c_int, [P, C.c_int64, C.c_double, C.POINTER(QueryInfo)]),
    "ibfs_server_release": (C.c_int, [P, C.c_int64]),
    "ibfs_server_destroy": (C.c_int, [P]),
    "ibfs_autotune": (C.c_int, [P, I32P, C.c_int64, C.POINTER(Config), C.c_int32, C.c_int32,
                                C.POINTER(Config), C.c_char_p, C.c_size_t]),
    "ibfs_autotune_batch": (C.c_int, [P, I32P, C.c_int64, C.POINTER(Config), C.c_int32, C.c_int32,
                                      C.c_int32, C.c_int32, C.POINTER(Config), I32P, I32P,
                                      C.c_char_p, C.c_size_t]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads the in-tree libibfs.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ibfs_amd.build`")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().ibfs_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc != IBFS_OK:
        raise IbfsError(rc, last_error())
