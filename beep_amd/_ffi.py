"""ctypes declarations of libbeep.so (include/beep.h)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# BEEP_LIB names another build of the library (e.g. for side-by-side timing).
LIB_PATH = os.environ.get("BEEP_LIB") or os.path.join(HERE, "libbeep.so")
MAXK = 10

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class GraphDesc(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64), ("max_degree", C.c_uint32),
                ("cutoff", C.c_uint32), ("mean", C.c_double), ("sd", C.c_double),
                ("rowptr", u64p), ("col", u32p)]


class PlanDesc(C.Structure):
    _fields_ = [("k", C.c_int), ("mode", C.c_int), ("lunmat", C.c_int), ("asc", C.c_int),
                ("max_qdeg", C.c_int), ("min_qdeg", C.c_int),
                ("seq", C.c_uint32 * MAXK), ("qdeg", C.c_uint32 * MAXK),
                ("qedge_ptr", C.c_uint32 * (MAXK + 1)), ("qedge", C.c_uint32 * 45),
                ("sym_ptr", C.c_uint32 * (MAXK + 1)), ("sym", C.c_uint32 * 45), ("upper_rows", C.c_int)]


class PrepDesc(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64), ("n_roots", C.c_uint64),
                ("q0_size", C.c_uint64), ("boundary", C.c_uint64), ("cutoff", C.c_uint32),
                ("max_degree", C.c_uint32), ("max_root_degree", C.c_uint32),
                ("peel_rounds", C.c_int), ("asc", C.c_int), ("alg_bytes", C.c_uint64),
                ("n_hub_roots", C.c_uint64), ("n_lds_batches", C.c_uint64),
                ("rowptr", u64p), ("col", u32p), ("ocol", u32p), ("nlow", u32p), ("roots", u32p),
                ("p1", C.c_int), ("p2", C.c_int), ("q0_head", C.c_uint32), ("q0_head_larger", C.c_uint32),
                ("lds_class_roots", C.c_uint64 * 8), ("lds_class_units", C.c_uint64 * 8),
                ("lds_class_bytes", C.c_uint64 * 8)]


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("emit", C.c_int), ("match_cap", C.c_uint64),
                ("match_buf", u32p), ("lanes", C.c_int), ("donate", C.c_int),
                ("donate_min", C.c_uint32), ("timing", C.c_int), ("deadline_ms", C.c_double),
                ("profile", C.c_int), ("build_mode", C.c_int)]


class Result(C.Structure):
    _fields_ = [("count", C.c_uint64), ("hash_sum", C.c_uint64), ("hash_xor", C.c_uint64),
                ("n_emitted", C.c_uint64), ("n_written", C.c_uint64), ("t_enum_ms", C.c_double),
                ("t_hub_ms", C.c_double), ("t_lds_ms", C.c_double), ("donations", C.c_uint64),
                ("stopped", C.c_uint64), ("balance_lds", C.c_double), ("balance_hub", C.c_double),
                ("donated_frac", C.c_double), ("lds_build_frac", C.c_double),
                ("t_lds_class_ms", C.c_double * 8), ("balance_cu", C.c_double), ("cus_busy", C.c_uint64)]


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int), ("lanes", C.c_int), ("pool_bytes", C.c_uint64),
                ("arena_cap", C.c_uint64), ("budget_ms", C.c_double), ("pilot_stride", C.c_uint32)]


class ServerDone(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int), ("result", Result),
                ("queue_ms", C.c_double), ("service_ms", C.c_double), ("latency_ms", C.c_double),
                ("footprint", C.c_uint64), ("predicted_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("submitted", "completed", "stopped", "failed", "pending",
                                          "running", "peak_running", "pool_bytes", "pool_used",
                                          "pool_peak", "variants", "root_queues", "refused",
                                          "pilots")]


# Every function declared in include/beep.h with its ctypes signature.
SIGNATURES = {
    "beep_graph_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    "beep_graph_from_edges": (C.c_int, [u32p, u32p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "beep_graph_rmat": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                  C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "beep_graph_chung_lu": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_uint64,
                                      C.POINTER(C.c_void_p)]),
    "beep_graph_describe": (C.c_int, [C.c_void_p, C.POINTER(GraphDesc)]),
    "beep_graph_write_bel": (C.c_int, [C.c_void_p, C.c_char_p]),
    "beep_graph_free": (None, [C.c_void_p]),
    "beep_convert": (C.c_int, [C.c_int, C.c_char_p, C.c_char_p]),
    "beep_plan_from_mtx": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    "beep_plan_from_edges": (C.c_int, [u32p, u32p, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    "beep_plan_describe": (C.c_int, [C.c_void_p, C.POINTER(PlanDesc)]),
    "beep_plan_free": (None, [C.c_void_p]),
    "beep_prepare": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "beep_prep_describe": (C.c_int, [C.c_void_p, C.POINTER(PrepDesc)]),
    "beep_prep_subset": (C.c_int, [C.c_void_p, C.c_void_p, u64p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "beep_prep_upload": (C.c_int, [C.c_void_p, C.c_int]),
    "beep_prep_free": (None, [C.c_void_p]),
    "beep_opts_default": (None, [C.POINTER(Opts)]),
    "beep_count": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(Result)]),
    "beep_count_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts), C.c_void_p, C.POINTER(Result)]),
    "beep_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(Opts), C.c_int,
                                    C.POINTER(C.c_void_p)]),
    "beep_batch_run": (C.c_int, [C.c_void_p, C.POINTER(Result), C.POINTER(C.c_double)]),
    "beep_batch_free": (None, [C.c_void_p]),
    "beep_server_opts_default": (None, [C.POINTER(ServerOpts)]),
    "beep_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    "beep_server_prepare": (C.c_int, [C.c_void_p, C.c_void_p]),
    "beep_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(C.c_uint64)]),
    "beep_server_poll": (C.c_int, [C.c_void_p, C.POINTER(ServerDone), C.c_int]),
    "beep_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "beep_server_free": (None, [C.c_void_p]),
    "beep_tuple_hash": (C.c_uint64, [u32p, C.c_int]),
    "beep_last_error": (C.c_char_p, []),
    "beep_version": (C.c_char_p, []),
}

_lib = None


def lib():
    """Loads libbeep.so; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; "
                               "g.build()'` or `make -C beep_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class BeepError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def check(rc):
    if rc != 0:
        raise BeepError(rc, lib().beep_last_error().decode(errors="replace"))
