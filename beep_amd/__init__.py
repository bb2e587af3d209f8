"""beep_amd -- MI355X-native subgraph enumeration (BEEP's path) over libbeep.so.

The Python layer mirrors the reference's in-process interface
``graph::SG_Match<uint>(task, by, dev, cutoff, ndev).run(dataGraph, patGraph)``
(subgraph_matching.cuh:113, 178) on top of the C ABI in include/beep.h.  All
compute runs in libbeep.so (host C++ preparation + HIP kernels); there is no
Python or CPU fallback for the enumeration.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import BeepError, check, lib

SGM, SGC = 0, 1
BUILD_AUTO, BUILD_SCAN, BUILD_SEARCH = 0, 1, 2
EMIT_NONE, EMIT_HASH, EMIT_BUFFER = 0, 1, 2

__all__ = ["Graph", "Plan", "Prepared", "Batch", "Server", "SG_Match", "SGM", "SGC", "BeepError",
           "EMIT_NONE", "EMIT_HASH", "EMIT_BUFFER", "ETIMEDOUT", "ERANGE", "tuple_hash", "mode_of", "convert"]

ETIMEDOUT = -110
ERANGE = -34


def mode_of(task):
    """Maps the reference's -m task names (Config.h:80-84) to a mode."""
    if task in (SGM, SGC):
        return task
    if task == "sgm":
        return SGM
    if task == "sgc":
        return SGC
    raise ValueError(f"Unrecognized -mt option (Main TASK): {task}")


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_ffi.u32p)


class Graph:
    """Unpeeled CSR with the reference's loader rules and its cutoff statistic."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        d = _ffi.GraphDesc()
        check(lib().beep_graph_describe(self._h, C.byref(d)))
        self.n, self.m = d.n, d.m
        self.max_degree, self.cutoff = d.max_degree, d.cutoff
        self.mean, self.sd = d.mean, d.sd

    @classmethod
    def load(cls, path, sort_edges=False):
        h = C.c_void_p()
        check(lib().beep_graph_load(str(path).encode(), 1 if sort_edges else 0, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, src, dst):
        s, sp = _u32(src)
        d, dp = _u32(dst)
        if len(s) != len(d):
            raise ValueError("src and dst differ in length")
        h = C.c_void_p()
        check(lib().beep_graph_from_edges(sp, dp, len(s), C.byref(h)))
        return cls(h.value)

    @classmethod
    def rmat(cls, scale, samples, a=0.57, b=0.19, c=0.19, seed=1, n_keep=0):
        h = C.c_void_p()
        check(lib().beep_graph_rmat(scale, samples, a, b, c, seed, n_keep, C.byref(h)))
        return cls(h.value)

    @classmethod
    def chung_lu(cls, n, samples, gamma, i0, seed):
        h = C.c_void_p()
        check(lib().beep_graph_chung_lu(n, samples, gamma, i0, seed, C.byref(h)))
        return cls(h.value)

    def csr(self):
        d = _ffi.GraphDesc()
        check(lib().beep_graph_describe(self._h, C.byref(d)))
        rp = np.ctypeslib.as_array(d.rowptr, shape=(d.n + 1,)).copy()
        col = np.ctypeslib.as_array(d.col, shape=(max(d.m, 1),))[:d.m].copy()
        return rp, col

    def write_bel(self, path):
        check(lib().beep_graph_write_bel(self._h, str(path).encode()))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().beep_graph_free(h)
            self._h = C.c_void_p()


class Plan:
    """Query plan (matching order, back edges, symmetry chain, LUNMAT, asc/desc)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        d = _ffi.PlanDesc()
        check(lib().beep_plan_describe(self._h, C.byref(d)))
        k = d.k
        self.k, self.mode, self.lunmat, self.asc = k, d.mode, d.lunmat, d.asc
        self.max_qdeg, self.min_qdeg = d.max_qdeg, d.min_qdeg
        self.seq = list(d.seq[:k])
        self.qdeg = list(d.qdeg[:k])
        self.qedge_ptr = list(d.qedge_ptr[:k + 1])
        self.sym_ptr = list(d.sym_ptr[:k + 1])
        self.qedge = [list(d.qedge[d.qedge_ptr[i]:d.qedge_ptr[i + 1]]) for i in range(k)]
        self.sym = [list(d.sym[d.sym_ptr[i]:d.sym_ptr[i + 1]]) for i in range(k)]
        self.width = k - self.lunmat
        self.upper_rows = bool(d.upper_rows)

    @classmethod
    def from_mtx(cls, path, mode):
        h = C.c_void_p()
        check(lib().beep_plan_from_mtx(str(path).encode(), mode_of(mode), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, src, dst, mode):
        s, sp = _u32(src)
        d, dp = _u32(dst)
        h = C.c_void_p()
        check(lib().beep_plan_from_edges(sp, dp, len(s), mode_of(mode), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_undirected(cls, edges, mode):
        """Pattern from undirected edges; both directions are listed, rows grouped."""
        e = sorted({(a, b) for a, b in edges} | {(b, a) for a, b in edges})
        return cls.from_edges([a for a, _ in e], [b for _, b in e], mode)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().beep_plan_free(h)
            self._h = C.c_void_p()


class Prepared:
    """Peeled + degree-oriented graph with the root queue, resident in HBM once uploaded."""

    def __init__(self, graph, plan, _handle=None, _parent=None):
        h = _handle
        if h is None:
            h = C.c_void_p()
            check(lib().beep_prepare(graph._h, plan._h, C.byref(h)))
        self._h = h
        self._parent = _parent  # keeps the shared variant's owner alive
        self.graph, self.plan = graph, plan
        d = self.describe()
        self.n, self.m, self.n_roots = d.n, d.m, d.n_roots
        self.alg_bytes = d.alg_bytes

    def subset(self, ranks):
        """The same query over the roots at the given (strictly ascending)
        ranks of this one's root list."""
        r = np.ascontiguousarray(np.asarray(ranks, dtype=np.uint64))
        h = C.c_void_p()
        check(lib().beep_prep_subset(self._h, self.plan._h, r.ctypes.data_as(_ffi.u64p), len(r), C.byref(h)))
        return Prepared(self.graph, self.plan, _handle=h, _parent=self)

    def describe(self):
        d = _ffi.PrepDesc()
        check(lib().beep_prep_describe(self._h, C.byref(d)))
        return d

    def arrays(self, copy=True):
        """Host arrays of the prepared graph; with copy=False, views that are
        valid while this object (and its graph variant) lives."""
        d = self.describe()
        n, m = d.n, d.m

        def A(ptr, k, dtype):  # host view of k entries (an empty vector may be NULL)
            if k == 0 or not ptr:
                return np.zeros(0, dtype=dtype)
            v = np.ctypeslib.as_array(ptr, shape=(k,))
            return v.copy() if copy else v

        return dict(rowptr=A(d.rowptr, n + 1, np.uint64),
                    col=A(d.col, m, np.uint32),
                    ocol=A(d.ocol, m, np.uint32),
                    nlow=A(d.nlow, n, np.uint32),
                    roots=A(d.roots, d.n_roots, np.uint32),
                    q0_size=d.q0_size, boundary=d.boundary, cutoff=d.cutoff,
                    max_deg=d.max_degree, peel_rounds=d.peel_rounds, asc=d.asc,
                    n_hub_roots=d.n_hub_roots, n_lds_batches=d.n_lds_batches,
                    alg_bytes=d.alg_bytes, max_root_degree=d.max_root_degree,
                    p1=bool(d.p1), p2=bool(d.p2), q0_head=d.q0_head, q0_head_larger=d.q0_head_larger,
                    lds_class_roots=list(d.lds_class_roots), lds_class_units=list(d.lds_class_units),
                    lds_class_bytes=list(d.lds_class_bytes))

    def preconditions(self):
        """P1 (no vertex above the root window) and P2 (the head skip drops
        nothing), with the figures they are decided on."""
        d = self.describe()
        return dict(p1=bool(d.p1), p2=bool(d.p2), max_degree=d.max_degree, cutoff=d.cutoff,
                    window_hi=100 * d.cutoff, boundary=d.boundary, asc=bool(d.asc),
                    q0_head=d.q0_head, q0_head_larger=d.q0_head_larger)

    def upload(self, device=0):
        check(lib().beep_prep_upload(self._h, device))

    def count(self, device=0, emit=EMIT_NONE, match_cap=0, lanes=0, donate=True, donate_min=0,
              deadline_ms=0.0, profile=False, build_mode=0):
        """Runs the enumeration.  With a deadline the result may be partial:
        then ``stopped`` is 1 (the C call's BEEP_ETIMEDOUT is not raised).
        With ``profile`` the balance figures are filled in."""
        o = _opts(device, emit, lanes, donate, donate_min, deadline_ms, profile, build_mode)
        buf = None
        if emit == EMIT_BUFFER and match_cap:
            buf = np.zeros(match_cap * self.plan.width, dtype=np.uint32)
            o.match_cap = match_cap
            o.match_buf = buf.ctypes.data_as(_ffi.u32p)
        r = _ffi.Result()
        rc = lib().beep_count(self._h, self.plan._h, C.byref(o), C.byref(r))
        if rc != ETIMEDOUT:
            check(rc)
        out = _result_dict(r)
        if buf is not None:
            out["matches"] = buf[:r.n_written * self.plan.width].reshape(-1, self.plan.width)
        return out

    def count_noalloc(self, device=0, emit=EMIT_NONE, match_cap=0, lanes=0, donate=True, donate_min=0,
                      deadline_ms=0.0, build_mode=0, stream=None):
        """beep_count_noalloc: the same enumeration with its scratch taken
        from the library's fixed device array (no device allocation); the
        prepared query must have been uploaded.  ``stream`` is a hipStream_t
        as an int (None: the null stream)."""
        o = _opts(device, emit, lanes, donate, donate_min, deadline_ms, False, build_mode)
        buf = None
        if emit == EMIT_BUFFER and match_cap:
            buf = np.zeros(match_cap * self.plan.width, dtype=np.uint32)
            o.match_cap = match_cap
            o.match_buf = buf.ctypes.data_as(_ffi.u32p)
        r = _ffi.Result()
        rc = lib().beep_count_noalloc(self._h, self.plan._h, C.byref(o), C.c_void_p(stream), C.byref(r))
        if rc != ETIMEDOUT:
            check(rc)
        out = _result_dict(r)
        if buf is not None:
            out["matches"] = buf[:r.n_written * self.plan.width].reshape(-1, self.plan.width)
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().beep_prep_free(h)
            self._h = C.c_void_p()


def _opts(device=0, emit=EMIT_NONE, lanes=0, donate=True, donate_min=0, deadline_ms=0.0, profile=False,
          build_mode=0):
    o = _ffi.Opts()
    lib().beep_opts_default(C.byref(o))
    o.device, o.emit, o.lanes = device, emit, lanes
    o.donate = 1 if donate else 0
    if donate_min:
        o.donate_min = donate_min
    o.deadline_ms = float(deadline_ms)
    o.profile = 1 if profile else 0
    o.build_mode = build_mode
    return o


def _result_dict(r):
    return dict(count=r.count, hash_sum=r.hash_sum, hash_xor=r.hash_xor,
                emitted=r.n_emitted, written=r.n_written, t_enum_ms=r.t_enum_ms,
                t_hub_ms=r.t_hub_ms, t_lds_ms=r.t_lds_ms, donations=r.donations,
                stopped=r.stopped, balance_lds=r.balance_lds, balance_hub=r.balance_hub,
                donated_frac=r.donated_frac, lds_build_frac=r.lds_build_frac,
                t_lds_class_ms=list(r.t_lds_class_ms), balance_cu=r.balance_cu, cus_busy=r.cus_busy)


class Batch:
    """Several prepared queries run as one captured hipGraph (one persistent
    kernel over all their LDS batches plus per-query hub branches)."""

    def __init__(self, prepared, device=0, emit=EMIT_NONE, tune=None):
        self.prepared = list(prepared)
        n = len(self.prepared)
        tune = tune or [{}] * n
        self._opts = (_ffi.Opts * n)(*[_opts(device=device, emit=emit, lanes=t.get("lanes", 0),
                                             donate_min=t.get("donate_min", 0),
                                             build_mode=t.get("build_mode", BUILD_AUTO)) for t in tune])
        preps = (C.c_void_p * n)(*[p._h.value for p in self.prepared])
        plans = (C.c_void_p * n)(*[p.plan._h.value for p in self.prepared])
        h = C.c_void_p()
        check(lib().beep_batch_create(preps, plans, self._opts, n, C.byref(h)))
        self._h = h

    def run(self):
        n = len(self.prepared)
        res = (_ffi.Result * n)()
        t = C.c_double()
        check(lib().beep_batch_run(self._h, res, C.byref(t)))
        return [_result_dict(r) for r in res], t.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().beep_batch_free(h)
            self._h = C.c_void_p()


class Server:
    """Serving loop over one resident data graph (libbeep's beep_server).

    ``submit`` returns a ticket at once; a worker thread in the library admits
    queries in arrival order as lanes and scratch memory free up and runs them
    concurrently; ``poll`` returns finished queries as dicts with the counts,
    ``status`` (0, ETIMEDOUT when the deadline stopped it) and the queue,
    service and end-to-end latencies in ms.
    """

    def __init__(self, graph, device=0, lanes=4, pool_bytes=0, arena_cap=0, budget_ms=0.0, pilot_stride=0):
        o = _ffi.ServerOpts()
        lib().beep_server_opts_default(C.byref(o))
        o.device, o.lanes, o.pool_bytes, o.arena_cap = device, lanes, pool_bytes, arena_cap
        o.budget_ms, o.pilot_stride = float(budget_ms), pilot_stride
        h = C.c_void_p()
        check(lib().beep_server_create(graph._h, C.byref(o), C.byref(h)))
        self._h = h
        self.graph = graph      # the library keeps a pointer to it
        self._keep = {}         # ticket -> (plan, opts, buffer) alive until polled

    def prepare(self, plan):
        check(lib().beep_server_prepare(self._h, plan._h))

    def submit(self, plan, emit=EMIT_NONE, match_cap=0, lanes=0, donate=True, donate_min=0,
               deadline_ms=0.0):
        o = _opts(0, emit, lanes, donate, donate_min, deadline_ms)
        buf = None
        if emit == EMIT_BUFFER and match_cap:
            buf = np.zeros(match_cap * plan.width, dtype=np.uint32)
            o.match_cap = match_cap
            o.match_buf = buf.ctypes.data_as(_ffi.u32p)
        t = C.c_uint64()
        check(lib().beep_server_submit(self._h, plan._h, C.byref(o), C.byref(t)))
        self._keep[t.value] = (plan, o, buf)
        return t.value

    def poll(self, timeout_ms=-1):
        """One finished query, or None when none finished within timeout_ms."""
        d = _ffi.ServerDone()
        rc = lib().beep_server_poll(self._h, C.byref(d), int(timeout_ms))
        if rc == ETIMEDOUT:
            return None
        check(rc)
        plan, _, buf = self._keep.pop(d.ticket, (None, None, None))
        out = _result_dict(d.result)
        out.update(ticket=d.ticket, status=d.status, queue_ms=d.queue_ms, service_ms=d.service_ms,
                   latency_ms=d.latency_ms, footprint=d.footprint, predicted_ms=d.predicted_ms)
        if buf is not None and plan is not None:
            out["matches"] = buf[:d.result.n_written * plan.width].reshape(-1, plan.width)
        return out

    def stats(self):
        s = _ffi.ServerStats()
        check(lib().beep_server_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _ffi.ServerStats._fields_}

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().beep_server_free(h)
            self._h = C.c_void_p()
        self._keep.clear()

    def __del__(self):
        self.close()


class SG_Match:
    """Reference-shaped driver: SG_Match(task, by, dev, cutoff, ndev).run(data, pattern).

    ``cutoff`` is accepted for signature parity; like the reference CLI
    (main.cu:132) the cutoff actually used is the one computed from the data
    graph.  ``ndev`` must be 1: this build drives one GPU.
    """

    def __init__(self, task="sgm", by="node", dev=0, cutoff=768, ndev=1):
        if ndev != 1:
            raise ValueError("this build drives exactly one GPU (ndev must be 1)")
        if by not in ("node", 0):
            raise ValueError("only process-by-node is supported (the reference's only path)")
        self.mode = mode_of(task)
        self.dev = dev
        self.last = None

    def run(self, data, pattern, emit=EMIT_NONE):
        plan = pattern if isinstance(pattern, Plan) else Plan.from_mtx(pattern, self.mode)
        graph = data if isinstance(data, Graph) else Graph.load(data)
        prep = Prepared(graph, plan)
        prep.upload(self.dev)
        self.last = prep.count(device=self.dev, emit=emit)
        return self.last["count"]


CONVERSIONS = {"mtx-bel": 1, "tsv-bel": 2, "tsv-mtx": 3, "bel-mtx": 4, "txt-bel": 5}


def convert(kind, src, dst):
    """The reference's -m <kind> -g src -r dst format conversion (MtB_Writer)."""
    check(lib().beep_convert(CONVERSIONS[kind], str(src).encode(), str(dst).encode()))


def tuple_hash(t):
    a, ap = _u32(t)
    return lib().beep_tuple_hash(ap, len(a))
