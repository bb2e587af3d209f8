"""ctypes binding of the CPU restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  The product package (beep_amd)
never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXK = 10


class Csr(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64),
                ("rowptr", C.POINTER(C.c_uint64)), ("col", C.POINTER(C.c_uint32))]


class Plan(C.Structure):
    _fields_ = [("k", C.c_int), ("sgc", C.c_int), ("lunmat", C.c_int), ("asc", C.c_int),
                ("max_qdeg", C.c_int), ("min_qdeg", C.c_int),
                ("seq", C.c_uint32 * MAXK), ("qdeg", C.c_uint32 * MAXK),
                ("qedge_ptr", C.c_uint32 * (MAXK + 1)), ("qedge", C.c_uint32 * 45),
                ("sym_ptr", C.c_uint32 * (MAXK + 1)), ("sym", C.c_uint32 * 45)]

    def lists(self):
        k = self.k
        qe = [list(self.qedge[self.qedge_ptr[i]:self.qedge_ptr[i + 1]]) for i in range(k)]
        sy = [list(self.sym[self.sym_ptr[i]:self.sym_ptr[i + 1]]) for i in range(k)]
        return dict(k=k, seq=list(self.seq[:k]), qedge=qe, sym=sy, lunmat=self.lunmat,
                    asc=self.asc, max_qdeg=self.max_qdeg, min_qdeg=self.min_qdeg,
                    qedge_ptr=list(self.qedge_ptr[:k + 1]), sym_ptr=list(self.sym_ptr[:k + 1]))


class Prep(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64),
                ("rowptr", C.POINTER(C.c_uint64)), ("col", C.POINTER(C.c_uint32)),
                ("ocol", C.POINTER(C.c_uint32)), ("nlow", C.POINTER(C.c_uint32)),
                ("roots", C.POINTER(C.c_uint32)), ("nroots", C.c_uint64),
                ("q0_size", C.c_uint64), ("boundary", C.c_uint64),
                ("cutoff", C.c_uint32), ("max_deg", C.c_uint32),
                ("peel_rounds", C.c_int), ("asc", C.c_int)]


class Result(C.Structure):
    _fields_ = [("count", C.c_uint64), ("hash_sum", C.c_uint64), ("hash_xor", C.c_uint64),
                ("n_emitted", C.c_uint64), ("n_written", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        u32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        L.or_read_edges.argtypes = [C.c_char_p, C.POINTER(u32p), C.POINTER(u32p)]
        L.or_read_edges.restype = C.c_int64
        L.or_free.argtypes = [C.c_void_p]
        L.or_csr_build.argtypes = [u32p, u32p, C.c_uint64, C.POINTER(Csr)]
        L.or_csr_free.argtypes = [C.POINTER(Csr)]
        L.or_cutoff.argtypes = [C.POINTER(Csr), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_cutoff.restype = C.c_uint32
        L.or_plan_build.argtypes = [u32p, u32p, C.c_uint64, C.c_int, C.POINTER(Plan)]
        L.or_prepare.argtypes = [C.POINTER(Csr), C.POINTER(Plan), C.c_uint32, C.POINTER(Prep)]
        L.or_prep_free.argtypes = [C.POINTER(Prep)]
        L.or_count.argtypes = [C.POINTER(Prep), C.POINTER(Plan), C.c_int, C.c_int, u32p,
                               C.c_uint64, C.POINTER(Result)]
        L.or_count_range.argtypes = [C.POINTER(Prep), C.POINTER(Plan), C.c_uint64, C.c_uint64,
                                     C.c_int, C.POINTER(Result)]
        L.or_count_ranks.argtypes = [C.POINTER(Prep), C.POINTER(Plan), u64p, C.c_uint64, C.c_int, C.c_int,
                                     C.POINTER(Result)]
        L.or_tuple_hash.argtypes = [u32p, C.c_int]
        L.or_tuple_hash.restype = C.c_uint64
        L.or_alg_bytes.argtypes = [C.POINTER(Prep)]
        L.or_alg_bytes.restype = C.c_uint64
        _lib = L
    return _lib


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


def read_edges(path):
    L = lib()
    s, d = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    n = L.or_read_edges(path.encode(), C.byref(s), C.byref(d))
    if n < 0:
        raise IOError(f"oracle cannot read {path} ({n})")
    src = np.ctypeslib.as_array(s, shape=(max(n, 1),))[:n].copy()
    dst = np.ctypeslib.as_array(d, shape=(max(n, 1),))[:n].copy()
    L.or_free(C.cast(s, C.c_void_p))
    L.or_free(C.cast(d, C.c_void_p))
    return src, dst


class Graph:
    """Unpeeled CSR built with the reference's from_edgelist rules."""

    def __init__(self, src, dst):
        L = lib()
        self._s, sp = _u32(src)
        self._d, dp = _u32(dst)
        self.c = Csr()
        r = L.or_csr_build(sp, dp, len(self._s), C.byref(self.c))
        if r:
            L.or_csr_free(C.byref(self.c))
            raise ValueError(f"or_csr_build failed ({r})")
        mean, sd = C.c_double(), C.c_double()
        self.cutoff = L.or_cutoff(C.byref(self.c), C.byref(mean), C.byref(sd))
        self.mean, self.sd = mean.value, sd.value

    @property
    def n(self):
        return self.c.n

    @property
    def m(self):
        return self.c.m

    def arrays(self):
        rp = np.ctypeslib.as_array(self.c.rowptr, shape=(self.c.n + 1,)).copy()
        col = np.ctypeslib.as_array(self.c.col, shape=(max(self.c.m, 1),))[:self.c.m].copy()
        return rp, col

    def __del__(self):
        try:
            lib().or_csr_free(C.byref(self.c))
        except Exception:
            pass


def plan(src, dst, sgc):
    L = lib()
    s, sp = _u32(src)
    d, dp = _u32(dst)
    p = Plan()
    r = L.or_plan_build(sp, dp, len(s), 1 if sgc else 0, C.byref(p))
    if r:
        raise ValueError(f"or_plan_build failed ({r})")
    return p


class Prepared:
    def __init__(self, g, p, cutoff=None):
        self.g, self.p = g, p
        self.pp = Prep()
        r = lib().or_prepare(C.byref(g.c), C.byref(p), g.cutoff if cutoff is None else cutoff,
                             C.byref(self.pp))
        if r:
            raise ValueError(f"or_prepare failed ({r})")

    def arrays(self):
        pp = self.pp
        n, m = pp.n, pp.m
        A = np.ctypeslib.as_array
        return dict(
            rowptr=A(pp.rowptr, shape=(n + 1,)).copy(),
            col=A(pp.col, shape=(max(m, 1),))[:m].copy(),
            ocol=A(pp.ocol, shape=(max(m, 1),))[:m].copy(),
            nlow=A(pp.nlow, shape=(max(n, 1),))[:n].copy(),
            roots=A(pp.roots, shape=(max(pp.nroots, 1),))[:pp.nroots].copy(),
            q0_size=pp.q0_size, boundary=pp.boundary, cutoff=pp.cutoff,
            max_deg=pp.max_deg, peel_rounds=pp.peel_rounds, asc=pp.asc)

    def count(self, nthreads=0, want_hash=False, match_cap=0):
        L = lib()
        res = Result()
        buf = None
        bp = None
        width = self.p.k - self.p.lunmat
        if match_cap:
            buf = np.zeros(match_cap * width, dtype=np.uint32)
            bp = buf.ctypes.data_as(C.POINTER(C.c_uint32))
        r = L.or_count(C.byref(self.pp), C.byref(self.p), nthreads, 1 if want_hash else 0, bp,
                       match_cap, C.byref(res))
        if r:
            raise RuntimeError("or_count failed")
        out = dict(count=res.count, hash_sum=res.hash_sum, hash_xor=res.hash_xor,
                   emitted=res.n_emitted, written=res.n_written)
        if match_cap:
            out["matches"] = buf[:res.n_written * width].reshape(-1, width)
        return out

    def count_range(self, r0, r1, nthreads=0):
        res = Result()
        if lib().or_count_range(C.byref(self.pp), C.byref(self.p), r0, r1, nthreads, C.byref(res)):
            raise RuntimeError("or_count_range failed")
        return res.count

    def count_ranks(self, ranks, nthreads=0, want_hash=False):
        """Count over the roots at the given ranks of the root list."""
        r = np.ascontiguousarray(np.asarray(ranks, dtype=np.uint64))
        res = Result()
        if lib().or_count_ranks(C.byref(self.pp), C.byref(self.p), r.ctypes.data_as(C.POINTER(C.c_uint64)),
                                len(r), nthreads, 1 if want_hash else 0, C.byref(res)):
            raise RuntimeError("or_count_ranks failed")
        return dict(count=res.count, hash_sum=res.hash_sum, hash_xor=res.hash_xor)

    def alg_bytes(self):
        return lib().or_alg_bytes(C.byref(self.pp))

    def __del__(self):
        try:
            lib().or_prep_free(C.byref(self.pp))
        except Exception:
            pass


class PreparedView(Prepared):
    """An or_prep over arrays prepared elsewhere (rowptr, col, ocol, nlow,
    roots as numpy arrays, kept alive here): the oracle's enumeration on a
    graph too large for its own single-threaded preparation.  Only the
    enumeration is then checked; the preparation is checked on smaller graphs."""

    def __init__(self, p, rowptr, col, ocol, nlow, roots, asc):
        self.g, self.p = None, p
        self._keep = [np.ascontiguousarray(rowptr, dtype=np.uint64), np.ascontiguousarray(col, dtype=np.uint32),
                      np.ascontiguousarray(ocol, dtype=np.uint32), np.ascontiguousarray(nlow, dtype=np.uint32),
                      np.ascontiguousarray(roots, dtype=np.uint32)]
        rp, cl, oc, nl, rt = self._keep
        pp = Prep()
        pp.n = len(rp) - 1
        pp.m = int(rp[-1])
        u32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        pp.rowptr = rp.ctypes.data_as(u64p)
        pp.col = cl.ctypes.data_as(u32p)
        pp.ocol = oc.ctypes.data_as(u32p)
        pp.nlow = nl.ctypes.data_as(u32p)
        pp.roots = rt.ctypes.data_as(u32p)
        pp.nroots = len(rt)
        pp.asc = 1 if asc else 0
        self.pp = pp

    def __del__(self):
        pass  # the arrays belong to this object, not to liboracle


def tuple_hash(t):
    a, ap = _u32(t)
    return lib().or_tuple_hash(ap, len(a))
