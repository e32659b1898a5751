"""ctypes binding of libgpm (include/gpm.h) for tests, bench and scripts.

The engine is the HIP library; this module only converts numpy arrays to
pointers and status codes to exceptions.  There is no fallback: if
lib/libgpm.so is missing, every call raises GpmError.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgpm.so")

# gpm.h constants
GPM_OK = 0
CSR_SYMMETRIZE, CSR_DEDUP, CSR_DROP_SELF, CSR_SORT = 0x1, 0x2, 0x4, 0x8
CSR_CLEAN = CSR_SYMMETRIZE | CSR_DEDUP | CSR_DROP_SELF | CSR_SORT
GRAPH_ORIENTED, GRAPH_SYMMETRIC = 0x1, 0x2
SCHED_DEGREE, SCHED_LEARNED, SCHED_DEGREE_SPLIT = 0, 1, 2
TC_EDGE, TC_VERTEX = 0, 1
FORM_G2MINER, FORM_PATTERN, FORM_BLOCK_PREFILTER = 0, 1, 2
PATTERNS = {"P1": 1, "P2": 2, "P3": 3, "P4": 4, "P5": 5, "P6": 6, "P7": 7,
            "P8": 8, "Motif4": 9}

_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f32p = C.POINTER(C.c_float)


class GpmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gpm status {status}: {msg}")
        self.status = status


class Sched(C.Structure):
    _fields_ = [("policy", C.c_int), ("chunk", C.c_uint32),
                ("root_begin", C.c_uint32), ("root_end", C.c_uint32),
                ("root_cost", _f32p), ("wave_cap", C.c_uint32),
                ("n_root_list", C.c_uint32), ("root_list", _u32p)]


class Query(C.Structure):
    _fields_ = [("n", C.c_uint32), ("n_edges", C.c_uint32),
                ("edges", _u32p), ("labels", _u32p)]


class Tuning(C.Structure):
    _fields_ = [("autotune", C.c_int), ("policy", C.c_int),
                ("chunk", C.c_uint32), ("use_graph", C.c_int),
                ("root_model", _f32p), ("set_variant", C.c_int), ("pass_budget", C.c_uint32)]


_lib = None


def lib() -> C.CDLL:
    """Loads libgpm.so once; raises GpmError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GpmError(-1, f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.gpm_last_error.restype = C.c_char_p
        L.gpm_free.argtypes = [C.c_void_p]
        L.gpm_free.restype = None
        _lib = L
    return _lib


def _check(st: int) -> None:
    if st != GPM_OK:
        raise GpmError(st, lib().gpm_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray, ctype=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(ctype))


def _take(p, count: int, dtype=np.uint32) -> np.ndarray:
    """Copies a library-owned array into numpy and frees it."""
    if count == 0:
        out = np.zeros(0, dtype=dtype)
    else:
        buf = (C.c_uint32 * count).from_address(C.cast(p, C.c_void_p).value)
        out = np.frombuffer(buf, dtype=dtype, count=count).copy()
    lib().gpm_free(C.cast(p, C.c_void_p))
    return out


@dataclass
class HostCSR:
    offsets: np.ndarray  # uint32, n+1
    adj: np.ndarray      # uint32, m

    @property
    def n(self) -> int:
        return len(self.offsets) - 1

    @property
    def m(self) -> int:
        return len(self.adj)

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets.astype(np.int64))


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def csr_from_edges(n: int, src, dst, flags: int = CSR_CLEAN) -> HostCSR:
    src, dst = _u32(src), _u32(dst)
    assert len(src) == len(dst)
    po, pa, m = _u32p(), _u32p(), C.c_uint64()
    _check(lib().gpm_csr_from_edges(C.c_uint32(n), _ptr(src), _ptr(dst),
                                    C.c_uint64(len(src)), C.c_uint32(flags),
                                    C.byref(po), C.byref(pa), C.byref(m)))
    return HostCSR(_take(po, n + 1), _take(pa, m.value))


def orient(g: HostCSR) -> HostCSR:
    po, pa, m = _u32p(), _u32p(), C.c_uint64()
    _check(lib().gpm_csr_orient(C.c_uint32(g.n), _ptr(g.offsets), _ptr(g.adj),
                                C.byref(po), C.byref(pa), C.byref(m)))
    return HostCSR(_take(po, g.n + 1), _take(pa, m.value))


def sort_by_degree(g: HostCSR):
    po, pa, pp = _u32p(), _u32p(), _u32p()
    _check(lib().gpm_csr_sort_by_degree(C.c_uint32(g.n), _ptr(g.offsets),
                                        _ptr(g.adj), C.byref(po), C.byref(pa),
                                        C.byref(pp)))
    return HostCSR(_take(po, g.n + 1), _take(pa, g.m)), _take(pp, g.n)


def load_mtx(path: str):
    n, ps, pd, ne = C.c_uint32(), _u32p(), _u32p(), C.c_uint64()
    _check(lib().gpm_load_mtx(path.encode(), C.byref(n), C.byref(ps),
                              C.byref(pd), C.byref(ne)))
    return n.value, _take(ps, ne.value), _take(pd, ne.value)


def load_labeled(path: str):
    """sub_ceci text graph -> (HostCSR symmetric sorted, labels, n_labels)."""
    n, po, pa, m, pl, nl = C.c_uint32(), _u32p(), _u32p(), C.c_uint64(), _u32p(), C.c_uint32()
    _check(lib().gpm_load_labeled(path.encode(), C.byref(n), C.byref(po), C.byref(pa),
                                  C.byref(m), C.byref(pl), C.byref(nl)))
    off = _take(po, n.value + 1)
    adj = _take(pa, m.value)
    lab = _take(pl, n.value)
    return HostCSR(off, adj), lab, nl.value


def load_ceci_bin(degree_path: str, adj_path: str, label_path: Optional[str] = None):
    """sub_ceci compressed graph -> (HostCSR with sorted lists, labels or None, n_labels)."""
    n, po, pa, m, pl, nl = C.c_uint32(), _u32p(), _u32p(), C.c_uint64(), _u32p(), C.c_uint32()
    _check(lib().gpm_load_ceci_bin(degree_path.encode(), adj_path.encode(),
                                   label_path.encode() if label_path else None, C.byref(n),
                                   C.byref(po), C.byref(pa), C.byref(m),
                                   C.byref(pl) if label_path else None,
                                   C.byref(nl) if label_path else None))
    g = HostCSR(_take(po, n.value + 1), _take(pa, m.value))
    lab = _take(pl, n.value) if label_path else None
    return g, lab, nl.value


def store_ceci_bin(degree_path: str, adj_path: str, g: HostCSR, labels=None,
                   label_path: Optional[str] = None) -> None:
    lab = _u32(labels) if labels is not None else None
    _check(lib().gpm_store_ceci_bin(degree_path.encode(), adj_path.encode(),
                                    label_path.encode() if label_path else None,
                                    C.c_uint32(g.n), _ptr(_u32(g.offsets)), _ptr(_u32(g.adj)),
                                    _ptr(lab) if lab is not None else None))


def load_gbin(directory: str):
    """GraphMiner binary directory -> (HostCSR in file order, max_degree)."""
    n, po, pa, m, md = C.c_uint32(), _u32p(), _u32p(), C.c_uint64(), C.c_uint32()
    _check(lib().gpm_load_gbin(directory.encode(), C.byref(n), C.byref(po), C.byref(pa),
                               C.byref(m), C.byref(md)))
    return HostCSR(_take(po, n.value + 1), _take(pa, m.value)), md.value


def store_gbin(directory: str, g: HostCSR) -> None:
    _check(lib().gpm_store_gbin(directory.encode(), C.c_uint32(g.n), _ptr(_u32(g.offsets)),
                                _ptr(_u32(g.adj))))


def gen_er(n: int, m: int, seed: int):
    ps, pd, ne = _u32p(), _u32p(), C.c_uint64()
    _check(lib().gpm_gen_er(C.c_uint32(n), C.c_uint64(m), C.c_uint64(seed),
                            C.byref(ps), C.byref(pd), C.byref(ne)))
    return _take(ps, ne.value), _take(pd, ne.value)


def gen_rmat(scale: int, edge_factor: int, seed: int, a=0.57, b=0.19, c=0.19):
    ps, pd, ne = _u32p(), _u32p(), C.c_uint64()
    _check(lib().gpm_gen_rmat(C.c_int(scale), C.c_int(edge_factor),
                              C.c_double(a), C.c_double(b), C.c_double(c),
                              C.c_uint64(seed), C.byref(ps), C.byref(pd),
                              C.byref(ne)))
    return _take(ps, ne.value), _take(pd, ne.value)


def gen_labels(n: int, n_labels: int, seed: int) -> np.ndarray:
    pl = _u32p()
    _check(lib().gpm_gen_labels(C.c_uint32(n), C.c_uint32(n_labels),
                                C.c_uint64(seed), C.byref(pl)))
    return _take(pl, n)


def er_graph(n: int, m: int, seed: int) -> HostCSR:
    """Symmetric, deduplicated, sorted G(n, M) CSR."""
    s, d = gen_er(n, m, seed)
    return csr_from_edges(n, s, d, CSR_CLEAN)


def rmat_graph(scale: int, edge_factor: int, seed: int, a=0.57, b=0.19,
               c=0.19) -> HostCSR:
    """Symmetrised, deduplicated, self-loop-free, sorted R-MAT CSR."""
    s, d = gen_rmat(scale, edge_factor, seed, a, b, c)
    return csr_from_edges(1 << scale, s, d, CSR_CLEAN)


class DeviceGraph:
    """A CSR resident in HBM (gpm_graph handle)."""

    def __init__(self, g: HostCSR, oriented: bool, labels=None,
                 n_labels: int = 0, _handle=None):
        self._h = C.c_void_p()
        self.oriented = oriented
        if _handle is not None:
            self._h = _handle
            return
        flags = GRAPH_ORIENTED if oriented else GRAPH_SYMMETRIC
        lab_ptr = None
        if labels is not None:
            self._labels = _u32(labels)
            lab_ptr = _ptr(self._labels)
            n_labels = n_labels or int(self._labels.max()) + 1
        _check(lib().gpm_graph_create(_ptr(_u32(g.offsets)), C.c_uint32(g.n),
                                      _ptr(_u32(g.adj)), C.c_uint64(g.m),
                                      lab_ptr, C.c_uint32(n_labels),
                                      C.c_uint32(flags), C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().gpm_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def info(self):
        n, m, d = C.c_uint32(), C.c_uint64(), C.c_uint32()
        _check(lib().gpm_graph_info(self._h, C.byref(n), C.byref(m), C.byref(d)))
        return n.value, m.value, d.value

    @classmethod
    def build(cls, n: int, src, dst, flags: int = CSR_CLEAN, oriented: bool = False,
              labels=None, n_labels: int = 0) -> "DeviceGraph":
        """Device graph built on the GPU from an edge list (gpm_graph_build)."""
        src, dst = _u32(src), _u32(dst)
        assert len(src) == len(dst)
        lab = _u32(labels) if labels is not None else None
        if lab is not None:
            n_labels = n_labels or (int(lab.max()) + 1 if len(lab) else 1)
        h = C.c_void_p()
        _check(lib().gpm_graph_build(C.c_uint32(n), _ptr(src), _ptr(dst), C.c_uint64(len(src)),
                                     C.c_uint32(flags), _ptr(lab) if lab is not None else None,
                                     C.c_uint32(n_labels),
                                     C.c_uint32(GRAPH_ORIENTED if oriented else GRAPH_SYMMETRIC),
                                     C.byref(h)))
        return cls(None, oriented, _handle=h)

    def relabel_by_degree(self):
        """(new DeviceGraph, perm) with perm[old] = new id (gpm_graph_relabel_by_degree)."""
        h, pp = C.c_void_p(), _u32p()
        _check(lib().gpm_graph_relabel_by_degree(self._h, C.byref(h), C.byref(pp)))
        n = self.info()[0]
        return DeviceGraph(None, self.oriented, _handle=h), _take(pp, n)

    def to_host(self) -> HostCSR:
        n, m, _ = self.info()
        off = np.zeros(n + 1, dtype=np.uint32)
        adj = np.zeros(max(m, 1), dtype=np.uint32)
        _check(lib().gpm_graph_copy_csr(self._h, _ptr(off), _ptr(adj)))
        return HostCSR(off, adj[:m])

    def orient(self) -> "DeviceGraph":
        h = C.c_void_p()
        _check(lib().gpm_graph_orient(self._h, C.byref(h)))
        return DeviceGraph(None, True, _handle=h)

    def last_kernel_ms(self) -> float:
        ms = C.c_double()
        _check(lib().gpm_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def kernel_breakdown(self):
        """[(kernel name, ms)] of the last call, in launch order."""
        ms = (C.c_double * 8)()
        names = (C.c_char_p * 8)()
        cnt = C.c_int()
        _check(lib().gpm_last_kernel_breakdown(self._h, ms, names, C.c_int(8), C.byref(cnt)))
        return [(names[i].decode(), ms[i]) for i in range(min(8, cnt.value))]

    def tc(self, variant: int = TC_VERTEX) -> int:
        out = C.c_uint64()
        _check(lib().gpm_tc(self._h, C.c_int(variant), C.byref(out)))
        return out.value

    def tc_repeat(self, variant: int, reps: int) -> int:
        """reps back-to-back counts; last_kernel_ms() is the mean per count."""
        out = C.c_uint64()
        _check(lib().gpm_tc_repeat(self._h, C.c_int(variant), C.c_int(reps), C.byref(out)))
        return out.value

    def tc_repeat_fused(self, reps: int) -> int:
        """reps counts, 32 per launch (grid y); last_kernel_ms() is the mean per count."""
        out = C.c_uint64()
        _check(lib().gpm_tc_repeat_fused(self._h, C.c_int(reps), C.byref(out)))
        return out.value

    def clique(self, k: int, policy: int = SCHED_DEGREE, chunk: int = 0,
               root_begin: int = 0, root_end: int = 0,
               cost: Optional[np.ndarray] = None, wave_cap: int = 0,
               roots: Optional[np.ndarray] = None) -> int:
        s = Sched()
        s.policy, s.chunk = policy, chunk
        s.root_begin, s.root_end, s.wave_cap = root_begin, root_end, wave_cap
        if cost is not None:
            self._cost = np.ascontiguousarray(cost, dtype=np.float32)
            s.root_cost = _ptr(self._cost, C.c_float)
        if roots is not None:
            self._roots = _u32(roots)
            s.n_root_list = len(self._roots)
            s.root_list = _ptr(self._roots)
        out = C.c_uint64()
        _check(lib().gpm_clique(self._h, C.c_int(k), C.byref(s), C.byref(out)))
        return out.value

    def clique_all(self, policy: int = SCHED_DEGREE, cost: Optional[np.ndarray] = None):
        """(k3, k4, k5) clique counts from one pass over the roots."""
        s = Sched()
        s.policy = policy
        if cost is not None:
            self._cost = np.ascontiguousarray(cost, dtype=np.float32)
            s.root_cost = _ptr(self._cost, C.c_float)
        out = (C.c_uint64 * 3)()
        _check(lib().gpm_clique_all(self._h, C.byref(s), out))
        return tuple(int(x) for x in out)

    def clique_plan_bytes(self, k: int) -> int:
        out = C.c_uint64()
        _check(lib().gpm_clique_plan_bytes(self._h, C.c_int(k), C.byref(out)))
        return out.value

    def clique_root_work(self, k: int) -> np.ndarray:
        n = self.info()[0]
        w = np.zeros(n, dtype=np.uint64)
        _check(lib().gpm_clique_root_work(self._h, C.c_int(k), _ptr(w, C.c_uint64)))
        return w

    def pattern(self, name: str, formulation: int = FORM_G2MINER):
        out = (C.c_uint64 * 6)()
        scalar = C.c_uint64()
        _check(lib().gpm_pattern(self._h, C.c_int(PATTERNS[name]),
                                 C.c_int(formulation), out, C.byref(scalar)))
        return [int(x) for x in out], scalar.value

    @staticmethod
    def _queries(queries):
        arr = (Query * len(queries))()
        keep = []
        for i, (qn, edges, labels) in enumerate(queries):
            e = _u32(np.asarray(edges, dtype=np.uint32).reshape(-1))
            l = _u32(labels)
            keep += [e, l]
            arr[i].n, arr[i].n_edges = qn, len(e) // 2
            arr[i].edges, arr[i].labels = _ptr(e), _ptr(l)
        return arr, keep

    def match_batch(self, queries: Sequence, autotune: int = 1,
                    policy: int = SCHED_DEGREE, chunk: int = 0,
                    use_graph: int = 0, model: Optional[np.ndarray] = None,
                    set_variant: int = 0, pass_budget: int = 0) -> np.ndarray:
        """queries: sequence of (n, edges [(u, v), ...], labels [..]); model: packed
        root-cost MLP (4553 float32, see MatchModel) for the LEARNED policy;
        set_variant: MATCH_SETV_AUTO / _BINARY / _MERGE; pass_budget: search
        steps between a running task's checks for a drained queue, after which
        it is cut finer and counted again (0: automatic, passes on short
        batches; PASS_BUDGET_OFF: one launch)."""
        arr, keep = self._queries(queries)
        t = Tuning()
        t.autotune, t.policy, t.chunk, t.use_graph = autotune, policy, chunk, use_graph
        t.set_variant, t.pass_budget = set_variant, pass_budget
        if model is not None:
            self._model = np.ascontiguousarray(model, dtype=np.float32)
            assert self._model.shape == (4553,)
            t.root_model = _ptr(self._model, C.c_float)
        out = np.zeros(len(queries), dtype=np.uint64)
        _check(lib().gpm_match_batch(self._h, arr, C.c_int(len(queries)),
                                     C.byref(t), _ptr(out, C.c_uint64)))
        return out

    def match_root_work(self, query, policy: int = SCHED_DEGREE) -> np.ndarray:
        """Per-vertex weighted DFS steps of one query's plan (gpm_match_root_work)."""
        arr, keep = self._queries([query])
        n = self.info()[0]
        w = np.zeros(n, dtype=np.uint64)
        _check(lib().gpm_match_root_work(self._h, arr, C.c_int(policy), _ptr(w, C.c_uint64)))
        return w

    def match_root_features(self, query, policy: int = SCHED_DEGREE) -> np.ndarray:
        """n x 4 root features of one query's plan (gpm_match_root_features)."""
        arr, keep = self._queries([query])
        n = self.info()[0]
        f = np.zeros(4 * n, dtype=np.float32)
        _check(lib().gpm_match_root_features(self._h, arr, C.c_int(policy), _ptr(f, C.c_float)))
        return f.reshape(n, 4)


def tc_edge_noalloc(d_offsets: int, d_adj: int, n: int, m: int, d_count: int,
                    stream: int = 0) -> None:
    """gpm_tc_edge_noalloc on device addresses (ints): enqueues the count of
    the oriented CSR at d_offsets/d_adj into the device uint64 at d_count on
    `stream` (a hipStream_t as an int) and returns without waiting."""
    _check(lib().gpm_tc_edge_noalloc(C.c_void_p(d_offsets), C.c_void_p(d_adj),
                                     C.c_uint32(n), C.c_uint64(m), C.c_void_p(d_count),
                                     C.c_void_p(stream)))


MATCH_SETV_AUTO, MATCH_SETV_BINARY, MATCH_SETV_MERGE = 0, 1, 2
PASS_BUDGET_OFF = 0xFFFFFFFF  # gpm_tuning.pass_budget: tuned batches in one launch
SETOP_ISECT_COUNT, SETOP_ISECT, SETOP_DIFF_COUNT, SETOP_DIFF, SETOP_FUSE = 0, 1, 2, 3, 4
SETV_BINARY, SETV_TWO_PHASE, SETV_MERGE, SETV_HASH, SETV_BITMAP = 0, 1, 2, 3, 4


def setops(op: int, variant: int, pairs, ub=None):
    """gpm_setops over a list of (A, B) pairs of strictly increasing uint32
    lists.  Returns counts (n, or n x 3 for FUSE) and, for ISECT / DIFF, the
    lists; for FUSE the triples (A \\ B, A & B, B \\ A)."""
    As = [_u32(a) for a, _ in pairs]
    Bs = [_u32(b) for _, b in pairs]
    a_off = np.zeros(len(pairs) + 1, dtype=np.uint64)
    b_off = np.zeros(len(pairs) + 1, dtype=np.uint64)
    a_off[1:] = np.cumsum([len(a) for a in As]) if pairs else []
    b_off[1:] = np.cumsum([len(b) for b in Bs]) if pairs else []
    A = _u32(np.concatenate(As)) if As else _u32([])
    B = _u32(np.concatenate(Bs)) if Bs else _u32([])
    fuse = op == SETOP_FUSE
    counts = np.zeros(len(pairs) * (3 if fuse else 1), dtype=np.uint64)
    oa = np.zeros(max(len(A), 1), dtype=np.uint32)
    oab = np.zeros(max(len(A), 1), dtype=np.uint32)
    ob = np.zeros(max(len(B), 1), dtype=np.uint32)
    ub_arr = _u32(ub) if ub is not None else None
    ubp = _ptr(ub_arr) if ub_arr is not None else None
    _check(lib().gpm_setops(C.c_int(op), C.c_int(variant), C.c_uint32(len(pairs)),
                            _ptr(a_off, C.c_uint64), _ptr(A), _ptr(b_off, C.c_uint64), _ptr(B),
                            ubp, _ptr(counts, C.c_uint64), _ptr(oa), _ptr(oab), _ptr(ob)))
    if fuse:
        c = counts.reshape(-1, 3)
        lists = [(oa[int(a_off[i]):int(a_off[i]) + int(c[i, 0])],
                  oab[int(a_off[i]):int(a_off[i]) + int(c[i, 1])],
                  ob[int(b_off[i]):int(b_off[i]) + int(c[i, 2])]) for i in range(len(pairs))]
        return c, lists
    if op in (SETOP_ISECT, SETOP_DIFF):
        return counts, [oa[int(a_off[i]):int(a_off[i]) + int(counts[i])] for i in range(len(pairs))]
    return counts, None


def clique_tier_caps():
    """(wave_cap, lds_cap, wide_cap) out-degree limits of the clique tiers."""
    a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().gpm_clique_tier_caps(C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


def vertex_features(sym: DeviceGraph, dag: Optional[DeviceGraph]) -> np.ndarray:
    n = sym.info()[0]
    f = np.zeros(4 * n, dtype=np.float32)
    _check(lib().gpm_vertex_features(sym.handle, dag.handle if dag else None,
                                     _ptr(f, C.c_float)))
    return f.reshape(n, 4)


def mlp_predict(feat: np.ndarray, mean, std, w1, b1, w2, b2, w3, b3) -> np.ndarray:
    feat = np.ascontiguousarray(feat, dtype=np.float32)
    arrs = [np.ascontiguousarray(x, dtype=np.float32)
            for x in (mean, std, w1, b1, w2, b2, w3, b3)]
    out = np.zeros(len(feat), dtype=np.float32)
    _check(lib().gpm_mlp_predict(_ptr(feat, C.c_float), C.c_uint32(len(feat)),
                                 *[_ptr(a, C.c_float) for a in arrs],
                                 _ptr(out, C.c_float)))
    return out
