"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check or time beside the HIP engine.  The
engine never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_CECI = os.path.join(_HERE, "_ref", "myceci")

_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        for name in ("orc_tc", "orc_clique", "orc_clique_list", "orc_plan_bytes",
                     "orc_pattern", "orc_ceci"):
            getattr(L, name).restype = C.c_uint64
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(t))


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def set_threads(t: int) -> None:
    lib().orc_set_threads(C.c_int(t))


def num_threads() -> int:
    return lib().orc_num_threads()


def tc(off, adj) -> int:
    off, adj = _u32(off), _u32(adj)
    return lib().orc_tc(C.c_uint32(len(off) - 1), _p(off), _p(adj))


def clique(off, adj, k: int, root_begin: int = 0, root_end: int = 0) -> int:
    off, adj = _u32(off), _u32(adj)
    return lib().orc_clique(C.c_uint32(len(off) - 1), _p(off), _p(adj),
                            C.c_int(k), C.c_uint32(root_begin),
                            C.c_uint32(root_end))


def clique_list(off, adj, k: int, roots) -> int:
    off, adj, roots = _u32(off), _u32(adj), _u32(roots)
    return lib().orc_clique_list(C.c_uint32(len(off) - 1), _p(off), _p(adj),
                                 C.c_int(k), _p(roots), C.c_uint32(len(roots)))


def plan_bytes(off, adj, k: int) -> int:
    off, adj = _u32(off), _u32(adj)
    return lib().orc_plan_bytes(C.c_uint32(len(off) - 1), _p(off), _p(adj),
                                C.c_int(k))


def motif4(off, adj):
    off, adj = _u32(off), _u32(adj)
    out = np.zeros(6, dtype=np.uint64)
    lib().orc_motif4(C.c_uint32(len(off) - 1), _p(off), _p(adj),
                     _p(out, C.c_uint64))
    return [int(x) for x in out]


def pattern(off, adj, pid: int) -> int:
    off, adj = _u32(off), _u32(adj)
    return lib().orc_pattern(C.c_uint32(len(off) - 1), _p(off), _p(adj),
                             C.c_int(pid))


def ceci(off, adj, labels, qn: int, qedges, qlabels) -> int:
    off, adj, labels = _u32(off), _u32(adj), _u32(labels)
    qe = _u32(np.asarray(qedges, dtype=np.uint32).reshape(-1))
    ql = _u32(qlabels)
    return lib().orc_ceci(C.c_uint32(len(off) - 1), _p(off), _p(adj),
                          _p(labels), C.c_uint32(qn), C.c_uint32(len(qe) // 2),
                          _p(qe), _p(ql))


def ceci_batch(off, adj, labels, queries) -> np.ndarray:
    """queries: list of (n, edges, labels); parallel across queries."""
    off, adj, labels = _u32(off), _u32(adj), _u32(labels)
    qn = _u32([q[0] for q in queries])
    es = [_u32(np.asarray(q[1], dtype=np.uint32).reshape(-1)) for q in queries]
    ls = [_u32(q[2]) for q in queries]
    qm = _u32([len(e) // 2 for e in es])
    eoff = np.zeros(len(queries), dtype=np.uint64)
    loff = np.zeros(len(queries), dtype=np.uint64)
    if queries:
        eoff[1:] = np.cumsum([len(e) for e in es])[:-1]
        loff[1:] = np.cumsum([len(l) for l in ls])[:-1]
    E = _u32(np.concatenate(es)) if es else _u32([])
    Lb = _u32(np.concatenate(ls)) if ls else _u32([])
    out = np.zeros(len(queries), dtype=np.uint64)
    lib().orc_ceci_batch(C.c_uint32(len(off) - 1), _p(off), _p(adj), _p(labels),
                         C.c_int(len(queries)), _p(qn), _p(qm),
                         _p(eoff, C.c_uint64), _p(E), _p(loff, C.c_uint64),
                         _p(Lb), _p(out, C.c_uint64))
    return out


def ref_ceci_available() -> bool:
    return os.path.exists(REF_CECI)


def ref_ceci(data_path: str, query_path: str, timeout: float = 120.0) -> int:
    """Runs the reference CECI matcher built from source (oracle/_ref)."""
    p = subprocess.run([REF_CECI, "-d", data_path, "-q", query_path,
                        "-filter", "CECI", "-order", "CECI", "-engine", "CECI"],
                       stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                       timeout=timeout, check=True)
    for line in p.stdout.decode(errors="replace").splitlines():
        if "#Embeddings" in line:
            return int(line.split(":")[1].strip())
    raise RuntimeError("no #Embeddings line in reference output")
