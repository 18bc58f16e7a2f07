"""ctypes binding of the oracle (oracle/liboracle.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it;
the product package never imports this module.
"""
import ctypes
import os

import numpy as np

import graphpi_amd as gp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

_o = None


def olib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        l = ctypes.CDLL(ORACLE_LIB)
        l.gpo_pattern_matching.restype = ctypes.c_int64
        l.gpo_pattern_matching.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                           ctypes.POINTER(gp.SchedFlat), ctypes.c_int, ctypes.c_int64,
                                           ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
        l.gpo_pattern_matching_v1.restype = ctypes.c_int64
        l.gpo_pattern_matching_v1.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                              ctypes.POINTER(gp.SchedFlat), ctypes.c_int, ctypes.c_int64,
                                              ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                              ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.POINTER(ctypes.c_uint64)]
        l.gpo_pattern_matching_timed.restype = ctypes.c_int64
        l.gpo_pattern_matching_timed.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                 ctypes.POINTER(gp.SchedFlat), ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64),
                                                 ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
        l.gpo_triangle_counting.restype = ctypes.c_int64
        l.gpo_triangle_counting.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int]
        l.gpo_brute_force.restype = ctypes.c_int64
        l.gpo_brute_force.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_void_p]
        _o = l
    return _o


def _csr(g):
    v = np.ascontiguousarray(g.vertex, dtype=np.uint32)
    e = np.ascontiguousarray(g.edge, dtype=np.int32)
    if e.size == 0:
        e = np.zeros(1, dtype=np.int32)
    return v, e


def pattern_matching(g, sched, threads=8, root_begin=0, root_end=-1, root_stride=1, v1_begin=0, v1_end=-1):
    """Returns dict(count, raw, balg, expanded).  v1_begin/v1_end keep only
    embeddings whose depth-1 vertex is in [v1_begin, v1_end) (test slicing)."""
    v, e = _csr(g)
    raw = ctypes.c_uint64()
    balg = ctypes.c_uint64()
    exp = ctypes.c_uint64()
    c = olib().gpo_pattern_matching_v1(e.ctypes.data, v.ctypes.data, g.v_cnt, ctypes.byref(sched), threads,
                                       root_begin, root_end, root_stride, v1_begin, v1_end, ctypes.byref(raw),
                                       ctypes.byref(balg), ctypes.byref(exp))
    return {"count": c, "raw": raw.value, "balg": balg.value, "expanded": exp.value}


def triangle_counting(g, threads=8):
    v, e = _csr(g)
    return olib().gpo_triangle_counting(e.ctypes.data, v.ctypes.data, g.v_cnt, threads)


def brute_force(g, adj):
    v, e = _csr(g)
    a = np.ascontiguousarray(np.asarray(adj, dtype=np.int32))
    return olib().gpo_brute_force(e.ctypes.data, v.ctypes.data, g.v_cnt, a.shape[0], a.ctypes.data)


def pattern_matching_timed(g, sched, roots, time_limit_s, threads=8):
    """The loop nest over the root ids in `roots` until time_limit_s passes
    (roots still running then are abandoned).  Returns dict(raw, roots_done,
    busy_s) over the roots that finished."""
    v, e = _csr(g)
    r = np.ascontiguousarray(np.asarray(roots, dtype=np.int32))
    raw = ctypes.c_uint64()
    done = ctypes.c_int64()
    busy = ctypes.c_double()
    olib().gpo_pattern_matching_timed(e.ctypes.data, v.ctypes.data, g.v_cnt, ctypes.byref(sched), threads,
                                      r.ctypes.data, r.size, time_limit_s, ctypes.byref(raw), ctypes.byref(done),
                                      ctypes.byref(busy))
    return {"raw": raw.value, "roots_done": done.value, "busy_s": busy.value}
