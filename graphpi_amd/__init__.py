"""graphpi_amd — MI355X-native GraphPi counting engine, Python binding.

A thin ctypes layer over the C ABI in include/graphpi_amd.h (the same entry
points a C++ `Graph::pattern_matching` drop-in uses).  It mirrors the
reference's objects: a host CSR (`HostGraph`, DataLoader semantics), the
planner (`build_schedule`, Schedule's constructor), the device-resident graph
(`DeviceGraph`, whose `pattern_matching` is Graph::pattern_matching) and the
motif generator.  The native library must be built (`python -m
graphpi_amd.build`); nothing here falls back to a CPU implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GRAPHPI_AMD_LIB points at another build of the library (same ABI)
LIB_PATH = os.environ.get("GRAPHPI_AMD_LIB") or os.path.join(_HERE, "libgraphpi_amd.so")

GP_MAX_SIZE = 8
GP_MAX_PREFIX = 28
GP_MAX_RESTRICT = 28
GP_MAX_IEP_K = 6
GP_MAX_IEP_GROUPS = 203


class GPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"graphpi_amd error {code}: {msg}")
        self.code = code


class GPTimeout(GPError):
    """GP_ETIME: gp_tuning.time_limit_s ran out between two launches.
    .partial is the CountResult of the task ranges counted before it (its
    count is None; stats.tasks of stats.tasks_total were counted)."""

    def __init__(self, code, msg, partial=None):
        super().__init__(code, msg)
        self.partial = partial


GP_ETIME = -7


class SchedFlat(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int32),
        ("total_prefix_num", ctypes.c_int32),
        ("father", ctypes.c_int32 * GP_MAX_PREFIX),
        ("next", ctypes.c_int32 * GP_MAX_PREFIX),
        ("last", ctypes.c_int32 * GP_MAX_SIZE),
        ("loop_set", ctypes.c_int32 * GP_MAX_SIZE),
        ("total_restrict_num", ctypes.c_int32),
        ("restrict_last", ctypes.c_int32 * GP_MAX_SIZE),
        ("restrict_next", ctypes.c_int32 * GP_MAX_RESTRICT),
        ("restrict_index", ctypes.c_int32 * GP_MAX_RESTRICT),
        ("iep_num", ctypes.c_int32),
        ("iep_group_num", ctypes.c_int32),
        ("iep_block", (ctypes.c_int8 * GP_MAX_SIZE) * GP_MAX_IEP_GROUPS),
        ("iep_val", ctypes.c_int64 * GP_MAX_IEP_GROUPS),
        ("redundancy", ctypes.c_int64),
        ("adj", ctypes.c_int32 * (GP_MAX_SIZE * GP_MAX_SIZE)),
        ("restrict_pair_num", ctypes.c_int32),
        ("restrict_pair", (ctypes.c_int32 * 2) * GP_MAX_RESTRICT),
        ("multiplicity", ctypes.c_int32),
    ]

    def to_dict(self):
        """Same field names as the reference schedule dumps."""
        n = self.size
        P = self.total_prefix_num
        R = self.total_restrict_num
        k = self.iep_num
        groups = []
        for g in range(self.iep_group_num):
            blocks = {}
            for i in range(k):
                blocks.setdefault(self.iep_block[g][i], []).append(i)
            groups.append([blocks[b] for b in sorted(blocks)])
        return {
            "size": n,
            "adj": "".join("1" if self.adj[i] else "0" for i in range(n * n)),
            "restrict_pair": [[self.restrict_pair[r][0], self.restrict_pair[r][1]]
                              for r in range(self.restrict_pair_num)],
            "iep_num": k,
            "iep_groups": groups,
            "iep_vals": [int(self.iep_val[g]) for g in range(self.iep_group_num)],
            "redundancy": int(self.redundancy),
            "total_prefix_num": P,
            "father": [self.father[p] for p in range(P)],
            "next": [self.next[p] for p in range(P)],
            "last": [self.last[d] for d in range(n)],
            "loop_set": [self.loop_set[d] for d in range(n)],
            "total_restrict_num": R,
            "restrict_last": [self.restrict_last[d] for d in range(n)],
            "restrict_next": [self.restrict_next[r] for r in range(R)],
            "restrict_index": [self.restrict_index[r] for r in range(R)],
            "multiplicity": self.multiplicity,
        }


class Tuning(ctypes.Structure):
    _fields_ = [
        ("waves_per_cu", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("lds_stage", ctypes.c_int32),
        ("no_truncate", ctypes.c_int32),
        ("collect_balg", ctypes.c_int32),
        ("no_batch", ctypes.c_int32),
        ("no_pair", ctypes.c_int32),
        ("split_min", ctypes.c_int32),
        ("task_begin", ctypes.c_int64),
        ("task_end", ctypes.c_int64),
        ("pair_map", ctypes.c_int32),
        ("probe_pct", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("no_shared_search", ctypes.c_int32),
        ("no_pair_v1", ctypes.c_int32),
        ("bitmap_ratio", ctypes.c_int32),
        ("slice_ms", ctypes.c_int32),
        ("no_special", ctypes.c_int32),
        ("time_limit_s", ctypes.c_double),
        ("merge_ratio", ctypes.c_int32),
        ("abort_s", ctypes.c_double),
        ("task_abort_s", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("balg_bytes", ctypes.c_uint64),
        ("tasks", ctypes.c_uint64),
        ("waves", ctypes.c_int32),
        ("kernel_variant", ctypes.c_int32),
        ("finish_p50_ms", ctypes.c_double),
        ("finish_p90_ms", ctypes.c_double),
        ("finish_max_ms", ctypes.c_double),
        ("split_tasks", ctypes.c_uint64),
        ("inplace_tasks", ctypes.c_uint64),
        ("launches", ctypes.c_int32),
        ("finished", ctypes.c_int32),
        ("tasks_total", ctypes.c_uint64),
        ("exact_count_valid", ctypes.c_int32),
        ("exact_count_lo", ctypes.c_uint64),
        ("exact_count_hi", ctypes.c_uint64),
    ]

    def exact_count(self):
        """The count without wrap-around (signed 128-bit) when a closed form
        computed it, else None."""
        if not self.exact_count_valid:
            return None
        v = (self.exact_count_hi << 64) | self.exact_count_lo
        return v - (1 << 128) if v >> 127 else v


class PlanInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("leaf", "iep_k", "root_tasks", "batched", "pair", "pair_grouped_v1",
                                               "shared_search", "blocks", "kind1_blocks", "scratch_slots",
                                               "shape")]


class _Csr(ctypes.Structure):
    _fields_ = [
        ("v_cnt", ctypes.c_int32),
        ("e_cnt", ctypes.c_uint32),
        ("max_degree", ctypes.c_int32),
        ("max_intersection_size", ctypes.c_int32),
        ("tri_cnt", ctypes.c_int64),
        ("edge", ctypes.POINTER(ctypes.c_int32)),
        ("vertex", ctypes.POINTER(ctypes.c_uint32)),
    ]


_lib = None

# exported C symbols and their signatures (restype, argtypes)
_SIGS = {
    "gp_graph_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "gp_graph_free": (ctypes.c_int, [ctypes.c_void_p]),
    "gp_graph_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "gp_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SchedFlat), ctypes.POINTER(Tuning),
                                ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(Stats)]),
    "gp_count_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SchedFlat), ctypes.POINTER(Tuning),
                                        ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(Stats)]),
    "gp_count_tasks": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SchedFlat), ctypes.POINTER(ctypes.c_uint64)]),
    "gp_count_footprint": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SchedFlat), ctypes.c_int32,
                                          ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_uint64)]),
    "gp_count_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SchedFlat), ctypes.c_int32,
                                      ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(Stats)]),
    "gp_count_batch_progress": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SchedFlat), ctypes.c_int32,
                                               ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_uint64),
                                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                               ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32),
                                               ctypes.POINTER(Stats)]),
    "gp_tc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(Stats)]),
    "gp_kernel_plan": (ctypes.c_int, [ctypes.POINTER(SchedFlat), ctypes.POINTER(PlanInfo)]),
    "gp_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)]),
    "gp_device_memory": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "gp_last_error": (ctypes.c_char_p, []),
    "gp_schedule_build": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                         ctypes.c_int64, ctypes.POINTER(SchedFlat),
                                         ctypes.POINTER(ctypes.c_int32)]),
    "gp_csr_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.POINTER(ctypes.POINTER(_Csr))]),
    "gp_csr_complete": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.POINTER(_Csr))]),
    "gp_csr_reorder": (ctypes.c_int, [ctypes.POINTER(_Csr), ctypes.c_int32, ctypes.POINTER(ctypes.POINTER(_Csr))]),
    "gp_csr_from_edges": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.POINTER(_Csr))]),
    "gp_csr_generate": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                       ctypes.POINTER(ctypes.POINTER(_Csr))]),
    "gp_csr_write_text": (ctypes.c_int, [ctypes.POINTER(_Csr), ctypes.c_char_p]),
    "gp_csr_write_binary": (ctypes.c_int, [ctypes.POINTER(_Csr), ctypes.c_char_p]),
    "gp_csr_hash": (ctypes.c_int, [ctypes.POINTER(_Csr), ctypes.POINTER(ctypes.c_uint64)]),
    "gp_csr_free": (ctypes.c_int, [ctypes.POINTER(_Csr)]),
    "gp_motifs": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                 ctypes.POINTER(ctypes.c_int32)]),
}


def lib():
    """The loaded native library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m graphpi_amd.build` first")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _check(rc):
    if rc != 0:
        raise GPError(rc, lib().gp_last_error().decode(errors="replace"))


def device_count():
    n = ctypes.c_int32(0)
    _check(lib().gp_device_count(ctypes.byref(n)))
    return n.value


def device_memory():
    """(free, total) bytes of the current HIP device's memory."""
    f, t = ctypes.c_uint64(), ctypes.c_uint64()
    _check(lib().gp_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


# ---------------------------------------------------------------------------
# host graph
# ---------------------------------------------------------------------------

class HostGraph:
    """Host CSR owned by the native library (DataLoader semantics)."""

    def __init__(self, ptr):
        self._p = ptr

    @classmethod
    def load_text(cls, path, oriented_type=0):
        p = ctypes.POINTER(_Csr)()
        _check(lib().gp_csr_load(os.fsencode(path), 0, oriented_type, ctypes.byref(p)))
        return cls(p)

    @classmethod
    def load_binary(cls, path):
        p = ctypes.POINTER(_Csr)()
        _check(lib().gp_csr_load(os.fsencode(path), 1, 0, ctypes.byref(p)))
        return cls(p)

    @classmethod
    def complete(cls, n):
        p = ctypes.POINTER(_Csr)()
        _check(lib().gp_csr_complete(n, ctypes.byref(p)))
        return cls(p)

    @classmethod
    def from_edges(cls, edges):
        e = np.ascontiguousarray(np.asarray(edges, dtype=np.int32).reshape(-1, 2))
        p = ctypes.POINTER(_Csr)()
        _check(lib().gp_csr_from_edges(e.ctypes.data, e.shape[0], ctypes.byref(p)))
        return cls(p)

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1):
        p = ctypes.POINTER(_Csr)()
        _check(lib().gp_csr_generate(0, scale, edge_factor, a, b, c, seed, ctypes.byref(p)))
        return cls(p)

    @classmethod
    def chung_lu(cls, n, edges, alpha=2.1, max_expected_degree=30000.0, seed=3):
        p = ctypes.POINTER(_Csr)()
        _check(lib().gp_csr_generate(1, n, edges, alpha, max_expected_degree, 0.0, seed, ctypes.byref(p)))
        return cls(p)

    def reorder(self, oriented_type=1):
        """A relabelled copy: vertices renumbered by degree as the text
        loader's oriented_type does (1 = high degree first, 2 = low first)."""
        p = ctypes.POINTER(_Csr)()
        _check(lib().gp_csr_reorder(self._p, oriented_type, ctypes.byref(p)))
        return HostGraph(p)

    @property
    def v_cnt(self):
        return self._p.contents.v_cnt

    @property
    def e_cnt(self):
        return self._p.contents.e_cnt

    @property
    def max_degree(self):
        return self._p.contents.max_degree

    @property
    def max_intersection_size(self):
        return self._p.contents.max_intersection_size

    @property
    def vertex(self):
        return np.ctypeslib.as_array(self._p.contents.vertex, shape=(self.v_cnt + 1,))

    @property
    def edge(self):
        if self.e_cnt == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(self._p.contents.edge, shape=(self.e_cnt,))

    def hash(self):
        h = ctypes.c_uint64()
        _check(lib().gp_csr_hash(self._p, ctypes.byref(h)))
        return f"{h.value:016x}"

    def write_text(self, path):
        _check(lib().gp_csr_write_text(self._p, os.fsencode(path)))

    def write_binary(self, path):
        _check(lib().gp_csr_write_binary(self._p, os.fsencode(path)))

    def free(self):
        if self._p:
            lib().gp_csr_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# planner
# ---------------------------------------------------------------------------

def adj_from_string(size, s):
    """'0110...' row-major string (the reference's Pattern(size, char*) form)."""
    a = np.array([1 if ch == "1" else 0 for ch in s], dtype=np.int32).reshape(size, size)
    return np.maximum(a, a.T)


def adj_from_edges(size, edges):
    a = np.zeros((size, size), dtype=np.int32)
    for x, y in edges:
        a[x, y] = a[y, x] = 1
    return a


# preset patterns: the reference's PatternType (src/pattern.cpp:45-133) plus
# the ones the benchmark configurations name
PATTERNS = {
    "triangle": adj_from_edges(3, [(0, 1), (0, 2), (1, 2)]),
    "rectangle": adj_from_edges(4, [(0, 1), (0, 2), (1, 3), (2, 3)]),
    "qg3": adj_from_edges(4, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3)]),
    "pentagon": adj_from_edges(5, [(0, 1), (0, 2), (1, 3), (2, 4), (3, 4)]),
    "house": adj_from_edges(5, [(0, 1), (0, 2), (0, 3), (1, 3), (1, 4), (2, 4)]),
    "hourglass": adj_from_edges(6, [(0, 1), (0, 2), (0, 4), (1, 2), (1, 5), (2, 3), (3, 4), (3, 5), (4, 5)]),
    "cycle_6_tri": adj_from_edges(6, [(0, 1), (0, 2), (1, 2), (1, 3), (1, 4), (2, 3), (2, 5), (0, 4), (0, 5)]),
    "clique_7_minus": adj_from_edges(7, [(j, i) for i in range(7) for j in range(i) if not (i == 6 and j == 5)]),
    "4clique": adj_from_edges(4, [(i, j) for i in range(4) for j in range(i)]),
    "5clique": adj_from_edges(5, [(i, j) for i in range(5) for j in range(i)]),
    # two triangles joined by an edge (6 vertices, 7 edges)
    "tri_edge_tri": adj_from_edges(6, [(0, 1), (0, 2), (1, 2), (2, 3), (3, 4), (3, 5), (4, 5)]),
    "wedge": adj_from_edges(3, [(0, 1), (0, 2)]),
    "edge": adj_from_edges(2, [(0, 1)]),
}
PATTERNS["5cycle"] = PATTERNS["pentagon"]
PATTERNS["6cycle"] = adj_from_edges(6, [(i, (i + 1) % 6) for i in range(6)])
PATTERNS["4path"] = adj_from_edges(4, [(0, 1), (1, 2), (2, 3)])
PATTERNS["5path"] = adj_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)])


def build_schedule(adj, performance_modeling_type=1, restricts_type=1, use_in_exclusion_optimize=True,
                   v_cnt=0, e_cnt=0, tri_cnt=0):
    """Schedule(pattern, valid, pm, rt, iep, v_cnt, e_cnt, tri_cnt) flattened.
    Returns (SchedFlat, is_pattern_valid)."""
    a = np.ascontiguousarray(np.asarray(adj, dtype=np.int32))
    n = a.shape[0]
    out = SchedFlat()
    valid = ctypes.c_int32(0)
    _check(lib().gp_schedule_build(n, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   performance_modeling_type, restricts_type,
                                   1 if use_in_exclusion_optimize else 0, v_cnt, e_cnt, tri_cnt,
                                   ctypes.byref(out), ctypes.byref(valid)))
    return out, bool(valid.value)


def kernel_plan(sched):
    """gp_kernel_plan: which kernel paths the schedule takes, as a dict."""
    info = PlanInfo()
    _check(lib().gp_kernel_plan(ctypes.byref(sched), ctypes.byref(info)))
    return {n: getattr(info, n) for n, _ in PlanInfo._fields_}


def motifs(n):
    """All connected n-vertex patterns in MotifGenerator order."""
    cnt = ctypes.c_int32(0)
    _check(lib().gp_motifs(n, None, 0, ctypes.byref(cnt)))
    buf = np.zeros(cnt.value * n * n, dtype=np.int32)
    _check(lib().gp_motifs(n, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cnt.value, ctypes.byref(cnt)))
    return [buf[i * n * n:(i + 1) * n * n].reshape(n, n).copy() for i in range(cnt.value)]


# ---------------------------------------------------------------------------
# device graph
# ---------------------------------------------------------------------------

class CountResult:
    __slots__ = ("count", "raw", "stats")

    def __init__(self, count, raw, stats):
        self.count, self.raw, self.stats = count, raw, stats


class DeviceGraph:
    """CSR resident in HBM; counting runs on the GPU through gp_count."""

    def __init__(self, host):
        self._h = ctypes.c_void_p()
        self.v_cnt = host.v_cnt
        self.e_cnt = host.e_cnt
        _check(lib().gp_graph_upload(host._p.contents.edge, host._p.contents.vertex, host.v_cnt, host.e_cnt,
                                     ctypes.byref(self._h)))

    def count(self, sched, tuning=None):
        """gp_count.  Raises GPTimeout (with the partial sum) when the
        tuning's time_limit_s runs out."""
        raw = ctypes.c_uint64()
        cnt = ctypes.c_int64()
        st = Stats()
        rc = lib().gp_count(self._h, ctypes.byref(sched), ctypes.byref(tuning) if tuning else None,
                            ctypes.byref(raw), ctypes.byref(cnt), ctypes.byref(st))
        if rc == GP_ETIME:
            raise GPTimeout(rc, lib().gp_last_error().decode(errors="replace"), CountResult(None, raw.value, st))
        _check(rc)
        return CountResult(cnt.value, raw.value, st)

    def count_noalloc(self, sched, tuning=None):
        """gp_count_noalloc: the same count with no device allocation (static
        scratch, one launch)."""
        raw = ctypes.c_uint64()
        cnt = ctypes.c_int64()
        st = Stats()
        _check(lib().gp_count_noalloc(self._h, ctypes.byref(sched), ctypes.byref(tuning) if tuning else None,
                                      ctypes.byref(raw), ctypes.byref(cnt), ctypes.byref(st)))
        return CountResult(cnt.value, raw.value, st)

    def count_batch(self, scheds, tuning=None):
        """gp_count_batch: several schedules in one persistent launch.
        Returns ([CountResult per schedule], Stats of the whole batch)."""
        n = len(scheds)
        arr = (SchedFlat * n)(*scheds)
        raws = (ctypes.c_uint64 * n)()
        cnts = (ctypes.c_int64 * n)()
        st = Stats()
        _check(lib().gp_count_batch(self._h, arr, n, ctypes.byref(tuning) if tuning else None, raws, cnts,
                                    ctypes.byref(st)))
        return [CountResult(cnts[i], raws[i], None) for i in range(n)], st

    def count_batch_progress(self, scheds, begin=None, tuning=None, abort_s=None):
        """gp_count_batch_progress: count each schedule's tasks from begin[i]
        (default 0) in bounded launches until done, tuning.time_limit_s runs
        out or a schedule's abort time (abort_s[i], seconds after the call
        starts) passes.  Returns (done, raws, stats, finished, aborted):
        done[i] is the first task not counted, raws[i] the u64 raw sum of
        tasks [begin[i], done[i]), aborted[i] whether schedule i's abort time
        stopped it; finished is True when every count is complete."""
        n = len(scheds)
        arr = (SchedFlat * n)(*scheds)
        b = (ctypes.c_uint64 * n)(*(begin or [0] * n))
        ab_s = None if abort_s is None else (ctypes.c_double * n)(*[x if x is not None else 0.0 for x in abort_s])
        done = (ctypes.c_uint64 * n)()
        raws = (ctypes.c_uint64 * n)()
        ab = (ctypes.c_int32 * n)()
        st = Stats()
        rc = lib().gp_count_batch_progress(self._h, arr, n, ctypes.byref(tuning) if tuning else None, b, ab_s, done,
                                           raws, ab, ctypes.byref(st))
        if rc not in (0, GP_ETIME):
            _check(rc)
        return list(done), list(raws), st, rc == 0, [bool(x) for x in ab]

    def footprint(self, scheds, tuning=None):
        """Device bytes (besides the graph) a count_batch of scheds would use."""
        n = len(scheds)
        arr = (SchedFlat * n)(*scheds)
        b = ctypes.c_uint64()
        _check(lib().gp_count_footprint(self._h, arr, n, ctypes.byref(tuning) if tuning else None, ctypes.byref(b)))
        return b.value

    def count_tasks(self, sched):
        t = ctypes.c_uint64()
        _check(lib().gp_count_tasks(self._h, ctypes.byref(sched), ctypes.byref(t)))
        return t.value

    def count_sliced(self, sched, slices, tuning=None, progress=None):
        """The same count as count(), run as `slices` launches over disjoint
        task ranges (raw sums add up); progress(i, result) after each."""
        total = self.count_tasks(sched)
        raw = 0
        ms = 0.0
        balg = 0
        t = Tuning()
        if tuning is not None:
            ctypes.memmove(ctypes.byref(t), ctypes.byref(tuning), ctypes.sizeof(Tuning))
        slices = max(1, min(slices, total if total else 1))
        for i in range(slices):
            t.task_begin = total * i // slices
            t.task_end = total * (i + 1) // slices
            r = self.count(sched, t)
            raw = (raw + r.raw) % (1 << 64)
            ms += r.stats.kernel_ms
            balg += r.stats.balg_bytes
            if progress:
                progress(i, r)
        signed = raw - (1 << 64) if raw >= (1 << 63) else raw
        red = sched.redundancy or 1
        st = Stats()
        st.kernel_ms = ms
        st.balg_bytes = balg
        q = abs(signed) // abs(red)
        count = q if (signed >= 0) == (red > 0) else -q
        return CountResult(count, raw, st)

    def pattern_matching(self, sched):
        """Graph::pattern_matching: the count divided by the IEP redundancy."""
        return self.count(sched).count

    def triangle_counting(self):
        t = ctypes.c_uint64()
        st = Stats()
        _check(lib().gp_tc(self._h, ctypes.byref(t), ctypes.byref(st)))
        return t.value

    def nbytes(self):
        b = ctypes.c_uint64()
        _check(lib().gp_graph_bytes(self._h, ctypes.byref(b)))
        return b.value

    def free(self):
        if self._h:
            lib().gp_graph_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
