"""ctypes binding of the C ABI in include/csm.h (libcsm.so built in-tree under lib/).

This is how Python callers (tests, the benchmark, smoke checks) drive the HIP library; the product
path is the C ABI itself.  Loading fails loudly when the shared library has not been built.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcsm.so")
BIN_DIR = os.path.join(HERE, "bin")

# Every symbol include/csm.h declares, in declaration order.
EXPORTED = [
    "csm_session_create", "csm_session_destroy", "csm_last_error", "csm_load_graph_file",
    "csm_load_query_file", "csm_open_update_file", "csm_build_index", "csm_plan",
    "csm_get_plan_text", "csm_get_summary", "csm_apply_batch", "csm_apply_range", "csm_apply_range_static", "csm_apply_edges", "csm_serve_start", "csm_serve_submit", "csm_serve_wait", "csm_serve_stop", "csm_arena_usage",
    "csm_get_relation_edge_counts",
    "csm_get_counts", "csm_pool_usage", "csm_read_candidates", "csm_read_adjacency",
    "csm_synchronize", "csm_snapshot", "csm_restore", "csm_get_tunables", "csm_set_tunables", "csm_tune",
    "csm_profile_enable", "csm_profile_read", "csm_calibration_copy",
]

STATUS = {0: "OK", 1: "ARGUMENT", 2: "IO", 3: "FORMAT", 4: "LIMIT", 5: "DEVICE", 6: "MEMORY", 7: "STATE",
          8: "PRECONDITION", 9: "MISMATCH"}


class Options(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("directed", ctypes.c_int32), ("pool_bytes", ctypes.c_uint64),
                ("ring_bytes", ctypes.c_uint64), ("cp_list_bytes", ctypes.c_uint64),
                ("max_batch_edges", ctypes.c_uint32), ("semi_bfs_threshold", ctypes.c_uint32), ("lanes", ctypes.c_uint32)]


class Technique(ctypes.Structure):
    _fields_ = [("matching_order", ctypes.c_int32), ("search_strategy", ctypes.c_int32),
                ("local_index", ctypes.c_int32), ("cartesian_product", ctypes.c_int32),
                ("heavy_vertex_avoidance", ctypes.c_int32), ("relation_switch", ctypes.c_int32)]


class QueryResult(ctypes.Structure):
    _fields_ = [("positive", ctypes.c_uint64), ("negative", ctypes.c_uint64),
                ("positive_hash", ctypes.c_uint64), ("negative_hash", ctypes.c_uint64)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("edges", ctypes.c_uint64), ("inserts", ctypes.c_uint64), ("deletes", ctypes.c_uint64),
                ("ms_total", ctypes.c_double), ("ms_index", ctypes.c_double), ("ms_match", ctypes.c_double)]


TUNABLES = ("semi_bfs_threshold", "waves_per_anchor", "stack_depth", "dfs_blocks", "local_blocks", "team", "level_blocks")


class Tunables(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint32) for name in TUNABLES]

    def as_dict(self):
        return {name: getattr(self, name) for name in TUNABLES}


class CsmError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"csm status {STATUS.get(status, status)}: {message}")
        self.status = status


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        lib = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
        P = ctypes.POINTER
        lib.csm_session_create.argtypes = [P(Options), P(vp)]
        lib.csm_session_destroy.argtypes = [vp]
        lib.csm_session_destroy.restype = None
        lib.csm_last_error.argtypes = [vp]
        lib.csm_last_error.restype = ctypes.c_char_p
        lib.csm_load_graph_file.argtypes = [vp, ctypes.c_char_p]
        lib.csm_load_query_file.argtypes = [vp, ctypes.c_char_p, P(u32)]
        lib.csm_open_update_file.argtypes = [vp, ctypes.c_char_p, u32, P(u64)]
        lib.csm_build_index.argtypes = [vp]
        lib.csm_plan.argtypes = [vp, u32, P(Technique)]
        lib.csm_get_plan_text.argtypes = [vp, u32, ctypes.c_char_p, ctypes.c_size_t, P(ctypes.c_size_t)]
        lib.csm_get_summary.argtypes = [vp, u32, P(u32), P(ctypes.c_float), u32, P(u32)]
        lib.csm_apply_batch.argtypes = [vp, u64, P(QueryResult), u32, P(BatchInfo)]
        lib.csm_apply_range.argtypes = [vp, u64, u64, P(QueryResult), u32, P(BatchInfo)]
        lib.csm_apply_range_static.argtypes = [vp, u64, u64, P(QueryResult), u32, P(BatchInfo)]
        lib.csm_apply_edges.argtypes = [vp, P(u32), P(u32), P(u32), P(ctypes.c_uint8), u64, P(QueryResult), u32, P(BatchInfo)]
        lib.csm_serve_start.argtypes = [vp]
        lib.csm_serve_submit.argtypes = [vp, P(u32), P(u32), P(u32), P(ctypes.c_uint8), u64, P(u64)]
        lib.csm_serve_wait.argtypes = [vp, u64, P(QueryResult), u32, P(BatchInfo), P(ctypes.c_double)]
        lib.csm_serve_stop.argtypes = [vp]
        lib.csm_arena_usage.argtypes = [vp, P(u64), P(u64)]
        lib.csm_get_relation_edge_counts.argtypes = [vp, u32, u64, P(u64), u32, P(u32)]
        lib.csm_get_counts.argtypes = [vp, P(u32), P(u32)]
        lib.csm_pool_usage.argtypes = [vp, P(u64), P(u64)]
        lib.csm_read_candidates.argtypes = [vp, u32, u32, P(u32), u32]
        lib.csm_read_adjacency.argtypes = [vp, u32, u32, P(u32), u32, P(u32)]
        lib.csm_synchronize.argtypes = [vp]
        lib.csm_snapshot.argtypes = [vp]
        lib.csm_restore.argtypes = [vp]
        lib.csm_get_tunables.argtypes = [vp, P(Tunables)]
        lib.csm_set_tunables.argtypes = [vp, P(Tunables)]
        lib.csm_tune.argtypes = [vp, u64, u64, u32, u32, P(Tunables), u32, P(Tunables), P(ctypes.c_double), u32, P(u32), P(u32)]
        lib.csm_profile_enable.argtypes = [vp, u32]
        lib.csm_profile_read.argtypes = [vp, u32, P(u64), P(ctypes.c_double), P(u64)]
        lib.csm_calibration_copy.argtypes = [vp, u64, P(ctypes.c_double)]
        for name in EXPORTED:
            f = getattr(lib, name)
            if f.restype is ctypes.c_int:
                f.restype = i32
        _lib = lib
    return _lib


def technique(matching_order=2, search_strategy=2, local_index=True, cartesian_product=True,
              heavy_vertex_avoidance=True, relation_switch=True):
    return Technique(matching_order, search_strategy, int(local_index), int(cartesian_product),
                     int(heavy_vertex_avoidance), int(relation_switch))


class Session:
    """One device session: graph, standing queries, update stream."""

    def __init__(self, device=0, directed=False, pool_bytes=0, ring_bytes=0, cp_list_bytes=0, max_batch_edges=0,
                 semi_bfs_threshold=0, lanes=0):
        self.lib = load()
        self.h = ctypes.c_void_p()
        o = Options(device, int(directed), pool_bytes, ring_bytes, cp_list_bytes, max_batch_edges, semi_bfs_threshold, lanes)
        st = self.lib.csm_session_create(ctypes.byref(o), ctypes.byref(self.h))
        if st != 0:
            msg = self.lib.csm_last_error(self.h).decode() if self.h else "session create failed"
            if self.h:
                self.lib.csm_session_destroy(self.h)
                self.h = None
            raise CsmError(st, msg)
        self.queries = []
        self.num_batches = 0

    def _check(self, st):
        if st != 0:
            raise CsmError(st, self.lib.csm_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.csm_session_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load_graph(self, path):
        self._check(self.lib.csm_load_graph_file(self.h, str(path).encode()))

    def add_query(self, path):
        qid = ctypes.c_uint32()
        self._check(self.lib.csm_load_query_file(self.h, str(path).encode(), ctypes.byref(qid)))
        self.queries.append(qid.value)
        return qid.value

    def open_updates(self, path, batch_size):
        nb = ctypes.c_uint64()
        self._check(self.lib.csm_open_update_file(self.h, str(path).encode(), batch_size, ctypes.byref(nb)))
        self.num_batches = nb.value
        return nb.value

    def build(self):
        self._check(self.lib.csm_build_index(self.h))

    def plan(self, qid, tech=None):
        t = tech or technique()
        self._check(self.lib.csm_plan(self.h, qid, ctypes.byref(t)))

    def plan_text(self, qid):
        need = ctypes.c_size_t()
        self._check(self.lib.csm_get_plan_text(self.h, qid, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        self._check(self.lib.csm_get_plan_text(self.h, qid, buf, need.value, ctypes.byref(need)))
        return buf.value.decode()

    def summary(self, qid):
        card = (ctypes.c_uint32 * 84)()
        avg = (ctypes.c_float * 84)()
        n = ctypes.c_uint32()
        self._check(self.lib.csm_get_summary(self.h, qid, card, avg, 84, ctypes.byref(n)))
        return list(card[:n.value]), list(avg[:n.value])

    def apply_batch(self, b):
        nq = len(self.queries)
        res = (QueryResult * max(1, nq))()
        info = BatchInfo()
        self._check(self.lib.csm_apply_batch(self.h, b, res, nq, ctypes.byref(info)))
        return [(r.positive, r.negative, r.positive_hash, r.negative_hash) for r in res[:nq]], info

    def apply_range(self, first, n):
        nq = len(self.queries)
        res = (QueryResult * max(1, nq))()
        info = BatchInfo()
        self._check(self.lib.csm_apply_range(self.h, first, n, res, nq, ctypes.byref(info)))
        return [(r.positive, r.negative, r.positive_hash, r.negative_hash) for r in res[:nq]], info

    def apply_range_static(self, first, n):
        """csm_apply_range with the partial-result rings in the library's static device ring."""
        nq = len(self.queries)
        res = (QueryResult * max(1, nq))()
        info = BatchInfo()
        self._check(self.lib.csm_apply_range_static(self.h, first, n, res, nq, ctypes.byref(info)))
        return [(r.positive, r.negative, r.positive_hash, r.negative_hash) for r in res[:nq]], info

    def apply_edges(self, src, dst, elabel, op):
        """One in-memory batch.  Lists, or contiguous numpy arrays (uint32 src/dst/elabel, uint8 op),
        which are passed without copying."""
        n = len(src)
        nq = len(self.queries)
        res = (QueryResult * max(1, nq))()
        info = BatchInfo()
        if hasattr(src, "ctypes"):
            P32, P8 = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint8)
            args = (src.ctypes.data_as(P32), dst.ctypes.data_as(P32), elabel.ctypes.data_as(P32), op.ctypes.data_as(P8))
        else:
            A = ctypes.c_uint32 * max(1, n)
            args = (A(*src), A(*dst), A(*elabel), (ctypes.c_uint8 * max(1, n))(*op))
        st = self.lib.csm_apply_edges(self.h, *args, n, res, nq, ctypes.byref(info))
        self._check(st)
        return [(r.positive, r.negative, r.positive_hash, r.negative_hash) for r in res[:nq]], info

    @staticmethod
    def _edge_args(src, dst, elabel, op):
        n = len(src)
        if hasattr(src, "ctypes"):
            P32, P8 = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint8)
            return (src.ctypes.data_as(P32), dst.ctypes.data_as(P32), elabel.ctypes.data_as(P32), op.ctypes.data_as(P8)), n
        A = ctypes.c_uint32 * max(1, n)
        return (A(*src), A(*dst), A(*elabel), (ctypes.c_uint8 * max(1, n))(*op)), n

    def serve_start(self):
        """Starts the request loop (csm_serve_start)."""
        self._check(self.lib.csm_serve_start(self.h))

    def serve_submit(self, src, dst, elabel, op):
        """Submits one batch; returns its ticket."""
        args, n = self._edge_args(src, dst, elabel, op)
        t = ctypes.c_uint64()
        self._check(self.lib.csm_serve_submit(self.h, *args, n, ctypes.byref(t)))
        return t.value

    def serve_wait(self, ticket):
        """Waits for a request: (results, info, latency_ms).  Raises CsmError if it failed."""
        nq = len(self.queries)
        res = (QueryResult * max(1, nq))()
        info = BatchInfo()
        ms = ctypes.c_double()
        self._check(self.lib.csm_serve_wait(self.h, ticket, res, nq, ctypes.byref(info), ctypes.byref(ms)))
        return [(r.positive, r.negative, r.positive_hash, r.negative_hash) for r in res[:nq]], info, ms.value

    def serve_stop(self):
        self._check(self.lib.csm_serve_stop(self.h))

    def arena_usage(self):
        u, c = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.lib.csm_arena_usage(self.h, ctypes.byref(u), ctypes.byref(c)))
        return u.value, c.value

    def counts(self):
        nv, nq = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self.lib.csm_get_counts(self.h, ctypes.byref(nv), ctypes.byref(nq)))
        return nv.value, nq.value

    def candidates(self, qid, u):
        nv, _ = self.counts()
        words = (nv + 31) // 32
        buf = (ctypes.c_uint32 * words)()
        self._check(self.lib.csm_read_candidates(self.h, qid, u, buf, words))
        return list(buf)

    def adjacency(self, v, direction=0):
        deg = ctypes.c_uint32()
        self._check(self.lib.csm_read_adjacency(self.h, v, direction, None, 0, ctypes.byref(deg)))
        buf = (ctypes.c_uint32 * max(1, deg.value))()
        self._check(self.lib.csm_read_adjacency(self.h, v, direction, buf, deg.value, ctypes.byref(deg)))
        return list(buf[:deg.value])

    def synchronize(self):
        self._check(self.lib.csm_synchronize(self.h))

    def snapshot(self):
        """Device copy of the graph state between batches (csm_snapshot)."""
        self._check(self.lib.csm_snapshot(self.h))

    def restore(self):
        """Back to the last snapshot's graph state (csm_restore)."""
        self._check(self.lib.csm_restore(self.h))

    def tunables(self):
        """The session's launch settings (csm_get_tunables) as a dict."""
        t = Tunables()
        self._check(self.lib.csm_get_tunables(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tunables(self, **kw):
        """Change launch settings between batches (csm_set_tunables); omitted ones are kept."""
        self._check(self.lib.csm_set_tunables(self.h, ctypes.byref(Tunables(**kw))))

    def tune(self, first, batch_edges, warmup=1, batches=2, candidates=None, cap=256):
        """The library's autotuner (csm_tune) over stream edges [first, first + (warmup + batches) *
        batch_edges).  candidates: dicts of settings (omitted ones keep the session's), or None for
        the coordinate search.  Returns (evaluations as [(settings dict, ms per batch)], index of the
        fastest); the session keeps the fastest settings."""
        cands = [Tunables(**c) for c in candidates or []]
        arr = (Tunables * max(1, len(cands)))(*cands)
        ev, ms = (Tunables * cap)(), (ctypes.c_double * cap)()
        n, best = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self.lib.csm_tune(self.h, first, batch_edges, warmup, batches, arr, len(cands), ev, ms, cap,
                                      ctypes.byref(n), ctypes.byref(best)))
        return [(ev[k].as_dict(), ms[k]) for k in range(min(n.value, cap))], best.value

    def profile_enable(self, mask):
        self._check(self.lib.csm_profile_enable(self.h, mask))

    def profile_read(self, family):
        n, ms, b = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_uint64()
        self._check(self.lib.csm_profile_read(self.h, family, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)))
        return n.value, ms.value, b.value

    def calibration_copy(self, words):
        ms = ctypes.c_double()
        self._check(self.lib.csm_calibration_copy(self.h, words, ctypes.byref(ms)))
        return ms.value

    def pool_usage(self):
        u, c = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.lib.csm_pool_usage(self.h, ctypes.byref(u), ctypes.byref(c)))
        return u.value, c.value
