"""Python mirror of efanna::FIndex<float> (efanna.hpp:9-75) over libefg.so.

Names, argument meaning and order follow the reference:
``FIndex(features, distance, params)``, ``buildIndex()``, ``buildTrees()``,
``knnSearch(k, query)``, ``setSearchParams(epochs, init_num, extend_to,
search_trees=0, search_lv=-1, search_method=0)``, ``saveGraph/loadGraph``,
``saveTrees/loadTrees``, ``saveIndex/loadIndex``, ``saveResults``,
``getGraphSize()``, ``getGraphRow(row)``.  Errors raise EfgError instead of
printing and exiting.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BuildParams, EfgError, check, lib

KDTREE_UB = 0
HASHING = 1


class Distance:
    """Marker for the reference's Distance<float> classes (general/distance.hpp)."""


class L2Distance(Distance):
    pass


class L2DistanceSSE(Distance):
    pass


class L2DistanceAVX(Distance):
    pass


class L2DistanceAVXr4(Distance):
    pass


_L2_CLASSES = (L2Distance, L2DistanceSSE, L2DistanceAVX, L2DistanceAVXr4)


class IndexParams:
    """IndexParams (general/params.hpp:19-33)."""

    def __init__(self):
        self.init_index_type = KDTREE_UB
        self.K = 10
        self.S = 10
        self.L = 30
        self.TNS = 10
        self.Check_K = 40
        self.build_epoches = 4
        self.extra_params = {}
        self.reverse_nn_used = True
        self.seed = 1998


class KDTreeUbIndexParams(IndexParams):
    """KDTreeUbIndexParams (algorithm/kdtreeub_index.hpp:14-35)."""

    def __init__(self, rnn_used, tree_num_total, merge_level=4, epoches=4, check=25, myL=30, building_use_k=10,
                 tree_num_build=0, myS=10):
        super().__init__()
        self.reverse_nn_used = rnn_used
        self.init_index_type = KDTREE_UB
        self.K = building_use_k
        self.build_epoches = epoches
        self.S = myS
        self.extra_params = {"trees": tree_num_total,
                             "treesb": tree_num_build if tree_num_build > 0 else tree_num_total,
                             "ml": merge_level}
        self.L = myL
        self.Check_K = check


class Matrix:
    """Matrix<float> (general/matrix.hpp): rows x cols over a caller array."""

    def __init__(self, rows, cols, data):
        a = np.asarray(data, dtype=np.float32).reshape(rows, -1)
        self.rows, self.cols = rows, cols
        self.data = np.ascontiguousarray(a[:, :cols])

    def get_rows(self):
        return self.rows

    def get_cols(self):
        return self.cols


def _p(a, t):
    return a.ctypes.data_as(t)


class BatchPlan:
    """Captured heterogeneous batch (efg_batch_plan_*); destroy before the index."""

    def __init__(self, index, q_ptr, nq, q_stride, pool, k, out_stride, ids_ptr, d_ptr, extend, tune, stream):
        pool = np.ascontiguousarray(pool, np.int32)
        k = np.ascontiguousarray(k, np.int32)
        ext = None if extend is None else np.ascontiguousarray(extend, np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        self._index = index
        self.h = ctypes.c_void_p()
        check(lib().efg_batch_plan_create(index.h, q_ptr, nq, q_stride, _p(pool, ip),
                                          _p(ext, ip) if ext is not None else None, _p(k, ip), int(out_stride),
                                          ids_ptr, d_ptr, int(tune), stream, ctypes.byref(self.h)))

    def run(self, stream=None):
        check(lib().efg_batch_plan_run(self.h, stream))

    def close(self):
        if self.h:
            lib().efg_batch_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FIndex:
    def __init__(self, features, distance: Distance, params: IndexParams):
        if not isinstance(distance, _L2_CLASSES):
            raise EfgError(-2, "only the squared L2 distances are supported")
        if params.init_index_type != KDTREE_UB:
            raise EfgError(-2, "only the KDTREE_UB index is supported")
        m = features if isinstance(features, Matrix) else Matrix(*np.asarray(features).shape, features)
        self.n, self.d = m.rows, m.cols
        self._probe_epochs = None
        bp = BuildParams()
        lib().efg_default_build_params(ctypes.byref(bp), 4)
        ep = params.extra_params
        bp.trees = int(ep.get("trees", 4))
        bp.trees_build = int(ep.get("treesb", bp.trees))
        bp.merge_level = int(ep.get("ml", -1))
        bp.epochs = int(params.build_epoches)
        bp.check = int(params.Check_K)
        bp.L = int(params.L)
        bp.K = int(params.K)
        bp.S = int(params.S)
        bp.tns = int(params.TNS)
        bp.seed = int(getattr(params, "seed", 1998))
        self.params = bp
        h = ctypes.c_void_p()
        check(lib().efg_create(_p(m.data, _lib.c_f32p), self.n, self.d, self.d, ctypes.byref(bp), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                lib().efg_destroy(h)
            except Exception:
                pass
            self.h = None

    # --- build
    def buildIndex(self):
        check(lib().efg_build_index(self.h))

    def buildTrees(self):
        check(lib().efg_build_trees(self.h))

    def build_stage(self, stage: int):
        check(lib().efg_build_stage(self.h, stage))

    # --- search
    def setSearchParams(self, epochs, init_num, extend_to, search_trees=0, search_lv=-1, search_method=0):
        check(lib().efg_set_search_params(self.h, epochs, init_num, extend_to, search_trees, search_lv,
                                          search_method))

    def setHashCodes(self, base_codes, query_codes, codelen, upbits, radius, lenshift=0):
        """Make this a hashing index (IEH seeds, hashing_index.hpp): tables x n
        base codes and tables x nq query codes (unshifted), as the reference's
        code files hold them; buildIndex() then builds the hash tables."""
        b = np.ascontiguousarray(base_codes, np.uint64)
        q = np.ascontiguousarray(query_codes, np.uint64)
        if b.ndim != 2 or q.ndim != 2 or b.shape[0] != q.shape[0]:
            raise ValueError("codes must be (tables, n) and (tables, nq)")
        check(lib().efg_hash_set_codes(self.h, b.shape[0], int(codelen), int(lenshift), int(upbits), int(radius),
                                       _p(b, _lib.c_u64p), _p(q, _lib.c_u64p), q.shape[1]))

    def loadHashCodes(self, tables, codelen, upbits, radius, base_prefix, query_prefix, lenshift=0):
        """HASHINGIndexParams(codelen, tables, upbits, radius, base, query, lenshift):
        code files "<prefix>_<t>", t = 1..tables."""
        check(lib().efg_hash_load_codes(self.h, int(tables), int(codelen), int(lenshift), int(upbits), int(radius),
                                        str(base_prefix).encode(), str(query_prefix).encode()))

    def setBuildPrecision(self, precision: int):
        """0 = FP32 tiles; 1 = BF16 / 2 = FP16 MFMA merge and join tiles, FP32 re-rank of sets and pools."""
        check(lib().efg_set_build_precision(self.h, int(precision)))

    def setJoinStagingMargin(self, margin: float):
        """Join staging room: margin x the last iteration's kept candidate fraction (default 1.25)."""
        check(lib().efg_set_join_staging_margin(self.h, float(margin)))

    def setJoinRouting(self, direct: bool = True):
        """Join candidates straight into their targets' buckets (default), or staged first."""
        check(lib().efg_set_join_routing(self.h, 1 if direct else 0))

    def setSearchPrecision(self, precision: int):
        """0 = FP32 rows; 1 = BF16 rows with FP32 re-rank of the best 2k."""
        check(lib().efg_set_search_precision(self.h, int(precision)))

    # --- heterogeneous batch (no counterpart in the reference)
    def setBatchForm(self, one_pass: bool = True):
        """Batched search: all classes in one pass, or one pass per class (default)."""
        check(lib().efg_set_batch_form(self.h, 1 if one_pass else 0))

    def batch_shape(self):
        """The one-pass form's shape: small group's largest pool (0: one group),
        threads per query of the small and the large group, small group's share."""
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().efg_get_batch_shape(self.h, *[ctypes.byref(x) for x in v]))
        return dict(split=v[0].value, t_small=v[1].value, t_large=v[2].value, share=v[3].value / 100.0)

    def knnSearchBatch(self, query, pool, k, extend=None, out_stride=None, tune=False):
        """Per-query pool (init_num), k and extend_to (default pool // 6); returns (ids, dists)
        with out_stride columns (default max k), -1 / FLT_MAX past each query's k."""
        q = query.data if isinstance(query, Matrix) else np.ascontiguousarray(query, np.float32)
        nq = q.shape[0]
        pool = np.ascontiguousarray(pool, np.int32)
        k = np.ascontiguousarray(k, np.int32)
        assert pool.shape == (nq,) and k.shape == (nq,)
        ext = None if extend is None else np.ascontiguousarray(extend, np.int32)
        stride = int(out_stride or (int(k.max()) if nq else 1))
        ids = np.empty((nq, stride), np.int32)
        d = np.empty((nq, stride), np.float32)
        ip = ctypes.POINTER(ctypes.c_int)
        check(lib().efg_knn_search_batch(self.h, _p(q, _lib.c_f32p), nq, q.shape[1], _p(pool, ip),
                                         _p(ext, ip) if ext is not None else None, _p(k, ip), stride,
                                         _p(ids, _lib.c_i32p), _p(d, _lib.c_f32p), int(tune)))
        return ids, d

    def knn_search_batch_device(self, q_ptr, nq, q_stride, pool, k, out_stride, ids_ptr, d_ptr=None, extend=None,
                                tune=False, stream=None):
        pool = np.ascontiguousarray(pool, np.int32)
        k = np.ascontiguousarray(k, np.int32)
        ext = None if extend is None else np.ascontiguousarray(extend, np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        check(lib().efg_knn_search_batch_device(self.h, q_ptr, nq, q_stride, _p(pool, ip),
                                                _p(ext, ip) if ext is not None else None, _p(k, ip), int(out_stride),
                                                ids_ptr, d_ptr, int(tune), stream))

    def batch_plan(self, q_ptr, nq, q_stride, pool, k, out_stride, ids_ptr, d_ptr=None, extend=None, tune=False,
                   stream=None):
        """A heterogeneous batch captured into a HIP graph over these device buffers: BatchPlan.run()
        replays it (write the next queries into q_ptr first)."""
        return BatchPlan(self, q_ptr, nq, q_stride, pool, k, out_stride, ids_ptr, d_ptr, extend, tune, stream)

    # --- launch tuning (no counterpart in the reference)
    def setSearchThreads(self, threads: int):
        """Threads per query: 64, 128, 256, or 0 for tuned / rule of thumb."""
        check(lib().efg_set_search_threads(self.h, int(threads)))

    def autotune(self, k, query):
        """Times 64/128/256 threads per query on these queries; returns (full, probe) choices."""
        q = query.data if isinstance(query, Matrix) else np.ascontiguousarray(query, np.float32)
        out = (ctypes.c_int * 2)()
        check(lib().efg_autotune_search(self.h, int(k), _p(q, _lib.c_f32p), q.shape[0], q.shape[1], out))
        return out[0], out[1]

    # --- learned stopping (no counterpart in the reference)
    def setProbeParams(self, pool=0, extend_to=0, epochs=0):
        """Probe search of the learned stopping rule (0 = derived from the search parameters)."""
        check(lib().efg_set_probe_params(self.h, int(pool), int(extend_to), int(epochs)))
        self._probe_epochs = int(epochs) if epochs > 0 else None

    def searchTraces(self, k, query, plain_ids):
        """Probe-search traces: features [nq, C, 16] and labels [nq, C] (255 = not reached)."""
        q = query.data if isinstance(query, Matrix) else np.ascontiguousarray(query, np.float32)
        plain = np.ascontiguousarray(plain_ids, np.int32)
        nq = q.shape[0]
        assert plain.shape == (nq, k)
        ep = self.getSearchParams()["epochs"] if self._probe_epochs is None else self._probe_epochs
        C = ep + 1
        feat = np.empty((nq, C, 16), np.float32)
        label = np.empty((nq, C), np.uint8)
        check(lib().efg_search_traces(self.h, int(k), _p(q, _lib.c_f32p), nq, q.shape[1], _p(plain, _lib.c_i32p),
                                      _p(feat, _lib.c_f32p), _p(label, _lib.c_u8p)))
        return feat, label

    def setStoppingNetwork(self, params, norm, threshold=0.9):
        params = np.ascontiguousarray(params, np.float32)
        norm = np.ascontiguousarray(norm, np.float32)
        assert params.shape == (lib().efg_mlp_param_count(),) and norm.shape == (32,)
        check(lib().efg_set_stopping_network(self.h, _p(params, _lib.c_f32p), _p(norm, _lib.c_f32p),
                                             float(threshold)))

    def setSearchMode(self, mode: int):
        """0 = plain search; 1 = probe search stopped by the network, the full
        search from scratch for the rest; 2 = the same probe, the rest continue
        in place into the full search."""
        check(lib().efg_set_search_mode(self.h, int(mode)))

    def learned_stats(self):
        """last_fallbacks: queries of the last learned search the network did
        not stop; last_restarts: those of them that ran the full search from
        scratch; running totals."""
        v = [ctypes.c_uint64() for _ in range(3)]
        check(lib().efg_get_learned_stats(self.h, *[ctypes.byref(x) for x in v]))
        r = ctypes.c_uint64()
        check(lib().efg_get_learned_restarts(self.h, ctypes.byref(r)))
        out = dict(zip(["last_fallbacks", "total_fallbacks", "total_queries"], [x.value for x in v]))
        out["last_restarts"] = r.value
        return out

    def getSearchParams(self):
        v = [ctypes.c_int() for _ in range(5)]
        check(lib().efg_get_search_params(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(["epochs", "init_num", "extend_to", "search_trees", "search_method"], [x.value for x in v]))

    def knnSearch(self, k, query):
        q = query.data if isinstance(query, Matrix) else np.ascontiguousarray(query, np.float32)
        check(lib().efg_knn_search(self.h, int(k), _p(q, _lib.c_f32p), q.shape[0], q.shape[1]))

    def results(self):
        nq = ctypes.c_size_t()
        k = ctypes.c_int()
        check(lib().efg_get_results(self.h, ctypes.byref(nq), ctypes.byref(k), None, None))
        ids = np.empty((nq.value, k.value), np.int32)
        d = np.empty((nq.value, k.value), np.float32)
        check(lib().efg_get_results(self.h, None, None, _p(ids, _lib.c_i32p), _p(d, _lib.c_f32p)))
        return ids, d

    def knn_search_device(self, k, q_ptr, nq, q_stride, ids_ptr, d_ptr=None, ndist_ptr=None, stream=None):
        check(lib().efg_knn_search_device(self.h, int(k), q_ptr, nq, q_stride, ids_ptr, d_ptr, ndist_ptr, stream))

    # --- persistence
    def saveGraph(self, fn):
        check(lib().efg_save_graph(self.h, str(fn).encode()))

    def loadGraph(self, fn):
        check(lib().efg_load_graph(self.h, str(fn).encode()))

    def saveTrees(self, fn):
        check(lib().efg_save_trees(self.h, str(fn).encode()))

    def loadTrees(self, fn):
        check(lib().efg_load_trees(self.h, str(fn).encode()))

    def saveIndex(self, fn):
        check(lib().efg_save_index(self.h, str(fn).encode()))

    def loadIndex(self, fn):
        check(lib().efg_load_index(self.h, str(fn).encode()))

    def saveImage(self, fn):
        """Device-loadable index image (trees + graph as laid out in device memory)."""
        check(lib().efg_save_image(self.h, str(fn).encode()))

    def loadImage(self, fn):
        check(lib().efg_load_image(self.h, str(fn).encode()))

    def saveResults(self, fn):
        check(lib().efg_save_results(self.h, str(fn).encode()))

    # --- graph access
    def getGraphSize(self):
        n = ctypes.c_size_t()
        check(lib().efg_graph_size(self.h, ctypes.byref(n)))
        return n.value

    def getGraphRow(self, row):
        ln = ctypes.c_size_t()
        check(lib().efg_graph_row(self.h, row, None, 0, ctypes.byref(ln)))
        out = np.empty(ln.value, np.uint32)
        if ln.value:
            check(lib().efg_graph_row(self.h, row, _p(out, _lib.c_u32p), ln.value, ctypes.byref(ln)))
        return out.tolist()

    def outputVisitBucketNum(self):
        """No-op for the KD-tree index (kdtreeub_index.hpp:1277)."""

    # --- inspection
    def trees(self):
        T = ctypes.c_int()
        M = ctypes.c_size_t()
        check(lib().efg_tree_shape(self.h, ctypes.byref(T), ctypes.byref(M)))
        T, M = T.value, M.value
        out = dict(dim=np.empty(M, np.int32), val=np.empty(M, np.float32), left=np.empty(M, np.int32),
                   right=np.empty(M, np.int32), start=np.empty(M, np.uint32), end=np.empty(M, np.uint32),
                   roots=np.empty(T, np.int32), leaf=np.empty((T, self.n), np.uint32))
        check(lib().efg_get_trees(self.h, _p(out["dim"], _lib.c_i32p), _p(out["val"], _lib.c_f32p),
                                  _p(out["left"], _lib.c_i32p), _p(out["right"], _lib.c_i32p),
                                  _p(out["start"], _lib.c_u32p), _p(out["end"], _lib.c_u32p),
                                  _p(out["roots"], _lib.c_i32p), _p(out["leaf"], _lib.c_u32p)))
        return out

    def set_trees(self, t):
        a = {k: np.ascontiguousarray(t[k], dt) for k, dt in
             dict(dim=np.int32, val=np.float32, left=np.int32, right=np.int32, start=np.uint32, end=np.uint32,
                  roots=np.int32, leaf=np.uint32).items()}
        check(lib().efg_set_trees(self.h, len(a["roots"]), len(a["dim"]), _p(a["dim"], _lib.c_i32p),
                                  _p(a["val"], _lib.c_f32p), _p(a["left"], _lib.c_i32p), _p(a["right"], _lib.c_i32p),
                                  _p(a["start"], _lib.c_u32p), _p(a["end"], _lib.c_u32p),
                                  _p(a["roots"], _lib.c_i32p), _p(a["leaf"], _lib.c_u32p)))

    def graph(self):
        n = self.n
        off = np.empty(n + 1, np.uint64)
        check(lib().efg_get_graph(self.h, _p(off, _lib.c_u64p), None, 0))
        ids = np.empty(int(off[n]), np.uint32)
        check(lib().efg_get_graph(self.h, _p(off, _lib.c_u64p), _p(ids, _lib.c_u32p), ids.size))
        return off, ids

    def graph_matrix(self):
        off, ids = self.graph()
        return ids.reshape(self.n, int(off[1] - off[0]))

    def set_graph(self, off, ids):
        off = np.ascontiguousarray(off, np.uint64)
        ids = np.ascontiguousarray(ids, np.uint32)
        check(lib().efg_set_graph(self.h, _p(off, _lib.c_u64p), _p(ids, _lib.c_u32p)))

    def counters(self):
        c = np.zeros(8, np.uint64)
        check(lib().efg_get_counters(self.h, _p(c, _lib.c_u64p), 8))
        return dict(search_dist=int(c[0]), merge=int(c[1]), init=int(c[2]), join=int(c[3]), overflow=int(c[4]),
                    overflow2=int(c[5]), join_rows=int(c[6]), join_reruns=int(c[7]))

    def join_pairs(self):
        out = np.zeros(64, np.uint64)
        m = ctypes.c_int()
        check(lib().efg_get_join_pairs(self.h, _p(out, _lib.c_u64p), 64, ctypes.byref(m)))
        return out[:m.value].tolist()

    def set_profiling(self, on: bool):
        check(lib().efg_set_profiling(self.h, int(on)))

    def kernel_times(self):
        """Summed ms [tree descent, search kernel, overflow re-run] and the search count."""
        out = (ctypes.c_double * 3)()
        m = ctypes.c_int()
        check(lib().efg_get_kernel_times(self.h, out, 3, ctypes.byref(m)))
        return list(out), m.value

    def search_work(self):
        """[distances, seed leaf ids read, nodes expanded] since the last call."""
        c = np.zeros(3, np.uint64)
        check(lib().efg_get_search_work(self.h, _p(c, _lib.c_u64p), 3))
        return [int(x) for x in c]

    def build_times(self):
        out = (ctypes.c_double * 4)()
        check(lib().efg_get_build_times(self.h, out, 4))
        return list(out)

    def pools(self):
        L = self.params.L
        keys = np.empty((self.n, L), np.uint64)
        flags = np.empty((self.n, L), np.uint8)
        fill = np.empty(self.n, np.uint32)
        rng = np.empty(self.n, np.uint32)
        check(lib().efg_get_pools(self.h, _p(keys, _lib.c_u64p), _p(flags, _lib.c_u8p), _p(fill, _lib.c_u32p),
                                  _p(rng, _lib.c_u32p)))
        return keys, flags, fill, rng

    def lists(self, which):
        off = np.empty(self.n + 1, np.uint64)
        check(lib().efg_get_lists(self.h, which, _p(off, _lib.c_u64p), None, 0))
        ids = np.empty(int(off[-1]), np.uint32)
        check(lib().efg_get_lists(self.h, which, _p(off, _lib.c_u64p), _p(ids, _lib.c_u32p), ids.size))
        return off, ids

    def device_data(self):
        p = ctypes.c_void_p()
        s = ctypes.c_size_t()
        check(lib().efg_device_data(self.h, ctypes.byref(p), ctypes.byref(s)))
        return p.value, s.value


def exact_knn_device(base_ptr, n, d, base_stride, q_ptr, nq, q_stride, k, exclude_self, ids_ptr, d_ptr=None,
                     stream=None, precision=0):
    """Exact kNN over device rows; precision 0 FP32, 1 BF16 / 2 FP16 MFMA tiles + FP32 re-rank of the best 2k."""
    check(lib().efg_exact_knn_device_prec(base_ptr, n, d, base_stride, q_ptr, nq, q_stride, k, int(exclude_self),
                                          int(precision), ids_ptr, d_ptr, stream))
