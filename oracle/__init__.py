"""ctypes binding of the CPU oracle (oracle/efg_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product library never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_sz = ctypes.c_size_t


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        sig = {
            "orc_last_error": (ctypes.c_char_p, []),
            "orc_hash4": (ctypes.c_uint64, [ctypes.c_uint64] * 4),
            "orc_mix64": (ctypes.c_uint64, [ctypes.c_uint64]),
            "orc_perm": (ctypes.c_uint64, [ctypes.c_uint64] * 4),
            "orc_l2": (ctypes.c_float, [_f32p, _f32p, _sz]),
            "orc_insert_into_knn": (ctypes.c_uint, [_u32p, _f32p, ctypes.c_uint, ctypes.c_uint32, ctypes.c_float]),
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_exact_knn": (None, [_f32p, _sz, _sz, _sz, _f32p, _sz, _sz, ctypes.c_int, ctypes.c_int, _i32p, _f32p]),
            "orc_new": (ctypes.c_void_p, [_f32p, _sz, _sz, _sz]),
            "orc_free": (None, [ctypes.c_void_p]),
            "orc_set_params": (None, [ctypes.c_void_p] + [ctypes.c_int] * 9 + [ctypes.c_uint64, ctypes.c_int]),
            "orc_gen_random": (None, [ctypes.c_uint64, ctypes.c_uint32, _u32p, ctypes.c_uint, ctypes.c_uint]),
            "orc_keep_rev": (None, [ctypes.c_void_p, ctypes.c_int]),
            "orc_rev_total": (_sz, [ctypes.c_void_p, ctypes.c_int]),
            "orc_get_rev": (None, [ctypes.c_void_p, ctypes.c_int, _u64p, _u32p]),
            "orc_search_seeds": (ctypes.c_int, [ctypes.c_void_p, _f32p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                                _u32p]),
            "orc_build_trees": (ctypes.c_int, [ctypes.c_void_p]),
            "orc_merge_init": (ctypes.c_int, [ctypes.c_void_p]),
            "orc_refine_iter": (ctypes.c_int, [ctypes.c_void_p]),
            "orc_finalize": (ctypes.c_int, [ctypes.c_void_p]),
            "orc_build_index": (ctypes.c_int, [ctypes.c_void_p]),
            "orc_node_count": (_sz, [ctypes.c_void_p]),
            "orc_tree_count": (ctypes.c_int, [ctypes.c_void_p]),
            "orc_get_nodes": (None, [ctypes.c_void_p, _i32p, _f32p, _i32p, _i32p, _u32p, _u32p, _i32p]),
            "orc_get_leaf_lists": (None, [ctypes.c_void_p, _u32p]),
            "orc_set_trees": (None, [ctypes.c_void_p, ctypes.c_int, _sz, _i32p, _f32p, _i32p, _i32p, _u32p, _u32p, _i32p, _u32p]),
            "orc_graph_total": (_sz, [ctypes.c_void_p]),
            "orc_get_graph": (None, [ctypes.c_void_p, _u64p, _u32p]),
            "orc_set_graph": (None, [ctypes.c_void_p, _u64p, _u32p]),
            "orc_get_pools": (None, [ctypes.c_void_p, _u64p, _u8p, _u32p, _u32p]),
            "orc_list_total": (_sz, [ctypes.c_void_p, ctypes.c_int]),
            "orc_get_lists": (None, [ctypes.c_void_p, ctypes.c_int, _u64p, _u32p]),
            "orc_get_found": (None, [ctypes.c_void_p, _u8p, _f32p]),
            "orc_get_counters": (None, [ctypes.c_void_p, _u64p]),
            "orc_join_pairs": (_sz, [ctypes.c_void_p, _u64p, _sz]),
            "orc_search": (ctypes.c_int, [ctypes.c_void_p, _f32p, _sz, _sz, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _i32p, _f32p, _u64p]),
            "orc_search_probe": (ctypes.c_int, [ctypes.c_void_p, _f32p, _sz, _sz] + [ctypes.c_int] * 9 +
                                 [_f32p, _f32p, ctypes.c_float, _i32p, _i32p, _f32p, _f32p, _u8p, _u8p]),
            "orc_mlp_logits": (None, [_f32p, _f32p, _f32p, _sz, _f32p]),
            "orc_mlp_grad64": (ctypes.c_double, [_f32p, _f32p, _f32p, _u8p, _sz, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_double)]),
            "orc_mlp_param_count": (ctypes.c_int, []),
            "orc_hash_set": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64p, _u64p,
                                            _sz]),
            "orc_hash_masks": (_sz, [ctypes.c_void_p, _u64p, _sz]),
            "orc_hash_seeds": (_sz, [ctypes.c_void_p, _sz, ctypes.c_int, ctypes.c_int, _u32p]),
            "orc_hash_search": (ctypes.c_int, [ctypes.c_void_p, _f32p, _sz, _sz, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, _i32p,
                                               _f32p, _u64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(f"oracle error {rc}: {lib().orc_last_error().decode()}")


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def mix64(x: int) -> int:
    return lib().orc_mix64(x)


def hash4(seed, stream, a, b) -> int:
    return lib().orc_hash4(seed, stream, a, b)


def perm(seed, tree, n, i) -> int:
    return lib().orc_perm(seed, tree, n, i)


def l2(a: np.ndarray, b: np.ndarray) -> float:
    """Squared L2 in the reference's AVX accumulation order."""
    d = a.shape[0]
    pad = (d + 7) // 8 * 8
    aa = np.zeros(pad, np.float32)
    bb = np.zeros(pad, np.float32)
    aa[:d] = a
    bb[:d] = b
    return float(lib().orc_l2(_p(aa, _f32p), _p(bb, _f32p), d))


def gen_random(seed, node, size, N):
    """GenRandom (kdtreeub_index.hpp:1121-1135) with the oracle's generator."""
    out = np.empty(size, np.uint32)
    lib().orc_gen_random(seed, node, _p(out, _u32p), size, N)
    return out


def insert_into_knn(ids, dists, K, nid, ndist):
    ids = np.ascontiguousarray(ids, np.uint32).copy()
    dists = np.ascontiguousarray(dists, np.float32).copy()
    r = lib().orc_insert_into_knn(_p(ids, _u32p), _p(dists, _f32p), K, nid, ndist)
    return r, ids, dists


def _pad8(a: np.ndarray) -> np.ndarray:
    """Rows zero-padded to a multiple of 8 floats (the reference Matrix stride)."""
    a = np.ascontiguousarray(a, np.float32)
    d = a.shape[-1]
    D = (d + 7) & ~7
    if D == d:
        return a
    out = np.zeros(a.shape[:-1] + (D,), np.float32)
    out[..., :d] = a
    return out


def exact_knn(base: np.ndarray, queries: np.ndarray, k: int, exclude_self: bool = False):
    n, d = np.asarray(base).shape
    base = _pad8(base)
    queries = _pad8(queries)
    D = base.shape[1]
    nq = queries.shape[0]
    ids = np.empty((nq, k), np.int32)
    dists = np.empty((nq, k), np.float32)
    lib().orc_exact_knn(_p(base, _f32p), n, d, D, _p(queries, _f32p), nq, D, k, int(exclude_self),
                        _p(ids, _i32p), _p(dists, _f32p))
    return ids, dists


class Index:
    """Oracle index: trees, merge/init, refine, search — mirrors FIndex."""

    def __init__(self, data: np.ndarray, trees=8, trees_build=0, ml=8, epochs=8, check=25, L=30, K=10, S=10,
                 tns=10, seed=1998, join_mode=0):
        self.data = np.ascontiguousarray(data, np.float32)
        self.n, self.d = self.data.shape
        self.h = lib().orc_new(_p(self.data, _f32p), self.n, self.d, self.d)
        self.params = dict(trees=trees, trees_build=trees_build or trees, ml=ml, epochs=epochs, check=check, L=L,
                           K=K, S=S, tns=tns, seed=seed, join_mode=join_mode)
        lib().orc_set_params(self.h, trees, trees_build or trees, ml, epochs, check, L, K, S, tns, seed, join_mode)
        self.init_num_state = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_free(self.h)
                self.h = None
        except Exception:
            pass

    # ---- build stages
    def build_trees(self):
        _check(lib().orc_build_trees(self.h))

    def merge_init(self):
        _check(lib().orc_merge_init(self.h))

    def refine_iter(self):
        _check(lib().orc_refine_iter(self.h))

    def finalize(self):
        _check(lib().orc_finalize(self.h))

    def build_index(self):
        _check(lib().orc_build_index(self.h))

    # ---- state
    def trees(self):
        m = lib().orc_node_count(self.h)
        T = lib().orc_tree_count(self.h)
        dim = np.empty(m, np.int32)
        val = np.empty(m, np.float32)
        left = np.empty(m, np.int32)
        right = np.empty(m, np.int32)
        start = np.empty(m, np.uint32)
        end = np.empty(m, np.uint32)
        roots = np.empty(T, np.int32)
        lib().orc_get_nodes(self.h, _p(dim, _i32p), _p(val, _f32p), _p(left, _i32p), _p(right, _i32p),
                            _p(start, _u32p), _p(end, _u32p), _p(roots, _i32p))
        leaf = np.empty((T, self.n), np.uint32)
        lib().orc_get_leaf_lists(self.h, _p(leaf, _u32p))
        return dict(dim=dim, val=val, left=left, right=right, start=start, end=end, roots=roots, leaf=leaf)

    def set_trees(self, t):
        T = len(t["roots"])
        m = len(t["dim"])
        arr = {k: np.ascontiguousarray(t[k], dt) for k, dt in
               dict(dim=np.int32, val=np.float32, left=np.int32, right=np.int32, start=np.uint32, end=np.uint32,
                    roots=np.int32, leaf=np.uint32).items()}
        lib().orc_set_trees(self.h, T, m, _p(arr["dim"], _i32p), _p(arr["val"], _f32p), _p(arr["left"], _i32p),
                            _p(arr["right"], _i32p), _p(arr["start"], _u32p), _p(arr["end"], _u32p),
                            _p(arr["roots"], _i32p), _p(arr["leaf"], _u32p))

    def graph(self):
        tot = lib().orc_graph_total(self.h)
        off = np.empty(self.n + 1, np.uint64)
        ids = np.empty(tot, np.uint32)
        lib().orc_get_graph(self.h, _p(off, _u64p), _p(ids, _u32p))
        return off, ids

    def graph_matrix(self):
        off, ids = self.graph()
        k = int(off[1] - off[0])
        return ids.reshape(self.n, k)

    def set_graph(self, off, ids):
        off = np.ascontiguousarray(off, np.uint64)
        ids = np.ascontiguousarray(ids, np.uint32)
        lib().orc_set_graph(self.h, _p(off, _u64p), _p(ids, _u32p))

    def pools(self):
        L = self.params["L"]
        keys = np.empty((self.n, L), np.uint64)
        flags = np.empty((self.n, L), np.uint8)
        fill = np.empty(self.n, np.uint32)
        rng = np.empty(self.n, np.uint32)
        lib().orc_get_pools(self.h, _p(keys, _u64p), _p(flags, _u8p), _p(fill, _u32p), _p(rng, _u32p))
        return keys, flags, fill, rng

    def lists(self, which):
        tot = lib().orc_list_total(self.h, which)
        off = np.empty(self.n + 1, np.uint64)
        ids = np.empty(tot, np.uint32)
        lib().orc_get_lists(self.h, which, _p(off, _u64p), _p(ids, _u32p))
        return off, ids

    def keep_reverse_lists(self, on=True):
        """Record the reverse lists of each update before and after the cap."""
        lib().orc_keep_rev(self.h, int(on))

    def reverse_lists(self, which):
        """which: 0 new before sampling, 1 old before, 2 new after, 3 old after (CSR, each row sorted)."""
        tot = lib().orc_rev_total(self.h, which)
        off = np.empty(self.n + 1, np.uint64)
        ids = np.empty(tot, np.uint32)
        lib().orc_get_rev(self.h, which, _p(off, _u64p), _p(ids, _u32p))
        return off, ids

    def search_seeds(self, query, qid, init_num, search_trees):
        """Seed pool of one query (method 0) in sequence order."""
        q = np.ascontiguousarray(query, np.float32)
        out = np.empty(init_num, np.uint32)
        _check(lib().orc_search_seeds(self.h, _p(q, _f32p), qid, init_num, search_trees, _p(out, _u32p)))
        return out

    def found(self):
        f = np.empty(self.n, np.uint8)
        r = np.empty(self.n, np.float32)
        lib().orc_get_found(self.h, _p(f, _u8p), _p(r, _f32p))
        return f, r

    def counters(self):
        c = np.empty(4, np.uint64)
        lib().orc_get_counters(self.h, _p(c, _u64p))
        return dict(merge=int(c[0]), init=int(c[1]), join=int(c[2]), error_flag=int(c[3]))

    def join_pairs(self):
        out = np.empty(64, np.uint64)
        m = lib().orc_join_pairs(self.h, _p(out, _u64p), 64)
        return out[:m].tolist()

    def search_probe(self, queries, k, epochs, init_num, extend_to, search_trees, p_pool, p_extend, p_epochs,
                     mode, params=None, norm=None, zthr=0.0, plain=None):
        """Learned stopping restated: mode 2 returns traces (features, labels);
        mode 1 returns (ids, dists, stopped) of the probe search with the
        network and the full search from scratch for the queries it does not
        stop; mode 3 the same with those queries continued in place."""
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        C = p_epochs + 1
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        feat = np.zeros((nq, C, 16), np.float32)
        label = np.full((nq, C), 255, np.uint8)
        stopped = np.zeros(nq, np.uint8)
        w = np.ascontiguousarray(params if params is not None else np.zeros(5313), np.float32)
        nm = np.ascontiguousarray(norm if norm is not None else np.zeros(32), np.float32)
        pl = np.ascontiguousarray(plain if plain is not None else np.zeros((nq, k)), np.int32)
        _check(lib().orc_search_probe(self.h, _p(q, _f32p), nq, q.shape[1], k, epochs, init_num, extend_to,
                                      search_trees, p_pool, p_extend, p_epochs, mode, _p(w, _f32p), _p(nm, _f32p),
                                      ctypes.c_float(zthr), _p(pl, _i32p), _p(ids, _i32p), _p(dists, _f32p),
                                      _p(feat, _f32p), _p(label, _u8p), _p(stopped, _u8p)))
        if mode == 2:
            return feat, label
        return ids, dists, stopped.astype(bool)

    def search(self, queries, k, epochs, init_num, extend_to, search_trees, method=0):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        nd = np.empty(nq, np.uint64)
        init = ctypes.c_int(int(init_num))
        _check(lib().orc_search(self.h, _p(q, _f32p), nq, q.shape[1], k, epochs, ctypes.byref(init), extend_to,
                                search_trees, method, _p(ids, _i32p), _p(dists, _f32p), _p(nd, _u64p)))
        self.init_num_state = init.value
        return ids, dists, nd

    # ---- hashing index (IEH seeds) ---------------------------------------
    def set_hash(self, base_codes, query_codes, bits, radius):
        """Tables from T x n base codes and T x nq query codes (already
        shifted), code length `bits`, Hamming radius `radius`."""
        b = np.ascontiguousarray(base_codes, np.uint64)
        qc = np.ascontiguousarray(query_codes, np.uint64)
        T = b.shape[0]
        _check(lib().orc_hash_set(self.h, T, int(bits), int(radius), _p(b, _u64p), _p(qc, _u64p), qc.shape[1]))

    def hash_masks(self):
        m = lib().orc_hash_masks(self.h, None, 0)
        out = np.empty(m, np.uint64)
        lib().orc_hash_masks(self.h, _p(out, _u64p), m)
        return out

    def hash_seeds(self, qi, init_num, topup_to=-1):
        out = np.empty(max(init_num, topup_to, 1), np.uint32)
        m = lib().orc_hash_seeds(self.h, int(qi), int(init_num), int(topup_to), _p(out, _u32p))
        return out[:m]

    def hash_search(self, queries, k, epochs, init_num, extend_to, method=1):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        nd = np.empty(nq, np.uint64)
        init = ctypes.c_int(int(init_num))
        _check(lib().orc_hash_search(self.h, _p(q, _f32p), nq, q.shape[1], k, epochs, ctypes.byref(init),
                                     extend_to, method, _p(ids, _i32p), _p(dists, _f32p), _p(nd, _u64p)))
        self.init_num_state = init.value
        return ids, dists, nd


def mlp_logits(params, norm, feat):
    """Logits of the stopping network (FP32, BF16-rounded parameters)."""
    w = np.ascontiguousarray(params, np.float32)
    nm = np.ascontiguousarray(norm, np.float32)
    f = np.ascontiguousarray(feat, np.float32).reshape(-1, 16)
    z = np.empty(f.shape[0], np.float32)
    lib().orc_mlp_logits(_p(w, _f32p), _p(nm, _f32p), _p(f, _f32p), f.shape[0], _p(z, _f32p))
    return z


def mlp_grad64(params, norm, feat, label, round_params=True):
    """(mean BCE loss, gradient) in float64."""
    w = np.ascontiguousarray(params, np.float32)
    nm = np.ascontiguousarray(norm, np.float32)
    f = np.ascontiguousarray(feat, np.float32).reshape(-1, 16)
    y = np.ascontiguousarray(label, np.uint8)
    g = np.empty(lib().orc_mlp_param_count(), np.float64)
    loss = lib().orc_mlp_grad64(_p(w, _f32p), _p(nm, _f32p), _p(f, _f32p), _p(y, _u8p), f.shape[0],
                                1 if round_params else 0, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return loss, g
