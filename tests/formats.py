"""Byte layouts of the reference's files, written from their description
(kdtreeub_index.hpp:197-312, 658-700; base_index.hpp:136-151;
samples/*.cc load_data), for fixtures and checks."""
from __future__ import annotations

import struct

import numpy as np


def write_vecs(path, a, dtype=np.float32):
    a = np.ascontiguousarray(a, dtype)
    n, d = a.shape
    rec = np.empty((n, d + 1), np.int32)
    rec[:, 0] = d
    rec[:, 1:] = a.view(np.int32)
    rec.tofile(path)


def read_ivecs(path):
    raw = np.fromfile(path, np.int32)
    d = raw[0]
    return raw.reshape(-1, d + 1)[:, 1:]


def graph_bytes(rows):
    out = bytearray()
    for r in rows:
        out += struct.pack("<I", len(r)) + np.asarray(r, "<u4").tobytes()
    return bytes(out)


def read_graph(data: bytes):
    rows, o = [], 0
    while o < len(data):
        (k,) = struct.unpack_from("<I", data, o)
        rows.append(np.frombuffer(data, "<u4", k, o + 4).copy())
        o += 4 + 4 * k
    return rows


def results_bytes(rows):
    out = bytearray()
    for r in rows:
        out += struct.pack("<i", len(r)) + np.asarray(r, "<i4").tobytes()
    return bytes(out)


def read_trees(data: bytes):
    """Header (u32 K, u32 tree_num, u64 num, u64 dim), per tree i32 node_num,
    i32 node_size, post-order 40-byte records {i32 DivDim, f32 DivVal, u64
    Start, u64 End, u64 L, u64 R}, then tree_num x num u32 leaf lists.  Links
    are rebuilt from Start/End adjacency as the reference's loader does."""
    K, T, num, dim = struct.unpack_from("<IIQQ", data, 0)
    o = 24
    trees = []
    for _ in range(T):
        node_num, node_size = struct.unpack_from("<ii", data, o)
        o += 8
        recs = []
        for _j in range(node_num):
            dd, dv, s, e, lp, rp = struct.unpack_from("<ifQQQQ", data, o)
            o += 40
            recs.append(dict(dim=dd, val=dv, start=s, end=e, lp=lp, rp=rp, left=None, right=None))
        st = []
        for i, r in enumerate(recs):
            if len(st) >= 2 and recs[st[-1]]["end"] == r["end"] and recs[st[-2]]["start"] == r["start"]:
                r["right"], r["left"] = st.pop(), st.pop()
            st.append(i)
        assert len(st) == 1
        trees.append(dict(node_size=node_size, nodes=recs, root=st[0]))
    leaf = np.frombuffer(data, "<u4", T * num, o).reshape(T, num).copy()
    o += 4 * T * num
    return dict(K=K, T=T, num=num, dim=dim, trees=trees, leaf=leaf, end=o)


def trees_bytes(K, num, dim, trees, leaf):
    """Inverse of read_trees for trees given as lists of post-order records."""
    out = bytearray(struct.pack("<IIQQ", K, len(trees), num, dim))
    for recs in trees:
        out += struct.pack("<ii", len(recs), 48)
        for r in recs:
            out += struct.pack("<ifQQQQ", r["dim"], r["val"], r["start"], r["end"], r.get("lp", 0), r.get("rp", 0))
    out += np.asarray(leaf, "<u4").tobytes()
    return bytes(out)
