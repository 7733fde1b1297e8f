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


# ---- index image (efg_save_image / efg_load_image) --------------------------
# header: 8-byte magic, u32 version, u32 sections, u64 n d D M nnz, u32 T,
# 8 x i32 (reg_deg, max_deg, leaf0_perm, tree_lb, tree_S, lazy, pad), u64
# seed, u64 check; then per section u32 id, u32 element size, u64 count,
# u64 offset, u64 check; sections at 4096-byte aligned offsets.
IMAGE_HEADER = struct.Struct("<8sII5QI7iQQ")
IMAGE_SECTION = struct.Struct("<IIQQQ")
M64 = (1 << 64) - 1


def image_hash(data: bytes) -> int:
    """The image's 64-bit multiply-xor checksum, written independently."""
    h = 0x9E3779B97F4A7C15
    n8 = len(data) // 8
    words = struct.unpack_from(f"<{n8}Q", data) if n8 else ()
    for w in words:
        h = ((h ^ w) * 0xFF51AFD7ED558CCD) & M64
        h ^= h >> 29
    for b in data[8 * n8:]:
        h = ((h ^ b) * 0x100000001B3) & M64
    return h


def read_image(data: bytes):
    hdr = IMAGE_HEADER.unpack_from(data, 0)
    nsec = hdr[2]
    secs = [IMAGE_SECTION.unpack_from(data, IMAGE_HEADER.size + i * IMAGE_SECTION.size) for i in range(nsec)]
    return hdr, secs


def resign_image(data: bytearray):
    """Recompute every section checksum and the header checksum in place."""
    hdr, secs = read_image(bytes(data))
    table = b""
    for i, (sid, elem, count, off, _) in enumerate(secs):
        chk = image_hash(bytes(data[off:off + elem * count]))
        rec = IMAGE_SECTION.pack(sid, elem, count, off, chk)
        data[IMAGE_HEADER.size + i * IMAGE_SECTION.size:IMAGE_HEADER.size + (i + 1) * IMAGE_SECTION.size] = rec
        table += rec
    h0 = IMAGE_HEADER.pack(*hdr[:-1], 0)
    data[:IMAGE_HEADER.size] = IMAGE_HEADER.pack(*hdr[:-1], image_hash(h0) ^ image_hash(table))
