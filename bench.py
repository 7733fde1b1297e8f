#!/usr/bin/env python3
"""Benchmark: EFANNA graph search and graph build on one MI355X.

Workload (mixture1m): N = 1,000,000 base vectors, d = 128, Gaussian mixture of
1,000 centres (uniform in [0,255]^128, sigma 20, clipped, rounded); graph built
with 16 trees (8 used for the graph), conquer depth 8, 8 NN-descent
iterations, L 30, check 25, K 10, S 10; 10,000 queries from the same mixture
searched with 16 trees, 4 epochs, pool 1200, extend 200, k 10.

One step = one search of all queries (inputs resident in HBM).  The graph
build is timed once before the steps.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "search QPS & graph-build s at recall@10≥0.95, chosen vs plain FP32 (1M×128)"
HBM_PEAK_GBS = 8000.0


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, min(n, 16))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-queries", type=int, default=10_000, help="queries in the CPU baseline sample")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run one 1 GiB device copy first (reference point for HBM counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(search shards by query, build by node range; see DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import oracle
    from efanna_amd import datasets
    from efanna_amd.index import FIndex, KDTreeUbIndexParams, L2DistanceAVX, exact_knn_device

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")

    trees, trees_build, ml, epochs, check, L, K, S = 16, 8, 8, 8, 25, 30, 10, 10
    s_epochs, s_pool, s_extend, s_trees, k = 4, 1200, 200, 16, 10

    t = time.perf_counter()
    base, q = datasets.workload("mixture1m", n=args.n, nq=args.nq)
    gen_s = time.perf_counter() - t
    n, d = base.shape
    nq = q.shape[0]

    gix = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, trees, ml, epochs, check, L, K, trees_build, S))
    torch.cuda.synchronize()
    t = time.perf_counter()
    gix.buildIndex()
    build_s = time.perf_counter() - t
    phase_ms = gix.build_times()
    bc = gix.counters()
    join_pairs = gix.join_pairs()

    gix.setSearchParams(s_epochs, s_pool, s_extend, s_trees, -1, 0)
    stream = torch.cuda.Stream()  # a real stream: the search is launched on it
    torch.cuda.set_stream(stream)
    if args.copy_calib:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst
    qd = torch.from_numpy(q).to(dev).contiguous()
    ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    dd = torch.empty((nq, k), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def step():
        gix.knn_search_device(k, qd.data_ptr(), nq, d, ids.data_ptr(), dd.data_ptr(), None, stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    gix.kernel_times()
    gix.search_work()
    gix.set_profiling(True)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gix.set_profiling(False)
    dev_ms = ev0.elapsed_time(ev1)
    kms, nsearch = gix.kernel_times()
    work = gix.search_work()
    ms_per_step = wall * 1e3 / args.steps
    qps = nq / (ms_per_step / 1e3)

    res_plain = ids.cpu().numpy()

    # the precision form: BF16 rows, FP32 re-rank of the best 2k
    gix.setSearchPrecision(1)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    bf_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    res_bf = ids.cpu().numpy()
    gix.setSearchPrecision(0)

    # method 1 (kgraph) at the same parameters, for reference
    gix.setSearchParams(s_epochs, s_pool, s_extend, s_trees, -1, 1)
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m1_steps = max(1, args.steps // 4)
    for _ in range(m1_steps):
        gix.setSearchParams(s_epochs, s_pool, s_extend, s_trees, -1, 1)
        step()
    torch.cuda.synchronize()
    m1_ms = (time.perf_counter() - t0) * 1e3 / m1_steps
    res_m1 = ids.cpu().numpy()
    gix.setSearchParams(s_epochs, s_pool, s_extend, s_trees, -1, 0)

    # recall@10 against exact kNN on the GPU (ties to the lower id)
    dptr, stride = gix.device_data()
    gt = torch.empty((nq, k), dtype=torch.int32, device=dev)
    exact_knn_device(dptr, n, d, stride, qd.data_ptr(), nq, d, k, False, gt.data_ptr(), None,
                     stream.cuda_stream)
    torch.cuda.synchronize()
    res = res_plain
    gtn = gt.cpu().numpy()

    def rec(r):
        return sum(len(set(a.tolist()) & set(b.tolist())) for a, b in zip(r, gtn)) / (nq * k)

    recall = rec(res)
    forms = {"bf16_rows": {"ms_per_step": round(bf_ms, 4), "qps": round(nq / (bf_ms / 1e3), 1),
                           "speedup_vs_plain": round(ms_per_step / bf_ms, 3), "recall_at_10": round(rec(res_bf), 5),
                           "identical_to_plain": bool(np.array_equal(res_bf, res))},
             "method1": {"ms_per_step": round(m1_ms, 4), "qps": round(nq / (m1_ms / 1e3), 1),
                         "recall_at_10": round(rec(res_m1), 5)}}

    # roofline of the search kernel: algorithmic bytes per launch (per query:
    # 4d per distance, 4 per seed id read, 4*K per expanded node)
    per_launch = [w / args.steps for w in work]
    alg_bytes = 4 * d * per_launch[0] + 4 * per_launch[1] + 4 * K * per_launch[2]
    kern_ms = kms[1] / max(1, nsearch)
    achieved = alg_bytes / (kern_ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "search_mixture1m_traffic.json")
    if os.path.exists(tpath) and n == 1_000_000 and nq == 10_000:
        with open(tpath) as fh:
            traffic = json.load(fh)["hbm_bytes_per_launch"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/search_mixture1m_traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE)"
                if traffic else None,
                "kernel": "k_search_nnexp", "kernel_ms": round(kern_ms, 4),
                "descend_ms": round(kms[0] / max(1, nsearch), 4),
                "bytes_per_launch": int(alg_bytes)}

    cpu = None
    if not args.no_cpu:
        nt = host_threads()
        oracle.set_threads(nt)
        oix = oracle.Index(base, trees=trees, trees_build=trees_build, ml=ml, epochs=epochs, check=check, L=L,
                           K=K, S=S)
        oix.set_trees(gix.trees())
        oix.set_graph(*gix.graph())
        m = min(args.cpu_queries, nq)
        t = time.perf_counter()
        cids, _, _ = oix.search(q[:m], k, s_epochs, s_pool, s_extend, s_trees, 0)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(m / cpu_s, 1), "unit": "queries/s", "cores": nt, "kind": "port",
               "sample": f"{m} of the {nq} queries, same index and parameters, OpenMP over queries",
               "matches_gpu": bool(np.array_equal(cids, res[:m]))}

    out = {
        "metric": METRIC,
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "mixture1m", "n": n, "d": d, "queries": nq, "k": k,
                   "build": {"trees": trees, "trees_build": trees_build, "ml": ml, "epochs": epochs,
                             "check": check, "L": L, "K": K, "S": S},
                   "search": {"trees": s_trees, "epochs": s_epochs, "pool": s_pool, "extend": s_extend},
                   "form": "plain fp32"},
        "recall_at_10": round(recall, 5),
        "forms": forms,
        "build_s": round(build_s, 3),
        "build_phase_ms": {"trees": round(phase_ms[0], 1), "merge_init": round(phase_ms[1], 1),
                           "join": round(phase_ms[2], 1), "update": round(phase_ms[3], 1)},
        "build_work": {"merge_distances": bc["merge"], "join_pairs": join_pairs},
        "search_work_per_query": {"distances": round(per_launch[0] / nq, 2),
                                  "seed_ids_read": round(per_launch[1] / nq, 2),
                                  "expanded": round(per_launch[2] / nq, 2)},
        "device_ms_per_step": round(dev_ms / args.steps, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "datagen_s": round(gen_s, 2),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
