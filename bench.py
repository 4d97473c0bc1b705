"""Throughput benchmark of the fused batched pipeline on one MI355X.

Workload (default): 1,000,000 synthetic molecules (10-40 atoms, seed 0) x 1,024 queries of 3-16
nodes (70% cut from the data graphs), K = 6 refinement steps, count-all.  One step = one
bms2_pipeline_run over the resident batch: query signatures, data signatures, filter, 6
refinement levels, GMCR, join, one u64 count per (graph, query) pair and the GMCR bit rows, all
written to HBM.

Two legs run on the same resident batch: plain (the reference's identity matching order) and
learned (matching order planned once for the query batch by the order network, trained on seeds
1-20 and shipped in bms2_amd/order_net.npz; planning traces 16,384 data graphs and runs bf16
inference, timed separately as plan_ms).  `value` is the learned leg's screened (query, graph)
pairs per second with the CSR batch already resident in HBM; both legs must produce the same
counts.  `roofline` prices the learned leg's device span (screen + join kernels of every chunk,
timed with HIP events on the pipeline's stream) against HBM bandwidth with the canonical
algorithmic bytes; `cpu_baseline` times the oracle's fused CPU path (OpenMP) on a bounded prefix
of the same data graphs with all queries, on this host, and checks the GPU counts of that prefix
against it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "(query,graph) pairs/s, exact counts; learned+batched vs plain HIP, 1 GPU"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(data, queries) -> float:
    """Compulsory HBM bytes of one run: CSR row offsets + column ids + labels of every data and
    query graph read once (4 + 5 n + 8 m per graph, m undirected edges) plus one u64 count per
    screened pair written once."""
    def side(g):
        return 4 * g.num_graphs + 5 * g.total_nodes + 4 * g.total_edges  # 8 m = 4 * (2 m)
    return float(side(data) + side(queries) + 8 * data.num_graphs * queries.num_graphs)


def host_cores() -> dict:
    """Cores of this host: nproc, the affinity mask, OMP_NUM_THREADS, and the number the CPU
    baseline uses (OMP_NUM_THREADS when set, which is the share of the host given to this job,
    else every core in the affinity mask)."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    omp_n = int(omp) if omp and omp.isdigit() and int(omp) > 0 else None
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": nproc, "affinity": aff, "omp": omp_n, "usable": omp_n or aff, "model": model}


def load_traffic() -> dict | None:
    p = ROOT / "profiles" / "pipeline_molecules-1m_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text())
        except Exception:
            return None
    return None


def cs_bench(args) -> int:
    """BMS2 --use-cs on one cluster batch (a 5-node prototype + 63 members, cluster_queries seed
    2: a prototype common in the molecules, ~5 partial matches per graph) over
    the first `--data-graphs` (default 200,000) molecules of the workload, K refinement steps,
    count-all: the two CS steps (joinPartialCandidates2 + joinWithPartialMatches) timed against the
    regular staged join on the same domains and GMCR, with HIP events.  Parity: per-GMCR-entry
    counts of the first --cpu-sample-graphs (default 20,000) data graphs against the oracle, whose
    CS join on that prefix (one thread) is the CPU baseline."""
    import oracle
    from bms2_amd import api, graphs

    D0 = args.data_graphs or 200_000
    data, _ = graphs.generate(args.workload, seed=0, num_data_graphs=D0)
    queries = graphs.cluster_queries(data, 63, seed=2, proto_nodes=(5, 5))
    K, Q, D = args.refinement_steps, queries.num_graphs, data.num_graphs
    end = int(queries.graph_offsets[1] - queries.graph_offsets[0])
    log = lambda m: print(f"[cs] {m}", file=sys.stderr, flush=True)  # noqa: E731
    log(f"{D} molecules x cluster of {Q} queries (prototype of {end} nodes), K={K}")
    q = api.Queue(0)
    dq = api.DeviceBatchedCSRGraph(q, queries)
    dd = api.DeviceBatchedCSRGraph(q, data)
    sig = api.Signature(q, dd.total_nodes, dq.total_nodes)
    path = api.PathSignature(q, dd.total_nodes, dq.total_nodes)
    cyc = api.CycleSignature(q, dd.total_nodes, dq.total_nodes)
    cand = api.Candidates(q, dq.total_nodes, dd.total_nodes)
    for g, fn in ((dd, "data"), (dq, "query")):
        getattr(sig, f"generate_{fn}_signatures")(g)
        getattr(path, f"generate_{fn}_path_signatures")(g)
        getattr(cyc, f"generate_{fn}_cycle_signatures")(g)
    api.filter_candidates(q, dq, dd, sig, path, cyc, cand)
    for k in range(1, K + 1):
        sig.refine_data_signatures(dd, k)
        sig.refine_query_signatures(dq, k)
        api.refine_candidates(q, dq, dd, sig, cand)
    gm = api.GMCR(q)
    gm.generate_gmcr(dq, dd, cand)
    q.wait()
    T = gm.total
    total = api.DeviceBuffer(q, 8, zero=True)
    pairs = api.DeviceBuffer(q, 8 * max(T, 1), zero=True)
    reg_ms, cs1_ms, cs2_ms = [], [], []
    part, n = api.join_partial_candidates(q, dq, dd, cand, end)  # sizes the buffer
    for it in range(args.warmup + args.steps):
        total.zero()
        e = api.join_candidates(q, dq, dd, cand, gm, total, False, pairs, event=True)
        q.wait()
        reg_total = int(total.to_numpy(np.uint64)[0])
        if it >= args.warmup:
            reg_ms.append(e.elapsed_ms())
        total.zero()
        part, n, e1 = api.join_partial_candidates(q, dq, dd, cand, end, capacity=n, event=True)
        e2 = api.join_with_partial_matches(q, dq, dd, cand, gm, end, part, n, total, pairs, False,
                                           event=True)
        q.wait()
        cs_total = int(total.to_numpy(np.uint64)[0])
        if it >= args.warmup:
            cs1_ms.append(e1.elapsed_ms())
            cs2_ms.append(e2.elapsed_ms())
    gpu_counts = pairs.to_numpy(np.uint64, T)
    off, _ = gm.host()
    log(f"regular join {np.mean(reg_ms):.2f} ms, CS partial {np.mean(cs1_ms):.2f} ms + extend "
        f"{np.mean(cs2_ms):.2f} ms; {n} partial matches, {T} GMCR pairs, totals {reg_total} / "
        f"{cs_total}")
    # parity and CPU baseline on a prefix
    S = min(args.cpu_sample_graphs or 20_000, D)
    sub = data.subset(0, S)
    ref = oracle.staged(queries, sub, K, keep_domains=False)
    t0 = time.perf_counter()
    rp = oracle.join_partial(queries, sub, ref.final_domains, end)
    rt, rc = oracle.join_with_partial(queries, sub, ref.final_domains, ref.gmcr_offsets,
                                      ref.gmcr_indices, end, rp)
    cpu_s = time.perf_counter() - t0
    ok = bool(np.array_equal(gpu_counts[:int(off[S])], rc))
    log(f"oracle CS join on the first {S} graphs: {cpu_s:.2f}s, counts equal: {ok}")
    cs_ms = float(np.mean(cs1_ms) + np.mean(cs2_ms))
    out = {
        "metric": "CS join (query,graph) pairs/s, exact counts, cluster batch, 1 GPU",
        "value": round(Q * D / (cs_ms / 1e3), 1), "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(cs_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (project generator, seed 0; cluster_queries seed 2, 5-node prototype)",
        "config": {"workload": f"{args.workload}: first {D} molecules x cluster of {Q} queries "
                               f"(prototype {end} nodes), K={K}, count-all, BMS2 --use-cs",
                   "data_graphs": D, "queries": Q, "refinement_steps": K},
        "partial_matches": n, "gmcr_pairs": T, "total": cs_total,
        "partial_ms": round(float(np.mean(cs1_ms)), 3), "extend_ms": round(float(np.mean(cs2_ms)), 3),
        "regular_join_ms": round(float(np.mean(reg_ms)), 3), "regular_total": reg_total,
        "cpu_baseline": {"value": round(Q * S / cpu_s, 1), "unit": "pairs/s", "cores": 1,
                         "kind": "port", "sample": f"oracle CS join (both steps) on the first {S} "
                                                   f"graphs, domains and GMCR from the oracle"},
        "parity_on_cpu_sample": ok,
    }
    print(json.dumps(out))
    return 0 if ok else 1


def ingest_bench(args) -> int:
    """Upload of a data batch (createDeviceCSRGraph, graph.hpp:251-310) two ways over the C ABI:
    bms2_device_csr_create (the batch packed on the host, then validated on the host: endpoints,
    self-loops, symmetry) and bms2_device_csr_pack (the graphs' local arrays uploaded as they are,
    offsets added and validation done on the GPU).  Both produce the same device arrays, which are
    compared.  Timed on the host clock around each call, inputs already in host memory."""
    import ctypes as C
    from bms2_amd import _lib, api, graphs

    D0 = args.data_graphs or 1_000_000
    data, _ = graphs.generate(args.workload, seed=0, num_data_graphs=D0)
    nodes, ro_local, ci_local = data.local_arrays()
    nl = np.ascontiguousarray(data.node_labels)
    el = np.ascontiguousarray(data.edge_labels)
    q = api.Queue(0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    host = data.to_c()
    create_ms, pack_ms = [], []
    for it in range(args.warmup + args.steps):
        d1, d2 = _lib.DeviceCsr(), _lib.DeviceCsr()
        t0 = time.perf_counter()
        _lib.check(_lib.lib().bms2_device_csr_create(q.h, C.byref(host), C.byref(d1)))
        t1 = time.perf_counter()
        _lib.check(_lib.lib().bms2_device_csr_pack(q.h, data.num_graphs, p(nodes), p(ro_local),
                                                   p(ci_local), p(nl), p(el), data.total_nodes,
                                                   len(ci_local), C.byref(d2)))
        t2 = time.perf_counter()
        if it >= args.warmup:
            create_ms.append(1e3 * (t1 - t0))
            pack_ms.append(1e3 * (t2 - t1))
        if it == 0:
            same = True
            for f, n, dt in (("graph_offsets", data.num_graphs + 1, np.uint32),
                             ("row_offsets", data.total_nodes + 1, np.uint32),
                             ("column_indices", len(ci_local), np.uint32),
                             ("node_labels", data.total_nodes, np.uint8),
                             ("edge_labels", len(ci_local), np.uint8)):
                a, b = np.empty(n, dt), np.empty(n, dt)
                _lib.check(_lib.lib().bms2_memcpy_d2h(q.h, p(a), C.c_void_p(getattr(d1, f)), a.nbytes))
                _lib.check(_lib.lib().bms2_memcpy_d2h(q.h, p(b), C.c_void_p(getattr(d2, f)), b.nbytes))
                same = same and bool(np.array_equal(a, b))
        _lib.lib().bms2_device_csr_destroy(q.h, C.byref(d1))
        _lib.lib().bms2_device_csr_destroy(q.h, C.byref(d2))
    cm, pm = float(np.mean(create_ms)), float(np.mean(pack_ms))
    print(f"[ingest] {data.num_graphs} graphs, {data.total_nodes} nodes, {len(ci_local)} CSR "
          f"entries: host-packed create {cm:.1f} ms, GPU pack {pm:.1f} ms, arrays equal: {same}",
          file=sys.stderr, flush=True)
    out = {
        "metric": "data batch upload (createDeviceCSRGraph) graphs/s, GPU-side packing, 1 GPU",
        "value": round(data.num_graphs / (pm / 1e3), 1), "unit": "graphs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(pm, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (project generator, seed 0)",
        "config": {"workload": f"{args.workload}: {data.num_graphs} molecules, "
                               f"{data.total_nodes} nodes, {len(ci_local)} CSR entries"},
        "host_create_ms": round(cm, 3), "gpu_pack_ms": round(pm, 3),
        "speedup": round(cm / pm, 2), "arrays_equal": same,
    }
    print(json.dumps(out))
    return 0 if same else 1


def staged_sequence(q, dq, dd, K: int, find_first: bool = False):
    """One run of the stage sequence of BMS2.cpp:131-234 through the staged entry points, on fresh
    stage objects (a Signature keeps the node sets its refinement levels counted,
    signature.hpp:340-385).  Returns (per-stage device ms, GMCR, per-pair counts, total buffer,
    candidate matrix bytes); the caller waits on the queue."""
    from bms2_amd import api
    sig = api.Signature(q, dd.total_nodes, dq.total_nodes)
    path = api.PathSignature(q, dd.total_nodes, dq.total_nodes)
    cyc = api.CycleSignature(q, dd.total_nodes, dq.total_nodes)
    cand = api.Candidates(q, dq.total_nodes, dd.total_nodes)
    gm = api.GMCR(q)
    total = api.DeviceBuffer(q, 8, zero=True)
    names = ("signatures", "filter", "refine signatures", "refine candidates", "gmcr", "join")
    t = dict.fromkeys(names, 0.0)
    evs = [("signatures", sig.generate_data_signatures(dd, event=True)),
           ("signatures", path.generate_data_path_signatures(dd, event=True)),
           ("signatures", cyc.generate_data_cycle_signatures(dd, event=True)),
           ("signatures", sig.generate_query_signatures(dq, event=True)),
           ("signatures", path.generate_query_path_signatures(dq, event=True)),
           ("signatures", cyc.generate_query_cycle_signatures(dq, event=True)),
           ("filter", api.filter_candidates(q, dq, dd, sig, path, cyc, cand, event=True))]
    for k in range(1, K + 1):
        evs.append(("refine signatures", sig.refine_data_signatures(dd, k, event=True)))
        evs.append(("refine signatures", sig.refine_query_signatures(dq, k, event=True)))
        evs.append(("refine candidates", api.refine_candidates(q, dq, dd, sig, cand, event=True)))
    evs.append(("gmcr", gm.generate_gmcr(dq, dd, cand, event=True)))
    per_pair = api.DeviceBuffer(q, 8 * max(gm.total, 1), zero=True)
    evs.append(("join", api.join_candidates(q, dq, dd, cand, gm, total, find_first, per_pair,
                                            event=True)))
    q.wait()
    for n, e in evs:
        t[n] += e.elapsed_ms()
    return t, gm, per_pair, total, dq.total_nodes * cand.W * 8


def staged_bench(args) -> int:
    """The stage sequence of BMS2.cpp:131-234 through the staged entry points (the reference's API
    boundary, one C-ABI call per stage) on a shard of the workload, count-all, identity order.
    Per-stage device times from each call's events (utils::BatchedEvent); the filter and the
    refinement passes against HBM bandwidth (the filter writes the Qn x Dn candidate bit matrix, a
    refinement step reads and rewrites it); per-pair counts compared with the fused pipeline's on
    the same shard."""
    from bms2_amd import api, graphs

    D0 = args.data_graphs or 100_000
    data, queries = graphs.generate(args.workload, seed=0, num_data_graphs=D0)
    K = args.refinement_steps
    q = api.Queue(0)
    dq = api.DeviceBatchedCSRGraph(q, queries)
    dd = api.DeviceBatchedCSRGraph(q, data)
    names = ("signatures", "filter", "refine signatures", "refine candidates", "gmcr", "join")
    per = {n: [] for n in names}
    for it in range(args.warmup + args.steps):
        t, gm, per_pair, total, matrix = staged_sequence(q, dq, dd, K)
        if it >= args.warmup:
            for n in names:
                per[n].append(t[n])
    ms = {n: float(np.mean(v)) for n, v in per.items()}
    run_ms = sum(ms.values())
    # per-pair counts against the fused pipeline on the same shard
    off, idx = gm.host()
    counts = per_pair.to_numpy(np.uint64, gm.total)
    staged = np.zeros((data.num_graphs, queries.num_graphs), np.uint64)
    rows = np.repeat(np.arange(data.num_graphs), np.diff(off.astype(np.int64)))
    staged[rows, idx] = counts
    fused, _, _ = api.run_pipeline(q, queries, data, K)
    same = bool(np.array_equal(staged, fused))
    filt_gbs = matrix / (ms["filter"] / 1e3) / 1e9
    ref_gbs = 2 * K * matrix / (ms["refine candidates"] / 1e3) / 1e9 if K else 0.0
    pairs = data.num_graphs * queries.num_graphs
    print(f"[staged] {data.num_graphs} graphs x {queries.num_graphs} queries, K={K}: "
          + ", ".join(f"{n} {v:.2f} ms" for n, v in ms.items())
          + f"; candidate matrix {matrix / 1e9:.2f} GB, filter {filt_gbs:.0f} GB/s, refinement "
          f"{ref_gbs:.0f} GB/s; counts equal to the fused pipeline: {same}",
          file=sys.stderr, flush=True)
    out = {
        "metric": "(query,graph) pairs/s, exact counts, staged entry points (one call per stage), 1 GPU",
        "value": round(pairs / (run_ms / 1e3), 1), "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(run_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (project generator, seed 0)",
        "config": {"workload": f"{args.workload}: {data.num_graphs} molecules x "
                               f"{queries.num_graphs} queries, K={K}, count-all, identity order, "
                               "staged calls"},
        "stage_ms": {n: round(v, 3) for n, v in ms.items()},
        "candidate_matrix_bytes": matrix,
        "roofline": {"bound": "hbm", "kernel": "refine_candidates (K steps)",
                     "achieved": round(ref_gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(ref_gbs / 8000.0, 4), "traffic": None,
                     "filter_achieved": round(filt_gbs, 1)},
        "gmcr_pairs": gm.total, "counts_equal_fused": same,
    }
    print(json.dumps(out))
    return 0 if same else 1


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="molecules-1m",
                    help="molecules-1m (1M molecules x 1,024 queries of 3-16 nodes) or uneven-1m "
                         "(lognormal sizes 5-250 nodes, 10%% fused ring systems, queries of 3-32 "
                         "nodes)")
    ap.add_argument("--seed", type=int, default=0, help="workload seed (100: held-out evaluation)")
    ap.add_argument("--data-graphs", type=int, default=0, help="override the number of data graphs")
    ap.add_argument("--refinement-steps", type=int, default=6)
    ap.add_argument("--find-first", action="store_true")
    ap.add_argument("--cpu-sample-graphs", type=int, default=0,
                    help="data graphs the CPU baseline runs on (0: all)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: every core this process may run on (OMP_NUM_THREADS when the host "
                         "sets it, else the affinity mask)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="time the plain leg only")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--no-dedup", action="store_true",
                    help="screen every query (default: one query per isomorphism class, the "
                         "others copy its results)")
    ap.add_argument("--skip-plain", action="store_true",
                    help="time the learned leg only (profiling runs)")
    ap.add_argument("--no-staged", action="store_true",
                    help="skip the staged leg (the reference's stage calls at full scale)")
    ap.add_argument("--staged-steps", type=int, default=2)
    ap.add_argument("--plan-graphs", type=int, default=16384,
                    help="data graphs traced to plan the learned order")
    ap.add_argument("--ingest", action="store_true",
                    help="time the data batch upload instead (host vs GPU packing)")
    ap.add_argument("--staged", action="store_true",
                    help="time the staged entry points stage by stage instead (100,000 graphs)")
    ap.add_argument("--cs", action="store_true",
                    help="time the CS (cluster) join instead (BMS2 --use-cs), see cs_bench")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: the pipeline "
              "would shard data graphs across GPUs with no cross-GPU traffic during compute, and "
              "this build runs and reports one GPU's share", file=sys.stderr)
        return 2

    if args.staged:
        return staged_bench(args)
    if args.cs:
        return cs_bench(args)
    if args.ingest:
        return ingest_bench(args)

    from bms2_amd import api, graphs

    t0 = time.perf_counter()
    over = {"num_data_graphs": args.data_graphs} if args.data_graphs else {}
    data, queries = graphs.generate(args.workload, seed=args.seed, **over)
    print(f"[bench] generated {data.num_graphs} graphs ({data.total_nodes} nodes) x "
          f"{queries.num_graphs} queries in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    K = args.refinement_steps
    q = api.Queue(0)
    dq = api.DeviceBatchedCSRGraph(q, queries)
    dd = api.DeviceBatchedCSRGraph(q, data)
    pipe = api.Pipeline(q, dq, dd, refinement_steps=K, find_first=args.find_first)
    # isomorphism classes of the query batch (host, computed when the batch is bound; timed here
    # on the same function)
    t0 = time.perf_counter()
    reps = api.query_classes(queries)
    classes_ms = 1000 * (time.perf_counter() - t0)
    ndup = pipe.set_dedup(not args.no_dedup)
    dedup_info = {"on": not args.no_dedup, "queries_copied": int((reps != np.arange(len(reps))).sum()),
                  "classes": int((reps == np.arange(len(reps))).sum()),
                  "classes_ms": round(classes_ms, 2)}
    counts, valid, stats = pipe.alloc_outputs()
    q.wait()
    print("[bench] resident batch uploaded; warming up", file=sys.stderr, flush=True)
    D, Q = data.num_graphs, queries.num_graphs
    pairs = D * Q

    def leg(name, steps=None):
        steps = steps or args.steps
        for _ in range(args.warmup):
            pipe.run(counts, valid, stats)
        q.wait()
        evs = []
        t0 = time.perf_counter()
        for _ in range(steps):
            _, kev = pipe.run(counts, valid, stats, kernel_event=True)
            evs.append(kev)
        q.wait()
        t = time.perf_counter() - t0
        st = stats.to_numpy(np.uint64, 8)
        kms = [e.elapsed_ms() for e in evs]
        print(f"[bench] {name}: {steps} steps, {1000 * t / steps:.2f} ms/step, "
              f"{pairs * steps / t:.4g} pairs/s, device span {np.mean(kms):.2f} ms, DFS nodes "
              f"{int(st[3])}", file=sys.stderr, flush=True)
        return t, kms, st

    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = pipe.autotune(counts, valid, runs=2)
        tune["seconds"] = round(time.perf_counter() - t0, 2)
        print(f"[bench] autotuned launch shape: {tune}", file=sys.stderr, flush=True)
    order = "identity"
    S = min(args.cpu_sample_graphs or D, D)
    plain_sample = None
    nodedup = None
    if not (args.skip_plain and not args.no_learned):
        if ndup and not args.no_dedup:
            # the plain pipeline screening every query (no class copies), for reference
            pipe.set_dedup(False)
            nsteps = min(args.steps, 3)
            t_nd, kms_nd, st_nd = leg("plain (identity order), every query screened", nsteps)
            nd_sample = pipe.results(counts, valid, rows=S)
            pipe.set_dedup(True)
            nodedup = {"value": round(pairs * nsteps / t_nd, 1),
                       "ms_per_step": round(1000 * t_nd / nsteps, 3),
                       "kernel_ms_avg": round(float(np.mean(kms_nd)), 3),
                       "dfs_nodes": int(st_nd[3]), "steps": nsteps}
        t_plain, kms_plain, st_plain = leg("plain (identity order)")
        # plain counts of the leading graphs; the learned leg must reproduce them (and the totals)
        plain_sample = pipe.results(counts, valid, rows=S)
        if nodedup is not None:
            same_nd = (np.array_equal(nd_sample[0], plain_sample[0]) and
                       np.array_equal(nd_sample[1], plain_sample[1]) and
                       all(int(st_nd[k]) == int(st_plain[k]) for k in range(3)))
            del nd_sample
            nodedup["counts_equal"] = bool(same_nd)
            if not same_nd:
                print("[bench] runs with and without class copies differ", file=sys.stderr, flush=True)
                return 1
    plan_info = None
    if not args.no_learned:
        from bms2_amd import learned
        net = learned.OrderNet.load(q)
        pl = learned.plan(pipe, net, queries, sample_graphs=args.plan_graphs)
        plan_info = dict(plan_ms=round(1000 * pl.seconds, 1), sample_graphs=pl.sample_graphs,
                         queries_reordered=int(pl.learned.sum()))
        print(f"[bench] learned plan: {plan_info}", file=sys.stderr, flush=True)
        order = "learned"
        if not args.no_autotune:
            # the launch shape of the learned leg, tuned under its own order (each leg its best)
            t0 = time.perf_counter()
            tune_l = pipe.autotune(counts, valid, runs=2)
            tune_l["seconds"] = round(time.perf_counter() - t0, 2)
            tune = {"plain": tune, "learned": tune_l}
            print(f"[bench] autotuned launch shape (learned order): {tune_l}", file=sys.stderr,
                  flush=True)
    t_total, kernel_ms, st = leg(f"{order} order")
    if plan_info is not None and plain_sample is not None:
        lc, lv = pipe.results(counts, valid, rows=plain_sample[0].shape[0])
        same = (np.array_equal(lc, plain_sample[0]) and np.array_equal(lv, plain_sample[1])
                and all(int(st[k]) == int(st_plain[k]) for k in range(3)))
        if not same:
            print("[bench] learned and plain counts differ", file=sys.stderr, flush=True)
            return 1

    ms_per_step = 1000.0 * t_total / args.steps
    value = pairs * args.steps / t_total
    plain_value = pairs * args.steps / t_plain if plain_sample is not None else None

    # the reference-structured path: the stage sequence through the staged entry points on the
    # same resident batch (one warm-up run, then timed runs); per-pair counts of the leading
    # graphs and the total must equal the fused pipeline's
    staged_info = None
    if not args.no_staged:
        st_ms, st_stage = [], None
        for it in range(1 + args.staged_steps):
            t_st, gm_s, pp_s, tot_s, _ = staged_sequence(q, dq, dd, K, args.find_first)
            if it:
                st_ms.append(sum(t_st.values()))
                st_stage = t_st
        off_s, idx_s = gm_s.host()
        S2 = min(S, D, 20000)
        n2 = int(off_s[S2])
        dense = np.zeros((S2, Q), np.uint64)
        rows_s = np.repeat(np.arange(S2), np.diff(off_s[:S2 + 1].astype(np.int64)))
        dense[rows_s, idx_s[:n2]] = pp_s.to_numpy(np.uint64, n2)
        fused_rows, _ = pipe.results(counts, valid, rows=S2)
        st_total = int(tot_s.to_numpy(np.uint64)[0])
        st_same = bool(np.array_equal(dense, fused_rows) and
                       st_total == int(st[1] if args.find_first else st[2]))
        del gm_s, pp_s, tot_s, dense, fused_rows
        st_avg = float(np.mean(st_ms))
        staged_info = {"value": round(pairs / (st_avg / 1e3), 1), "ms_per_step": round(st_avg, 3),
                       "stage_ms": {k: round(v, 3) for k, v in st_stage.items()},
                       "steps": args.staged_steps, "counts_equal_fused": st_same}
        print(f"[bench] staged entry points (reference stage order): {st_avg:.2f} ms/run, "
              f"counts equal: {st_same}", file=sys.stderr, flush=True)
        if not st_same:
            return 1

    abytes = algorithmic_bytes(data, queries)
    k_avg_s = float(np.mean(kernel_ms)) / 1000.0
    achieved = abytes / k_avg_s / 1e9
    traffic = load_traffic()
    # the measured traffic belongs to one configuration: the default molecules-1m learned run
    measured_here = (traffic is not None and args.workload == "molecules-1m" and args.seed == 0
                     and K == 6 and not args.find_first and not args.data_graphs
                     and order == traffic.get("order", "learned"))
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic.get("bytes_per_launch") if measured_here else None,
        "traffic_source": (
            "profiles/pipeline_molecules-1m_traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE of one "
            f"{traffic.get('order', 'learned')}-order run, tree {traffic.get('tree', 'unrecorded')}, "
            f"{traffic.get('measured', 'date unrecorded')} (not this run)") if measured_here else
            "not measured for this configuration",
        "kernel": f"screen_kernel<{K}> + joins + class copies (device span per run, {order} order)",
        "kernel_ms_avg": round(float(np.mean(kernel_ms)), 3),
        "algorithmic_bytes_per_launch": abytes,
        "bytes_per_pair": round(abytes / pairs, 4),
    }

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        import oracle
        host = host_cores()
        threads = args.cpu_threads or host["usable"]
        sub = data.subset(0, S)
        print(f"[bench] CPU baseline: oracle fused path on the first {S} graphs x {Q} queries, "
              f"{threads} threads", file=sys.stderr, flush=True)
        t = time.perf_counter()
        rc, rv, rs = oracle.pipeline(queries, sub, K, args.find_first, threads=threads)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(S * Q / cpu_s, 1), "unit": "pairs/s", "cores": threads,
               "kind": "port", "seconds": round(cpu_s, 2),
               "sample": (f"all {D}" if S == D else f"first {S} of {D}") +
                         f" data graphs x all {Q} queries, same K and mode, from host CSR",
               "nproc": host["nproc"], "affinity_cpus": host["affinity"],
               "omp_num_threads": host["omp"], "cpu_model": host["model"],
               "build": "g++ -O3 -march=native -fopenmp, OpenMP dynamic schedule over data graphs"}
        gc, gv = pipe.results(counts, valid, rows=S)
        parity = bool(np.array_equal(gc, rc) and np.array_equal(gv, rv))
        if plain_sample is not None:
            parity = parity and bool(np.array_equal(plain_sample[0][:S], rc))
        del rc, rv, gc, gv
        print(f"[bench] CPU {cpu['value']:.4g} pairs/s in {cpu_s:.1f}s; GPU/oracle counts on the "
              f"sample equal: {parity}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": f"synthetic (project generator, seed {args.seed})",
        "config": {
            "workload": f"{args.workload}: {D} data graphs x {Q} queries, K={K}, "
                        f"{'first-match' if args.find_first else 'count-all'}, {order} matching "
                        "order, batched HIP pipeline",
            "data_graphs": D, "queries": Q, "refinement_steps": K, "seed": args.seed,
            "mode": "first-match" if args.find_first else "count-all", "order": order,
        },
        "matched_pairs_per_s": round(float(st[1]) * args.steps / t_total, 1),
        "gmcr_pairs": int(st[0]), "matched_pairs": int(st[1]), "embeddings": int(st[2]),
        "dfs_nodes": int(st[3]),
        "plain": None if plain_value is None else {
            "value": round(plain_value, 1), "ms_per_step": round(1000 * t_plain / args.steps, 3),
            "kernel_ms_avg": round(float(np.mean(kms_plain)), 3), "dfs_nodes": int(st_plain[3])},
        "learned_vs_plain": None if plain_value is None else round(value / plain_value, 4),
        # isomorphic queries of the batch: one per class is screened and joined, the others copy
        "dedup": dedup_info,
        "plain_every_query": nodedup,
        "learned_vs_plain_every_query": None if nodedup is None else round(value / nodedup["value"], 4),
        # one pass over the batch with the planning included (the plan is made once per query
        # batch and reused by every later run, which `value` measures)
        "learned_single_pass": None if plan_info is None else {
            "value": round(pairs / ((plan_info["plan_ms"] + classes_ms + ms_per_step) / 1e3), 1),
            "ms": round(plan_info["plan_ms"] + classes_ms + ms_per_step, 3),
            "vs_plain": None if plain_value is None else round(
                (pairs / ((plan_info["plan_ms"] + classes_ms + ms_per_step) / 1e3)) / plain_value, 4),
            "break_even_runs": None if plain_value is None or value <= plain_value else round(
                plan_info["plan_ms"] / (1000 * t_plain / args.steps - ms_per_step), 1)},
        "learned_plan": plan_info,
        "staged": staged_info,
        "learned_vs_staged": None if staged_info is None else round(value / staged_info["value"], 4),
        "autotune": tune,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": None if cpu is None else round(value / cpu["value"], 1),
        "parity_on_cpu_sample": parity,
    }
    print(json.dumps(line))
    return 0 if parity in (None, True) else 1


if __name__ == "__main__":
    sys.exit(main())
