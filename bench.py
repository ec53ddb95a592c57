#!/usr/bin/env python3
"""Benchmark: kNN-graph construction (GNN-Descent) + NSG pruning on one MI355X.

Workload (the 1M configuration): N=1,000,000 x 128 integer-valued clipped 1024-component
Gaussian mixture, generated on the device from a fixed seed; GNN-Descent K=96; CFS
path/dist/1.0 pruning (NSG) with FINAL_DEGREE=32, m=64.  One step = kNN build + prune on
data already resident in HBM.

Three paths are timed in the same run:
  * plain FP32: FP32 MFMA screen + canonical FP32 re-score, random init, 10 iterations
    (the configuration the reference runs: pybind.cu:101 with iter=10);
  * headline:   learned initial lists + BF16 MFMA screen + canonical FP32 re-score,
    2 iterations (--init, --precision, --iters); the encoder is trained for one epoch on a
    25k-vector sample (--epochs, --sample); its sample graph, training, encoding and
    bucketing are inside the timed step (on integer data the screen's products run on int8
    MFMA, exact);
  * ablation:   random init + BF16 screen at the fewest iterations that reach the recall
    target (--ablation-iters), to separate the two ingredients.
`value` is the headline's vectors/s over the timed steps; the other paths' figures and the
ratios are reported beside it, the plain path also at its own fewest iterations reaching the
recall target.  Every kNN graph is judged by FP64 recall@10 on a seeded sample; the NSG index
by a beam search with held-out queries.  The roofline entry is the pruning forward kernel,
the step's largest, timed live with HIP events on the library's stream: its compulsory HBM
bytes against the HBM peak, the measured HBM traffic, and its vector-ALU issue fraction from
the committed counter run.  The CPU oracle runs the same pipeline on a bounded sample.
`end_to_end` times, once and by wall clock, the same rows uploaded from host memory and
prepared, the headline build and prune, and the NSG index file written.

Other workloads (--workload): small (10k x 128 in full, the GPU against the host CPU, graphs
compared), unit (1M x 96 unit vectors + NSSG), wide (5M x 960 + DPG/CAGRA), clusters (the
large-index pipeline on one GPU's cluster share).

    python bench.py [--gpus 1] [--steps 2] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measure(st, args, precision, iters, init, label, graph_iters=False, graph=False):
    import tagore_amd

    def step():
        st.knn_build(args.k, iters, args.seed, precision, init=init, graph_iters=graph_iters,
                     graph=graph, sample=args.sample, epochs=args.epochs)
        st.prune("path", "dist", "dist", args.fd, args.topm, 1.0,
                 args.prune_rows if precision == "bf16" else "fp32")

    for w in range(args.warmup):
        t = time.perf_counter()
        step()
        log(f"{label}: warmup {w} done in {time.perf_counter() - t:.3f}s")
    sec_knn = {k: 0.0 for k in tagore_amd.KNN_SECTIONS}
    sec_prune = {k: 0.0 for k in tagore_amd.PRUNE_SECTIONS}
    knn_ms, prune_ms = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        s = st.stats()
        knn_ms.append(s["knn_ms"])
        prune_ms.append(s["prune_ms"])
        for k, v in s["knn_kernel_ms"].items():
            sec_knn[k] += v / args.steps
        for k, v in s["prune_kernel_ms"].items():
            sec_prune[k] += v / args.steps
        log(f"{label}: step {i}: knn {s['knn_ms']:.1f} ms, prune {s['prune_ms']:.1f} ms")
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    last = st.stats()
    rec = st.recall(min(args.recall_queries, st.N), 10, 0x5A11)
    # the stored FP32 distances of the same judged rows against FP64 (bound 1e-5 relative)
    derr = st.dist_check(min(args.recall_queries, st.N), 0x5A11)
    log(f"{label}: {ms:.1f} ms/step, recall@10 = {rec:.4f}, stored distances within "
        f"{derr:.2e} of FP64")
    return dict(ms_per_step=ms, vectors_per_s=st.N / (ms / 1e3), recall_at_10=rec,
                dist_max_rel_err=derr, dist_within_1e5=bool(derr <= 1e-5),
                precision=precision, iters=iters, init=init,
                learn_ms=last["learn_ms"], learn_loss=last["learn_loss"],
                knn_ms=sum(knn_ms) / len(knn_ms), prune_ms=sum(prune_ms) / len(prune_ms),
                knn_sections=sec_knn, prune_sections=sec_prune,
                knn_evals=last["knn_evals"], knn_rescored=last["knn_rescored"],
                prune_counters=last["prune_counters"])


def reachable_fraction(st):
    """Fraction of the vectors reachable from the pruned graph's entry point (host BFS)."""
    import numpy as np

    deg, ids, ep = st.get_pruned()
    N = deg.shape[0]
    cols = np.arange(ids.shape[1])[None, :]
    seen = np.zeros(N, bool)
    seen[ep] = True
    front = np.array([ep], np.int64)
    while front.size:
        nb = ids[front][cols < deg[front][:, None]].astype(np.int64)
        nb = np.unique(nb[~seen[nb]])
        seen[nb] = True
        front = nb
    return float(seen.mean())


def _profile_record(name, N, D, K, FD, TOPM):
    """A committed counter record under profiles/ for this workload, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    w = rec.get("workload", {})
    if (w.get("N"), w.get("dim"), w.get("K"), w.get("final_degree"), w.get("m")) != (
            N, D, K, FD, TOPM):
        return None
    return rec


def measured_traffic(N, D, K, FD, TOPM):
    """HBM bytes per k_forward launch from the committed rocprofv3 counter runs of the same
    workload (FETCH_SIZE and WRITE_SIZE, each scaled by its copy-kernel calibration), or
    (None, None) when no run of this workload is recorded."""
    rec = _profile_record("k_forward_traffic.json", N, D, K, FD, TOPM)
    if rec is None:
        return None, None
    # the calibration copy moved `scale` times the bytes the counter reported
    b = rec["fetch_kib"] * 1024 * rec["fetch_scale"] + rec["write_kib"] * 1024 * rec["write_scale"]
    return b, "profiles/k_forward_traffic.json"


def issue_bound(N, D, K, FD, TOPM, kernel_ms):
    """What bounds k_forward: its vector-ALU issue, from the committed SQ counter run of the
    same workload.  valu_issue_frac = VALU wave-instructions x 4 cycles (one wave64 VALU
    instruction occupies a SIMD for 4 cycles) / (SIMDs x cycles of the launch at the
    measured clock)."""
    rec = _profile_record("k_forward_counters.json", N, D, K, FD, TOPM)
    if rec is None:
        return None
    c = rec["counters"]
    simds = rec.get("simds", 1024)
    clk_ghz = rec.get("clock_ghz", 2.4)
    busy = simds * clk_ghz * 1e9 * kernel_ms / 1e3
    out = {"source": "profiles/k_forward_counters.json", "simds": simds, "clock_ghz": clk_ghz,
           "valu_insts": c["SQ_INSTS_VALU"], "valu_insts_per_node": c["SQ_INSTS_VALU"] / N,
           "valu_issue_frac": c["SQ_INSTS_VALU"] * 4 / busy}
    if "distance_valu_insts" in rec:
        out["distance_valu_share"] = rec["distance_valu_insts"] / c["SQ_INSTS_VALU"]
    return out


def weak_components(deg, ids):
    """Number of weakly connected components of a directed graph given as rows
    (deg [N], ids [N, width]) and the fraction of vertices in the largest one."""
    import numpy as np
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    N = deg.shape[0]
    valid = np.arange(ids.shape[1])[None, :] < deg[:, None]
    rows = np.broadcast_to(np.arange(N, dtype=np.int32)[:, None], ids.shape)[valid]
    cols = ids[valid].astype(np.int32)
    A = csr_matrix((np.ones(rows.size, np.int8), (rows, cols)), shape=(N, N))
    n, lab = connected_components(A, directed=True, connection="weak")
    return int(n), float(np.bincount(lab).max() / N)


def write_cluster_files(d, nclusters, median, D, seed):
    """The clustered large-index input at one GPU's share: cluster sizes (rows) log-normal
    around `median` (sigma 0.945, so the mean is 1.5625x the median), clipped to
    [median/8, 5 x median]; every point lies in two clusters (a primary and a secondary one,
    as the two-centre file of pybind.cu:311-355 lists them), so the points number half the
    rows.  Vectors: unit-norm, around their primary cluster's centre.  Writes the
    reference's files: cluster_<c> (u32 count, then the members' rows in ascending point
    order) and centres.bin (u32 first, u32 second centre per point)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    size = np.clip(np.exp(rng.normal(np.log(median), 0.945, nclusters)), median / 8, median * 5)
    prim = np.maximum((size / 2).astype(np.int64), 1)
    npts = int(prim.sum())
    cen = np.repeat(np.arange(nclusters, dtype=np.uint32), prim)
    cen = cen[rng.permutation(npts)]
    # secondary centre: drawn in proportion to the clusters' primaries, never the primary
    cen2 = rng.choice(nclusters, npts, p=prim / npts).astype(np.uint32)
    while True:
        same = np.nonzero(cen2 == cen)[0]
        if same.size == 0:
            break
        cen2[same] = rng.choice(nclusters, same.size, p=prim / npts).astype(np.uint32)
    centres = rng.standard_normal((nclusters, D)).astype(np.float32)
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    X = np.empty((npts, D), np.float32)
    for a in range(0, npts, 1 << 20):
        b = min(npts, a + (1 << 20))
        x = 3.0 * centres[cen[a:b]] + rng.standard_normal((b - a, D), dtype=np.float32)
        X[a:b] = x / np.linalg.norm(x, axis=1, keepdims=True)
    np.stack([cen, cen2], 1).tofile(os.path.join(d, "centres.bin"))
    rows = 0
    for c in range(nclusters):
        mem = np.nonzero((cen == c) | (cen2 == c))[0]
        with open(os.path.join(d, f"cluster_{c}"), "wb") as f:
            np.asarray([mem.size], np.uint32).tofile(f)
            X[mem].tofile(f)
        rows += mem.size
    return npts, rows, int(size.min()), int(size.max())


def bench_clusters(args):
    """One GPU's share of the clustered large-index pipeline through the drop-in
    largeIndex([0], ...) (pybind.cu:297-391): 64 clusters with log-normal sizes (median
    400k rows, ~40M rows in total, 96-dim unit vectors), per cluster GNN-Descent K=64 with
    8+8 iterations and Vamana (alpha 1.2, FD=32, m=64), then the streaming union merge into
    one index of max degree 22.  Batched: every cluster of a batch (batches sized from free
    device memory) in the same launches, with the kernel variant the autotuner picks on
    the first batch.  Per cluster: one launch sequence per cluster (the reference's loop,
    large_index.cu:173-277), same variant.  Both must write byte-identical index files."""
    import filecmp
    import shutil

    import tagore_amd

    d = args.cluster_dir
    os.makedirs(d, exist_ok=True)
    need = int(args.cluster_median * 1.6 * args.clusters * 96 * 4 * 1.6)
    free = shutil.disk_usage(d).free
    if free < need:
        print(f"bench.py: {d} has {free / 1e9:.1f} GB free, the cluster files need about "
              f"{need / 1e9:.1f} GB", file=sys.stderr)
        sys.exit(2)
    t = time.perf_counter()
    npts, rows, smin, smax = write_cluster_files(d, args.clusters, args.cluster_median, 96,
                                                 0x7A600004)
    gen_s = time.perf_counter() - t
    log(f"clusters: {args.clusters} files, {rows} rows ({npts} points), cluster sizes "
        f"{smin}..{smax}, written in {gen_s:.1f}s")
    K, FD, TOPM, md, B, thre = 64, 32, 64, 22, 4, 1.2
    common = dict(DIM=96, K=K, cluster_num=args.clusters, max_points=10 ** 8,
                  file_prefix=os.path.join(d, "cluster_"),
                  local_index_path_prefix=os.path.join(d, "local_"),
                  max_degree=md, buffer_size=B, all_node_num=npts,
                  centers_path=os.path.join(d, "centres.bin"), thre=thre, FINAL_DEGREE=FD,
                  TOPM=TOPM, iters=16, verbose=True)
    idx_b = os.path.join(d, "index_batched.bin")
    idx_c = os.path.join(d, "index_per_cluster.bin")
    log("largeIndex: batched launches (autotuned variant)")
    sb = tagore_amd.large_index(index_store_path=idx_b, precision="bf16", batch_rows=0,
                                tune=True, **common)
    log(f"batched: {sb}")
    log("largeIndex: one launch sequence per cluster")
    sc = tagore_amd.large_index(index_store_path=idx_c, precision=sb["precision"],
                                tile=sb["tile"], batch_rows=1, tune=False, **common)
    log(f"per cluster: {sc}")
    same = filecmp.cmp(idx_b, idx_c, shallow=False)
    for f in os.listdir(d):
        os.remove(os.path.join(d, f))
    out = {
        "metric": "clustered large-index rows/s (GNN-Descent + Vamana per cluster + merge), "
                  "batched launches vs one launch sequence per cluster, 1 GPU",
        "value": rows / (sb["device_ms"] / 1e3), "unit": "rows/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": sb["device_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": sb["precision"], "data": "synthetic",
        "config": {"workload": f"one GPU's cluster share: {args.clusters} log-normal clusters "
                               f"(median {args.cluster_median} rows, {smin}..{smax}), {rows} rows "
                               f"of {npts} 96-dim unit vectors; GNN-Descent K=64 8+8 "
                               f"iterations + Vamana alpha 1.2 FD=32 m=64, merge max_degree 22",
                   "clusters": args.clusters, "rows": rows, "points": npts, "dim": 96, "K": K},
        "value_definition": "rows over the summed device windows of the batched builds and "
                            "prunes; the host I/O and the merge are reported beside it",
        "batched": sb, "per_cluster": sc,
        "device_speedup_vs_per_cluster": sc["device_ms"] / sb["device_ms"],
        "build_speedup_vs_per_cluster": sc["build_ms"] / sb["build_ms"],
        "index_files_identical": same,
        "input_generation_s": gen_s,
    }
    print(json.dumps(out), flush=True)


def bench_iterations(args, N, D, K, kind, ncomp, seed, precision, prunes, sweep, epochs=0):
    """Iterations to recall@10 >= 0.95, learned vs random initial lists, then the prunes.
    Each build's device window is timed; recall is the FP64 judge on a seeded sample.
    precision "auto": the autotuner's fastest kNN variant (all give the same graph).
    epochs: training epochs of the learned lists' encoder (0: the library's default, 2)."""
    import tagore_amd

    log(f"workload N={N} D={D} K={K} data kind {kind} precision {precision}")
    st = tagore_amd.State(N, D, 0)
    st.gen_data(kind, ncomp, seed, 20.0)
    tile = 0
    if precision == "auto":
        (precision, tile), cands = st.autotune(K, seed)
        log(f"autotune: {cands} -> {precision} tile {tile}")
    st.set_profiling(True)
    rows = {}
    st.knn_build(K, 2, seed, precision, init="learned", tile=tile, epochs=epochs)  # warm-up
    nq = min(args.recall_queries, N, 2000 if D > 256 else 10000)
    for init, iters in sweep:
        st.knn_build(K, iters, seed, precision, init=init, tile=tile, epochs=epochs)
        s = st.stats()
        best = s["knn_ms"]
        rec = st.recall(nq, 10, 0x5A11)
        rows[f"{init}_{iters}"] = {"knn_ms": best, "recall_at_10": rec,
                                   "sections": s["knn_kernel_ms"]}
        log(f"{init} init, {iters} iterations: {best:.1f} ms, recall@10 {rec:.4f}")
    first = {}
    for init in ("learned", "random"):
        ok = [(int(k.split("_")[1]), v) for k, v in rows.items()
              if k.startswith(init) and v["recall_at_10"] >= 0.95]
        first[init] = min(ok, key=lambda x: x[0]) if ok else None
    # prunes on the learned graph at its first iteration count reaching the target
    it = first["learned"][0] if first["learned"] else max(i for n, i in sweep if n == "learned")
    st.knn_build(K, it, seed, precision, init="learned", tile=tile, epochs=epochs)
    knn_ms = st.stats()["knn_ms"]
    pr = {}
    for name, (kind_, fr, rr, th, pprecs) in prunes.items():
        for pp in pprecs:
            st.prune(kind_, fr, rr, 32, 64, th, pp)
            s = st.stats()
            pr[f"{name}_{pp}"] = {"prune_ms": s["prune_ms"], "sections": s["prune_kernel_ms"],
                                  "counters": s["prune_counters"]}
            log(f"prune {name} ({pp}): {s['prune_ms']:.1f} ms")
    st.close()
    # each prune at its fastest row precision (the results are identical)
    best = {}
    for k, v in pr.items():
        n = k.rsplit("_", 1)[0]
        best[n] = min(best.get(n, 1e30), v["prune_ms"])
    step_ms = knn_ms + sum(best.values())
    return rows, first, knn_ms, pr, step_ms, (precision, tile), it


def bench_unit(args):
    """1M x 96 unit-norm Gaussian mixture (512 components), K=64: learned (encoder trained
    5 epochs on the 50k sample's pairs) vs random initial lists (iterations to recall@10 >=
    0.95), then NSSG (2-hop / angle 60 degrees)."""
    N = args.n
    rows, first, knn_ms, pr, step_ms, var, it = bench_iterations(
        args, N, 96, 64, 1, 512, 0x7A600003, args.precision,
        {"NSSG": ("2-hop", "angle", "angle", 0.5, ["fp32"])},
        [("learned", 2), ("learned", 4), ("random", 2), ("random", 4), ("random", 6),
         ("random", 8)], epochs=5)
    out = {
        "metric": "kNN+NSSG vectors/s at recall@10>=0.95, learned/BF16, 1 GPU",
        "value": N / (step_ms / 1e3), "unit": "vectors/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"GNN-Descent K=64 + NSSG FD=32 angle 60 over {N}x96 unit "
                               "GMM(512), learned lists from a 5-epoch encoder",
                   "N": N, "dim": 96, "K": 64, "epochs": 5},
        "iterations_to_target": {k: (v[0] if v else None) for k, v in first.items()},
        "timed_iterations": it, "knn_variant": list(var),
        "builds": rows, "prunes": pr,
    }
    print(json.dumps(out), flush=True)


def bench_wide(args):
    """5M x 960 non-negative vectors (|Gaussian mixture| of 256 components scaled to [0,1]),
    K=64, FP8 e4m3 screen with FP32 re-score, learned vs random initial lists, then DPG and
    CFS 1-hop/rank (CAGRA) with FD=32."""
    N = args.n if args.n != 1_000_000 else 5_000_000
    rows, first, knn_ms, pr, step_ms, var, it = bench_iterations(
        args, N, 960, 64, 2, 256, 0x7A600005, "auto",
        {"DPG": ("dpg", "dist", "rank", 0.0, ["fp32", "bf16", "fp8"]),
         "CAGRA": ("cagra", "dist", "rank", 0.0, ["fp32"])},
        [("learned", 2), ("learned", 4), ("learned", 8), ("learned", 12), ("learned", 16),
         ("random", 4), ("random", 8), ("random", 16)])
    out = {
        "metric": "kNN+DPG+CAGRA vectors/s, learned init, autotuned screen, 1 GPU",
        "value": N / (step_ms / 1e3), "unit": "vectors/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": var[0], "data": "synthetic",
        "config": {"workload": f"GNN-Descent K=64 (autotuned screen) + DPG + CAGRA FD=32 over "
                               f"{N}x960 |GMM(256)| in [0,1]", "N": N, "dim": 960, "K": 64},
        "iterations_to_target": {k: (v[0] if v else None) for k, v in first.items()},
        "timed_iterations": it, "knn_variant": list(var),
        "builds": rows, "prunes": pr,
    }
    print(json.dumps(out), flush=True)


def bench_small(args):
    """The CPU-runnable configuration in full: 10k x 128 integer GMM(64), sigma 20, GNN-Descent
    K=32 for 10 iterations from random lists and NSG (FD=32, m=64), plain FP32, on the GPU and
    on the host CPU (the OpenMP oracle); the two pruned graphs are compared."""
    import numpy as np
    import oracle
    import tagore_amd

    N, D, K, FD, M, it, seed = 10000, 128, 32, 32, 64, 10, 0x7A600001
    st = tagore_amd.State(N, D, 0)
    st.gen_data(0, 64, seed, 20.0)
    X = st.get_data()

    def step():
        st.knn_build(K, it, seed, "fp32")
        st.prune("path", "dist", "dist", FD, M, 1.0, "fp32")

    for _ in range(max(1, args.warmup)):
        step()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    gpu_s = (time.perf_counter() - t) / args.steps
    deg, ids, ep = st.get_pruned()
    st.close()
    t = time.perf_counter()
    G, GD, _ = oracle.gnn_descent(X, K, it, seed)
    ref = oracle.prune(X, G, GD, FD, M, 0, 0, 0, 1.0)
    cpu_s = time.perf_counter() - t
    same = (np.array_equal(deg, ref["deg"]) and ep == ref["ep"] and
            all(np.array_equal(ids[j, :deg[j]], ref["ids"][j, :deg[j]]) for j in range(N)))
    log(f"small: GPU {gpu_s * 1e3:.1f} ms per step, CPU {cpu_s:.2f} s, identical {same}")
    out = {
        "metric": "kNN+NSG vectors/s, plain FP32, 1 GPU (10k x 128 in full, against the host CPU)",
        "value": N / gpu_s, "unit": "vectors/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": gpu_s * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"GNN-Descent K={K} {it} iterations (random init) + NSG FD={FD} "
                               f"m={M} over {N}x{D} integer GMM(64), sigma 20",
                   "N": N, "dim": D, "K": K, "iters": it},
        "cpu_baseline": {"value": N / cpu_s, "unit": "vectors/s", "cores": oracle.num_threads(),
                         "kind": "port", "seconds": cpu_s, "sample": "the whole configuration"},
        "identical_to_cpu": bool(same),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--k", type=int, default=96)
    ap.add_argument("--iters", type=int, default=2, help="headline iterations")
    ap.add_argument("--init", default="learned", choices=["learned", "random"],
                    help="headline initial lists")
    ap.add_argument("--sample", type=int, default=25000,
                    help="learned init: vectors in the encoder's training sample")
    ap.add_argument("--epochs", type=int, default=1,
                    help="learned init: training epochs of the encoder")
    ap.add_argument("--plain-iters", type=int, default=10)
    ap.add_argument("--ablation-iters", type=int, default=6,
                    help="random-init BF16 iterations of the ablation line (0 skips it)")
    ap.add_argument("--fd", type=int, default=32)
    ap.add_argument("--topm", type=int, default=64)
    ap.add_argument("--ncomp", type=int, default=1024)
    ap.add_argument("--precision", default="bf16", help="headline precision")
    ap.add_argument("--plain-precision", default="fp32")
    ap.add_argument("--prune-rows", default="fp32", choices=["fp32", "bf16"],
                    help="rows the headline prune reads (bf16 only when the data is exact in BF16)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--search-queries", type=int, default=10000,
                    help="held-out queries for the beam-search check of the NSG index (0: off)")
    ap.add_argument("--end-to-end", type=int, default=1,
                    help="also time upload + build + prune + index file once (wall clock)")
    ap.add_argument("--components", type=int, default=1,
                    help="count the weakly connected components of the kNN and NSG graphs")
    ap.add_argument("--phase-clocks", action="store_true",
                    help="also report the prune forward kernel's per-phase shader cycles")
    ap.add_argument("--seed", type=int, default=0x7A600002)
    ap.add_argument("--recall-queries", type=int, default=10000)
    ap.add_argument("--cpu-sample", type=int, default=100000,
                    help="vectors in the CPU-oracle baseline sample (0 skips it)")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "small", "clusters", "unit", "wide"],
                    help="small: 10k x 128 in full, GPU against the host CPU; "
                         "clusters: the large-index pipeline on one GPU's cluster share; "
                         "unit: 1M x 96 unit vectors + NSSG; wide: 5M x 960 FP8 + DPG/CAGRA")
    ap.add_argument("--clusters", type=int, default=64)
    ap.add_argument("--cluster-median", type=int, default=400000,
                    help="median cluster size in rows (clusters workload)")
    ap.add_argument("--cluster-dir", default=os.path.join(os.environ.get("TMPDIR", "/tmp"),
                                                          "tagore_clusters"),
                    help="scratch directory of the clusters workload's files")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the design notes describe how clusters would shard across GPUs).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "small":
        bench_small(args)
        return
    if args.workload == "clusters":
        bench_clusters(args)
        return
    if args.workload == "unit":
        bench_unit(args)
        return
    if args.workload == "wide":
        bench_wide(args)
        return

    import numpy as np

    import tagore_amd

    N, D, K = args.n, args.dim, args.k
    log(f"workload N={N} D={D} K={K} NSG FD={args.fd} m={args.topm}")
    st = tagore_amd.State(N, D, 0)
    st.gen_data(0, args.ncomp, args.seed, 20.0)
    st.set_profiling(True)

    plain = plain_target = ablation = None
    if not args.no_plain:
        plain = measure(st, args, args.plain_precision, args.plain_iters, "random", "plain")
        # the plain path at the fewest iterations that reach the recall target
        sweep = []
        for it in range(2, args.plain_iters + 1):
            st.knn_build(K, it, args.seed, args.plain_precision, init="random")
            r = st.recall(min(args.recall_queries, N), 10, 0x5A11)
            sweep.append({"iters": it, "knn_ms": st.stats()["knn_ms"], "recall_at_10": r})
            log(f"plain sweep: {it} iterations, recall@10 {r:.4f}")
            if r >= 0.95:
                break
        hit = [x for x in sweep if x["recall_at_10"] >= 0.95]
        if hit:
            plain_target = measure(st, args, args.plain_precision, hit[0]["iters"], "random",
                                   "plain at target")
            plain_target["sweep"] = sweep
        if args.ablation_iters > 0:
            ablation = measure(st, args, args.precision, args.ablation_iters, "random",
                               "ablation")
    head = measure(st, args, args.precision, args.iters, args.init, "headline")
    # the same step with the encoder's training loop and the GNN-Descent iterations replayed
    # from captured HIP graphs
    head_graph = measure(st, args, args.precision, args.iters, args.init,
                         "headline, graph-captured training and iterations", graph_iters=True,
                         graph=True)
    fdeg, _, _ = st.get_forward()
    fwd_edges = int(fdeg.astype(np.int64).sum())
    if args.phase_clocks:
        # one more prune of the last headline graph with the forward kernel's phase counters
        # (instrumented, so not part of any timing above)
        st.set_profiling(2)
        st.prune("path", "dist", "dist", args.fd, args.topm, 1.0,
                 args.prune_rows if args.precision == "bf16" else "fp32")
        pcs = st.stats()["prune_counters"]
        head["phase_cycles_per_node"] = {k: pcs[k] / N for k in
                                         ("collect_cycles", "filter_cycles", "pool_cycles",
                                          "eval_cycles", "beam_pre_cycles", "beam_sort_cycles",
                                          "beam_merge_cycles", "beam_hash_cycles",
                                          "pool_sort_cycles")}
        log(f"phase cycles per node: {head['phase_cycles_per_node']}")
        st.set_profiling(True)

    # index quality: beam search (get_neighbors) over the headline's NSG graph with held-out
    # queries, the next vectors N.. of the same generator
    search = None
    if args.search_queries > 0:
        nq = args.search_queries
        with tagore_amd.State(N + nq, D, 0) as gq:
            gq.gen_data(0, args.ncomp, args.seed, 20.0)
            Q = gq.get_data(N, nq)
        search = {"reachable_from_entry": reachable_fraction(st), "sweep": []}
        log(f"search: {search['reachable_from_entry']:.4f} of the vectors reachable from the "
            "entry point")
        for L in (10, 16, 32, 64, 128):
            st.search(Q[:min(nq, 1000)], k=10, L=L)  # warm-up
            ids, _, ss = st.search(Q, k=10, L=L)
            rec = st.search_recall(Q, ids, 10)
            row = {"L": L, "qps": nq / (ss["ms"] / 1e3), "ms": ss["ms"], "recall_at_10": rec,
                   "evals_per_query": ss["evals"] / nq,
                   "expansions_per_query": ss["expansions"] / nq}
            search["sweep"].append(row)
            log(f"search L={L}: {row['qps']:.0f} queries/s, recall@10 {rec:.4f}")

    # roofline of the pruning forward kernel (Collect + Filter + Store), the step's largest.
    # Algorithmic bytes are the compulsory ones: every input byte read once (the FP32 rows,
    # the kNN lists' ids and distances) and every output byte written once (degrees, forward
    # ids and distances, one 8-byte reverse push per forward edge).  The gathers the rule
    # performs (one row per canonical distance evaluation) re-read rows that L2 and the
    # Infinity Cache serve; they are reported beside it as row_reads_bytes.
    pc = head["prune_counters"]
    fwd_ms = head["prune_sections"]["forward"]
    evals = pc["collect"] + pc["filter"]
    alg_bytes = N * D * 4 + N * K * 8 + N * 4 + N * args.fd * 8 + fwd_edges * 8
    achieved = alg_bytes / (fwd_ms / 1e3) / 1e9 if fwd_ms > 0 else None
    traffic, traffic_src = measured_traffic(N, D, K, args.fd, args.topm)
    roof = {"bound": "hbm", "kernel": "k_forward (NSG path collect + filter + store)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
            "traffic_source": traffic_src,
            "hbm_frac_measured": (traffic / (fwd_ms / 1e3) / 1e9 / HBM_PEAK_GBS)
            if traffic and fwd_ms > 0 else None,
            "algorithmic_bytes": alg_bytes, "kernel_ms": fwd_ms, "evaluations": evals,
            "row_reads_bytes": evals * D * 4 + pc["adj_rows"] * K * 4,
            "issue": issue_bound(N, D, K, args.fd, args.topm, fwd_ms)}

    # graph structure: weakly connected components of the kNN graph and of the NSG index
    comps = None
    if args.components:
        Gk, _ = st.get_knn()
        kc = weak_components(np.full(N, K, np.uint32), Gk)
        del Gk
        d_, i_, _ = st.get_pruned()
        nc = weak_components(d_, i_)
        comps = {"knn_weak_components": kc[0], "knn_largest_component_frac": kc[1],
                 "nsg_weak_components": nc[0], "nsg_largest_component_frac": nc[1]}
        log(f"components: kNN {kc}, NSG {nc}")

    # the other prunes of the workload on the headline graph (device windows)
    other_prunes = {}
    for name, (kind_, fr, rr, th) in {"Vamana": ("path", "dist", "dist", 1.2),
                                      "CAGRA": ("cagra", "dist", "rank", 0.0)}.items():
        ms = []
        rows = args.prune_rows if args.precision == "bf16" else "fp32"
        for _ in range(max(1, args.steps)):
            st.prune(kind_, fr, rr, args.fd, args.topm, th, rows)
            ms.append(st.stats()["prune_ms"])
        s_ = st.stats()
        other_prunes[name] = {"prune_ms": min(ms), "step_ms": head["knn_ms"] + min(ms),
                              "sections": s_["prune_kernel_ms"],
                              "counters": {k: s_["prune_counters"][k] for k in
                                           ("collect", "filter", "reverse", "adj_rows")}}
        if args.phase_clocks and kind_ == "path":
            st.set_profiling(2)
            st.prune(kind_, fr, rr, args.fd, args.topm, th, rows)
            pcs = st.stats()["prune_counters"]
            other_prunes[name]["phase_cycles_per_node"] = {
                k: pcs[k] / N for k in ("collect_cycles", "filter_cycles", "pool_cycles",
                                        "eval_cycles", "beam_pre_cycles", "beam_sort_cycles",
                                        "beam_merge_cycles", "beam_hash_cycles",
                                        "pool_sort_cycles")}
            st.set_profiling(True)
        log(f"prune {name}: {min(ms):.1f} ms, {other_prunes[name]}")

    # The end-to-end window (the reference's load -> build -> prune -> store): the same rows
    # uploaded from host memory and prepared (u8 copy, norms), the headline build and prune,
    # the NSG index file written.  Wall clock around the library calls, once.
    e2e = None
    if args.end_to_end:
        Xh = st.get_data()
        path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"tagore_e2e_{os.getpid()}.nsg")
        t0 = time.perf_counter()
        st.set_data(Xh)
        t1 = time.perf_counter()
        st.knn_build(K, args.iters, args.seed, args.precision, init=args.init,
                     sample=args.sample, epochs=args.epochs)
        st.prune("path", "dist", "dist", args.fd, args.topm, 1.0,
                 args.prune_rows if args.precision == "bf16" else "fp32")
        t2 = time.perf_counter()
        st.store_index(path, "NSG")
        t3 = time.perf_counter()
        e2e = {"upload_prepare_ms": (t1 - t0) * 1e3, "build_prune_ms": (t2 - t1) * 1e3,
               "store_ms": (t3 - t2) * 1e3, "total_ms": (t3 - t0) * 1e3,
               "vectors_per_s": N / (t3 - t0), "index_bytes": os.path.getsize(path),
               "host_rows_bytes": int(Xh.nbytes)}
        os.remove(path)
        del Xh
        log(f"end to end: {e2e}")

    cpu = None
    if args.cpu_sample > 0:
        import oracle

        n = min(args.cpu_sample, N)
        Xs = st.get_data(0, n)
        t = time.perf_counter()
        G, GD, _ = oracle.gnn_descent(Xs, K, args.plain_iters, args.seed)
        oracle.prune(Xs, G, GD, args.fd, args.topm, 0, 0, 0, 1.0)
        dt = time.perf_counter() - t
        cpu = {"value": n / dt, "unit": "vectors/s", "cores": oracle.num_threads(),
               "kind": "port", "seconds": dt,
               "sample": f"first {n} vectors of the same generated set; GNN-Descent K={K} "
                         f"{args.plain_iters} iterations + NSG FD={args.fd} m={args.topm}"}
        log(f"cpu baseline: {n} vectors in {dt:.1f}s")
    st.close()

    out = {
        "metric": "kNN+prune vectors/s at recall@10>=0.95: learned/BF16 vs plain FP32, 1 GPU",
        "value": head["vectors_per_s"], "unit": "vectors/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"GNN-Descent K={K} + NSG prune FD={args.fd} m={args.topm} over "
                               f"{N}x{D} integer GMM({args.ncomp}), sigma 20",
                   "N": N, "dim": D, "K": K, "iters": args.iters, "final_degree": args.fd,
                   "m": args.topm, "precision": args.precision, "init": args.init,
                   "learned_sample": args.sample, "learned_epochs": args.epochs},
        "recall_at_10": head["recall_at_10"],
        "recall_gate": "recall@10 of the kNN graph against FP64 brute force on a seeded "
                       "sample of the vectors; the NSG index's own beam-search recall is "
                       "index_search_recall_at_10",
        "recall_target_met": bool(head["recall_at_10"] >= 0.95
                                  and (plain is None or plain["recall_at_10"] >= 0.95)),
        "index_search_recall_at_10": next((r["recall_at_10"] for r in search["sweep"]
                                           if r["L"] == 64), None) if search else None,
        "headline": head,
        "end_to_end": e2e,
        "headline_graph_iterations": {k: head_graph[k] for k in
                                      ("ms_per_step", "vectors_per_s", "knn_ms", "prune_ms",
                                       "recall_at_10")},
        "plain_fp32": plain,
        "plain_fp32_at_target": plain_target,
        "ablation_random_bf16": ablation,
        "speedup_vs_plain": (plain["ms_per_step"] / head["ms_per_step"]) if plain else None,
        "speedup_vs_plain_at_target": (plain_target["ms_per_step"] / head["ms_per_step"])
        if plain_target else None,
        "other_prunes": other_prunes,
        "graph_structure": comps,
        "roofline": roof,
        "cpu_baseline": cpu,
        "search_nsg": search,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
