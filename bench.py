#!/usr/bin/env python3
"""Benchmark: kNN-graph construction (GNN-Descent) + NSG pruning on one MI355X.

One step = GNN-Descent (K=96, 10 iterations) over N=1,000,000 x 128 vectors resident in HBM,
followed by CFS path/dist/1.0 pruning (NSG) with FINAL_DEGREE=32, m=64.  Data: a seeded
integer-valued clipped 1024-component Gaussian mixture generated on the device.  The JSON
line reports vectors/s over the timed steps, the roofline position of the dominant kernel
(timed live with HIP events on the library's stream) and the CPU oracle timed on a bounded
sample of the same workload.

    python bench.py [--gpus 1] [--steps 2] [--warmup 1] [--precision exact]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--k", type=int, default=96)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--fd", type=int, default=32)
    ap.add_argument("--topm", type=int, default=64)
    ap.add_argument("--ncomp", type=int, default=1024)
    ap.add_argument("--precision", default="exact")
    ap.add_argument("--seed", type=int, default=0x7A600002)
    ap.add_argument("--recall-queries", type=int, default=10000)
    ap.add_argument("--cpu-sample", type=int, default=30000,
                    help="vectors in the CPU-oracle baseline sample (0 skips it)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(clusters would be sharded per GPU; see the design notes).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import tagore_amd

    N, D, K = args.n, args.dim, args.k
    log(f"workload N={N} D={D} K={K} iters={args.iters} NSG FD={args.fd} m={args.topm} "
        f"precision={args.precision}")
    st = tagore_amd.State(N, D, 0)
    st.gen_data(0, args.ncomp, args.seed, 20.0)
    st.set_profiling(True)

    def step():
        st.knn_build(K, args.iters, args.seed, args.precision)
        st.prune("path", "dist", "dist", args.fd, args.topm, 1.0)

    for w in range(args.warmup):
        t = time.perf_counter()
        step()
        log(f"warmup {w} done in {time.perf_counter() - t:.3f}s")

    knn_ms, prune_ms = [], []
    sec_knn = {k: 0.0 for k in tagore_amd.KNN_SECTIONS}
    sec_prune = {k: 0.0 for k in tagore_amd.PRUNE_SECTIONS}
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        s = st.stats()
        knn_ms.append(s["knn_ms"])
        prune_ms.append(s["prune_ms"])
        for k, v in s["knn_kernel_ms"].items():
            sec_knn[k] += v
        for k, v in s["prune_kernel_ms"].items():
            sec_prune[k] += v
        log(f"step {i}: knn {s['knn_ms']:.1f} ms, prune {s['prune_ms']:.1f} ms")
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = N / (ms_per_step / 1e3)
    last = st.stats()

    rec = st.recall(min(args.recall_queries, N), 10, 0x5A11)
    log(f"recall@10 = {rec:.4f}")

    # ---- roofline of the pruning forward kernel (Collect + Filter), HBM bound
    pc = last["prune_counters"]
    fwd_ms = sec_prune["forward"] / args.steps
    evals = pc["collect"] + pc["filter"]
    alg_bytes = evals * D * 4 + pc["adj_rows"] * K * 4 + N * K * 8 + N * args.fd * 8
    achieved = alg_bytes / (fwd_ms / 1e3) / 1e9 if fwd_ms > 0 else None
    roof = {"bound": "hbm", "kernel": "prune_forward_path", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
            "algorithmic_bytes": alg_bytes, "kernel_ms": fwd_ms}

    # ---- CPU oracle on a bounded sample of the same workload
    cpu = None
    if args.cpu_sample > 0:
        import oracle

        n = min(args.cpu_sample, N)
        Xs = st.get_data(0, n)
        t = time.perf_counter()
        G, GD, _ = oracle.gnn_descent(Xs, K, args.iters, args.seed)
        oracle.prune(Xs, G, GD, args.fd, args.topm, 0, 0, 0, 1.0)
        dt = time.perf_counter() - t
        cpu = {"value": n / dt, "unit": "vectors/s", "cores": oracle.num_threads(),
               "kind": "port", "seconds": dt,
               "sample": f"first {n} vectors of the same generated set, same K/iters/NSG params"}
        log(f"cpu baseline: {n} vectors in {dt:.1f}s")
    st.close()

    out = {
        "metric": "kNN+prune vectors/s at recall@10>=0.95: learned/BF16 vs plain FP32, 1 GPU",
        "value": value, "unit": "vectors/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"GNN-Descent K={K} iters={args.iters} + NSG prune FD={args.fd} "
                               f"m={args.topm} over {N}x{D} integer GMM({args.ncomp})",
                   "N": N, "dim": D, "K": K, "iters": args.iters, "final_degree": args.fd,
                   "m": args.topm, "precision": args.precision, "init": "random"},
        "recall_at_10": rec,
        "device_ms": {"knn": sum(knn_ms) / len(knn_ms), "prune": sum(prune_ms) / len(prune_ms),
                      "knn_sections": {k: v / args.steps for k, v in sec_knn.items()},
                      "prune_sections": {k: v / args.steps for k, v in sec_prune.items()}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
