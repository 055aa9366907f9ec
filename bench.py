"""Benchmark of the RapidBoW hot path on one MI355X.

Default workload (a "step"): the batched vocabulary build -- hierarchical
k-means, branch 16, depth 4 (65,536 words), 2,000,000 SIFT-like uint8[128]
descriptors, 10 Lloyd passes per node, k-means++ seeds, every node of a level
in the same launches.  `value` is the k-means assign throughput in tera-ops/s
(2*16*128 int8 multiply-adds per point per pass) of the I8-MFMA kernel, from
HIP events around the assign launches inside the timed builds.  Beside it:
the same build with the plain scalar kernel, descriptor lookup throughput,
and the oracle on the host cores as the CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "leaf lookups/s & k-means assign TOPS (MFMA vs plain); query p50/p99 latency"
I8_PEAK_TOPS = 5000.0     # MI355X dense int8 MFMA, 5.0 POPS
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_once(rb, d, n, kernel):
    voc = rb.build_vocabulary(d, n, 1000, 4, max_iters=10, seed_mode=1, kernel=kernel)
    return voc, voc.build_stats()


def cpu_baseline(rb, oracle, threads):
    """Oracle k-means build on a bounded sample (same shape, 1/10 of the points)."""
    n = 200_000
    desc = rb.synth_descriptors(n, seed=1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    t = time.perf_counter()
    res = oracle.build(desc, 1000, 4, max_iters=10, seed_mode=1)
    dt = time.perf_counter() - t
    # point-passes: node size x passes; node sizes from the parent occupancy
    sizes = np.concatenate([[n], res["occ"].reshape(-1)[: res["passes"].size - 1]])
    ops = float((sizes.astype(np.float64) * res["passes"]).sum()) * 2 * 16 * 128
    return {"value": ops / dt / 1e12, "unit": "TOPS", "cores": threads, "kind": "port",
            "sample": f"oracle hierarchical k-means b16 L4, {n} descriptors, 10 passes, k-means++ "
                      f"({dt:.1f} s wall)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is design-only (see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import rapidbow_amd as rb
    from rapidbow_amd import hip

    N = 2_000_000
    log(f"bench: generating {N} SIFT-like descriptors")
    desc = rb.synth_descriptors(N, seed=1)
    d = rb.DeviceArray.from_numpy(desc)

    log("bench: warmup builds")
    for _ in range(args.warmup):
        build_once(rb, d, N, 0)
    hip.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        voc, st = build_once(rb, d, N, 0)
        stats.append(st)
        log(f"bench: step {i}: build {st['total_ms']:.1f} ms, assign {st['assign_ms']:.2f} ms")
    hip.synchronize()
    wall = time.perf_counter() - t0
    assign_ms = float(np.median([s["assign_ms"] for s in stats]))
    ops = stats[-1]["assign_ops"]
    tops = ops / (assign_ms * 1e-3) / 1e12
    build_ms = float(np.median([s["total_ms"] for s in stats]))

    log("bench: plain-kernel build")
    _, stp = build_once(rb, d, N, 1)
    plain_tops = stp["assign_ops"] / (stp["assign_ms"] * 1e-3) / 1e12

    # lookup on the built tree (b16 L4 uint8), 10M queries resident in HBM
    NQ = 10_000_000
    log(f"bench: lookup {NQ} queries")
    q = rb.DeviceArray.from_numpy(rb.synth_descriptors(NQ, seed=2))
    out = rb.DeviceArray((NQ,), np.uint32)
    look = {}
    for name, k in (("mfma", rb.LOOKUP_MFMA), ("plain", rb.LOOKUP_PLAIN)):
        voc.lookup(q, NQ, out, k)
        hip.synchronize()
        e0, e1 = rb.Event(), rb.Event()
        e0.record()
        reps = 3
        for _ in range(reps):
            voc.lookup(q, NQ, out, k)
        e1.record()
        e1.synchronize()
        look[name] = NQ * reps / (e0.elapsed_ms(e1) * 1e-3)
        log(f"bench: lookup {name}: {look[name] / 1e9:.2f} G desc/s")

    line = {
        "metric": METRIC, "value": round(tops, 3), "unit": "TOPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int8", "data": "synthetic",
        "config": {"workload": "batched vocab build: b16 L4 (65,536 words), 2M SIFT-like uint8[128], "
                               "10 Lloyd passes/node, k-means++, one launch per level pass",
                   "nb_desc": N, "branch": 16, "levels": 4, "passes": 10},
        "roofline": {"bound": "mfma", "achieved": round(tops, 3), "peak": I8_PEAK_TOPS, "unit": "TFLOP/s",
                     "frac": round(tops / I8_PEAK_TOPS, 5), "traffic": None},
        "assign_tops_plain": round(plain_tops, 3),
        "build_ms_mfma": round(build_ms, 2), "build_ms_plain": round(stp["total_ms"], 2),
        "lookup_b16L4_u8_desc_per_s": {k: round(v, 1) for k, v in look.items()},
    }
    if not args.no_cpu:
        import oracle
        threads = min(os.cpu_count() or 1, 16)
        log(f"bench: CPU baseline on {threads} threads")
        line["cpu_baseline"] = cpu_baseline(rb, oracle, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
