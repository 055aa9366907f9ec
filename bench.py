"""Benchmark of the RapidBoW hot path on one MI355X.

Default workload (a "step"): the batched vocabulary build -- hierarchical
k-means, branch 16, depth 4 (65,536 words), 2,000,000 SIFT-like uint8[128]
descriptors, 10 Lloyd passes per node, k-means++ seeds, every node of a level
in the same launches.  `value` is the k-means assign throughput in tera-ops/s
(2*16*128 int8 multiply-adds per point per pass) of the I8-MFMA assign step
(assign kernel + reduction of its per-chunk partials, and the persistent
launches that run small tasks' whole Lloyd loops), from HIP events that those
launches' dispatches stamp at kernel start and end inside the timed builds
(build option timing=1).  The roofline object is the per-pass assign
kernel's: 129 algorithmic bytes per point-pass (128 in + 1 assignment) over its
average launch time, against HBM (its intensity, 32 int8 ops per byte, is far
below the I8 MFMA ridge of ~625 ops/byte).

Beside it on the same JSON line:
  * the same build with the plain scalar kernel (assign TOPS, "MFMA vs plain");
  * the build of a 1,048,576-word tree (branch 32, depth 4, float32
    centroids) on the 2M descriptors, k-means with the int8-split MFMA
    candidate filter vs plain FP32 chains, with the identity check of the trees;
  * leaf lookups/s on that tree for 10,000,000
    queries: int8-split MFMA candidate filter + FP32 re-score vs the plain FP32
    kernel, with the identity check of the leaf ids and the near-tie count;
  * leaf lookups/s on a 4,096-word uint8 tree (branch 16, depth 3) for 50,000
    queries (I8 MFMA vs plain);
  * batched retrieval: 100,000 images (lognormal descriptor counts, median
    2,000, clamped to [200, 8,000]) on the 1M-word tree, batches of 64
    re-noised query images, top-50: batch latency and queries/s as a separate
    lookup + batch query, and through the query pipeline as direct launches and
    as ONE captured HIP graph launch per batch; with the CPU leg, ranked-list
    identity and max score error vs the oracle;
  * serving: 1,000,000 images (about 2e9 postings) resident in HBM, Poisson
    arrivals at 200 / 1,000 / 5,000 QPS of queries with 300-6,000 descriptors,
    p50/p99 latency and queries/s of the continuous-batching server (one graph
    launch per batch, and with direct launches) vs one query at a time with the
    plain kernels;
  * the CPU baseline: the oracle, on the host cores, on the same build (median
    of 5 runs), whose tree is also compared with the timed GPU build
    (tree_equals_oracle); with it the error report: centroids vs the binary64
    means of their points, lookup distances vs binary64, scores vs the FP64
    ideal;
  * the queryAndAddDoc loop (query, then insert) on the 100k- and 1M-image
    files: per-call latency, which the delta segment keeps independent of the
    file's size, and the cost of folding the delta into the main CSR.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--skip-retrieval] [--skip-serving]
       python bench.py --profile-builds N | --profile-lookups N   (short runs for a profiler)
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "leaf lookups/s & k-means assign TOPS (MFMA vs plain); query p50/p99 latency"
I8_PEAK_TOPS = 5000.0     # MI355X dense int8 MFMA, 5.0 POPS
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, 8 TB/s
LOOKUP_BYTES = 132        # algorithmic bytes per descriptor looked up: 128 in + 4 out
ASSIGN_BYTES = 129        # algorithmic bytes per point and Lloyd pass: 128 in + 1 assignment
COPY_BYTES = 1 << 30      # device-to-device copy that calibrates the achievable HBM rate
CPU_RUNS = 5              # timed runs of the CPU baseline (median reported)

# seeds of the synthetic inputs
SEED_TRAIN, SEED_QUERY, SEED_IMAGES, SEED_ARRIVALS = 1, 2, 3, 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_lookup(rb, voc, q, n, out, kernel, reps, ties=None, rounds=3, warm_s=0.1):
    """Steady-state rate: one lookup (it also counts near ties), then back-to-back warm
    lookups for about `warm_s` seconds, then `rounds` rounds of `reps` lookups; the median
    round's rate.  The GPU's clocks take a few tens of ms of continuous load to settle:
    rounds right after one warm call, or after warm calls separated by host waits, read up
    to 10 % low."""
    voc.lookup(q, n, out, kernel, ties)
    rb.hip.synchronize()
    t0 = time.perf_counter()
    voc.lookup(q, n, out, kernel)
    rb.hip.synchronize()
    one = max(time.perf_counter() - t0, 1e-6)
    for _ in range(max(2, min(1000, int(warm_s / one)))):
        voc.lookup(q, n, out, kernel)
    rb.hip.synchronize()
    e0, e1 = rb.Event(), rb.Event()
    rates = []
    for _ in range(rounds):
        e0.record()
        for _ in range(reps):
            voc.lookup(q, n, out, kernel)
        e1.record()
        e1.synchronize()
        rates.append(n * reps / (e0.elapsed_ms(e1) * 1e-3))
    return float(np.median(rates))


def copy_rate(rb, reps=5):
    """Achievable HBM rate: a 1 GiB device-to-device copy (read + write bytes / time)."""
    src = rb.DeviceArray((COPY_BYTES,), np.uint8)
    dst = rb.DeviceArray((COPY_BYTES,), np.uint8)
    dst.copy_from(src)
    rb.hip.synchronize()
    e0, e1 = rb.Event(), rb.Event()
    e0.record()
    for _ in range(reps):
        dst.copy_from(src)
    e1.record()
    e1.synchronize()
    return 2.0 * COPY_BYTES * reps / (e0.elapsed_ms(e1) * 1e-3) / 1e9


def pmc_traffic():
    """HBM bytes per assign launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", "kmeans_assign_b16L4_2M_counters.md")
    try:
        for ln in open(path):
            if ln.startswith("traffic_bytes_per_launch:"):
                return float(ln.split(":")[1])
    except OSError:
        pass
    return None


def host_info():
    """The host the CPU legs ran on: logical CPUs, the ones this process may use, the model."""
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)), "model": model}


def cpu_threads():
    """Every host core this process may use (capped by OMP_NUM_THREADS when set)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(rb, oracle, desc, threads):
    """Oracle build of the same workload (b16 L4, 10 passes, k-means++) on the host cores:
    one warm-up run, then the median of CPU_RUNS timed runs."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    n = desc.shape[0]
    res = oracle.build(desc, 1000, 4, max_iters=10, seed_mode=1)
    times = []
    for _ in range(CPU_RUNS):
        t = time.perf_counter()
        oracle.build(desc, 1000, 4, max_iters=10, seed_mode=1)
        times.append(time.perf_counter() - t)
    dt = float(np.median(times))
    sizes = np.concatenate([[n], res["occ"].reshape(-1)[: res["passes"].size - 1]])
    ops = float((sizes.astype(np.float64) * res["passes"]).sum()) * 2 * 16 * 128
    return res, {"value": round(ops / dt / 1e12, 5), "unit": "TOPS", "cores": threads, "kind": "port",
                 "runs": CPU_RUNS, "run_s": [round(x, 3) for x in times], "host": host_info(),
                 "sample": f"oracle hierarchical k-means b16 L4 on the same {n} descriptors, 10 passes, "
                           f"k-means++, median of {CPU_RUNS} runs ({dt:.2f} s wall each; includes the host "
                           f"regroup)"}


def tree_vs_oracle_u8(voc, res):
    """Nodes, masks, IDF weights and pass counts of a uint8 b16 tree against the oracle's build."""
    blob = voc.serialize()
    n = res["nodes"].shape[0]
    nodes = np.frombuffer(blob, np.uint8, n * 2048, 4).reshape(n, 2048)
    masks = np.frombuffer(blob, np.uint16, n, 4 + n * 2048)
    w = np.frombuffer(blob, np.float32, n * 16, 4 + n * 2050).reshape(n, 16)
    out = {"nodes": bool(np.array_equal(nodes, res["nodes"])),
           "masks": bool(np.array_equal(masks, res["masks"].astype(np.uint16))),
           "weights": bool(np.array_equal(w.view(np.uint32), res["weights"].view(np.uint32))),
           "passes": bool(np.array_equal(voc.build_passes(), res["passes"])),
           "differing_nodes": int((nodes != res["nodes"]).any(1).sum())}
    out["all"] = out["nodes"] and out["masks"] and out["weights"] and out["passes"]
    return out


def tree_vs_oracle_f32(voc, res, b, L):
    out = {"differing_nodes": 0}
    for l in range(L):
        b0 = (b ** l - 1) // (b - 1)
        out["differing_nodes"] += int((voc.centroids(l) != res["nodes"][b0:b0 + b ** l]).any(axis=(1, 2)).sum())
    out["nodes"] = out["differing_nodes"] == 0
    out["weights"] = bool(np.array_equal(voc.weights().reshape(-1, b).view(np.uint32),
                                         res["weights"].view(np.uint32)))
    out["passes"] = bool(np.array_equal(voc.build_passes(), res["passes"]))
    out["all"] = out["nodes"] and out["weights"] and out["passes"]
    return out


def query_and_add_loop(db, leaf_sets, first_doc, K, odb=None, check=0):
    """queryAndAddDoc over new documents: per-call latency; the first `check` calls against the oracle."""
    lat, same = [], 0
    for i, lf in enumerate(leaf_sets):
        t = time.perf_counter()
        got = db.query_and_add(first_doc + i, lf, K)
        lat.append((time.perf_counter() - t) * 1e3)
        if i < check:
            od, _ = odb.query(lf, K)
            odb.add(first_doc + i, lf)
            same += int(np.array_equal(got, od))
    lat = np.array(lat)
    out = {"calls": len(leaf_sets), "postings_per_call_mean": round(float(np.mean([x.size for x in leaf_sets])), 1),
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "delta_postings_after": int(db.nb_delta_postings)}
    if check:
        out["ranked_equal_oracle"] = f"{same}/{check}"
    return out


def lognormal_counts(n, seed, median, sigma=0.5, lo=200, hi=8000):
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(np.exp(rng.normal(np.log(median), sigma, n))), lo, hi).astype(np.int64)


def image_descriptors(rb, syn, first_image, counts, desc_seed, noise_seed, noise_first, out):
    """Descriptors of images first_image.. (one centre subset of 64 per image) into `out` (device rows)."""
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    n = int(off[-1])
    doff = rb.DeviceArray.from_numpy(off)
    ids = syn.image_centre_ids(doff, len(counts), first_image, SEED_IMAGES, desc_seed, n)
    syn.descriptors(n, noise_seed, first=noise_first, dev_centre_ids=ids, out=out)
    return off


def bench_retrieval(rb, syn, v32, with_cpu, threads, graphs=True):
    """100k images on the 1M-word tree; batches of 64 re-noised query images, top-50."""
    from rapidbow_amd import hip
    NI, K, B, NB = 100_000, 50, 64, 16
    cnt = lognormal_counts(NI, SEED_IMAGES, 2000)
    off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint64)
    total = int(off[-1])
    log(f"bench: retrieval DB: {NI} images, {total} descriptors")
    db = rb.DataBase(v32)
    CH = 10_000
    maxn = max(int(off[min(i + CH, NI)] - off[i]) for i in range(0, NI, CH))
    dbuf = rb.DeviceArray((maxn, 128), np.uint8)
    lbuf = rb.DeviceArray((maxn,), np.uint32)
    host_leaf = np.empty(total, np.uint32) if with_cpu else None
    t = time.perf_counter()
    for i0 in range(0, NI, CH):
        i1 = min(i0 + CH, NI)
        o = image_descriptors(rb, syn, i0, cnt[i0:i1], desc_seed=31, noise_seed=SEED_IMAGES,
                              noise_first=int(off[i0]), out=dbuf)
        n = int(o[-1])
        v32.lookup(dbuf, n, lbuf, rb.LOOKUP_MFMA)
        db.add_docs_device(np.arange(i0, i1, dtype=np.uint32), lbuf, o)
        if with_cpu:
            host_leaf[int(off[i0]):int(off[i1])] = lbuf.numpy(n)
    hip.synchronize()
    build_s = time.perf_counter() - t
    del dbuf, lbuf
    # queries: re-noised copies of random DB images (same centre draws, fresh noise)
    rng = np.random.default_rng(SEED_QUERY)
    qimg = rng.choice(NI, B * NB, replace=False)
    qcnt = cnt[qimg]
    qoff = np.concatenate([[0], np.cumsum(qcnt)]).astype(np.uint64)
    qdesc = rb.DeviceArray((int(qoff[-1]), 128), np.uint8)
    for i, g in enumerate(qimg):
        image_descriptors(rb, syn, int(g), qcnt[i:i + 1], desc_seed=31, noise_seed=SEED_QUERY + 100,
                          noise_first=int(off[g]), out=qdesc.view(int(qoff[i]), int(qcnt[i])))
    qleaf = rb.DeviceArray((int(qoff[-1]),), np.uint32)
    docs = rb.DeviceArray((B, K), np.uint32)
    sc = rb.DeviceArray((B, K), np.float32)
    cn = rb.DeviceArray((B,), np.uint32)
    maxb = max(int(qoff[(bi + 1) * B] - qoff[bi * B]) for bi in range(NB))
    pipe = rb.QueryPipeline(v32, db, B, maxb, K)
    # scoring tile shape (walk depth) tuned on the first batch, kept for the database
    o = qoff[0:B + 1]
    v32.lookup(qdesc.view(int(o[0]), int(o[-1] - o[0])), int(o[-1] - o[0]), qleaf.view(int(o[0]), int(o[-1] - o[0])),
               rb.LOOKUP_MFMA)
    walk_steps, walk_ms = db.tune_scoring(qleaf, o, K)
    log(f"bench: scoring tile shape: walk steps {walk_steps}, batch ms per candidate {walk_ms}")
    hip.synchronize()
    lat, res_docs, res_sc = [], [], []
    lat_p = {False: [], True: []}
    same_pipe = True
    for rep in range(2):  # first pass warms (and captures the graphs)
        for bi in range(NB):
            q0, q1 = bi * B, (bi + 1) * B
            o = qoff[q0:q1 + 1]
            t0 = time.perf_counter()
            v32.lookup(qdesc.view(int(o[0]), int(o[-1] - o[0])), int(o[-1] - o[0]),
                       qleaf.view(int(o[0]), int(o[-1] - o[0])), rb.LOOKUP_MFMA)
            db.query_batch(qleaf, o, K, docs, sc, cn)  # returns when done
            dt = time.perf_counter() - t0
            c = cn.numpy()
            D, S = docs.numpy(), sc.numpy()
            if rep:
                lat.append(dt * 1e3)
                res_docs += [D[q, :c[q]].copy() for q in range(B)]
                res_sc += [S[q, :c[q]].copy() for q in range(B)]
            # the same batch through the pipeline: direct launches, then one graph launch
            for g in ((False, True) if graphs else (False,)):
                t0 = time.perf_counter()
                pipe.run(qdesc, o, docs, sc, cn, lookup_kernel=rb.LOOKUP_MFMA, use_graph=g)  # returns when done
                dt = time.perf_counter() - t0
                if rep:
                    lat_p[g].append(dt * 1e3)
                    same_pipe &= bool(np.array_equal(cn.numpy(), c) and np.array_equal(docs.numpy(), D)
                                      and np.array_equal(sc.numpy(), S))
    lat = np.array(lat)
    out = {"images": NI, "postings": int(db.nb_postings), "db_build_s": round(build_s, 2),
           "batch": B, "top_k": K, "batches": NB, "query_desc_mean": round(float(qcnt.mean()), 1),
           "batch_ms_p50": round(float(np.percentile(lat, 50)), 3),
           "batch_ms_p99": round(float(np.percentile(lat, 99)), 3),
           "queries_per_s": round(B * NB / (lat.sum() * 1e-3), 1),
           "pipeline": {"direct_batch_ms_p50": round(float(np.percentile(lat_p[False], 50)), 3),
                        "direct_batch_ms_p99": round(float(np.percentile(lat_p[False], 99)), 3),
                        "graph_batch_ms_p50": round(float(np.percentile(lat_p[True], 50)), 3) if graphs else None,
                        "graph_batch_ms_p99": round(float(np.percentile(lat_p[True], 99)), 3) if graphs else None,
                        "graph_queries_per_s": round(B * NB / (np.sum(lat_p[True]) * 1e-3), 1) if graphs else None,
                        "graphs_captured": int(pipe.graphs), "same_results": same_pipe},
           "self_first": float(np.mean([d.size > 0 and d[0] == g for d, g in zip(res_docs, qimg)])),
           "scoring_tile_tuned": {"walk_steps": walk_steps,
                                  "batch_ms_per_candidate": {str(c): round(m, 3) for c, m in walk_ms.items()}}}
    if with_cpu:
        import oracle
        log("bench: retrieval oracle DB (host)")
        w = v32.weights()
        odb = oracle.DataBase(32, 4, w)
        for i in range(NI):
            odb.add(i, host_leaf[int(off[i]):int(off[i + 1])])
            if i % 20_000 == 19_999:
                log(f"bench:   oracle DB {i + 1} images")
        ql = qleaf.numpy()
        sample = range(0, B * NB, 8)
        qs = [ql[int(qoff[i]):int(qoff[i + 1])] for i in sample]
        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            ref = list(ex.map(lambda q: odb.query(q, K), qs))
        cpu_s = time.perf_counter() - t
        with ThreadPoolExecutor(threads) as ex:
            ref64 = list(ex.map(lambda q: odb.query_fp64(q, K), qs))
        exact = swaps = 0
        err = 0.0
        for (od, os_), (fd, fs), i in zip(ref, ref64, sample):
            gd, gs = res_docs[i], res_sc[i]
            if np.array_equal(gd, od):
                exact += 1
            elif np.array_equal(np.sort(gd), np.sort(od)) and np.max(np.abs(gs - os_)) < 2e-6:
                swaps += 1  # neighbours with scores within the tolerance swapped
            f = {int(d): s for d, s in zip(fd, fs)}
            err = max(err, max((abs(float(s) - f[int(d)]) for d, s in zip(gd, gs) if int(d) in f), default=0.0))
        out["oracle_check"] = {"queries": len(qs), "ranked_exact": exact, "tie_swaps_only": swaps,
                               "max_abs_score_err_vs_fp64": err}
        out["cpu_queries_per_s"] = {"value": round(len(qs) / cpu_s, 2), "cores": threads, "kind": "port",
                                    "sample": f"oracle query of {len(qs)} of the leaf-id queries, one per thread"}
    # queryAndAddDoc (DataBase.cpp:192-197): the query images' leaves come back as new documents
    ql = qleaf.numpy()
    sets = [ql[int(qoff[i]):int(qoff[i + 1])].copy() for i in range(B * NB)][:200]
    out["query_and_add"] = query_and_add_loop(db, sets, NI, K, odb if with_cpu else None, 20 if with_cpu else 0)
    t = time.perf_counter()
    db.compact()
    out["query_and_add"]["compact_ms"] = round((time.perf_counter() - t) * 1e3, 2)
    log(f"bench: queryAndAddDoc at {NI} images {out['query_and_add']}")
    return out


def bench_serving(rb, syn, v32, rates, duration_s):
    """1M-image inverted file (postings synthesised from leaf clouds); Poisson replay per rate."""
    from rapidbow_amd import hip
    NI, K, CLOUD = 1_000_000, 50, 1024
    # leaf clouds: the looked-up leaves of CLOUD samples around each synthetic centre
    ids = rb.DeviceArray.from_numpy(np.repeat(np.arange(syn.nb_centres, dtype=np.uint32), CLOUD))
    cd = syn.descriptors(syn.nb_centres * CLOUD, SEED_IMAGES + 50, dev_centre_ids=ids)
    cloud = rb.DeviceArray((syn.nb_centres * CLOUD,), np.uint32)
    v32.lookup(cd, syn.nb_centres * CLOUD, cloud, rb.LOOKUP_MFMA)
    del cd, ids
    cnt = lognormal_counts(NI, SEED_IMAGES + 1, 1765)  # mean ~2,000 postings per image
    off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint64)
    total = int(off[-1])
    log(f"bench: serving DB: {NI} images, {total} postings")
    t = time.perf_counter()
    leaf = rb.DeviceArray((total,), np.uint32)
    doff = rb.DeviceArray.from_numpy(off)
    syn.image_postings(doff, NI, 0, SEED_IMAGES, SEED_IMAGES + 2, cloud, CLOUD, total, out=leaf)
    db = rb.DataBase(v32)
    db.add_docs_device(np.arange(NI, dtype=np.uint32), leaf, off)
    hip.synchronize()
    del leaf, doff
    build_s = time.perf_counter() - t
    # query pool: 300-6000 descriptors (uniform), re-observations of random DB images
    rng = np.random.default_rng(SEED_QUERY + 7)
    NQ = max(1000, int(max(rates) * duration_s))
    qimg = rng.integers(0, NI, NQ)
    qn = rng.integers(300, 6001, NQ)
    qoff = np.concatenate([[0], np.cumsum(qn)]).astype(np.uint64)
    qd = rb.DeviceArray((int(qoff[-1]), 128), np.uint8)
    for i in range(NQ):
        v = qd.view(int(qoff[i]), int(qn[i]))
        o = np.array([0, qn[i]], np.uint64)
        doff = rb.DeviceArray.from_numpy(o)
        cid = syn.image_centre_ids(doff, 1, int(qimg[i]), SEED_IMAGES, 77, int(qn[i]))
        syn.descriptors(int(qn[i]), SEED_QUERY + 7, first=int(qoff[i]), dev_centre_ids=cid, out=v)
    hip.synchronize()
    qhost = qd.numpy()
    del qd
    out = {"images": NI, "postings": int(db.nb_postings), "db_build_s": round(build_s, 1), "top_k": K,
           "query_desc": "uniform 300-6000", "rates": []}

    def run(srv, n, qps):
        o = qoff[:n + 1]
        r = srv.replay(qhost[: int(o[-1])], o, qps, seed=SEED_ARRIVALS)
        lat = r["latency_ms"]
        return r, {"queries": n, "achieved_qps": round(n / r["elapsed_s"], 1),
                   "p50_ms": round(float(np.percentile(lat, 50)), 3),
                   "p99_ms": round(float(np.percentile(lat, 99)), 3),
                   "mean_batch": round(float(r["batch_sizes"].mean()), 2)}

    batched = rb.Server(v32, db, max_batch=64, nb_streams=3, mem_budget=8 << 30, max_results=K,
                        lookup_kernel=rb.LOOKUP_AUTO)
    direct = rb.Server(v32, db, max_batch=64, nb_streams=3, mem_budget=8 << 30, max_results=K,
                       lookup_kernel=rb.LOOKUP_AUTO, use_graphs=False)
    single = rb.Server(v32, db, max_batch=1, nb_streams=1, mem_budget=8 << 30, max_results=K,
                       lookup_kernel=rb.LOOKUP_PLAIN)
    run(batched, 64, 1000.0)  # warm them
    run(direct, 64, 1000.0)
    _, w1 = run(single, 16, 50.0)
    # one-at-a-time service time (arrivals 20 ms apart), to bound the overloaded baseline runs to ~10 s each
    svc_ms = max(w1["p50_ms"], 0.05)
    out["one_at_a_time_service_ms"] = svc_ms
    for qps in rates:
        n = min(NQ, max(1000, int(qps * duration_s)))
        log(f"bench: serving {qps} QPS, {n} queries")
        rb_, m = run(batched, n, qps)
        _, m_direct = run(direct, n, qps)
        nb = min(n, max(200, int(10e3 / svc_ms)))
        rs, s = run(single, nb, qps)
        same = 0
        for i in range(nb):
            c = int(rs["counts"][i])
            same += int(c == int(rb_["counts"][i]) and np.array_equal(rs["docs"][i, :c], rb_["docs"][i, :c]))
        out["rates"].append({"offered_qps": qps, "batched": m, "batched_without_graphs": m_direct,
                             "one_at_a_time_plain": s,
                             "same_ranked_ids": f"{same}/{nb}"})
        log(f"bench:   batched p50 {m['p50_ms']} ms p99 {m['p99_ms']} ms {m['achieved_qps']} q/s "
            f"(no graphs p50 {m_direct['p50_ms']} p99 {m_direct['p99_ms']}) | "
            f"single p50 {s['p50_ms']} p99 {s['p99_ms']} {s['achieved_qps']} q/s")
    batched.close()
    direct.close()
    single.close()
    # queryAndAddDoc on the 1M-image file: new documents are re-observations of pool queries
    lk = rb.DeviceArray((int(qoff[200]),), np.uint32)
    qd = rb.DeviceArray.from_numpy(qhost[: int(qoff[200])])
    v32.lookup(qd, int(qoff[200]), lk, rb.LOOKUP_MFMA)
    hl = lk.numpy()
    sets = [hl[int(qoff[i]):int(qoff[i + 1])].copy() for i in range(200)]
    out["query_and_add"] = query_and_add_loop(db, sets, NI, K)
    t = time.perf_counter()
    db.compact()
    out["query_and_add"]["compact_ms"] = round((time.perf_counter() - t) * 1e3, 2)
    log(f"bench: queryAndAddDoc at {NI} images {out['query_and_add']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-retrieval", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--serving-seconds", type=float, default=1.5)
    ap.add_argument("--profile-builds", type=int, default=0,
                    help="run this many b16 L4 builds without per-launch timing (after the warmup), "
                         "print their times and stop: a short command for a profiler")
    ap.add_argument("--no-graphs", action="store_true",
                    help="retrieval through the query pipeline with direct launches only (for counter profiling)")
    ap.add_argument("--profile-lookups", type=int, default=0,
                    help="build the b32 L4 float tree, run three warm and then this many 10M-query MFMA "
                         "lookups, print their times and stop: a short command for a profiler")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this library runs on one GPU only; sharding over several GPUs is not implemented",
              file=sys.stderr)
        sys.exit(2)

    import rapidbow_amd as rb
    from rapidbow_amd import hip

    syn = rb.DeviceSynth()
    N = 2_000_000
    log(f"bench: generating {N} SIFT-like descriptors in HBM")
    d = syn.descriptors(N, seed=SEED_TRAIN)
    hip.synchronize()
    if args.profile_lookups:
        v32 = rb.build_vocabulary(d, N, 1000, 4, branch_factor=32, max_iters=10, seed_mode=1, centroid_type=1,
                                  kernel=0)
        nq = 10_000_000
        q = syn.descriptors(nq, seed=SEED_QUERY)
        out = rb.DeviceArray((nq,), np.uint32)
        for _ in range(3):
            v32.lookup(q, nq, out, rb.LOOKUP_MFMA)
        e0, e1 = rb.Event(), rb.Event()
        for i in range(args.profile_lookups):
            e0.record()
            v32.lookup(q, nq, out, rb.LOOKUP_MFMA)
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_ms(e1)
            log(f"bench: lookup {i}: {ms:.3f} ms, {nq / (ms * 1e-3) / 1e9:.3f} G desc/s")
        return

    # ---------------- batched vocabulary build (the timed steps) ----------------
    log("bench: warmup builds")
    for _ in range(args.warmup):
        rb.build_vocabulary(d, N, 1000, 4, max_iters=10, seed_mode=1, kernel=0, timing=True)
    hip.synchronize()
    if args.profile_builds:
        for i in range(args.profile_builds):
            t = time.perf_counter()
            v = rb.build_vocabulary(d, N, 1000, 4, max_iters=10, seed_mode=1, kernel=0)
            hip.synchronize()
            log(f"bench: build {i}: {(time.perf_counter() - t) * 1e3:.2f} ms wall, "
                f"{v.build_stats()['total_ms']:.2f} ms library")
            del v
        return
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        voc = rb.build_vocabulary(d, N, 1000, 4, max_iters=10, seed_mode=1, kernel=0, timing=True)
        st = voc.build_stats()
        stats.append(st)
        log(f"bench: step {i}: build {st['total_ms']:.1f} ms, assign {st['assign_ms']:.2f} ms")
    hip.synchronize()
    wall = time.perf_counter() - t0
    # the assign step = the assign kernel + the reduction of its chunk partials
    step_ms = float(np.median([s["assign_ms"] + s["reduce_ms"] for s in stats]))
    ops = stats[-1]["assign_ops"]
    tops = ops / (step_ms * 1e-3) / 1e12
    log(f"bench: assign step {step_ms:.3f} ms, {tops:.1f} TOPS")
    # roofline of the dominant kernel (the assign kernel alone): algorithmic bytes / its time
    st = min(stats, key=lambda x: x["stream_ms"])
    launch_us = st["stream_ms"] * 1e3 / st["assign_launches"]
    bytes_per_launch = st["stream_point_passes"] * ASSIGN_BYTES / st["assign_launches"]
    assign_gbs = bytes_per_launch / (launch_us * 1e-6) / 1e9
    copy_gbs = copy_rate(rb)
    log(f"bench: assign kernel {launch_us:.1f} us/launch, {assign_gbs:.0f} GB/s algorithmic; copy {copy_gbs:.0f} GB/s")
    # the same build without per-launch events (the library's default): its wall time
    prod_ms = []
    for _ in range(max(args.steps, 1)):
        vq = rb.build_vocabulary(d, N, 1000, 4, max_iters=10, seed_mode=1, kernel=0)
        prod_ms.append(vq.build_stats()["total_ms"])
        same_prod = vq.serialize() == voc.serialize()
        del vq
    log(f"bench: build without per-launch timing {np.median(prod_ms):.2f} ms, identical={same_prod}")
    # the same MFMA build with every node streamed (no persistent on-chip Lloyd loops)
    vstream = rb.build_vocabulary(d, N, 1000, 4, max_iters=10, seed_mode=1, kernel=0, persist_max_tiles=0,
                                  timing=True)
    sts = vstream.build_stats()
    stream_tops = sts["assign_ops"] / ((sts["assign_ms"] + sts["reduce_ms"]) * 1e-3) / 1e12
    same_stream = vstream.serialize() == voc.serialize()
    del vstream
    log(f"bench: stream-only build {sts['total_ms']:.1f} ms, assign {sts['assign_ms']:.2f} ms, "
        f"{stream_tops:.1f} TOPS, identical={same_stream}")
    log("bench: plain-kernel build")
    vplain = rb.build_vocabulary(d, N, 1000, 4, max_iters=10, seed_mode=1, kernel=1, timing=True)
    stp = vplain.build_stats()
    plain_tops = stp["assign_ops"] / ((stp["assign_ms"] + stp["reduce_ms"]) * 1e-3) / 1e12
    same_tree = vplain.serialize() == voc.serialize()

    # ---------------- 1M-word float32 tree, 10M lookups ----------------
    log("bench: building the b32 L4 float32 tree (1,048,576 words) on the 2M descriptors")
    # one warm float build first: the timed one then carries no first-use costs (workspaces, code objects)
    rb.build_vocabulary(d, N, 1000, 4, branch_factor=32, max_iters=10, seed_mode=1, centroid_type=1, kernel=0)
    v32 = rb.build_vocabulary(d, N, 1000, 4, branch_factor=32, max_iters=10, seed_mode=1, centroid_type=1, kernel=0)
    st32 = v32.build_stats()
    v32p = rb.build_vocabulary(d, N, 1000, 4, branch_factor=32, max_iters=10, seed_mode=1, centroid_type=1, kernel=1)
    st32p = v32p.build_stats()
    same32 = v32p.serialize() == v32.serialize()
    del v32p
    log(f"bench: b32 L4 float build: int8-split MFMA filter {st32['total_ms']:.1f} ms, plain {st32p['total_ms']:.1f} ms, "
        f"identical={same32}")
    NQ = 10_000_000
    q = syn.descriptors(NQ, seed=SEED_QUERY)
    out_m = rb.DeviceArray((NQ,), np.uint32)
    out_p = rb.DeviceArray((NQ,), np.uint32)
    ties = rb.DeviceArray.zeros((1,), np.uint32)
    log("bench: lookups b32 L4 float32")
    r_m = timed_lookup(rb, v32, q, NQ, out_m, rb.LOOKUP_MFMA, 5, ties)
    r_p = timed_lookup(rb, v32, q, NQ, out_p, rb.LOOKUP_PLAIN, 1)
    lm, lp = out_m.numpy(), out_p.numpy()
    identical = bool(np.array_equal(lm, lp))
    near_ties = int(ties.numpy()[0])
    log(f"bench: b32 lookups mfma {r_m / 1e9:.3f} G/s plain {r_p / 1e9:.3f} G/s identical={identical} "
        f"near_ties={near_ties}")
    qh = q.view(0, 200_000).numpy() if not args.no_cpu else None
    del q, out_p

    # ---------------- 4,096-word uint8 tree, 50k lookups ----------------
    d1 = syn.descriptors(200_000, seed=SEED_TRAIN + 10)
    v1 = rb.build_vocabulary(d1, 200_000, 1000, 3, max_iters=8, seed_mode=1)
    q1 = syn.descriptors(50_000, seed=SEED_QUERY + 10)
    o1 = rb.DeviceArray((50_000,), np.uint32)
    o1p = rb.DeviceArray((50_000,), np.uint32)
    r1_m = timed_lookup(rb, v1, q1, 50_000, o1, rb.LOOKUP_MFMA, 20)
    r1_p = timed_lookup(rb, v1, q1, 50_000, o1p, rb.LOOKUP_PLAIN, 20)
    id1 = bool(np.array_equal(o1.numpy(), o1p.numpy()))
    # autotuned AUTO crossovers (plain below, bucketed MFMA from there on)
    q_tune = syn.descriptors(1 << 20, seed=SEED_QUERY + 20)
    cross_u8 = v1.tune_lookup(q_tune, 1 << 20)
    fused_u8 = v1.lookup_fused_max
    cross_f32 = v32.tune_lookup(q_tune, 1 << 20)
    del q_tune
    log(f"bench: tuned AUTO crossovers: b16 L3 uint8 {cross_u8} (fused up to {fused_u8}), b32 L4 float {cross_f32}")
    # build autotuner: the persistent on-chip Lloyd loop's size limit for this workload
    tuned_tiles, tune_ms = rb.tune_build_options(d, N, 1000, 4, max_iters=10, seed_mode=1, kernel=0)
    log(f"bench: build autotuner: persist_max_tiles {tuned_tiles}, build ms per candidate {tune_ms}")

    line = {
        "metric": METRIC, "value": round(tops, 3), "unit": "TOPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int8", "data": "synthetic",
        "config": {"workload": "batched vocab build: b16 L4 (65,536 words), 2M SIFT-like uint8[128], "
                               "10 Lloyd passes/node, k-means++, every node of a level per launch",
                   "nb_desc": N, "branch": 16, "levels": 4, "passes": 10},
        "roofline": {"bound": "hbm", "achieved": round(assign_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(assign_gbs / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(),
                     "kernel": "k_assignU8Mfma", "avg_launch_us": round(launch_us, 2),
                     "algorithmic_bytes_per_launch": round(bytes_per_launch), "copy_gbs": round(copy_gbs, 1)},
        "assign_tops_frac_of_i8_peak": round(tops / I8_PEAK_TOPS, 5),
        "build_tops_whole": round(ops / (float(np.median(prod_ms)) * 1e-3) / 1e12, 3),
        "assign_step_ms": {"assign": round(st["assign_ms"], 3), "of_which_persistent": round(st["persist_ms"], 3),
                           "reduce": round(st["reduce_ms"], 3), "update": round(st["update_ms"], 3),
                           "streamed_launches": st["assign_launches"]},
        "assign_tops_plain": round(plain_tops, 3),
        "assign_stream_only": {"tops": round(stream_tops, 3), "assign_ms": round(sts["assign_ms"], 3),
                               "reduce_ms": round(sts["reduce_ms"], 3), "build_ms": round(sts["total_ms"], 2),
                               "identical_tree": same_stream},
        "build_ms": {"mfma": round(float(np.median(prod_ms)), 2),
                     "mfma_with_launch_timing": round(float(np.median([s["total_ms"] for s in stats])), 2),
                     "plain_with_launch_timing": round(stp["total_ms"], 2), "identical_trees": same_tree and same_prod,
                     "b32L4_f32": round(st32["total_ms"], 1), "b32L4_f32_plain": round(st32p["total_ms"], 1),
                     "b32L4_f32_identical": same32},
        "lookup_b32L4_f32_10M": {"mfma_desc_per_s": round(r_m, 1), "plain_desc_per_s": round(r_p, 1),
                                 "identical_leaf_ids": identical, "near_ties": near_ties,
                                 "hbm_gbs_mfma": round(r_m * LOOKUP_BYTES / 1e9, 1),
                                 "hbm_frac_mfma": round(r_m * LOOKUP_BYTES / 1e9 / HBM_PEAK_GBS, 4)},
        "lookup_b16L3_u8_50k": {"mfma_desc_per_s": round(r1_m, 1), "plain_desc_per_s": round(r1_p, 1),
                                "identical_leaf_ids": id1},
        "lookup_auto_crossover_tuned": {"b16L3_u8": cross_u8, "b32L4_f32": cross_f32,
                                        "b16L3_u8_fused_max": fused_u8,
                                        "b32L4_f32_leaf_item_rows": v32.lookup_leaf_item_rows,
                                        "b16L3_u8_leaf_item_rows": v1.lookup_leaf_item_rows},
        "build_autotuned": {"persist_max_tiles": tuned_tiles,
                            "build_ms_per_candidate": {str(c): round(m, 3) for c, m in tune_ms.items()}},
    }
    threads = cpu_threads()
    if not args.skip_retrieval:
        line["retrieval_100k_b32L4"] = bench_retrieval(rb, syn, v32, not args.no_cpu, threads, not args.no_graphs)
        log(f"bench: retrieval {line['retrieval_100k_b32L4']}")
    if not args.skip_serving:
        line["serving_1M_b32L4"] = bench_serving(rb, syn, v32, (200.0, 1000.0, 5000.0), args.serving_seconds)
    if not args.no_cpu:
        import oracle
        log(f"bench: CPU baseline on {threads} threads")
        host = d.numpy()
        res, cb = cpu_baseline(rb, oracle, host, threads)
        line["cpu_baseline"] = cb
        line["tree_equals_oracle"] = tree_vs_oracle_u8(voc, res)
        log(f"bench: timed b16 L4 build vs oracle: {line['tree_equals_oracle']}")
        log("bench: oracle b32 L4 float build on the same descriptors")
        res32 = oracle.build(host, 1000, 4, b=32, max_iters=10, seed_mode=1, centroid_type=1)
        line["tree_equals_oracle_b32L4_f32"] = tree_vs_oracle_f32(v32, res32, 32, 4)
        log(f"bench: b32 L4 float build vs oracle: {line['tree_equals_oracle_b32L4_f32']}")
        nq = qh.shape[0]
        tr = np.zeros((v32.nb_nodes, 32, 128), np.float32)
        for l in range(4):
            b0 = (32 ** l - 1) // 31
            tr[b0:b0 + 32 ** l] = v32.centroids(l)
        t = time.perf_counter()
        lc = oracle.lookup_f32(tr, 32, 4, qh, threads)
        dt = time.perf_counter() - t
        line["cpu_lookup_b32L4_f32"] = {"desc_per_s": round(nq / dt, 1), "cores": threads, "kind": "port",
                                        "sample": f"oracle descent of the first {nq} queries",
                                        "matches_gpu": bool(np.array_equal(lc, lm[:nq]))}
        d32, d64 = oracle.path_dist_f32(tr, 32, 4, qh, lm[:nq], threads)
        derr = np.abs(d32.astype(np.float64) - d64)
        line["error_report_vs_fp64"] = {
            "centroid_max_abs_b16L4_u8": res["centroid_err"],
            "centroid_max_abs_b32L4_f32": res32["centroid_err"],
            "lookup_distance_max_abs_b32L4_f32": float(derr.max()),
            "lookup_distance_max_rel_b32L4_f32": float((derr / np.maximum(d64, 1.0)).max()),
            "lookup_distance_sample": f"the winning child at each of the 4 levels of the first {nq} queries",
            "score_max_abs": (line.get("retrieval_100k_b32L4", {}).get("oracle_check", {})
                              .get("max_abs_score_err_vs_fp64"))}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
