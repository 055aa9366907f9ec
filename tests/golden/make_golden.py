"""Writes the golden fixtures of this directory with the CPU oracle (oracle/).

Run from the repository root once the oracle is built (``oracle/Makefile``):

    python tests/golden/make_golden.py

Each fixture is a compressed ``.npz`` of plain arrays (no pickled objects, so
``numpy.load`` with its default ``allow_pickle=False`` reads it) holding the
inputs together with the expected outputs; the inputs come from numpy's seeded
PCG64, so reading a fixture needs neither the oracle nor the library.

  assign_known_answer.npz  the four known-answer tasks of the reference's
                           k-means assign/update suite (KmeansAssignCenterTests.cpp:
                           147-206, 266-291), inputs re-created from the
                           default-seeded std::mt19937_64 as that suite draws them
  build_ref_b16L2.npz      a reference-seed-mode vocabulary (b16, L2, pass cap
                           1000: kmeans.cpp:141-196 seeds, Vocabulary.cpp:32-121)
                           on 640 SIFT-like descriptors, and 100 looked-up queries
  db_b16L2.npz             an inverted file of 30 images on that vocabulary and
                           6 top-10 queries (DataBase.cpp:39-190)
  lookup_f32.npz           float-centroid trees (b32 L1, b16 L1) with planted
                           equal and 1e-3-apart children, and their leaf ids
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

ASSIGN_TASKS = (478, 15, 39, 134)  # KmeansAssignCenterTests.cpp:268


def sift_like(rng, n, centres):
    """Cluster centre + noise (sigma 0.3 of the centre norm), |.|, L2-normalise,
    clip at 0.2, renormalise, x512, round, saturate (the SIFT-like recipe)."""
    c = centres[rng.integers(0, centres.shape[0], n)]
    sigma = 0.3 * np.linalg.norm(c, axis=1, keepdims=True) / np.sqrt(128.0)
    x = np.abs(c + sigma * rng.normal(size=c.shape))
    v = x / np.linalg.norm(x, axis=1, keepdims=True)
    v = np.minimum(v, 0.2)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return np.clip(np.rint(512.0 * v), 0, 255).astype(np.uint8)


def assign_fixture():
    out = {}
    for n in ASSIGN_TASKS:
        tiles, cwords, prev = oracle.ref_assign_test_inputs(n)
        pts = oracle.tiles_to_rows(tiles)[:n]
        centers = oracle.centers_words_to_rows(cwords)
        a, counts, sums, changed = oracle.assign_u8(pts, centers, prev)
        out.update({f"n{n}_points": pts, f"n{n}_centers": centers, f"n{n}_prev": prev, f"n{n}_assign": a,
                    f"n{n}_counts": counts, f"n{n}_sums": sums.astype(np.int32), f"n{n}_changed": np.uint32(changed),
                    f"n{n}_updated": oracle.update_u8(centers, counts, sums)})
    return out


def build_fixture(rng):
    centres = rng.normal(size=(40, 128))
    desc = sift_like(rng, 640, centres)
    nb_doc = 20
    res = oracle.build(desc, nb_doc, 2, b=16, max_iters=1000, seed_mode=0, seed=5489)
    queries = sift_like(rng, 100, centres)
    leaf = oracle.lookup_u8(res["nodes"], 2, queries)
    return centres, {"descriptors": desc, "nb_doc": np.uint32(nb_doc), "nodes": res["nodes"], "masks": res["masks"],
                     "weights": res["weights"], "occurrences": res["occ"], "passes": res["passes"],
                     "queries": queries, "query_leaf": leaf}


def db_fixture(rng, centres, build):
    nodes, weights = build["nodes"], build["weights"]
    db = oracle.DataBase(16, 2, weights)
    sizes = rng.integers(20, 90, 30)
    leaf_lists = []
    for d, n in enumerate(sizes):
        img = sift_like(rng, int(n), centres[rng.choice(40, 4, replace=False)])
        leaf = oracle.lookup_u8(nodes, 2, img)
        leaf_lists.append(leaf)
        assert db.add(d, leaf) == 0
    K = 10
    qs = [leaf_lists[i][::2] for i in (0, 7, 19, 29)] + [leaf_lists[3], oracle.lookup_u8(nodes, 2, sift_like(rng, 50, centres))]
    out = {"doc_offsets": np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64),
           "doc_leaf": np.concatenate(leaf_lists).astype(np.uint32),
           "doc_rcp": np.array([db.rcp(d) for d in range(30)], np.float32),
           "query_offsets": np.concatenate([[0], np.cumsum([q.size for q in qs])]).astype(np.uint64),
           "query_leaf": np.concatenate(qs).astype(np.uint32), "K": np.uint32(K)}
    docs, scores, counts = [], [], []
    for q in qs:
        d, s = db.query(q, K)
        docs.append(np.pad(d, (0, K - d.size), constant_values=0xFFFFFFFF))
        scores.append(np.pad(s, (0, K - s.size)))
        counts.append(d.size)
    out.update({"ranked_docs": np.array(docs, np.uint32), "ranked_scores": np.array(scores, np.float32),
                "ranked_counts": np.array(counts, np.uint32)})
    return out


def lookup_f32_fixture(rng):
    out = {}
    for b in (32, 16):
        base = sift_like(rng, b, rng.normal(size=(8, 128))).astype(np.float32)
        rows = (base + rng.normal(0.0, 3.0, base.shape)).astype(np.float32)[None]  # one node: the root
        rows[0, 5] = rows[0, 2]                            # equal children: the lower index wins
        rows[0, 9] = rows[0, 8] + np.float32(1e-3)         # 1e-3 apart: decided by the plain FP32 chain
        q = np.concatenate([np.clip(np.rint(rows[0]), 0, 255).astype(np.uint8),
                            np.clip(np.rint(base + rng.normal(0.0, 4.0, base.shape)), 0, 255).astype(np.uint8),
                            sift_like(rng, 100, rng.normal(size=(8, 128)))])
        out[f"b{b}_rows"] = rows
        out[f"b{b}_queries"] = q
        out[f"b{b}_leaf"] = oracle.lookup_f32(rows, b, 1, q)
    return out


def main():
    rng = np.random.default_rng(20261204)
    np.savez_compressed(os.path.join(HERE, "assign_known_answer.npz"), **assign_fixture())
    centres, build = build_fixture(rng)
    np.savez_compressed(os.path.join(HERE, "build_ref_b16L2.npz"), **build)
    np.savez_compressed(os.path.join(HERE, "db_b16L2.npz"), **db_fixture(rng, centres, build))
    np.savez_compressed(os.path.join(HERE, "lookup_f32.npz"), **lookup_f32_fixture(rng))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
