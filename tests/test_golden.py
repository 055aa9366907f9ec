"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle).

CPU tests: the oracle reproduces every fixture bit for bit, so a change of the
oracle or of the toolchain it is built with (its reference-mode seeds come from
libstdc++'s std::mt19937_64 / std::shuffle) shows up here; the fixture lookups
are also re-derived by an independent numpy descent.  GPU tests: the library,
through the C ABI, reproduces the fixtures without running the oracle.
"""
import os

import numpy as np
import pytest

import oracle
import rapidbow_amd as rb
from tests.helpers import assert_ranked_equal, brute_descent, ref_blob, rbw2_blob

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ASSIGN_TASKS = (478, 15, 39, 134)  # KmeansAssignCenterTests.cpp:268


def load(name):
    with np.load(os.path.join(GOLDEN, name)) as f:  # allow_pickle stays False
        return {k: f[k] for k in f.files}


# ------------------------------------------------------------------ CPU ----
def test_golden_assign_inputs_and_oracle():
    g = load("assign_known_answer.npz")
    for n in ASSIGN_TASKS:
        tiles, cwords, prev = oracle.ref_assign_test_inputs(n)  # the suite's std::mt19937_64{} draws
        np.testing.assert_array_equal(oracle.tiles_to_rows(tiles)[:n], g[f"n{n}_points"])
        np.testing.assert_array_equal(oracle.centers_words_to_rows(cwords), g[f"n{n}_centers"])
        np.testing.assert_array_equal(prev, g[f"n{n}_prev"])
        a, counts, sums, changed = oracle.assign_u8(g[f"n{n}_points"], g[f"n{n}_centers"], g[f"n{n}_prev"])
        np.testing.assert_array_equal(a, g[f"n{n}_assign"])
        np.testing.assert_array_equal(counts, g[f"n{n}_counts"])
        np.testing.assert_array_equal(sums, g[f"n{n}_sums"])
        assert changed == int(g[f"n{n}_changed"])
        np.testing.assert_array_equal(oracle.update_u8(g[f"n{n}_centers"], counts, sums), g[f"n{n}_updated"])


def test_golden_build_oracle():
    g = load("build_ref_b16L2.npz")
    res = oracle.build(g["descriptors"], int(g["nb_doc"]), 2, b=16, max_iters=1000, seed_mode=0, seed=5489)
    for k, f in (("nodes", "nodes"), ("masks", "masks"), ("weights", "weights"), ("occ", "occurrences"),
                 ("passes", "passes")):
        np.testing.assert_array_equal(res[k], g[f], err_msg=k)
    np.testing.assert_array_equal(oracle.lookup_u8(g["nodes"], 2, g["queries"]), g["query_leaf"])
    rows = oracle.ref_nodes_to_rows(g["nodes"])
    np.testing.assert_array_equal(brute_descent(rows, 16, 2, g["queries"]), g["query_leaf"])


def test_golden_db_oracle():
    g, b = load("db_b16L2.npz"), load("build_ref_b16L2.npz")
    db = oracle.DataBase(16, 2, b["weights"])
    off = g["doc_offsets"]
    for d in range(off.size - 1):
        assert db.add(d, g["doc_leaf"][off[d]:off[d + 1]]) == 0
    np.testing.assert_array_equal(np.array([db.rcp(d) for d in range(off.size - 1)], np.float32), g["doc_rcp"])
    qo, K = g["query_offsets"], int(g["K"])
    for q in range(qo.size - 1):
        d, s = db.query(g["query_leaf"][qo[q]:qo[q + 1]], K)
        c = int(g["ranked_counts"][q])
        np.testing.assert_array_equal(d, g["ranked_docs"][q, :c])
        np.testing.assert_array_equal(s, g["ranked_scores"][q, :c])


def test_golden_lookup_f32_oracle():
    g = load("lookup_f32.npz")
    for b in (32, 16):
        np.testing.assert_array_equal(oracle.lookup_f32(g[f"b{b}_rows"], b, 1, g[f"b{b}_queries"]), g[f"b{b}_leaf"])
        # rows 2 and 5 are equal: their queries take the lower index
        assert g[f"b{b}_leaf"][5] == 2


# ------------------------------------------------------------------ GPU ----
@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
def test_golden_assign_gpu(kernel):
    g = load("assign_known_answer.npz")
    pts = [g[f"n{n}_points"] for n in ASSIGN_TASKS]
    X = rb.DeviceArray.from_numpy(np.concatenate(pts))
    C = rb.DeviceArray.from_numpy(np.stack([g[f"n{n}_centers"] for n in ASSIGN_TASKS]))
    A = rb.DeviceArray.from_numpy(np.concatenate([g[f"n{n}_prev"] for n in ASSIGN_TASKS]))
    cnt = rb.DeviceArray.zeros((4, 16), np.uint32)
    sums = rb.DeviceArray.zeros((4, 16, 128), np.int64)
    ch = rb.DeviceArray.zeros((4,), np.uint32)
    rb.kmeans_assign(X, list(ASSIGN_TASKS), 16, 0, C, A, cnt, sums, ch, kernel=kernel)
    ga, gc, gs, gch = A.numpy(), cnt.numpy(), sums.numpy(), ch.numpy()
    off = 0
    for t, n in enumerate(ASSIGN_TASKS):
        np.testing.assert_array_equal(ga[off:off + n], g[f"n{n}_assign"])
        np.testing.assert_array_equal(gc[t], g[f"n{n}_counts"])
        np.testing.assert_array_equal(gs[t], g[f"n{n}_sums"])
        assert gch[t] == int(g[f"n{n}_changed"])
        off += n
    rb.kmeans_update(list(ASSIGN_TASKS), 16, 0, C, cnt, sums, ch)
    np.testing.assert_array_equal(C.numpy(), np.stack([g[f"n{n}_updated"] for n in ASSIGN_TASKS]))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
def test_golden_build_gpu(kernel):
    g = load("build_ref_b16L2.npz")
    desc = g["descriptors"]
    voc = rb.build_vocabulary(rb.DeviceArray.from_numpy(desc), desc.shape[0], int(g["nb_doc"]), 2, max_iters=1000,
                              seed_mode=0, seed=5489, kernel=kernel)
    blob = voc.serialize()
    n = g["nodes"].shape[0]
    np.testing.assert_array_equal(np.frombuffer(blob, np.uint8, n * 2048, 4).reshape(n, 2048), g["nodes"])
    np.testing.assert_array_equal(np.frombuffer(blob, np.uint16, n, 4 + n * 2048), g["masks"].astype(np.uint16))
    np.testing.assert_array_equal(np.frombuffer(blob, np.float32, n * 16, 4 + n * 2050).reshape(n, 16), g["weights"])
    np.testing.assert_array_equal(voc.build_passes(), g["passes"])
    for k in (rb.LOOKUP_MFMA, rb.LOOKUP_PLAIN):
        np.testing.assert_array_equal(voc.lookup_host(g["queries"], k), g["query_leaf"])


@pytest.mark.gpu
def test_golden_db_gpu():
    g, b = load("db_b16L2.npz"), load("build_ref_b16L2.npz")
    voc = rb.deserialize_vocabulary(ref_blob(2, b["nodes"], b["masks"], b["weights"]))
    db = rb.DataBase(voc)
    off = g["doc_offsets"]
    for d in range(off.size - 1):
        db.add_doc(d, g["doc_leaf"][off[d]:off[d + 1]])
    qo, K = g["query_offsets"], int(g["K"])
    for q in range(qo.size - 1):
        gd, gs = db.query(g["query_leaf"][qo[q]:qo[q + 1]], K)
        c = int(g["ranked_counts"][q])
        assert_ranked_equal(gd, gs, g["ranked_docs"][q, :c], g["ranked_scores"][q, :c])


@pytest.mark.gpu
def test_golden_lookup_f32_gpu():
    g = load("lookup_f32.npz")
    for b in (32, 16):
        rows = g[f"b{b}_rows"]
        voc = rb.deserialize_vocabulary(rbw2_blob(b, 1, 1, rows, np.full(1, 0xFFFFFFFF), np.ones((1, b))))
        for k in (rb.LOOKUP_MFMA, rb.LOOKUP_PLAIN):
            np.testing.assert_array_equal(voc.lookup_host(g[f"b{b}_queries"], k), g[f"b{b}_leaf"])
