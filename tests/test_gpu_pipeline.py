"""Query batches through the pipeline (lookup + scoring + top-K), as one launch
of a captured HIP graph and as direct launches, against the separate lookup and
batch query and against the oracle (Vocabulary.cpp:162-173, DataBase.cpp:71-190)."""
import numpy as np
import pytest

import oracle
import rapidbow_amd as rb
from rapidbow_amd import hip
from tests.helpers import assert_ranked_equal

pytestmark = pytest.mark.gpu


def _setup(b, L, ct, nb_docs, seed):
    train = rb.synth_descriptors(40_000, seed=seed)
    voc = rb.build_vocabulary(rb.DeviceArray.from_numpy(train), train.shape[0], nb_docs, L, branch_factor=b,
                              max_iters=6, seed_mode=1, centroid_type=ct)
    rng = np.random.default_rng(seed)
    images = [rb.synth_descriptors(int(rng.integers(100, 900)), seed=seed * 1000 + i) for i in range(nb_docs)]
    db = rb.DataBase(voc)
    for i, img in enumerate(images):
        db.add_doc(i, voc.lookup_host(img))
    return voc, db, images


def _batch(images, rng, nq, empty=()):
    qs = []
    for i in range(nq):
        if i in empty:
            qs.append(np.zeros((0, 128), np.uint8))
            continue
        img = images[int(rng.integers(0, len(images)))]
        qs.append(img[rng.random(img.shape[0]) < 0.7])
    off = np.concatenate([[0], np.cumsum([q.shape[0] for q in qs])]).astype(np.uint64)
    cat = np.concatenate(qs) if off[-1] else np.zeros((1, 128), np.uint8)
    return qs, cat, off


def _run(pipe, cat, off, use_graph, kernel=rb.LOOKUP_AUTO):
    nq = off.size - 1
    d = rb.DeviceArray.from_numpy(np.ascontiguousarray(cat))
    docs = rb.DeviceArray((nq, pipe.K), np.uint32)
    sc = rb.DeviceArray((nq, pipe.K), np.float32)
    cn = rb.DeviceArray((nq,), np.uint32)
    pipe.run(d, off, docs, sc, cn, lookup_kernel=kernel, use_graph=use_graph)
    c, D, S = cn.numpy(), docs.numpy(), sc.numpy()
    return [(D[q, :c[q]].copy(), S[q, :c[q]].copy()) for q in range(nq)]


def _separate(voc, db, cat, off, K, kernel=rb.LOOKUP_AUTO):
    n = int(off[-1])
    d = rb.DeviceArray.from_numpy(np.ascontiguousarray(cat))
    leaf = rb.DeviceArray((max(n, 1),), np.uint32)
    if n:
        voc.lookup(d, n, leaf, kernel)
    nq = off.size - 1
    docs = rb.DeviceArray((nq, K), np.uint32)
    sc = rb.DeviceArray((nq, K), np.float32)
    cn = rb.DeviceArray((nq,), np.uint32)
    db.query_batch(leaf, off, K, docs, sc, cn)
    c, D, S = cn.numpy(), docs.numpy(), sc.numpy()
    return [(D[q, :c[q]].copy(), S[q, :c[q]].copy()) for q in range(nq)]


def _same(a, b):
    assert len(a) == len(b)
    for (ad, as_), (bd, bs) in zip(a, b):
        np.testing.assert_array_equal(ad, bd)
        np.testing.assert_array_equal(as_, bs)


def test_graph_replays_equal_direct_launches_and_the_oracle():
    """Batches of 1 to 64 images, empty ones included, each as a graph replay and
    as direct launches with eager work in between: identical results, equal to
    the separate lookup + batch query, and to the oracle on a sample; graphs are
    reused within a size class."""
    voc, db, images = _setup(16, 3, 0, 300, seed=51)
    K = 20
    pipe = rb.QueryPipeline(voc, db, 64, 200_000, K)
    w = voc.weights()
    odb = oracle.DataBase(16, 3, w)
    res_nodes = np.zeros((voc.nb_nodes, 2048), np.uint8)
    for l in range(3):
        b0 = (16 ** l - 1) // 15
        res_nodes[b0:b0 + 16 ** l] = oracle.rows_to_ref_nodes(voc.centroids(l).astype(np.uint8))
    for i, img in enumerate(images):
        odb.add(i, oracle.lookup_u8(res_nodes, 3, img))
    rng = np.random.default_rng(52)
    for rep, nq in enumerate([1, 3, 3, 64, 17, 64, 2, 1, 40]):
        qs, cat, off = _batch(images, rng, nq, empty=(1,) if nq > 2 else ())
        g = _run(pipe, cat, off, True)
        hip.synchronize()
        d = _run(pipe, cat, off, False)  # eager work between the replays
        _same(g, d)
        _same(g, _separate(voc, db, cat, off, K))
        for q in range(0, nq, 7):
            od, os_ = odb.query(oracle.lookup_u8(res_nodes, 3, qs[q]), K)
            assert_ranked_equal(g[q][0], g[q][1], od, os_)
    assert 0 < pipe.graphs <= 8  # size classes were reused


def test_graph_recaptured_after_the_database_changes():
    voc, db, images = _setup(16, 2, 0, 120, seed=53)
    K = 10
    pipe = rb.QueryPipeline(voc, db, 8, 50_000, K)
    rng = np.random.default_rng(54)
    qs, cat, off = _batch(images, rng, 5)
    before = _run(pipe, cat, off, True)
    n0 = pipe.graphs
    db.add_doc(1000, voc.lookup_host(qs[2]))  # a document equal to query 2
    after = _run(pipe, cat, off, True)
    assert pipe.graphs > n0
    _same(after, _run(pipe, cat, off, False))
    assert after[2][0][0] == 1000 and before[2][0][0] != 1000


@pytest.mark.parametrize("kernel", [rb.LOOKUP_MFMA, rb.LOOKUP_PLAIN])
def test_graph_float_tree(kernel):
    """A float b32 L2 tree: the graph holds the bucketed BF16 lookup (or the
    plain one) with the scoring; replays equal direct launches."""
    voc, db, images = _setup(32, 2, 1, 150, seed=55)
    K = 15
    pipe = rb.QueryPipeline(voc, db, 16, 100_000, K)
    rng = np.random.default_rng(56)
    for nq in (16, 9, 16):
        qs, cat, off = _batch(images, rng, nq)
        g = _run(pipe, cat, off, True, kernel)
        _same(g, _run(pipe, cat, off, False, kernel))
        _same(g, _separate(voc, db, cat, off, K, kernel))


def test_server_with_and_without_graphs():
    voc, db, images = _setup(16, 3, 0, 200, seed=57)
    rng = np.random.default_rng(58)
    qs, _, _ = _batch(images, rng, 40)
    out = []
    for use_graphs in (True, False):
        srv = rb.Server(voc, db, max_batch=8, nb_streams=2, max_results=12, use_graphs=use_graphs)
        tickets = [srv.submit(q) for q in qs]
        out.append([srv.wait(t)[:2] for t in tickets])
        srv.close()
    for (gd, gs), (dd, ds) in zip(*out):
        np.testing.assert_array_equal(gd, dd)
        np.testing.assert_array_equal(gs, ds)


def test_pipeline_capacity_errors():
    voc, db, images = _setup(16, 2, 0, 20, seed=59)
    pipe = rb.QueryPipeline(voc, db, 4, 20_000, 10)
    rng = np.random.default_rng(60)
    _, cat, off = _batch(images, rng, 5)
    with pytest.raises(rb.RbowError) as e:
        _run(pipe, cat, off, True)
    assert e.value.status == rb.ERR_CAPACITY
    big = np.zeros((20_001, 128), np.uint8)
    with pytest.raises(rb.RbowError) as e:
        _run(pipe, big, np.array([0, 10_000, 20_001], np.uint64), True)
    assert e.value.status == rb.ERR_CAPACITY
