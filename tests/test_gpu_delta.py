"""Incremental inserts through the delta segment, the device-side leaf-id checks
and the device IDF weights, against the oracle (DataBase.cpp:39-197,
Vocabulary.cpp:103-108)."""
import ctypes

import numpy as np
import pytest

import oracle
import rapidbow_amd as rb
from rapidbow_amd import hip
from tests.helpers import assert_ranked_equal
from tests.test_gpu_db import GpuDb, _random_docs, _voc_and_weights

pytestmark = pytest.mark.gpu

DELTA_MAX = 1 << 24  # delta postings that trigger the merge into the main CSR


def _query_and_add(g, doc, leaf, K):
    leaf = np.ascontiguousarray(leaf, np.uint32)
    out = np.zeros(max(K, 1), np.uint32)
    n = ctypes.c_uint32()
    rb.check(rb.lib.rbow_db_query_and_add_doc(g.h, ctypes.c_uint32(doc), leaf.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.c_uint32(leaf.size), ctypes.c_uint32(K),
                                              out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
    return out[:n.value]


def _delta(g):
    return rb.lib.rbow_db_nb_delta_postings(g.h)


def _compact(g):
    rb.check(rb.lib.rbow_db_compact(g.h, None))


def _concat(docs, keys):
    cat = np.concatenate([docs[d] for d in keys]).astype(np.uint32)
    off = np.concatenate([[0], np.cumsum([docs[d].size for d in keys])]).astype(np.uint64)
    return cat, off


def test_idf_weights_exhaustive():
    """Every occurrence count up to 2^20 and a sample up to 2^32, at six document
    counts: the device weights equal the oracle's bit for bit."""
    rng = np.random.default_rng(1)
    occ = np.concatenate([np.arange(1 << 20), rng.integers(1 << 20, 1 << 32, 1 << 16),
                          [0xFFFFFFFF, 0xFFFFFFFE]]).astype(np.uint32)
    d = rb.DeviceArray.from_numpy(occ)
    w = rb.DeviceArray((occ.size,), np.float32)
    for nb_doc in (1, 7, 1000, 100_000, 1_000_000, 0xFFFFFFFF):
        rb.idf_weights(d, occ.size, nb_doc, w)
        hip.synchronize()
        got = w.numpy()
        exp = oracle.idf_weights(occ, nb_doc)
        bad = np.nonzero(got.view(np.uint32) != exp.view(np.uint32))[0]
        assert bad.size == 0, f"nb_doc {nb_doc}: occ {occ[bad[:5]]} got {got[bad[:5]]} exp {exp[bad[:5]]}"


def test_query_and_add_loop_through_the_delta():
    """A reference-style queryAndAddDoc loop (DataBase.cpp:192-197) on a
    batch-built file: every call's ranking equals the oracle's query before the
    add, the inserts stay in the delta (the main CSR is not rebuilt per call),
    and an explicit compaction changes no result."""
    voc, w = _voc_and_weights(3, 16, seed=21)
    rng = np.random.default_rng(22)
    nleaves = 16 ** 3
    docs = _random_docs(rng, 6300, nleaves, lo=20, hi=200, hot=rng.integers(0, nleaves, 12))
    g = GpuDb(voc)
    o = oracle.DataBase(16, 3, w)
    cat, off = _concat(docs, range(6000))
    g.add_batch(np.arange(6000, dtype=np.uint32), cat, off)
    for d in range(6000):
        o.add(d, docs[d])
    assert _delta(g) == int(off[-1])  # a batch below the threshold lands in the delta
    _compact(g)
    assert _delta(g) == 0
    added = 0
    for d in range(6000, 6300):
        K = 20 if d % 3 else 60
        got = _query_and_add(g, d, docs[d], K)
        od, os_ = o.query(docs[d], K)
        np.testing.assert_array_equal(got, od, err_msg=f"queryAndAddDoc {d}")
        o.add(d, docs[d])
        added += docs[d].size
        if d % 50 == 0:  # scores too, through the plain query
            gd, gs = g.query(docs[d], 30)
            od, os_ = o.query(docs[d], 30)
            assert_ranked_equal(gd, gs, od, os_)
            assert gd[0] == d
    assert _delta(g) == added  # the added documents stayed in the delta: the main CSR was not rebuilt
    qs = [docs[int(i)] for i in rng.integers(0, 6300, 10)]
    before = [g.query(q, 40) for q in qs]
    _compact(g)
    assert _delta(g) == 0
    for q, (bd, bs) in zip(qs, before):
        gd, gs = g.query(q, 40)
        np.testing.assert_array_equal(gd, bd)
        np.testing.assert_array_equal(gs, bs)
        od, os_ = o.query(q, 40)
        assert_ranked_equal(gd, gs, od, os_)
    for leaf in rng.integers(0, nleaves, 20):
        np.testing.assert_array_equal(g.postings(int(leaf)), o.postings(int(leaf)))


def test_delta_straddles_ranges_and_inner_nodes():
    """Main CSR of three full document ranges and part of a fourth; the delta
    starts inside that range and fills two more (host inserts, then a small
    device batch on top); queries reach leaves and sparse inner nodes
    (DataBase.cpp:125-164) in both segments."""
    voc, w = _voc_and_weights(3, 16, seed=23)
    rng = np.random.default_rng(24)
    nb_main = 3 * 4096 + 100
    nb = nb_main + 5000 + 3000
    hot = rng.integers(0, 1024, 12)
    docs = {}
    for d in range(nb):
        n = int(rng.integers(20, 80))
        u = rng.random(n)
        x = np.where(u < 0.15, rng.choice(hot, n),
                     np.where(u < 0.9, rng.integers(0, 1024, n), rng.integers(1024, 4096, n)))
        if d % 7 == 0:
            x[: n // 2] = x[0]  # a run of duplicates
        docs[d] = x.astype(np.uint32)
    ids = np.array([d * 5 + 2 for d in range(nb)], np.uint32)
    g = GpuDb(voc)
    o = oracle.DataBase(16, 3, w)
    cat, off = _concat(docs, range(nb_main))
    g.add_batch(ids[:nb_main], cat, off)
    _compact(g)  # these documents form the main CSR
    for d in range(nb_main, nb_main + 5000):
        g.add(int(ids[d]), docs[d])
    cat, off = _concat(docs, range(nb_main + 5000, nb))
    g.add_batch(ids[nb_main + 5000:], cat, off)
    for d in range(nb):
        o.add(int(ids[d]), docs[d])
    assert _delta(g) > 0
    qs = [docs[int(i)] for i in rng.integers(0, nb, 10)]
    qs.append(rng.integers(1024, 4096, 40).astype(np.uint32))  # cold region: inner nodes scored
    qs.append(np.concatenate([rng.choice(hot, 300), rng.integers(0, 4096, 500)]).astype(np.uint32))
    for qi, q in enumerate(qs):
        for K in (60, 7, 1024 if qi % 4 == 0 else 300):
            gd, gs = g.query(q, K)
            od, os_ = o.query(q, K)
            assert_ranked_equal(gd, gs, od, os_)
    cat, off = _concat({i: q for i, q in enumerate(qs)}, range(len(qs)))
    res = g.query_batch(cat, off, 60)
    for q, (gd, gs) in zip(qs, res):
        od, os_ = o.query(q, 60)
        assert_ranked_equal(gd, gs, od, os_)
    for leaf in list(hot[:3]) + list(rng.integers(0, 4096, 20)):
        np.testing.assert_array_equal(g.postings(int(leaf)), o.postings(int(leaf)))
    for d in (nb_main - 1, nb_main, nb - 1):
        assert g.rcp(int(ids[d])) == o.rcp(int(ids[d]))


def test_device_leaf_ids_are_checked():
    voc, w = _voc_and_weights(2, 16, seed=25)
    g = GpuDb(voc)
    rng = np.random.default_rng(26)
    docs = _random_docs(rng, 40, 256)
    cat, off = _concat(docs, range(40))
    g.add_batch(np.arange(40, dtype=np.uint32), cat, off)
    bad = cat.copy()
    bad[17] = 256  # = b^L
    with pytest.raises(rb.RbowError) as e:
        g.add_batch(np.arange(100, 140, dtype=np.uint32), bad, off)
    assert e.value.status == rb.ERR_INVALID_ARG and str(e.value) == "leaf id out of range"
    assert rb.lib.rbow_db_nb_docs(g.h) == 40  # nothing was inserted
    g.add_batch(np.arange(100, 140, dtype=np.uint32), cat, off)  # the same docs, valid ids
    assert rb.lib.rbow_db_nb_docs(g.h) == 80
    for v in (256, 0xFFFFFFFF):
        q = np.concatenate([docs[1], docs[2]]).astype(np.uint32)
        q[-1] = v
        qoff = np.array([0, docs[1].size, q.size], np.uint64)
        with pytest.raises(rb.RbowError) as e:
            g.query_batch(q, qoff, 10)
        assert e.value.status == rb.ERR_INVALID_ARG
    o = oracle.DataBase(16, 2, w)
    for d in range(40):
        o.add(d, docs[d])
        o.add(100 + d, docs[d])
    gd, gs = g.query(docs[5], 10)  # the database is intact after the rejected calls
    od, os_ = o.query(docs[5], 10)
    assert_ranked_equal(gd, gs, od, os_)


def test_hundred_thousand_images_batch_query():
    """100,000 images (25 document ranges) built two ways -- one bulk device
    insert (above the delta threshold: straight into the main CSR) and four
    device batches through the delta with a merge -- then a 64-query batch,
    each against the oracle's DataBase."""
    b, L = 32, 3
    voc, w = _voc_and_weights(L, b, seed=27)
    rng = np.random.default_rng(28)
    nleaves, NI = b ** L, 100_000
    cnt = rng.integers(100, 250, NI)
    total = int(cnt.sum())
    assert total > DELTA_MAX
    hot = rng.integers(0, nleaves, 64)
    leaf = np.where(rng.random(total) < 0.2, rng.choice(hot, total), rng.integers(0, nleaves, total)).astype(np.uint32)
    off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint64)
    ids = np.arange(NI, dtype=np.uint32)
    ga = GpuDb(voc)
    ga.add_batch(ids, leaf, off)
    assert _delta(ga) == 0
    gb = GpuDb(voc)
    for c in range(4):
        i0, i1 = c * NI // 4, (c + 1) * NI // 4
        gb.add_batch(ids[i0:i1], leaf[int(off[i0]):int(off[i1])], off[i0:i1 + 1] - off[i0])
    assert _delta(gb) == 0  # the fourth batch pushed the delta over the threshold
    o = oracle.DataBase(b, L, w)
    for i in range(NI):
        o.add(i, leaf[int(off[i]):int(off[i + 1])])
    assert rb.lib.rbow_db_nb_postings(ga.h) == o.nb_postings() == total
    qimg = rng.choice(NI, 64, replace=False)
    qs = []
    for i in qimg:  # a re-observation of an image: most of its leaves plus noise
        x = leaf[int(off[i]):int(off[i + 1])]
        keep = x[rng.random(x.size) < 0.8]
        qs.append(np.concatenate([keep, rng.integers(0, nleaves, 30)]).astype(np.uint32))
    qcat, qoff = _concat(dict(enumerate(qs)), range(64))
    ra, rb_ = ga.query_batch(qcat, qoff, 50), gb.query_batch(qcat, qoff, 50)
    for i, q in enumerate(qs):
        od, os_ = o.query(q, 50)
        assert_ranked_equal(ra[i][0], ra[i][1], od, os_)
        np.testing.assert_array_equal(ra[i][0], rb_[i][0])
        np.testing.assert_array_equal(ra[i][1], rb_[i][1])
        assert ra[i][0][0] == qimg[i]
    for d in rng.integers(0, NI, 20):  # batch-inserted norms equal the reference-order host sums
        assert ga.rcp(int(d)) == o.rcp(int(d))
