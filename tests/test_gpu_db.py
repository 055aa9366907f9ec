"""Inverted-file database on the GPU vs the oracle's restatement of DataBase.cpp."""
import ctypes

import numpy as np
import pytest

import oracle
import rapidbow_amd as rb
from tests.helpers import assert_ranked_equal, level_beg, ref_blob, rbw2_blob

pytestmark = pytest.mark.gpu



class GpuDb:
    def __init__(self, voc):
        h = ctypes.c_void_p()
        rb.check(rb.lib.rbow_db_create(ctypes.c_void_p(voc.handle), ctypes.byref(h)))
        self.h = h
        self.voc = voc

    def __del__(self):
        rb.lib.rbow_db_destroy(self.h)

    def add(self, doc, leaf):
        leaf = np.ascontiguousarray(leaf, np.uint32)
        rb.check(rb.lib.rbow_db_add_doc(self.h, ctypes.c_uint32(doc), leaf.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.c_uint32(leaf.size)))

    def query(self, leaf, K):
        leaf = np.ascontiguousarray(leaf, np.uint32)
        docs = np.zeros(max(K, 1), np.uint32)
        sc = np.zeros(max(K, 1), np.float32)
        n = ctypes.c_uint32()
        rb.check(rb.lib.rbow_db_query(self.h, leaf.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(leaf.size),
                                      ctypes.c_uint32(K), docs.ctypes.data_as(ctypes.c_void_p),
                                      sc.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
        return docs[:n.value], sc[:n.value]

    def postings(self, leaf):
        n = ctypes.c_uint32()
        rb.check(rb.lib.rbow_db_postings(self.h, ctypes.c_uint32(leaf), None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, np.uint32)
        rb.check(rb.lib.rbow_db_postings(self.h, ctypes.c_uint32(leaf), out.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_uint32(n.value), ctypes.byref(n)))
        return out

    def rcp(self, doc):
        r = ctypes.c_float()
        rb.check(rb.lib.rbow_db_doc_rcp(self.h, ctypes.c_uint32(doc), ctypes.byref(r)))
        return r.value

    def add_batch(self, docs, leaf_concat, offsets):
        d = rb.DeviceArray.from_numpy(np.ascontiguousarray(leaf_concat, np.uint32))
        ids = np.ascontiguousarray(docs, np.uint32)
        off = np.ascontiguousarray(offsets, np.uint64)
        rb.check(rb.lib.rbow_db_add_docs_device(self.h, ids.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(ids.size),
                                                ctypes.c_void_p(d.ptr), off.ctypes.data_as(ctypes.c_void_p), None))

    def query_batch(self, leaf_concat, offsets, K):
        d = rb.DeviceArray.from_numpy(np.ascontiguousarray(leaf_concat, np.uint32))
        off = np.ascontiguousarray(offsets, np.uint64)
        B = off.size - 1
        docs = rb.DeviceArray((B, K), np.uint32)
        sc = rb.DeviceArray((B, K), np.float32)
        cnt = rb.DeviceArray((B,), np.uint32)
        rb.check(rb.lib.rbow_db_query_batch(self.h, ctypes.c_void_p(d.ptr), off.ctypes.data_as(ctypes.c_void_p),
                                            ctypes.c_uint32(B), ctypes.c_uint32(K), ctypes.c_void_p(docs.ptr),
                                            ctypes.c_void_p(sc.ptr), ctypes.c_void_p(cnt.ptr), None))
        c = cnt.numpy()
        D, S = docs.numpy(), sc.numpy()
        return [(D[q, :c[q]], S[q, :c[q]]) for q in range(B)]


def _voc_and_weights(L=2, b=16, seed=0):
    rng = np.random.default_rng(seed)
    n = level_beg(b, L)
    w = (rng.random((n, b)) * 4).astype(np.float32)
    w[rng.random(w.shape) < 0.15] = np.finfo(np.float32).tiny  # FLT_MIN weights exercise the flush-to-zero path
    if b == 16:
        rows = rng.integers(0, 256, (n, 16, 128), dtype=np.uint8)
        blob = ref_blob(L, oracle.rows_to_ref_nodes(rows), np.full(n, 0xFFFF), w)
    else:
        rows = rng.random((n, b, 128)).astype(np.float32) * 255
        blob = rbw2_blob(b, L, 1, rows, np.full(n, 0xFFFFFFFF), w)
    return rb.deserialize_vocabulary(blob), w


def _random_docs(rng, nb, nleaves, lo=1, hi=300, hot=None):
    docs = {}
    for d in range(nb):
        n = int(rng.integers(lo, hi))
        if hot is not None:  # skewed: a few hot leaves with long posting lists
            x = np.where(rng.random(n) < 0.3, rng.choice(hot, n), rng.integers(0, nleaves, n))
        else:
            x = rng.integers(0, nleaves, n)
        docs[d] = x.astype(np.uint32)
    return docs


@pytest.mark.parametrize("b,L", [(16, 2), (16, 3), (32, 2)])
def test_db_reference_api_parity(b, L):
    voc, w = _voc_and_weights(L, b, seed=L + b)
    rng = np.random.default_rng(1)
    nleaves = b ** L
    docs = _random_docs(rng, 300, nleaves, hot=rng.integers(0, nleaves, 8))
    g = GpuDb(voc)
    o = oracle.DataBase(b, L, w)
    for d, lf in docs.items():
        g.add(d * 7 + 3, lf)  # sparse doc ids
        o.add(d * 7 + 3, lf)
    for leaf in rng.integers(0, nleaves, 30):
        np.testing.assert_array_equal(g.postings(int(leaf)), o.postings(int(leaf)))
    for d in list(docs)[:20]:
        assert g.rcp(d * 7 + 3) == o.rcp(d * 7 + 3)
    for qi, K in ((5, 10), (17, 50), (40, 1), (3, 300), (9, 1024)):
        q = docs[qi]
        gd, gs = g.query(q, K)
        od, os_ = o.query(q, K)
        assert_ranked_equal(gd, gs, od, os_)
        assert gd[0] == qi * 7 + 3  # self-query first (ApiTests.cpp:136)
    # a query sharing nothing with the DB, and an empty query
    gd, _ = g.query(np.array([], np.uint32), 10)
    assert gd.size == 0


def test_db_query_sizes_up_to_the_maximum():
    """Queries of 1 to 16,384 leaves: every size class of the query preparation's
    register sort (1,024 x E keys, E = 1..16), one past each power of two and the
    maximum, at K = 50 and K = 1,024, against the oracle; 16,385 leaves are rejected."""
    voc, w = _voc_and_weights(3, 16, seed=12)
    rng = np.random.default_rng(13)
    nleaves = 16 ** 3
    docs = _random_docs(rng, 300, nleaves, lo=50, hi=2000, hot=rng.integers(0, nleaves, 16))
    g = GpuDb(voc)
    o = oracle.DataBase(16, 3, w)
    for d, lf in docs.items():
        g.add(d, lf)
        o.add(d, lf)
    for n in (1, 1023, 1024, 1025, 2048, 4097, 8192, 8193, 16384):
        q = np.concatenate([docs[int(i)] for i in rng.integers(0, 300, 40)])
        q = np.resize(q, n).astype(np.uint32)  # duplicates included
        for K in (50, 1024):
            gd, gs = g.query(q, K)
            od, os_ = o.query(q, K)
            assert_ranked_equal(gd, gs, od, os_)
    with pytest.raises(rb.RbowError) as e:
        g.query(np.zeros(16385, np.uint32), 10)
    assert e.value.status == rb.ERR_CAPACITY


def test_db_errors_and_query_and_add():
    voc, w = _voc_and_weights(2, 16, seed=3)
    g = GpuDb(voc)
    o = oracle.DataBase(16, 2, w)
    rng = np.random.default_rng(2)
    docs = _random_docs(rng, 50, 256)
    for d, lf in docs.items():
        g.add(d, lf)
        o.add(d, lf)
    with pytest.raises(rb.RbowError) as e:
        g.add(3, docs[3])
    assert e.value.status == rb.ERR_DUPLICATE_DOC and str(e.value) == "Duplicate document index"
    # queryAndAddDoc = query (without the new doc) then add
    q = rng.integers(0, 256, 40).astype(np.uint32)
    out = np.zeros(20, np.uint32)
    n = ctypes.c_uint32()
    rb.check(rb.lib.rbow_db_query_and_add_doc(g.h, ctypes.c_uint32(999), q.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.c_uint32(q.size), ctypes.c_uint32(20),
                                              out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
    od, _ = o.query(q, 20)
    o.add(999, q)
    np.testing.assert_array_equal(out[:n.value], od)
    gd, _ = g.query(q, 5)
    assert gd[0] == 999


def test_db_batch_add_and_batch_query():
    voc, w = _voc_and_weights(3, 16, seed=4)
    rng = np.random.default_rng(5)
    nleaves = 16 ** 3
    docs = _random_docs(rng, 400, nleaves, lo=50, hi=600, hot=rng.integers(0, nleaves, 16))
    ids = np.array([d * 3 for d in docs], np.uint32)
    cat = np.concatenate([docs[d] for d in docs])
    off = np.concatenate([[0], np.cumsum([docs[d].size for d in docs])]).astype(np.uint64)
    g = GpuDb(voc)
    # first 100 docs one at a time (host path), the rest in two device batches
    for d in range(100):
        g.add(d * 3, docs[d])
    g.add_batch(ids[100:250], cat, off[100:251])
    g.add_batch(ids[250:], cat, off[250:])
    o = oracle.DataBase(16, 3, w)
    for d, lf in docs.items():
        o.add(d * 3, lf)
    for leaf in rng.integers(0, nleaves, 40):
        np.testing.assert_array_equal(g.postings(int(leaf)), o.postings(int(leaf)))
    qs = [docs[i] for i in rng.integers(0, 400, 64)]
    qs[3] = np.array([], np.uint32)
    qcat = np.concatenate(qs)
    qoff = np.concatenate([[0], np.cumsum([x.size for x in qs])]).astype(np.uint64)
    res = g.query_batch(qcat, qoff, 50)
    for q, (gd, gs) in zip(qs, res):
        od, os_ = o.query(q, 50)
        assert_ranked_equal(gd, gs, od, os_)
    # each result is the same alone as in the batch
    for i in (0, 7, 63):
        gd, gs = g.query(qs[i], 50)
        np.testing.assert_array_equal(gd, res[i][0])
        np.testing.assert_array_equal(gs, res[i][1])


def test_db_score_error_vs_fp64():
    voc, w = _voc_and_weights(2, 32, seed=6)
    rng = np.random.default_rng(7)
    docs = _random_docs(rng, 200, 1024, lo=100, hi=400)
    g = GpuDb(voc)
    o = oracle.DataBase(32, 2, w)
    for d, lf in docs.items():
        g.add(d, lf)
        o.add(d, lf)
    worst = 0.0
    for qi in range(0, 200, 10):
        gd, gs = g.query(docs[qi], 50)
        fd, fs = o.query_fp64(docs[qi], 50)
        ref = {int(d): s for d, s in zip(fd, fs)}
        for d, s in zip(gd, gs):
            if int(d) in ref:
                worst = max(worst, abs(float(s) - ref[int(d)]))
    print(f"max |GPU score - FP64 ideal| = {worst:.3e}")
    assert worst < 1e-5


def test_db_many_document_ranges():
    """More documents than one scoring tile holds (several document ranges), long
    skewed posting lists with duplicates, and sparse subtrees whose inner nodes
    fall under the 16+4K threshold (DataBase.cpp:125-164)."""
    voc, w = _voc_and_weights(3, 16, seed=8)
    rng = np.random.default_rng(9)
    nb = 3 * 4096 + 123  # four ranges, the last one partial
    hot = rng.integers(0, 1024, 12)
    docs = {}
    for d in range(nb):
        n = int(rng.integers(20, 80))
        u = rng.random(n)
        x = np.where(u < 0.15, rng.choice(hot, n),
                     np.where(u < 0.9, rng.integers(0, 1024, n), rng.integers(1024, 4096, n)))
        if d % 7 == 0:
            x[: n // 2] = x[0]  # a run of duplicates of one leaf
        if d % 1000 == 5:  # runs longer than a wave step and than an unrolled iteration
            x = np.concatenate([x, np.full(70 + d // 10, hot[d % 3])])
        docs[d] = x.astype(np.uint32)
    g = GpuDb(voc)
    o = oracle.DataBase(16, 3, w)
    ids = np.array([d * 3 + 1 for d in range(nb)], np.uint32)
    for d in range(0, 1000):
        g.add(int(ids[d]), docs[d])
    cat = np.concatenate([docs[d] for d in range(1000, nb)])
    off = np.concatenate([[0], np.cumsum([docs[d].size for d in range(1000, nb)])]).astype(np.uint64)
    g.add_batch(ids[1000:], cat, off)
    for d in range(nb):
        o.add(int(ids[d]), docs[d])
    for leaf in list(hot[:3]) + list(rng.integers(0, 4096, 20)):
        np.testing.assert_array_equal(g.postings(int(leaf)), o.postings(int(leaf)))
    qs = [docs[int(i)] for i in rng.integers(0, nb, 12)]
    qs.append(rng.integers(1024, 4096, 40).astype(np.uint32))  # cold region: inner nodes scored
    qs.append(np.concatenate([rng.choice(hot, 300), rng.integers(0, 4096, 500)]).astype(np.uint32))
    for qi, q in enumerate(qs):
        for K in ((60, 7, 300, 1024) if qi % 4 == 0 else (60, 7, 300)):
            gd, gs = g.query(q, K)
            od, os_ = o.query(q, K)
            assert_ranked_equal(gd, gs, od, os_)
    qcat = np.concatenate(qs)
    qoff = np.concatenate([[0], np.cumsum([x.size for x in qs])]).astype(np.uint64)
    assert rb.lib.rbow_db_walk_steps(g.h) == 6  # the default scoring tile shape
    for steps in (6, 4):  # both scoring tile shapes: 6 and 4 steps of 64 postings per wave iteration
        rb.check(rb.lib.rbow_db_set_walk_steps(g.h, steps))
        assert rb.lib.rbow_db_walk_steps(g.h) == steps
        res = g.query_batch(qcat, qoff, 60)
        for q, (gd, gs) in zip(qs, res):
            od, os_ = o.query(q, 60)
            assert_ranked_equal(gd, gs, od, os_)
    # the tuner times both on this batch, checks they agree, and keeps the faster
    d = rb.DeviceArray.from_numpy(qcat)
    best, ms = ctypes.c_uint32(), (ctypes.c_float * 2)()
    rb.check(rb.lib.rbow_db_tune_scoring(g.h, ctypes.c_void_p(d.ptr), qoff.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_uint32(len(qs)), ctypes.c_uint32(60), ctypes.byref(best), ms, None))
    assert best.value in (4, 6) and rb.lib.rbow_db_walk_steps(g.h) == best.value and ms[0] > 0 and ms[1] > 0
    assert best.value == (4 if ms[0] < ms[1] else 6)
    assert rb.lib.rbow_db_set_walk_steps(g.h, 8) == rb.ERR_INVALID_ARG
    rb.check(rb.lib.rbow_db_set_walk_steps(g.h, 0))
    assert rb.lib.rbow_db_walk_steps(g.h) == 6


def test_db_subnormal_weights_read_as_zero():
    """IDF weights below FLT_MIN (possible in a deserialised blob) act as 0, as the
    reference's DAZ host build reads them (CMakeLists.txt:16)."""
    rng = np.random.default_rng(11)
    n = level_beg(16, 2)
    w = (rng.random((n, 16)) * 4).astype(np.float32)
    w[rng.random(w.shape) < 0.3] = np.float32(1e-40)  # subnormal
    rows = rng.integers(0, 256, (n, 16, 128), dtype=np.uint8)
    voc = rb.deserialize_vocabulary(ref_blob(2, oracle.rows_to_ref_nodes(rows), np.full(n, 0xFFFF), w))
    g = GpuDb(voc)
    o = oracle.DataBase(16, 2, w)
    docs = _random_docs(rng, 120, 256)
    for d, lf in docs.items():
        g.add(d, lf)
        o.add(d, lf)
        assert g.rcp(d) == o.rcp(d)
    for qi in (0, 5, 77):
        gd, gs = g.query(docs[qi], 40)
        od, os_ = o.query(docs[qi], 40)
        assert_ranked_equal(gd, gs, od, os_)
