"""Continuous-batching server: every result is the same alone as in a mixed batch,
and the same as the reference pipeline (vocTreeLookUp.cu:48-160 leaf ids, then
DataBase.cpp:71-190 scoring) restated by the oracle."""
import numpy as np
import pytest

import oracle
import rapidbow_amd as rb
from tests.helpers import assert_ranked_equal, level_beg, ref_blob, rbw2_blob

pytestmark = pytest.mark.gpu



def _tree(b, L, seed):
    rng = np.random.default_rng(seed)
    n = level_beg(b, L)
    w = (rng.random((n, b)) * 4).astype(np.float32)
    if b == 16:
        desc = rb.synth_descriptors(n * 16, seed=seed + 100)
        rows = desc.reshape(n, 16, 128)
        voc = rb.deserialize_vocabulary(ref_blob(L, oracle.rows_to_ref_nodes(rows), np.full(n, 0xFFFF), w))
        ref_leaf = lambda d: oracle.lookup_u8(oracle.rows_to_ref_nodes(rows), L, d)
    else:
        rows = rb.synth_descriptors(n * b, seed=seed + 100).astype(np.float32).reshape(n, b, 128)
        rows += rng.random(rows.shape).astype(np.float32) - 0.5
        voc = rb.deserialize_vocabulary(rbw2_blob(b, L, 1, rows, np.full(n, 0xFFFFFFFF), w))
        ref_leaf = lambda d: oracle.lookup_f32(rows, b, L, d)
    return voc, w, ref_leaf


def _setup(b, L, nb_docs=120, seed=0):
    voc, w, ref_leaf = _tree(b, L, seed)
    rng = np.random.default_rng(seed + 1)
    db = rb.DataBase(voc)
    odb = oracle.DataBase(b, L, w)
    images = []
    for d in range(nb_docs):
        n = int(rng.integers(300, 1200))
        desc = rb.synth_descriptors(n, seed=1000 + d)
        images.append(desc)
        leaf = ref_leaf(desc)
        db.add_doc(d, leaf)
        odb.add(d, leaf)
    return voc, db, odb, ref_leaf, images


def _queries(images, rng, nq):
    qs = []
    for i in range(nq):
        src = images[int(rng.integers(0, len(images)))]
        keep = rng.random(src.shape[0]) < 0.7  # a partial re-observation of a stored image
        qs.append(np.ascontiguousarray(src[keep]))
    qs[2] = np.zeros((0, 128), np.uint8)  # an empty query
    return qs


@pytest.mark.parametrize("b,L,kernel", [(16, 3, rb.LOOKUP_MFMA), (16, 3, rb.LOOKUP_PLAIN), (32, 2, rb.LOOKUP_MFMA)])
def test_server_alone_vs_mix_vs_oracle(b, L, kernel):
    voc, db, odb, ref_leaf, images = _setup(b, L, seed=b + L)
    rng = np.random.default_rng(7)
    qs = _queries(images, rng, 40)
    K = 20
    # the mix: everything submitted at once, batched up to 16 per launch on 3 streams
    srv = rb.Server(voc, db, max_batch=16, nb_streams=3, max_results=K, lookup_kernel=kernel)
    tickets = [srv.submit(q) for q in qs]
    mix = [srv.wait(t) for t in tickets]
    srv.close()
    # alone: one query per batch, one stream
    alone_srv = rb.Server(voc, db, max_batch=1, nb_streams=1, max_results=K, lookup_kernel=kernel)
    for i, q in enumerate(qs):
        ad, as_, _ = alone_srv.wait(alone_srv.submit(q))
        md, ms, lat = mix[i]
        assert_ranked_equal(md, ms, ad, as_)
        assert lat >= 0
        od, os_ = odb.query(ref_leaf(q), K)
        assert_ranked_equal(md, ms, od, os_)
        if q.shape[0] == 0:
            assert md.size == 0
    alone_srv.close()


def test_server_staging_ring_wraps_and_waits():
    # a 2 MiB staging ring (the minimum) for ~4 MiB of queries submitted at once:
    # submits wait for batches to complete, requests wrap around the ring, and
    # every result still equals the query alone
    voc, db, odb, ref_leaf, images = _setup(16, 3, nb_docs=60, seed=4)
    rng = np.random.default_rng(13)
    qs = _queries(images, rng, 60)
    K = 10
    srv = rb.Server(voc, db, max_batch=8, nb_streams=2, max_results=K, staging_bytes=2 << 20)
    tickets = [srv.submit(q) for q in qs]
    mix = [srv.wait(t) for t in tickets]
    srv.close()
    assert sum(q.nbytes for q in qs) > 3 << 20
    for i, q in enumerate(qs):
        od, os_ = odb.query(ref_leaf(q), K)
        md, ms, _ = mix[i]
        assert_ranked_equal(md, ms, od, os_)
    with pytest.raises(rb.RbowError) as e:
        rb.Server(voc, db, staging_bytes=(2 << 20) - 1)
    assert e.value.status == rb.ERR_INVALID_ARG


def test_server_replay_poisson():
    voc, db, odb, ref_leaf, images = _setup(16, 3, nb_docs=80, seed=3)
    rng = np.random.default_rng(9)
    qs = _queries(images, rng, 60)
    desc = np.concatenate(qs)
    off = np.concatenate([[0], np.cumsum([q.shape[0] for q in qs])]).astype(np.uint64)
    K = 10
    srv = rb.Server(voc, db, max_batch=8, nb_streams=2, max_results=K)
    r = srv.replay(desc, off, qps=2000.0, seed=11)
    srv.close()
    assert r["elapsed_s"] > 0 and np.all(r["latency_ms"] > 0)
    assert np.all(r["batch_sizes"] >= 1) and np.all(r["batch_sizes"] <= 8)
    for i, q in enumerate(qs):  # the replayed ranking itself, against the oracle
        od, os_ = odb.query(ref_leaf(q), K)
        c = int(r["counts"][i])
        assert c == od.size
        assert_ranked_equal(r["docs"][i, :c], r["scores"][i, :c], od, os_)


def test_server_errors():
    voc, db, _, _, _ = _setup(16, 2, nb_docs=10, seed=5)
    with pytest.raises(rb.RbowError) as e:
        rb.Server(voc, db, max_batch=0)
    assert e.value.status == rb.ERR_INVALID_ARG
    with pytest.raises(rb.RbowError) as e:
        rb.Server(voc, db, max_results=2000)
    assert e.value.status == rb.ERR_CAPACITY
    srv = rb.Server(voc, db, mem_budget=1 << 20)
    with pytest.raises(rb.RbowError) as e:
        srv.submit(np.zeros((16385, 128), np.uint8))
    assert e.value.status == rb.ERR_CAPACITY
    with pytest.raises(rb.RbowError) as e:  # scratch of one query above the budget
        srv.submit(np.zeros((16000, 128), np.uint8))
    assert e.value.status == rb.ERR_CAPACITY
    d, s, _ = srv.wait(srv.submit(rb.synth_descriptors(500, seed=1)))
    assert d.size == min(10, db.nb_docs) and np.all(np.diff(s) >= 0)  # ascending: 2 - 2*sum(min)
    srv.close()
