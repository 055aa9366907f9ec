"""Lookup parity on the GPU: plain and MFMA kernels vs the oracle, bit-exact leaf ids."""
import numpy as np
import pytest

import oracle
import rapidbow_amd as rb
from tests.helpers import level_beg, rbw2_blob, ref_blob

pytestmark = pytest.mark.gpu


def _u8_voc_from_oracle(desc, L, nb_doc=50, iters=8, seed_mode=1):
    res = oracle.build(desc, nb_doc, L, max_iters=iters, seed_mode=seed_mode)
    blob = ref_blob(L, res["nodes"], res["masks"], res["weights"])
    return rb.deserialize_vocabulary(blob), res, blob


@pytest.mark.parametrize("kernel", [rb.LOOKUP_PLAIN, rb.LOOKUP_MFMA])
def test_lookup_u8_cfg1_shape_parity(kernel):
    # b=16, L=3 (4,096 words) built on SIFT-like data; 50k queries, ragged tail
    desc = rb.synth_descriptors(60000, seed=1)
    voc, res, _ = _u8_voc_from_oracle(desc, 3)
    q = rb.synth_descriptors(50001, seed=2)
    expect = oracle.lookup_u8(res["nodes"], 3, q)
    got = voc.lookup_host(q, kernel)
    np.testing.assert_array_equal(got, expect)


@pytest.mark.parametrize("kernel", [rb.LOOKUP_PLAIN, rb.LOOKUP_MFMA])
def test_lookup_u8_ties_and_edges(kernel):
    rng = np.random.default_rng(3)
    for L in (1, 2, 4):
        n = level_beg(16, L)
        rows = rng.integers(0, 256, (n, 16, 128), dtype=np.uint8)
        rows[:, 11] = rows[:, 4]  # duplicate children: ties must go to the lower index
        rows[:, 14] = rows[:, 1]  # (across the MFMA's 4-child lane groups: 4 / 11, 1 / 14)
        rows[:, 6] = rows[:, 5]   # (within one group: 5 / 6)
        blob = ref_blob(L, oracle.rows_to_ref_nodes(rows), np.full(n, 0xFFFF), np.ones((n, 16), np.float32))
        voc = rb.deserialize_vocabulary(blob)
        for nq in (1, 15, 16, 17, 129, 1000):
            q = rng.integers(0, 256, (nq, 128), dtype=np.uint8)
            q[: min(nq, 16)] = rows[0, : min(nq, 16)]
            got = voc.lookup_host(q, kernel)
            np.testing.assert_array_equal(got, oracle.lookup_u8(oracle.rows_to_ref_nodes(rows), L, q))


def test_lookup_empty_is_noop():
    rng = np.random.default_rng(4)
    rows = rng.integers(0, 256, (17, 16, 128), dtype=np.uint8)
    voc = rb.deserialize_vocabulary(ref_blob(2, oracle.rows_to_ref_nodes(rows), np.zeros(17), np.zeros((17, 16))))
    out = rb.DeviceArray((4,), np.uint32)
    out.upload(np.array([7, 7, 7, 7], np.uint32))
    voc.lookup(None, 0, out)
    rb.hip.synchronize()
    np.testing.assert_array_equal(out.numpy(), [7, 7, 7, 7])


def _f32_tree(rng, b, L, desc, near_ties=True):
    n = level_beg(b, L)
    # centroids near data points so that distances are realistic
    # (means of uint8 data lie in [0, 255]; out-of-range centroids have a test of their own)
    pick = desc[rng.integers(0, desc.shape[0], n * b)].astype(np.float32).reshape(n, b, 128)
    rows = np.clip(pick + rng.normal(0, 3, pick.shape).astype(np.float32), 0, 254)
    if near_ties:
        rows[:, 5] = rows[:, 2]                       # exact duplicate
        rows[:, 9] = rows[:, 8] + np.float32(1e-3)    # near duplicate: forces the FP32 re-score
    return rows


@pytest.mark.parametrize("b,L", [(32, 3), (16, 3), (32, 1)])
def test_lookup_f32_parity_and_near_ties(b, L):
    rng = np.random.default_rng(5)
    desc = rb.synth_descriptors(20000, seed=7)
    rows = _f32_tree(rng, b, L, desc)
    n = rows.shape[0]
    voc = rb.deserialize_vocabulary(rbw2_blob(b, L, 1, rows, np.full(n, 0xFFFFFFFF), np.ones((n, b))))
    q = rb.synth_descriptors(30000, seed=8)
    q[:b] = np.clip(np.rint(rows[0]), 0, 255).astype(np.uint8)
    expect = oracle.lookup_f32(rows, b, L, q)
    plain = voc.lookup_host(q, rb.LOOKUP_PLAIN)
    np.testing.assert_array_equal(plain, expect)
    mfma, ties = voc.lookup_host(q, rb.LOOKUP_MFMA, return_near_ties=True)
    np.testing.assert_array_equal(mfma, expect)
    print(f"b={b} L={L} near ties re-scored: {ties}")
    assert ties > 0


@pytest.mark.parametrize("b", [32, 16])
def test_lookup_f32_candidate_band(b):
    """Siblings that differ from child 0 by perturbations from 1e-3 to 30 (in
    norm) put the plain argmin inside, at and outside the filter's certified
    band; queries sit on and between them.  MFMA leaf ids equal the plain
    kernel's and the oracle's."""
    rng = np.random.default_rng(11)
    L = 2
    desc = rb.synth_descriptors(4000, seed=9)
    rows = _f32_tree(rng, b, L, desc, near_ties=False)
    n = rows.shape[0]
    scales = np.logspace(-3, 1.5, b - 1).astype(np.float32)
    for v in range(n):
        u = rng.normal(0, 1, (b - 1, 128)).astype(np.float32)
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        rows[v, 1:] = rows[v, 0] + scales[:, None] * u
    voc = rb.deserialize_vocabulary(rbw2_blob(b, L, 1, rows, np.full(n, 0xFFFFFFFF), np.ones((n, b))))
    # queries: rounded centroids and points between child 0 and each sibling
    q = [np.clip(np.rint(rows[v, j]), 0, 255) for v in range(n) for j in range(b)]
    for v in range(n):
        for j in range(1, b):
            t = rng.uniform(0.3, 0.7)
            q.append(np.clip(np.rint((1 - t) * rows[v, 0] + t * rows[v, j] + rng.normal(0, 0.5, 128)), 0, 255))
    q = np.array(q, np.uint8)
    expect = oracle.lookup_f32(rows, b, L, q)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_PLAIN), expect)
    mfma, ties = voc.lookup_host(q, rb.LOOKUP_MFMA, return_near_ties=True)
    np.testing.assert_array_equal(mfma, expect)
    print(f"b={b}: {q.shape[0]} queries, near ties re-scored: {ties}")


@pytest.mark.parametrize("b", [32, 16])
@pytest.mark.parametrize("exact", [False, True])
def test_lookup_f32_all_rows_tie(b, exact):
    """Every query sits on a child with a twin: child b-1 equals child 3 in every
    node but for one value one ulp apart (exact=False), or bit for bit (exact=True).
    Near twins make every row of every work item a near tie at the root: the
    per-wave tie queue fills and is resolved several times within an item.  An
    exact twin has the same plain distance as child 3, so it can never win (ties go
    to the lowest child): it is no candidate at all, and the rows are not near ties.
    Leaf ids equal the plain kernel's and the oracle's."""
    rng = np.random.default_rng(21)
    desc = rb.synth_descriptors(2000, seed=22)
    rows = _f32_tree(rng, b, 2, desc, near_ties=False)
    n = rows.shape[0]
    for v in range(n):
        rows[v, b - 1] = rows[v, 3]
        if not exact:
            rows[v, b - 1, 17] = np.nextafter(rows[v, 3, 17], np.float32(255))
    voc = rb.deserialize_vocabulary(rbw2_blob(b, 2, 1, rows, np.full(n, 0xFFFFFFFF), np.ones((n, b))))
    target = np.clip(np.rint(rows[0, 3]), 0, 255).astype(np.uint8)
    q = np.repeat(target[None], 1000, axis=0)
    expect = oracle.lookup_f32(rows, b, 2, q)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_PLAIN), expect)
    mfma, ties = voc.lookup_host(q, rb.LOOKUP_MFMA, return_near_ties=True)
    np.testing.assert_array_equal(mfma, expect)
    if exact:
        assert ties < 1000  # the twins are no candidates
    else:
        assert ties >= 1000  # at least the root level ties for every row


@pytest.mark.parametrize("b", [32, 16])
def test_lookup_f32_out_of_range_centroids(b):
    """Float trees deserialized from blobs may hold centroid values outside the
    [0, 255] of any mean of uint8 data (negative, above 255, huge): such a node
    has no int8 split, and every row at it is decided by the plain FP32 chain.
    Nodes with in-range values in the same tree keep the MFMA filter.  Leaf ids
    equal the plain kernel's and the oracle's."""
    rng = np.random.default_rng(31)
    desc = rb.synth_descriptors(6000, seed=32)
    rows = _f32_tree(rng, b, 2, desc, near_ties=False)
    n = rows.shape[0]
    rows[0, 1, 7] = -3.5            # the root: one value below 0
    rows[2, 4, 100] = 255.75        # a level-1 node: one value above 255
    rows[5, 0, :] += 1e4            # another: a far child
    voc = rb.deserialize_vocabulary(rbw2_blob(b, 2, 1, rows, np.full(n, 0xFFFFFFFF), np.ones((n, b))))
    q = rb.synth_descriptors(8000, seed=33)
    expect = oracle.lookup_f32(rows, b, 2, q)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_PLAIN), expect)
    mfma, ties = voc.lookup_host(q, rb.LOOKUP_MFMA, return_near_ties=True)
    np.testing.assert_array_equal(mfma, expect)
    assert ties >= q.shape[0]  # every row is decided by the plain chain at the root


def test_blob_roundtrip_u8():
    desc = rb.synth_descriptors(8000, seed=9)
    voc, res, blob = _u8_voc_from_oracle(desc, 2)
    out = voc.serialize()
    assert out == blob
    n = level_beg(16, 2)
    assert len(out) == 4 + 2114 * n  # Vocabulary.cpp:177-182
    np.testing.assert_array_equal(voc.weights().reshape(-1, 16), res["weights"])
    voc2 = rb.deserialize_vocabulary(out)
    assert voc2.serialize() == blob


def test_lookup_autotuner_keeps_leaf_ids():
    desc = rb.synth_descriptors(60000, seed=5)
    voc, res, _ = _u8_voc_from_oracle(desc, 3)
    assert voc.lookup_crossover == 12288  # built-in default for uint8 trees
    q = rb.synth_descriptors(70000, seed=6)
    expect = oracle.lookup_u8(res["nodes"], 3, q)
    d = rb.DeviceArray.from_numpy(q)
    c = voc.tune_lookup(d, q.shape[0])
    assert c == voc.lookup_crossover and (c == 70001 or (c >= 4096 and c & (c - 1) == 0))
    f = voc.lookup_fused_max  # the tuned fused/bucketed boundary of this b16 uint8 tree
    assert f >= 4096 and f & (f - 1) == 0 and f <= 65536
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_AUTO), expect)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_MFMA), expect)
    for c in (1, 1 << 30):  # force either kernel through AUTO
        voc.lookup_crossover = c
        np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_AUTO), expect)
    voc.lookup_crossover = 0
    assert voc.lookup_crossover == 12288


@pytest.mark.parametrize("b,L,ct", [(32, 3, 1), (16, 3, 1), (16, 3, 0)])
def test_lookup_leaf_item_rows(b, L, ct):
    """The bucketed descent's leaf-level tile shape (256- or 512-row work items)
    gives the oracle's leaf ids either way, on float and uint8 trees, with leaf
    nodes holding from a few rows to more than one item; the autotuner keeps
    one of the two; other values are rejected and 0 restores the default."""
    rng = np.random.default_rng(41 + b + ct)
    desc = rb.synth_descriptors(20000, seed=9)
    n = level_beg(b, L)
    if ct:
        rows = _f32_tree(rng, b, L, desc)
        voc = rb.deserialize_vocabulary(rbw2_blob(b, L, 1, rows, np.full(n, 0xFFFFFFFF), np.ones((n, b))))
    else:
        rows = desc[rng.integers(0, desc.shape[0], n * 16)].reshape(n, 16, 128)
        nodes = oracle.rows_to_ref_nodes(rows)
        voc = rb.deserialize_vocabulary(ref_blob(L, nodes, np.full(n, 0xFFFF), np.ones((n, 16), np.float32)))
    assert voc.lookup_leaf_item_rows == (512 if ct else 256)
    q = rb.synth_descriptors(150_000, seed=10)  # ~37 rows per leaf at b32 L3, ~37 per node below level 2 at b16
    q[:2000] = q[0]                            # one leaf node holding more than 512 rows
    expect = oracle.lookup_f32(rows, b, L, q) if ct else oracle.lookup_u8(nodes, L, q)
    voc.lookup_fused_max = 1  # the bucketed descent on uint8 b16 trees too
    for rows_per_item in (256, 512):
        voc.lookup_leaf_item_rows = rows_per_item
        assert voc.lookup_leaf_item_rows == rows_per_item
        np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_MFMA), expect)
    d = rb.DeviceArray.from_numpy(q)
    voc.tune_lookup(d, q.shape[0])
    assert voc.lookup_leaf_item_rows in (256, 512)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_MFMA), expect)
    with pytest.raises(rb.RbowError) as e:
        voc.lookup_leaf_item_rows = 384
    assert e.value.status == rb.ERR_INVALID_ARG
    voc.lookup_leaf_item_rows = 0
    assert voc.lookup_leaf_item_rows == (512 if ct else 256)


def test_lookup_u8_fused_bucketed_plain_agree():
    """The MFMA form's two descents on uint8 b16 trees -- the fused single launch
    (small batches) and the level-synchronous bucketed one -- and the plain
    kernel give the oracle's leaf ids, with planted duplicate children, on trees
    of 1 to 4 levels and ragged batch sizes; 0 restores the default crossover."""
    rng = np.random.default_rng(31)
    for L in (1, 2, 3, 4):
        n = level_beg(16, L)
        rows = rng.integers(0, 256, (n, 16, 128), dtype=np.uint8)
        rows[:, 9] = rows[:, 3]   # ties across lane groups
        rows[:, 7] = rows[:, 6]   # and within one
        nodes = oracle.rows_to_ref_nodes(rows)
        voc = rb.deserialize_vocabulary(ref_blob(L, nodes, np.full(n, 0xFFFF), np.ones((n, 16), np.float32)))
        assert voc.lookup_fused_max == 196608
        for nq in (1, 16, 33, 1000, 40_001):
            q = rng.integers(0, 256, (nq, 128), dtype=np.uint8)
            q[: min(nq, 64)] = rows[rng.integers(0, n, min(nq, 64)), rng.integers(0, 16, min(nq, 64))]  # on a child
            exp = oracle.lookup_u8(nodes, L, q)
            voc.lookup_fused_max = 0
            fused = voc.lookup_host(q, rb.LOOKUP_MFMA)
            voc.lookup_fused_max = 1 if nq > 1 else 0  # every batch above one descriptor bucketed
            bucketed = voc.lookup_host(q, rb.LOOKUP_MFMA)
            voc.lookup_fused_max = 0
            np.testing.assert_array_equal(fused, exp, err_msg=f"fused L={L} n={nq}")
            np.testing.assert_array_equal(bucketed, exp, err_msg=f"bucketed L={L} n={nq}")
            np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_PLAIN), exp)


def test_lookup_u8_deep_tree_bucketed():
    """A b16 L5 tree (1,048,576 words): its leaf level has 65,536 parents, more than
    the one-block scan of the per-child item counts takes, so that level's counts
    go through the device-wide scan; the levels above use the one-block scan.
    Leaf ids of the bucketed MFMA descent equal the plain kernel's and the oracle's."""
    rng = np.random.default_rng(41)
    L = 5
    n = level_beg(16, L)
    rows = rng.integers(0, 256, (n, 16, 128), dtype=np.uint8)
    rows[:, 13] = rows[:, 2]  # duplicate children: ties go to the lower index
    nodes = oracle.rows_to_ref_nodes(rows)
    voc = rb.deserialize_vocabulary(ref_blob(L, nodes, np.full(n, 0xFFFF), np.ones((n, 16), np.float32)))
    voc.lookup_fused_max = 1  # every batch above one descriptor bucketed
    q = rng.integers(0, 256, (70_001, 128), dtype=np.uint8)
    q[:64] = rows[rng.integers(0, n, 64), rng.integers(0, 16, 64)]  # on a child
    exp = oracle.lookup_u8(nodes, L, q)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_MFMA), exp)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_PLAIN), exp)
