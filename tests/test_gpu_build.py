"""Vocabulary build on the GPU vs the oracle: nodes, masks, weights and Lloyd passes bit-exact."""
import numpy as np
import pytest

import oracle
import rapidbow_amd as rb
from tests.helpers import level_beg

pytestmark = pytest.mark.gpu


def _gpu_build(desc, nb_doc, L, **kw):
    d = rb.DeviceArray.from_numpy(desc)
    return rb.build_vocabulary(d, desc.shape[0], nb_doc, L, **kw)


def _check_u8(voc, res, L):
    blob = voc.serialize()
    n = level_beg(16, L)
    nodes = np.frombuffer(blob, np.uint8, n * 2048, 4).reshape(n, 2048)
    masks = np.frombuffer(blob, np.uint16, n, 4 + n * 2048)
    w = np.frombuffer(blob, np.float32, n * 16, 4 + n * 2050).reshape(n, 16)
    bad = np.nonzero((nodes != res["nodes"]).any(1))[0]
    assert bad.size == 0, f"nodes differ: {bad[:10]}"
    np.testing.assert_array_equal(masks, res["masks"].astype(np.uint16))
    np.testing.assert_array_equal(w, res["weights"])
    np.testing.assert_array_equal(voc.build_passes(), res["passes"])


@pytest.mark.parametrize("seed_mode", [0, 1])
@pytest.mark.parametrize("kernel", [0, 1])
def test_build_u8_parity(seed_mode, kernel):
    desc = rb.synth_descriptors(30000, seed=1)
    L, iters, nb_doc = 3, 8, 100
    res = oracle.build(desc, nb_doc, L, max_iters=iters, seed_mode=seed_mode)
    voc = _gpu_build(desc, nb_doc, L, max_iters=iters, seed_mode=seed_mode, kernel=kernel)
    _check_u8(voc, res, L)


def test_build_u8_reference_defaults_converge():
    # default options = reference: random picks, up to 1000 passes, polled every 4
    desc = rb.synth_descriptors(6000, seed=3, nb_centres=300)
    res = oracle.build(desc, 20, 2)
    voc = _gpu_build(desc, 20, 2)
    _check_u8(voc, res, 2)
    assert res["passes"].max() > 8


def test_build_voc_tests_invariants():
    # VocTests.cpp:86-132 (RMS == 0) and :38-84 (RMS <= 10) on the GPU build
    for which, L, nb_doc, lim in ((1, 3, 4, 0), (0, 2, 16, 10)):
        desc = oracle.ref_voc_test_inputs(which)
        voc = _gpu_build(desc, nb_doc, L)
        leaf = voc.lookup_host(desc)
        cent = voc.centroids(L - 1).reshape(-1, 128)[leaf]
        inertia = int(np.sum((cent.astype(np.int64) - desc) ** 2))
        assert np.sqrt(inertia // desc.shape[0]) <= lim
        res = oracle.build(desc, nb_doc, L)
        _check_u8(voc, res, L)


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("b,L", [(32, 2), (16, 2)])
def test_build_f32_parity(b, L, kernel):
    # kernel 0: BF16 MFMA candidate filter + FP32 re-score of near ties; 1: plain FP32 chains
    desc = rb.synth_descriptors(20000, seed=4)
    res = oracle.build(desc, 64, L, b=b, max_iters=10, seed_mode=1, centroid_type=1)
    voc = _gpu_build(desc, 64, L, branch_factor=b, max_iters=10, seed_mode=1, centroid_type=1, kernel=kernel)
    for l in range(L):
        got = voc.centroids(l)
        exp = res["nodes"][level_beg(b, l):level_beg(b, l + 1)]
        bad = np.nonzero((got != exp).any(axis=(1, 2)))[0]
        assert bad.size == 0, f"level {l}: nodes differ {bad[:10]}"
    np.testing.assert_array_equal(voc.weights().reshape(-1, b), res["weights"])
    np.testing.assert_array_equal(voc.build_passes(), res["passes"])
    q = rb.synth_descriptors(5000, seed=5)
    exp = oracle.lookup_f32(res["nodes"], b, L, q)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_MFMA), exp)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_PLAIN), exp)
    blob = voc.serialize()
    assert rb.deserialize_vocabulary(blob).serialize() == blob


@pytest.mark.parametrize("b", [32, 16])
def test_build_f32_mfma_equals_plain_three_levels(b):
    # 120k points, 3 levels (uneven tasks, many passes): the MFMA-filtered float
    # build must give the plain build's centroids bit for bit
    desc = rb.synth_descriptors(120000, seed=9, nb_centres=500)
    kw = dict(branch_factor=b, max_iters=12, seed_mode=1, centroid_type=1)
    vm = _gpu_build(desc, 100, 3, kernel=0, **kw)
    vp = _gpu_build(desc, 100, 3, kernel=1, **kw)
    for l in range(3):
        np.testing.assert_array_equal(vm.centroids(l), vp.centroids(l))
    np.testing.assert_array_equal(vm.build_passes(), vp.build_passes())
    assert vm.serialize() == vp.serialize()


def test_build_errors():
    d = rb.DeviceArray.from_numpy(np.zeros((15, 128), np.uint8))
    with pytest.raises(rb.RbowError) as e:
        rb.build_vocabulary(d, 15, 1, 2)
    assert str(e.value) == "Insufficient descriptors"


@pytest.mark.parametrize("iters", [3, 40])
def test_build_u8_persistent_and_streamed_tasks(iters):
    # 400k points, L=4: most of the 4,096 tasks of level 3 are small enough for the
    # on-chip persistent loop, in both LDS classes (<= 8 and 9..16 tiles of 64
    # points); the larger ones (17..128 tiles) run in the same persistent launch with
    # their tiles streamed through L2 each pass; levels 0-2 stream one launch per
    # pass; the pass cap stops some tasks early (3) or lets most converge (40)
    desc = rb.synth_descriptors(400000, seed=7, nb_centres=30)
    L, nb_doc = 4, 500
    res = oracle.build(desc, nb_doc, L, max_iters=iters, seed_mode=1)
    tiles = (res["occ"][level_beg(16, 2):level_beg(16, 3)].ravel().astype(np.int64) + 63) // 64  # level-3 tasks
    assert ((tiles >= 1) & (tiles <= 8)).sum() >= 256 and ((tiles > 8) & (tiles <= 16)).any()
    # with the cap at 3, level 3 also has tasks above 16 tiles, which stream in the persistent launch
    assert iters != 3 or ((tiles > 16).any() and tiles.max() <= 128)
    voc = _gpu_build(desc, nb_doc, L, max_iters=iters, seed_mode=1, kernel=0, timing=True)
    _check_u8(voc, res, L)
    st = voc.build_stats()
    assert st["persist_ms"] > 0 and st["assign_launches"] > 0  # both paths ran
    # the streamed assign processed exactly levels 0-2 (each holds a task above 128 tiles and
    # fewer small tasks than the device has CUs); level 3 ran in the persistent launches
    sizes = [np.array([desc.shape[0]])] + [res["occ"][level_beg(16, l - 1):level_beg(16, l)].ravel().astype(np.int64)
                                           for l in (1, 2)]
    per_level = [int((sizes[l] * res["passes"][level_beg(16, l):level_beg(16, l + 1)]).sum()) for l in range(3)]
    assert all(((z + 63) // 64).max() > 128 for z in sizes)
    assert st["stream_point_passes"] == sum(per_level) < st["point_passes"]
    # the same tree with every task streamed, one launch per pass
    voc0 = _gpu_build(desc, nb_doc, L, max_iters=iters, seed_mode=1, kernel=0, persist_max_tiles=0, timing=True)
    st0 = voc0.build_stats()
    assert st0["persist_ms"] == 0 and st0["stream_point_passes"] == st0["point_passes"] == st["point_passes"]
    assert voc0.serialize() == voc.serialize()


def test_build_timing_option():
    # per-launch events only with timing=1; the tree is the same either way
    desc = rb.synth_descriptors(120000, seed=11, nb_centres=40)
    timed = _gpu_build(desc, 300, 3, max_iters=6, seed_mode=1, kernel=0, timing=True)
    plain = _gpu_build(desc, 300, 3, max_iters=6, seed_mode=1, kernel=0)
    assert timed.serialize() == plain.serialize()
    st, s0 = timed.build_stats(), plain.build_stats()
    assert st["assign_ms"] > 0 and st["update_ms"] > 0 and st["total_ms"] > 0
    assert s0["assign_ms"] == 0 and s0["update_ms"] == 0 and s0["reduce_ms"] == 0 and s0["total_ms"] > 0
    # counts do not depend on timing
    for f in ("point_passes", "assign_ops", "assign_launches", "stream_point_passes"):
        assert st[f] == s0[f], f


def test_build_autotuner():
    """rbow_tune_build_options times the persistent-loop size limits on the
    caller's data and keeps the fastest; every candidate builds the same tree,
    which equals the oracle's."""
    desc = rb.synth_descriptors(150000, seed=41, nb_centres=40)
    d = rb.DeviceArray.from_numpy(desc)
    best, ms = rb.tune_build_options(d, desc.shape[0], 300, 4, max_iters=6, seed_mode=1)
    assert best in (0, 4, 8, 12, 16) and all(m > 0 for m in ms.values())
    assert ms[best] == min(ms.values())
    tuned = rb.build_vocabulary(d, desc.shape[0], 300, 4, max_iters=6, seed_mode=1, persist_max_tiles=best)
    res = oracle.build(desc, 300, 4, max_iters=6, seed_mode=1)
    _check_u8(tuned, res, 4)
    # float centroids: the option has no effect and is left as given
    best_f, ms_f = rb.tune_build_options(d, 20000, 50, 2, branch_factor=32, centroid_type=1, max_iters=3,
                                         persist_max_tiles=7)
    assert best_f == 7 and all(m == 0 for m in ms_f.values())


def test_build_workspace_reuse():
    """Builds one after another reuse the work buffers a completed build leaves
    with the device (larger than the next build needs, holding the last build's
    data): a larger build, a smaller one with the other seed mode, a float tree,
    then the first again, each against the oracle; after the release a fresh
    build gives the same tree."""
    big = rb.synth_descriptors(60000, seed=21)
    small = rb.synth_descriptors(9000, seed=22)
    res_big = oracle.build(big, 200, 3, max_iters=6, seed_mode=1)
    res_small = oracle.build(small, 50, 3, max_iters=6, seed_mode=0)
    first = _gpu_build(big, 200, 3, max_iters=6, seed_mode=1, kernel=0)
    _check_u8(first, res_big, 3)
    _check_u8(_gpu_build(small, 50, 3, max_iters=6, seed_mode=0, kernel=0), res_small, 3)
    res_f = oracle.build(small, 50, 2, b=32, max_iters=5, seed_mode=1, centroid_type=1)
    vf = _gpu_build(small, 50, 2, branch_factor=32, max_iters=5, seed_mode=1, centroid_type=1, kernel=0)
    for l in range(2):
        np.testing.assert_array_equal(vf.centroids(l), res_f["nodes"][level_beg(32, l):level_beg(32, l + 1)])
    again = _gpu_build(big, 200, 3, max_iters=6, seed_mode=1, kernel=0)
    assert again.serialize() == first.serialize()
    rb.release_build_workspaces()
    fresh = _gpu_build(big, 200, 3, max_iters=6, seed_mode=1, kernel=0)
    assert fresh.serialize() == first.serialize()


@pytest.mark.parametrize("seed_mode", [0, 1])
def test_build_and_lookup_small_tree_full_size(seed_mode):
    """The small-tree workload at its own size: 200,000 SIFT-like descriptors,
    b16 L3 (4,096 words), 8 passes per node, both seed modes; then 50,000 queries
    through both lookup kernels, all against the oracle."""
    desc = rb.synth_descriptors(200_000, seed=1)
    L, nb_doc = 3, 1000
    res = oracle.build(desc, nb_doc, L, max_iters=8, seed_mode=seed_mode)
    voc = _gpu_build(desc, nb_doc, L, max_iters=8, seed_mode=seed_mode, kernel=0)
    _check_u8(voc, res, L)
    q = rb.synth_descriptors(50_000, seed=2)
    exp = oracle.lookup_u8(res["nodes"], L, q)
    for k in (rb.LOOKUP_MFMA, rb.LOOKUP_PLAIN, rb.LOOKUP_AUTO):
        np.testing.assert_array_equal(voc.lookup_host(q, k), exp)


def test_build_and_lookup_b32_L4_float_tree():
    """A 1,048,576-word tree (b32 L4, float centroids, k-means++, 10 passes) on
    300,000 descriptors: every level's centroids, the weights and the pass
    counts against the oracle; 200,000 queries through both lookup kernels."""
    desc = rb.synth_descriptors(300_000, seed=3)
    b, L, nb_doc = 32, 4, 1000
    res = oracle.build(desc, nb_doc, L, b=b, max_iters=10, seed_mode=1, centroid_type=1)
    voc = _gpu_build(desc, nb_doc, L, branch_factor=b, max_iters=10, seed_mode=1, centroid_type=1, kernel=0)
    for l in range(L):
        got = voc.centroids(l)
        exp = res["nodes"][level_beg(b, l):level_beg(b, l + 1)]
        bad = np.nonzero((got != exp).any(axis=(1, 2)))[0]
        assert bad.size == 0, f"level {l}: nodes differ {bad[:10]}"
    np.testing.assert_array_equal(voc.weights().reshape(-1, b), res["weights"])
    np.testing.assert_array_equal(voc.build_passes(), res["passes"])
    q = rb.synth_descriptors(200_000, seed=4)
    exp = oracle.lookup_f32(res["nodes"], b, L, q)
    got, ties = voc.lookup_host(q, rb.LOOKUP_MFMA, return_near_ties=True)
    np.testing.assert_array_equal(got, exp)
    np.testing.assert_array_equal(voc.lookup_host(q, rb.LOOKUP_PLAIN), exp)
    print(f"near ties re-scored in FP32: {ties}")
