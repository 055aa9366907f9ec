"""Batched k-means kernels vs the oracle (re-runs of the reference's known-answer suites on the GPU)."""
import numpy as np
import pytest

import oracle
import rapidbow_amd as rb

pytestmark = pytest.mark.gpu

TASKS = [478, 15, 39, 134]  # KmeansAssignCenterTests.cpp:268


def _ref_inputs():
    pts, cents, prevs = [], [], []
    for n in TASKS:
        tiles, cwords, prev = oracle.ref_assign_test_inputs(n)
        pts.append(oracle.tiles_to_rows(tiles)[:n])
        cents.append(oracle.centers_words_to_rows(cwords))
        prevs.append(prev)
    return pts, cents, prevs


@pytest.mark.parametrize("kernel", [0, 1])
def test_assign_and_update_known_answer(kernel):
    pts, cents, prevs = _ref_inputs()
    X = rb.DeviceArray.from_numpy(np.concatenate(pts))
    C = rb.DeviceArray.from_numpy(np.stack(cents))
    A = rb.DeviceArray.from_numpy(np.concatenate(prevs))
    cnt = rb.DeviceArray.zeros((4, 16), np.uint32)
    sums = rb.DeviceArray.zeros((4, 16, 128), np.int64)
    ch = rb.DeviceArray.zeros((4,), np.uint32)
    rb.kmeans_assign(X, TASKS, 16, 0, C, A, cnt, sums, ch, kernel=kernel)
    ga, gc, gs, gch = A.numpy(), cnt.numpy(), sums.numpy(), ch.numpy()
    off = 0
    for t, n in enumerate(TASKS):
        ea, ec, es, ech = oracle.assign_u8(pts[t], cents[t], prevs[t])
        np.testing.assert_array_equal(ga[off:off + n], ea)
        np.testing.assert_array_equal(gc[t], ec)
        np.testing.assert_array_equal(gs[t], es)
        assert gch[t] == ech
        off += n
    rb.kmeans_update(TASKS, 16, 0, C, cnt, sums, ch)
    gcent = C.numpy()
    for t in range(4):
        ea, ec, es, _ = oracle.assign_u8(pts[t], cents[t], prevs[t])
        np.testing.assert_array_equal(gcent[t], oracle.update_u8(cents[t], ec, es))
    # accumulators are cleared for the next pass
    assert not cnt.numpy().any() and not sums.numpy().any() and not ch.numpy().any()


@pytest.mark.parametrize("b", [16, 32])
def test_assign_f32_parity(b):
    rng = np.random.default_rng(6)
    sizes = [1, 63, 64, 65, 700, 0, 2000]
    pts = [rb.synth_descriptors(n, seed=20 + i) for i, n in enumerate(sizes)]
    cents = [(rb.synth_descriptors(b, seed=50 + i).astype(np.float32) + rng.normal(0, 2, (b, 128)).astype(np.float32))
             for i in range(len(sizes))]
    cents[3][4] = cents[3][1]  # duplicate center: the lower index wins
    prevs = [rng.integers(0, b, n).astype(np.uint8) for n in sizes]
    X = rb.DeviceArray.from_numpy(np.concatenate(pts))
    C = rb.DeviceArray.from_numpy(np.stack(cents))
    A = rb.DeviceArray.from_numpy(np.concatenate(prevs))
    T = len(sizes)
    cnt = rb.DeviceArray.zeros((T, b), np.uint32)
    sums = rb.DeviceArray.zeros((T, b, 128), np.int64)
    ch = rb.DeviceArray.zeros((T,), np.uint32)
    rb.kmeans_assign(X, sizes, b, 1, C, A, cnt, sums, ch, kernel=1)
    ga, gc, gs, gch = A.numpy(), cnt.numpy(), sums.numpy(), ch.numpy()
    off = 0
    for t, n in enumerate(sizes):
        ea, ec, es, ech = oracle.assign_f32(pts[t], cents[t], prevs[t])
        np.testing.assert_array_equal(ga[off:off + n], ea)
        np.testing.assert_array_equal(gc[t], ec)
        np.testing.assert_array_equal(gs[t], es)
        assert gch[t] == ech
        off += n
    rb.kmeans_update(sizes, b, 1, C, cnt, sums, ch)
    gcent = C.numpy()
    for t, n in enumerate(sizes):
        ea, ec, es, ech = oracle.assign_f32(pts[t], cents[t], prevs[t])
        exp = oracle.update_f32(cents[t], ec, es) if ech else cents[t]
        np.testing.assert_array_equal(gcent[t], exp)


@pytest.mark.parametrize("tile_size", [32, 16])  # ReformatTests.cpp:105-107,115-117
def test_reformat_known_answer(tile_size):
    desc, idx = oracle.ref_reformat_test_inputs(136, 3, tile_size)
    D = rb.DeviceArray.from_numpy(desc)
    I = rb.DeviceArray.from_numpy(idx)
    T = rb.DeviceArray((3, 32, tile_size), np.uint32)
    rb.reformat_tiles(D, I, 3, tile_size, T)
    rb.hip.synchronize()
    np.testing.assert_array_equal(T.numpy(), oracle.reformat(desc, idx, tile_size))


def test_gather_rows_with_invalid():
    rng = np.random.default_rng(7)
    desc = rng.integers(0, 256, (500, 128), dtype=np.uint8)
    idx = rng.integers(0, 500, 1000).astype(np.uint32)
    idx[::7] = 0xFFFFFFFF
    out = rb.DeviceArray((1000, 128), np.uint8)
    rb.gather_rows(rb.DeviceArray.from_numpy(desc), rb.DeviceArray.from_numpy(idx), 1000, out)
    rb.hip.synchronize()
    exp = np.where((idx == 0xFFFFFFFF)[:, None], 255, desc[np.minimum(idx, 499)])
    np.testing.assert_array_equal(out.numpy(), exp)
