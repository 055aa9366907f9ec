"""Shared helpers of the test suite: blob construction and brute-force references."""
import struct

import numpy as np

RBW2_MAGIC = 0x32574252


def ref_blob(L, nodes_ref, masks, weights):
    """Reference vocabulary blob (Vocabulary.cpp:175-197): [u32 L][Node x n][u16 x n][f32[16] x n]."""
    n = nodes_ref.shape[0]
    assert nodes_ref.shape == (n, 2048) and weights.shape == (n, 16)
    return (struct.pack("<I", L) + np.ascontiguousarray(nodes_ref, np.uint8).tobytes()
            + np.asarray(masks, np.uint16).tobytes() + np.ascontiguousarray(weights, np.float32).tobytes())


def rbw2_blob(b, L, centroid_type, rows, masks, weights):
    hdr = struct.pack("<6I", RBW2_MAGIC, 1, b, L, centroid_type, 0)
    dt = np.uint8 if centroid_type == 0 else np.float32
    return (hdr + np.ascontiguousarray(rows, dt).tobytes() + np.asarray(masks, np.uint32).tobytes()
            + np.ascontiguousarray(weights, np.float32).tobytes())


def level_beg(b, l):
    return (b ** l - 1) // (b - 1)


def brute_descent(rows, b, L, desc, exact_int=True):
    """Independent numpy descent: rows [node][b][128]; distances in int64 (or float64); first minimum wins."""
    desc = np.asarray(desc)
    node = np.zeros(desc.shape[0], np.int64)
    x = desc.astype(np.int64 if exact_int else np.float64)
    for l in range(L):
        c = rows[level_beg(b, l) + node].astype(np.int64 if exact_int else np.float64)  # [n][b][128]
        d = ((x[:, None, :] - c) ** 2).sum(-1)
        node = node * b + np.argmin(d, axis=1)  # argmin returns the first minimum
    return node.astype(np.uint32)


# |GPU float score - reference-semantics float score|: both accumulate in binary64,
# the GPU with atomics in any order
SCORE_TOL = 2e-6


def assert_ranked_equal(gd, gs, od, os_, tol=SCORE_TOL):
    """Ranked ids must be identical; the only allowed difference is an order swap between
    neighbours whose scores differ by less than tol."""
    assert len(gd) == len(od)
    np.testing.assert_allclose(gs, os_, atol=tol, rtol=0)
    if not np.array_equal(gd, od):
        for i in np.nonzero(gd != od)[0]:
            j = int(np.nonzero(od == gd[i])[0][0]) if gd[i] in od else -1
            assert j >= 0 and abs(float(os_[j]) - float(gs[i])) < tol, (i, gd[i], od[i])
