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


# |GPU float score - reference-semantics float score|: both accumulate the same
# binary32 terms in binary64, the GPU with atomics in another order, so the
# binary64 sums differ by a few binary64 ulps and the rounded binary32 scores
# by at most one binary32 ulp.
SCORE_ULPS = 1


def _ulp(x):
    return np.spacing(np.abs(np.asarray(x, np.float32))).astype(np.float64)


def assert_ranked_equal(gd, gs, od, os_, ulps=SCORE_ULPS):
    """The GPU's ranked list must be the oracle's.  Every document in both lists
    has its oracle score to within `ulps` binary32 ulps; the GPU list is itself
    ranked by (score, doc id); and ids may differ from the oracle's only through
    those ulp-level differences: a document present in only one of the lists
    scores within `ulps` ulps of the other list's last score."""
    gd, od = np.asarray(gd, np.int64), np.asarray(od, np.int64)
    gs, os_ = np.asarray(gs, np.float32), np.asarray(os_, np.float32)
    assert len(gd) == len(od), (len(gd), len(od))
    if len(gd) == 0:
        return
    key = list(zip(gs.tolist(), gd.tolist()))
    assert key == sorted(key), "GPU list not ranked by (score, doc id)"
    oref = {int(d): float(s) for d, s in zip(od, os_)}
    for i, (d, s) in enumerate(zip(gd.tolist(), gs.astype(np.float64).tolist())):
        if d in oref:
            assert abs(s - oref[d]) <= ulps * _ulp(max(abs(s), abs(oref[d]))), (i, d, s, oref[d])
        else:  # entered at the K-th place through an ulp-level difference
            assert abs(s - float(os_[-1])) <= ulps * _ulp(max(abs(s), abs(float(os_[-1])))), (i, d, s, os_[-1])
    gset = set(gd.tolist())
    for j, d in enumerate(od.tolist()):
        if d not in gset:
            s = float(os_[j])
            assert abs(s - float(gs[-1])) <= ulps * _ulp(max(abs(s), abs(float(gs[-1])))), (j, d, s, gs[-1])
