"""Pins the CPU oracle (CPU only).

Each oracle function is checked against something that does not share its
code: the reference tests' known-answer checks re-run on their own regenerated
inputs (brute force in numpy, as computeRef / refCheck do), closed forms, the
C++ standard's known answer for std::mt19937_64, invariants, and independent
numpy re-implementations on small inputs.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import brute_descent, level_beg


def test_mt19937_64_known_answer():
    # [rand.predef]: the 10000th invocation of a default-constructed mt19937_64 yields 9981545732273789042
    r = oracle.Rng()
    v = [r.next() for _ in range(10000)]
    assert v[-1] == 9981545732273789042


def test_level_beg_closed_form():
    for b in (16, 32):
        for l in range(6):
            assert oracle.level_beg(b, l) == (b ** l - 1) // (b - 1)
    # FullTree.h:54-56 static_assert
    assert oracle.level_beg(16, 3) == 1 + 16 + 256


def test_dist_u8_brute():
    rng = np.random.default_rng(0)
    for _ in range(50):
        x = rng.integers(0, 256, 128, dtype=np.uint8)
        c = rng.integers(0, 256, 128, dtype=np.uint8)
        assert oracle.dist_u8(x, c) == int(((x.astype(np.int64) - c) ** 2).sum())
    assert oracle.dist_u8(np.zeros(128, np.uint8), np.full(128, 255, np.uint8)) == 128 * 255 * 255


def test_dist_f32_exact_cases_and_bound():
    rng = np.random.default_rng(1)
    # integer-valued centroids: every intermediate is an integer < 2^24, so the fma chain is exact
    for _ in range(20):
        x = rng.integers(0, 256, 128, dtype=np.uint8)
        c = rng.integers(0, 256, 128).astype(np.float32)
        assert oracle.dist_f32(x, c) == float(((x.astype(np.int64) - c.astype(np.int64)) ** 2).sum())
    # general centroids: within the forward-error bound of a 128-term chain
    for _ in range(50):
        x = rng.integers(0, 256, 128, dtype=np.uint8)
        c = (rng.random(128) * 255).astype(np.float32)
        exact = float(((x.astype(np.float64) - c.astype(np.float64)) ** 2).sum())
        assert abs(oracle.dist_f32(x, c) - exact) <= 2.0 ** -16 * exact + 1e-3


def _random_ref_tree(rng, L, dup=True):
    n = level_beg(16, L)
    rows = rng.integers(0, 256, (n, 16, 128), dtype=np.uint8)
    if dup:  # duplicate children force ties; the lower index must win
        rows[:, 9] = rows[:, 3]
        rows[:, 15] = rows[:, 0]
    return rows


def test_lookup_u8_brute_with_ties():
    rng = np.random.default_rng(2)
    for L in (1, 2, 3):
        rows = _random_ref_tree(rng, L)
        desc = rng.integers(0, 256, (700, 128), dtype=np.uint8)
        desc[:16] = rows[0, :, :]  # exact hits on root children (incl. duplicates)
        got = oracle.lookup_u8(oracle.rows_to_ref_nodes(rows), L, desc)
        np.testing.assert_array_equal(got, brute_descent(rows, 16, L, desc))
        assert not np.any((got >> (4 * (L - 1))) == 9) or L > 1  # duplicate child 9 never chosen at the root


def test_ref_layout_roundtrip():
    rng = np.random.default_rng(3)
    rows = rng.integers(0, 256, (5, 16, 128), dtype=np.uint8)
    ref = oracle.rows_to_ref_nodes(rows)
    # Vocabulary.h:42: byte(dim k, child j) = (k/4)*64 + 4j + k%4
    for (i, j, k) in [(0, 0, 0), (1, 5, 7), (4, 15, 127), (2, 3, 64)]:
        assert ref[i, (k // 4) * 64 + 4 * j + k % 4] == rows[i, j, k]
    np.testing.assert_array_equal(oracle.ref_nodes_to_rows(ref), rows)


def test_lookup_f32_integer_tree_brute():
    rng = np.random.default_rng(4)
    for b, L in ((32, 2), (16, 3)):
        n = level_beg(b, L)
        rows = rng.integers(0, 256, (n, b, 128)).astype(np.float32)
        rows[:, 7] = rows[:, 2]
        desc = rng.integers(0, 256, (500, 128), dtype=np.uint8)
        got = oracle.lookup_f32(rows, b, L, desc)
        np.testing.assert_array_equal(got, brute_descent(rows.astype(np.int64), b, L, desc))


def _compute_ref(pts, centers, prev):
    """tests/KmeansAssignCenterTests.cpp:147-189 computeRef, in numpy."""
    d = ((pts.astype(np.int64)[:, None, :] - centers.astype(np.int64)[None]) ** 2).sum(-1)
    a = np.argmin(d, axis=1)  # std::min_element: first minimum
    changed = int((a != prev).sum())
    counts = np.bincount(a, minlength=16).astype(np.uint32)
    sums = np.zeros((16, 128), np.int64)
    np.add.at(sums, a, pts.astype(np.int64))
    return a.astype(np.uint8), counts, sums, changed


@pytest.mark.parametrize("nb_desc", [478, 15, 39, 134])  # KmeansAssignCenterTests.cpp:268
def test_ref_assign_known_answer(nb_desc):
    tiles, cwords, prev = oracle.ref_assign_test_inputs(nb_desc)
    pts = oracle.tiles_to_rows(tiles)[:nb_desc]
    centers = oracle.centers_words_to_rows(cwords)
    # centers are descriptors min(i, n-1) (KmeansAssignCenterTests.cpp:59-64)
    for i in range(16):
        np.testing.assert_array_equal(centers[i], pts[min(i, nb_desc - 1)])
    a, counts, sums, ch = oracle.assign_u8(pts, centers, prev)
    ra, rc, rs, rch = _compute_ref(pts, centers, prev)
    np.testing.assert_array_equal(a, ra)
    np.testing.assert_array_equal(counts, rc)
    np.testing.assert_array_equal(sums, rs)
    assert ch == rch


@pytest.mark.parametrize("nb_desc", [478, 15, 39, 134])  # KmeansAssignCenterTests.cpp:285-291
def test_ref_update_known_answer(nb_desc):
    tiles, cwords, prev = oracle.ref_assign_test_inputs(nb_desc)
    pts = oracle.tiles_to_rows(tiles)[:nb_desc]
    centers = oracle.centers_words_to_rows(cwords)
    a, counts, sums, _ = oracle.assign_u8(pts, centers, prev)
    upd = oracle.update_u8(centers, counts, sums)
    q = sums.astype(np.float32) / np.maximum(counts, 1).astype(np.float32)[:, None]  # IEEE binary32 division
    expect = np.floor(q.astype(np.float64) + 0.5).astype(np.uint8)  # round half away from zero (q >= 0)
    expect[counts == 0] = centers[counts == 0]
    np.testing.assert_array_equal(upd, expect)


def test_update_rounding_identity():
    """uint8(roundf(float(s)/float(c))) (kmeans.cu:367, IEEE binary32 division, half away
    from zero) equals floor((2s + c) / (2c)) for 0 < c < 2^16 and 0 <= s <= 255 c: the
    persistent on-chip Lloyd kernel updates its centers with the integer form.  Every s
    for c <= 64, and for every c < 2^16 the sums next to each k + 1/2 boundary."""
    def ref(sv, cv):
        q = (sv.astype(np.float32) / cv.astype(np.float32)).astype(np.float64)  # binary32 quotient
        return np.floor(q + 0.5)  # roundf for q >= 0 (q + 0.5 is exact in binary64)

    for c in range(1, 65):
        sv = np.arange(0, 255 * c + 1, dtype=np.int64)
        cv = np.full_like(sv, c)
        np.testing.assert_array_equal((2 * sv + cv) // (2 * cv), ref(sv, cv))
    cv = np.arange(1, 1 << 16, dtype=np.int64)
    for k in range(0, 255):
        mid = (2 * k + 1) * cv // 2  # floor((k + 1/2) c)
        for d in (-1, 0, 1, 2):
            sv = np.clip(mid + d, 0, 255 * cv)
            np.testing.assert_array_equal((2 * sv + cv) // (2 * cv), ref(sv, cv), err_msg=f"k={k} d={d}")


@pytest.mark.parametrize("tile_size", [32, 16])  # ReformatTests.cpp:105-107,115-117
def test_ref_reformat_known_answer(tile_size):
    desc, idx = oracle.ref_reformat_test_inputs(136, 3, tile_size)
    tiles = oracle.reformat(desc, idx, tile_size)
    words = desc.view(np.uint32).reshape(136, 32)
    inv = idx == 0xFFFFFFFF
    assert 0 < inv.sum() < idx.size
    for t in range(3):
        for s in range(tile_size):
            exp = np.full(32, 0xFFFFFFFF, np.uint32) if inv[t, s] else words[idx[t, s]]
            np.testing.assert_array_equal(tiles[t, :, s], exp)


@pytest.mark.parametrize("n", [0, 1, 7, 15, 16, 17, 100, 255, 256, 257, 4000, 10 ** 6])
def test_rand_pick_invariants(n):
    r = oracle.Rng()
    for _ in range(3):
        dst, nb = r.rand_pick(n, 16)
        assert nb == min(n, 16)
        picks = dst[:nb]
        assert np.all(np.diff(picks.astype(np.int64)) > 0)  # sorted, distinct
        assert np.all(picks < max(n, 1))
        pad = dst[nb:]
        assert np.all(pad == (picks[-1] if nb else 0))


def test_rand_pick_rejection_branch_sequence():
    # n > 256: draws rng() % n until 16 distinct (kmeans.cpp:148-153): with a fresh engine the
    # picks are exactly the first 16 distinct residues of the engine's outputs.
    n = 10 ** 6
    r, ref = oracle.Rng(), oracle.Rng()
    dst, nb = r.rand_pick(n, 16)
    seen = []
    while len(seen) < 16:
        v = ref.next() % n
        if v not in seen:
            seen.append(v)
    np.testing.assert_array_equal(dst, np.sort(np.array(seen, np.uint32)))


def test_kmeanspp_invariants():
    rng = np.random.default_rng(5)
    desc = rng.integers(0, 256, (3000, 128), dtype=np.uint8)
    for n in (1, 5, 16, 40, 256, 2999):
        idx = np.arange(n, dtype=np.uint32) + 1
        pos, nv = oracle.kmeanspp_pick(desc, idx, 16, 7, 1, 3)
        assert nv == min(n, 16)
        assert len(set(pos[:nv].tolist())) == nv and np.all(pos < n)
        pos2, _ = oracle.kmeanspp_pick(desc, idx, 16, 7, 1, 3)
        np.testing.assert_array_equal(pos, pos2)
    # duplicates are never picked twice while a positive-weight point remains
    dup = np.repeat(desc[:4], 50, axis=0)
    pos, nv = oracle.kmeanspp_pick(dup, np.arange(200, dtype=np.uint32), 16, 3, 0, 0)
    assert nv == 4


def _numpy_build(desc, nb_doc, L, max_iters):
    """Independent re-implementation of Vocabulary.cpp:32-121 + kmeans.cpp (reference seed mode)."""
    N = desc.shape[0]
    r = oracle.Rng()
    n_nodes = level_beg(16, L)
    rows = np.zeros((n_nodes, 16, 128), np.uint8)
    masks = np.zeros(n_nodes, np.uint32)
    occ = np.zeros((n_nodes, 16), np.uint32)
    passes = np.zeros(n_nodes, np.uint32)
    groups = [np.arange(N)]
    for level in range(L):
        base = level_beg(16, level)
        seeds = []
        for g, grp in enumerate(groups):
            dst, nv = r.rand_pick(len(grp), 16)
            if len(grp):
                seeds.append(desc[grp[dst]].copy())
            else:
                nxt = next((gg[0] for gg in groups[g + 1:] if len(gg)), None)
                seeds.append(np.repeat((desc[nxt] if nxt is not None else np.full(128, 255, np.uint8))[None], 16, 0))
            masks[base + g] = (1 << nv) - 1
        new_groups = []
        for g, grp in enumerate(groups):
            c = seeds[g].astype(np.int64)
            x = desc[grp].astype(np.int64)
            a = np.full(len(grp), 255)
            for it in range(max_iters):
                passes[base + g] += 1
                na = np.argmin(((x[:, None] - c[None]) ** 2).sum(-1), axis=1) if len(grp) else a
                changed = int((na != a).sum())
                a = na
                if changed == 0:
                    break
                for j in range(16):
                    m = a == j
                    if m.any():
                        q = np.float32(x[m].sum(0)) / np.float32(m.sum())
                        c[j] = np.floor(q.astype(np.float64) + 0.5)
            rows[base + g] = c.astype(np.uint8)
            for j in range(16):
                child = grp[a == j] if len(grp) else grp
                occ[base + g, j] = len(child)
                new_groups.append(child)
        groups = new_groups
    q = np.float32(nb_doc) / np.maximum(occ, 1).astype(np.float32)
    w = np.log(q.astype(np.float64)).astype(np.float32)
    w = np.maximum(w, np.finfo(np.float32).tiny)
    w[occ == 0] = np.finfo(np.float32).tiny
    return rows, masks, occ, passes, w


def test_build_matches_independent_numpy():
    import rapidbow_amd as rb
    desc = rb.synth_descriptors(1500, seed=11, nb_centres=40)
    for L, iters in ((1, 5), (2, 6)):
        res = oracle.build(desc, 37, L, max_iters=iters)
        rows, masks, occ, passes, w = _numpy_build(desc, 37, L, iters)
        np.testing.assert_array_equal(oracle.ref_nodes_to_rows(res["nodes"]), rows)
        np.testing.assert_array_equal(res["masks"], masks)
        np.testing.assert_array_equal(res["occ"], occ)
        np.testing.assert_array_equal(res["passes"], passes)
        np.testing.assert_array_equal(res["weights"], w)


def test_build_voc_ref_check_full():
    # VocTests.cpp:86-132: 32 random descriptors, L=3 -> every descriptor's leaf centroid equals it
    desc = oracle.ref_voc_test_inputs(1)
    res = oracle.build(desc, 4, 3)
    leaf = oracle.lookup_u8(res["nodes"], 3, desc)
    rows = oracle.ref_nodes_to_rows(res["nodes"])
    cent = rows[level_beg(16, 2) + leaf // 16, leaf % 16]
    assert np.sum((cent.astype(np.int64) - desc) ** 2) == 0


def test_build_voc_ref_check_2d():
    # VocTests.cpp:38-84: RMS quantisation error <= 10 ("should be ~9.95")
    desc = oracle.ref_voc_test_inputs(0)
    res = oracle.build(desc, 16, 2)
    leaf = oracle.lookup_u8(res["nodes"], 2, desc)
    rows = oracle.ref_nodes_to_rows(res["nodes"])
    cent = rows[level_beg(16, 1) + leaf // 16, leaf % 16]
    inertia = int(np.sum((cent.astype(np.int64) - desc) ** 2))
    assert np.sqrt(inertia // 4096) <= 10


def test_build_f32_invariants():
    import rapidbow_amd as rb
    desc = rb.synth_descriptors(3000, seed=12, nb_centres=60)
    res = oracle.build(desc, 50, 1, b=32, max_iters=200, seed_mode=1, centroid_type=1)
    c = res["nodes"][0]
    assert res["passes"][0] < 200  # converged: the final centers are the means of their points
    leaf = oracle.lookup_f32(res["nodes"], 32, 1, desc)
    for j in range(32):
        m = leaf == j
        if m.any():
            np.testing.assert_array_equal(c[j], (desc[m].astype(np.int64).sum(0).astype(np.float32)
                                                 / np.float32(m.sum())))
    assert res["occ"][0].sum() == 3000
    np.testing.assert_array_equal(np.bincount(leaf, minlength=32), res["occ"][0])


def test_reject_bad_build_args():
    desc = np.zeros((15, 128), np.uint8)
    with pytest.raises(ValueError):
        oracle.build(desc, 1, 2)  # Insufficient descriptors
    with pytest.raises(ValueError):
        oracle.build(np.zeros((20, 128), np.uint8), 0, 2)  # nbDoc must be positive


# --------------------------------------------------------------- database --
def _tiny_db(seed, nb_docs=40, L=2, b=16):
    rng = np.random.default_rng(seed)
    n_nodes = level_beg(b, L)
    w = (rng.random((n_nodes, b)) * 3).astype(np.float32)
    w[rng.random(w.shape) < 0.1] = np.finfo(np.float32).tiny
    docs = {d: rng.integers(0, b ** L, rng.integers(1, 60)).astype(np.uint32) for d in range(nb_docs)}
    return w, docs


def _bow(leaf, w, b, L):
    """Explicit L1-normalised BoW over all tree nodes (the paper's vector), float64."""
    v = {}
    for t in range(L):
        lvl = L - 1 - t
        for x in leaf:
            node = (lvl, int(x) >> (t * int(np.log2(b))))
            v[node] = v.get(node, 0.0) + float(w[level_beg(b, lvl) + (node[1] >> int(np.log2(b))), node[1] % b])
    s = sum(v.values())
    return {k: val / s for k, val in v.items()}


def test_db_l1_identity_all_nodes_scored():
    w, docs = _tiny_db(21)
    db = oracle.DataBase(16, 2, w)
    for d, lf in docs.items():
        assert db.add(d, lf) == 0
    q = docs[7]
    K = 10 ** 6  # threshold 16 + 4K never binds: every shared node is scored
    got_d, got_s = db.query_fp64(q, K)
    qv = _bow(q, w, 16, 2)
    for d, s in zip(got_d, got_s):
        dv = _bow(docs[int(d)], w, 16, 2)
        l1 = sum(abs(qv.get(k, 0) - dv.get(k, 0)) for k in set(qv) | set(dv))
        assert abs(s - l1) < 1e-9
    d32, s32 = db.query(q, K)
    np.testing.assert_allclose(s32, got_s, atol=2e-6)
    assert d32[0] == 7 and got_d[0] == 7  # self-query ranks first (ApiTests.cpp:136)


def test_db_threshold_and_postings():
    w, docs = _tiny_db(22, nb_docs=80)
    db = oracle.DataBase(16, 2, w)
    for d, lf in docs.items():
        db.add(d, lf)
    assert db.add(3, docs[3]) == -1  # Duplicate document index
    allp = sum(len(v) for v in docs.values())
    assert db.nb_postings() == allp
    for leaf in (0, 17, 255):
        exp = sorted(d for d, lf in docs.items() for x in lf if x == leaf)
        np.testing.assert_array_equal(db.postings(leaf), np.array(exp, np.uint32))
    # with K=1 the inner-node threshold is 20 postings: nodes with >= 20 postings are skipped
    q = docs[5]
    K = 1
    s_all, s64, ntouch = db.query_all(q, K, 80)
    qs = np.sort(q)
    thr = 16 + 4 * K
    qrcp_l1 = sum(float(w[level_beg(16, 1) + (x >> 4), x % 16]) + float(w[0, x >> 4]) for x in qs)
    for d in range(80):
        lf = docs[d]
        acc = 0.0
        dl1 = sum(float(w[level_beg(16, 1) + (x >> 4), x % 16]) + float(w[0, x >> 4]) for x in lf)
        touched = False
        for t in range(2):
            for v in np.unique(qs >> (4 * t)):
                lo, hi = int(v) << (4 * t), (int(v) + 1) << (4 * t)
                P = sum(int(((docs[e] >= lo) & (docs[e] < hi)).sum()) for e in docs)
                if t == 1 and P >= thr:
                    continue
                m = int(((lf >= lo) & (lf < hi)).sum())
                if m == 0:
                    continue
                touched = True
                ww = float(w[0, v]) if t == 1 else float(w[level_beg(16, 1) + (v >> 4), v % 16])
                qE = ((qs >> (4 * t)) == v).sum() * ww / qrcp_l1
                tE = m * ww / dl1
                acc += abs(qE - tE) - (qE + tE)
        if touched:
            assert abs(s64[d] - (acc + 2)) < 1e-9, d
        else:
            assert np.isnan(s64[d])


def test_idf_weights_pinned():
    """Vocabulary.cpp:103-108 weights: the closed forms at occ = 0, occ >= nbDoc
    (log <= 0 -> FLT_MIN), and numpy's binary64 log of the binary32 quotient over
    every count to 2^16; a binary64 quotient or a binary32 log would differ."""
    tiny = np.finfo(np.float32).tiny
    occ = np.arange(1 << 16, dtype=np.uint32)
    for nb_doc in (1, 3, 1000, 65535, 1_000_000):
        w = oracle.idf_weights(occ, nb_doc)
        assert w[0] == tiny
        assert np.all(w[occ >= nb_doc] == tiny)
        o = occ[1:].astype(np.float32)
        q32 = (np.float32(nb_doc) / o).astype(np.float32)
        exp = np.maximum(tiny, np.log(q32.astype(np.float64)).astype(np.float32))
        np.testing.assert_array_equal(w[1:], exp)
        if nb_doc == 1000:
            q64 = np.float64(nb_doc) / o.astype(np.float64)
            assert np.any(np.maximum(tiny, np.log(q64).astype(np.float32)) != exp)  # the quotient's precision shows
            assert np.any(np.maximum(tiny, np.log(q32)) != exp)  # and so does the log's
    assert oracle.idf_weights(np.array([1], np.uint32), 1000)[0] == np.float32(np.log(1000.0))


def test_path_dist_pinned():
    """Distances along lookup paths: binary64 equals numpy's (to summation order), binary32 equals the
    FP32 chain of orc_dist_f32, and the child on the path is the one the leaf id
    names (level-l node leaf >> lg2b*(L-1-l), child its low bits)."""
    rng = np.random.default_rng(5)
    b, L = 4, 3
    n_nodes = (b ** L - 1) // (b - 1)
    nodes = (rng.random((n_nodes, b, 128)) * 255).astype(np.float32)
    desc = rng.integers(0, 256, (50, 128), dtype=np.uint8)
    leaf = rng.integers(0, b ** L, 50).astype(np.uint32)
    d32, d64 = oracle.path_dist_f32(nodes, b, L, desc, leaf)
    for i in range(50):
        for l in range(L):
            path = int(leaf[i]) >> (2 * (L - 1 - l))
            c = nodes[level_beg(b, l) + (path >> 2), path & 3]
            ref = float(((desc[i].astype(np.float64) - c.astype(np.float64)) ** 2).sum())  # pairwise order
            assert abs(d64[i, l] - ref) <= 1e-13 * ref
            assert d32[i, l] == oracle.dist_f32(desc[i], c)


def test_assert_ranked_equal_is_strict():
    """The ranking check accepts one-ulp score differences and the swaps they
    cause, and rejects a two-ulp error, a swap of distinct scores, an unranked
    list and a foreign document."""
    from tests.helpers import assert_ranked_equal
    od = np.array([5, 9, 2, 7], np.uint32)
    os_ = np.array([0.25, 0.5, 0.75, 1.0], np.float32)
    up = lambda x, k=1: np.float32(x) + np.float32(k) * np.spacing(np.float32(x))
    assert_ranked_equal(od, os_, od, os_)
    assert_ranked_equal(od, np.array([0.25, up(0.5), 0.75, 1.0], np.float32), od, os_)
    tied = np.array([0.25, 0.5, up(0.5), 1.0], np.float32)  # oracle: docs 9 and 2 one ulp apart
    assert_ranked_equal(np.array([5, 2, 9, 7], np.uint32), np.array([0.25, 0.5, 0.5, 1.0], np.float32),
                        np.array([5, 9, 2, 7], np.uint32), tied)
    bad = [
        (od, np.array([0.25, up(0.5, 2), 0.75, 1.0], np.float32)),                 # two ulps off
        (np.array([5, 2, 9, 7], np.uint32), np.array([0.25, 0.5, 0.75, 1.0], np.float32)),  # swap of distinct scores
        (np.array([9, 5, 2, 7], np.uint32), np.array([0.5, 0.25, 0.75, 1.0], np.float32)),  # not ranked
        (np.array([5, 9, 8, 7], np.uint32), os_),                                  # foreign doc at a distinct score
    ]
    for gd, gs in bad:
        with pytest.raises(AssertionError):
            assert_ranked_equal(gd, gs, od, os_)


@pytest.mark.parametrize("centroid_type", [0, 1])
def test_build_centroid_error_pinned(centroid_type):
    """centroid_err = max |centroid - binary64 mean of its points|: on a converged
    one-level build the members are the points the final centers' descent sends
    to each child, so numpy recomputes it from exact int64 sums."""
    rng = np.random.default_rng(31)
    desc = rng.integers(0, 256, (3000, 128), dtype=np.uint8)
    res = oracle.build(desc, 10, 1, max_iters=1000, seed_mode=1, centroid_type=centroid_type)
    assert res["passes"][0] < 1000  # converged: the assignment is the descent on the final centers
    if centroid_type == 0:
        cent = oracle.ref_nodes_to_rows(res["nodes"])[0].astype(np.float64)
        leaf = oracle.lookup_u8(res["nodes"], 1, desc)
    else:
        cent = res["nodes"][0].astype(np.float64)
        leaf = oracle.lookup_f32(res["nodes"], 16, 1, desc)
    err = 0.0
    for j in range(16):
        m = desc[leaf == j].astype(np.int64)
        if len(m):
            err = max(err, float(np.abs(cent[j] - m.sum(0) / len(m)).max()))
    assert res["centroid_err"] == err
    assert (err <= 0.5) if centroid_type == 0 else (0 < err < 1e-4)
