"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Used by tests/, the smoke check and the CPU
baseline leg of bench.py as the checker; the product never imports it.

Each C function cites the reference file:line it restates; the tests that
pin it are the golden fixtures (tests/golden), the known-answer and
closed-form checks of tests/test_oracle.py and the brute-force restatements
in tests/helpers.py.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
lib = ctypes.CDLL(LIB_PATH)

_vp, _u32, _u64, _i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype, f.argtypes = res, args


_sig("orc_level_beg", _u64, [_u32, _u32])
_sig("orc_dist_u8", _u32, [_vp, _vp])
_sig("orc_dist_f32", ctypes.c_float, [_vp, _vp])
_sig("orc_lookup_u8", None, [_vp, _u32, _vp, _u64, _vp, _i32])
_sig("orc_lookup_f32", None, [_vp, _u32, _u32, _vp, _u64, _vp, _i32])
_sig("orc_assign_u8", _u32, [_vp, _u32, _vp, _vp, _vp, _vp])
_sig("orc_update_u8", None, [_vp, _vp, _vp])
_sig("orc_assign_f32", _u32, [_vp, _u32, _vp, _u32, _vp, _vp, _vp])
_sig("orc_update_f32", None, [_vp, _u32, _vp, _vp])
_sig("orc_reformat", None, [_vp, _vp, _u32, _u32, _vp])
_sig("orc_rng_new", _vp, [_u64])
_sig("orc_rng_free", None, [_vp])
_sig("orc_rng_next", _u64, [_vp])
_sig("orc_rand_pick", _u32, [_vp, _vp, _u32, _u32])
_sig("orc_kmeanspp_pick", _u32, [_vp, _vp, _u32, _u32, _u64, _u32, _u32, _vp])
_sig("orc_hash", _u64, [_u64, _u64, _u64, _u64])
_sig("orc_build", _i32, [_vp, _u32, _u32, _u32, _u32, _u32, _i32, _u64, _i32, _vp, _vp, _vp, _vp, _vp, _vp])
_sig("orc_idf_weights", None, [_vp, _u64, _u32, _vp])
_sig("orc_path_dist_f32", None, [_vp, _u32, _u32, _vp, _u64, _vp, _vp, _vp, _i32])
_sig("orc_db_new", _vp, [_u32, _u32, _vp])
_sig("orc_db_free", None, [_vp])
_sig("orc_db_add", _i32, [_vp, _u32, _vp, _u32])
_sig("orc_db_rcp", ctypes.c_float, [_vp, _u32])
_sig("orc_db_nb_postings", _u64, [_vp])
_sig("orc_db_postings", _u32, [_vp, _u32, _vp, _u32])
_sig("orc_db_query", _u32, [_vp, _vp, _u32, _u32, _vp, _vp, _vp])
_sig("orc_db_query_fp64", _u32, [_vp, _vp, _u32, _u32, _vp, _vp])
_sig("orc_db_query_all", _u32, [_vp, _vp, _u32, _u32, _vp, _vp, _u32])
_sig("orc_ref_assign_test_inputs", None, [_u32, _vp, _vp, _vp])
_sig("orc_ref_reformat_test_inputs", None, [_u32, _u32, _u32, _vp, _vp])
_sig("orc_ref_voc_test_inputs", None, [_i32, _vp])


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def level_beg(b, level):
    return lib.orc_level_beg(b, level)


def nb_nodes(b, L):
    return level_beg(b, L)


def dist_u8(x, c):
    x, c = _c(x, np.uint8), _c(c, np.uint8)
    return lib.orc_dist_u8(x.ctypes.data, c.ctypes.data)


def dist_f32(x, c):
    x, c = _c(x, np.uint8), _c(c, np.float32)
    return lib.orc_dist_f32(x.ctypes.data, c.ctypes.data)


def lookup_u8(nodes_ref, L, desc, nthreads=0):
    nodes_ref, desc = _c(nodes_ref, np.uint8), _c(desc, np.uint8)
    out = np.empty(desc.shape[0], np.uint32)
    lib.orc_lookup_u8(nodes_ref.ctypes.data, L, desc.ctypes.data, desc.shape[0], out.ctypes.data, nthreads)
    return out


def lookup_f32(nodes, b, L, desc, nthreads=0):
    nodes, desc = _c(nodes, np.float32), _c(desc, np.uint8)
    out = np.empty(desc.shape[0], np.uint32)
    lib.orc_lookup_f32(nodes.ctypes.data, b, L, desc.ctypes.data, desc.shape[0], out.ctypes.data, nthreads)
    return out


def assign_u8(pts, centers, prev):
    pts, centers = _c(pts, np.uint8), _c(centers, np.uint8)
    a = _c(prev, np.uint8).copy()
    counts = np.zeros(16, np.uint32)
    sums = np.zeros((16, 128), np.int64)
    ch = lib.orc_assign_u8(pts.ctypes.data, pts.shape[0], centers.ctypes.data, a.ctypes.data, counts.ctypes.data,
                           sums.ctypes.data)
    return a, counts, sums, ch


def update_u8(centers, counts, sums):
    c = _c(centers, np.uint8).copy()
    lib.orc_update_u8(c.ctypes.data, _c(counts, np.uint32).ctypes.data, _c(sums, np.int64).ctypes.data)
    return c


def assign_f32(pts, centers, prev):
    pts, centers = _c(pts, np.uint8), _c(centers, np.float32)
    k = centers.shape[0]
    a = _c(prev, np.uint8).copy()
    counts = np.zeros(k, np.uint32)
    sums = np.zeros((k, 128), np.int64)
    ch = lib.orc_assign_f32(pts.ctypes.data, pts.shape[0], centers.ctypes.data, k, a.ctypes.data,
                            counts.ctypes.data, sums.ctypes.data)
    return a, counts, sums, ch


def update_f32(centers, counts, sums):
    c = _c(centers, np.float32).copy()
    counts, sums = _c(counts, np.uint32), _c(sums, np.int64)
    lib.orc_update_f32(c.ctypes.data, c.shape[0], counts.ctypes.data, sums.ctypes.data)
    return c


def reformat(desc, idx, tile_size):
    desc, idx = _c(desc, np.uint8), _c(idx, np.uint32)
    nb_tiles = idx.size // tile_size
    out = np.empty((nb_tiles, 32, tile_size), np.uint32)
    lib.orc_reformat(desc.ctypes.data, idx.ctypes.data, nb_tiles, tile_size, out.ctypes.data)
    return out


class Rng:
    """std::mt19937_64 shared across randPick calls (kmeans.cpp:145)."""

    def __init__(self, seed=5489):
        self.h = lib.orc_rng_new(seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_rng_free(self.h)

    def next(self):
        return lib.orc_rng_next(self.h)

    def rand_pick(self, rand_space, nb_to_pick=16):
        dst = np.zeros(nb_to_pick, np.uint32)
        n = lib.orc_rand_pick(self.h, dst.ctypes.data, rand_space, nb_to_pick)
        return dst, n


def kmeanspp_pick(desc, grp_idx, k, seed, level, grp):
    desc, grp_idx = _c(desc, np.uint8), _c(grp_idx, np.uint32)
    out = np.zeros(k, np.uint32)
    n = lib.orc_kmeanspp_pick(desc.ctypes.data, grp_idx.ctypes.data, grp_idx.size, k, seed, level, grp,
                              out.ctypes.data)
    return out, n


def hash4(seed, a, b, c):
    return lib.orc_hash(seed, a, b, c)


def build(desc, nb_doc, L, b=16, max_iters=1000, seed_mode=0, seed=5489, centroid_type=0):
    """Returns dict(nodes, masks, weights, occ, passes, centroid_err); u8 nodes in the reference
    layout [node][2048], float nodes [node][b][128]; centroid_err = max |centroid - binary64 mean|."""
    desc = _c(desc, np.uint8)
    n = nb_nodes(b, L)
    nodes = np.zeros((n, 2048), np.uint8) if centroid_type == 0 else np.zeros((n, b, 128), np.float32)
    masks = np.zeros(n, np.uint32)
    weights = np.zeros((n, b), np.float32)
    occ = np.zeros((n, b), np.uint32)
    passes = np.zeros(n, np.uint32)
    err = ctypes.c_double(0)
    r = lib.orc_build(desc.ctypes.data, desc.shape[0], nb_doc, b, L, max_iters, seed_mode, seed, centroid_type,
                      nodes.ctypes.data, masks.ctypes.data, weights.ctypes.data, occ.ctypes.data,
                      passes.ctypes.data, ctypes.byref(err))
    if r != 0:
        raise ValueError("orc_build rejected its arguments")
    return dict(nodes=nodes, masks=masks, weights=weights, occ=occ, passes=passes, centroid_err=err.value)


def idf_weights(occ, nb_doc):
    """Vocabulary.cpp:103-108 weights of the given occurrence counts."""
    occ = _c(occ, np.uint32)
    w = np.zeros(occ.size, np.float32)
    lib.orc_idf_weights(occ.ctypes.data, occ.size, nb_doc, w.ctypes.data)
    return w


def path_dist_f32(nodes, b, L, desc, leaf, nthreads=0):
    """(d32, d64) [n][L]: FP32-chain and binary64 distances to the child on each descriptor's path."""
    nodes, desc, leaf = _c(nodes, np.float32), _c(desc, np.uint8), _c(leaf, np.uint32)
    n = desc.shape[0]
    d32 = np.zeros((n, L), np.float32)
    d64 = np.zeros((n, L), np.float64)
    lib.orc_path_dist_f32(nodes.ctypes.data, b, L, desc.ctypes.data, n, leaf.ctypes.data, d32.ctypes.data,
                          d64.ctypes.data, nthreads)
    return d32, d64


def ref_nodes_to_rows(nodes_ref):
    """[node][2048] reference Node layout -> [node][16][128] child-major rows."""
    a = _c(nodes_ref, np.uint8).reshape(-1, 32, 16, 4)  # [node][word][child][byte]
    return a.transpose(0, 2, 1, 3).reshape(-1, 16, 128)


def rows_to_ref_nodes(rows):
    a = _c(rows, np.uint8).reshape(-1, 16, 32, 4)
    return a.transpose(0, 2, 1, 3).reshape(-1, 2048)


class DataBase:
    """DataBase (DataBase.h:35-64) restated on the CPU."""

    def __init__(self, b, L, weights):
        w = _c(weights, np.float32)
        self.b, self.L = b, L
        self.h = lib.orc_db_new(b, L, w.ctypes.data)
        self._keep = w

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_db_free(self.h)

    def add(self, doc, leaf):
        leaf = _c(leaf, np.uint32)
        return lib.orc_db_add(self.h, doc, leaf.ctypes.data, leaf.size)

    def rcp(self, doc):
        return lib.orc_db_rcp(self.h, doc)

    def nb_postings(self):
        return lib.orc_db_nb_postings(self.h)

    def postings(self, leaf):
        n = lib.orc_db_postings(self.h, leaf, None, 0)
        out = np.zeros(n, np.uint32)
        lib.orc_db_postings(self.h, leaf, out.ctypes.data, n)
        return out

    def query(self, leaf, K, with_touched=False):
        leaf = _c(leaf, np.uint32)
        docs = np.zeros(max(K, 1), np.uint32)
        scores = np.zeros(max(K, 1), np.float32)
        touched = ctypes.c_uint32(0)
        n = lib.orc_db_query(self.h, leaf.ctypes.data, leaf.size, K, docs.ctypes.data, scores.ctypes.data,
                             ctypes.byref(touched))
        if with_touched:
            return docs[:n], scores[:n], touched.value
        return docs[:n], scores[:n]

    def query_fp64(self, leaf, K):
        leaf = _c(leaf, np.uint32)
        docs = np.zeros(max(K, 1), np.uint32)
        scores = np.zeros(max(K, 1), np.float64)
        n = lib.orc_db_query_fp64(self.h, leaf.ctypes.data, leaf.size, K, docs.ctypes.data, scores.ctypes.data)
        return docs[:n], scores[:n]

    def query_all(self, leaf, K, nb_docs):
        leaf = _c(leaf, np.uint32)
        s32 = np.zeros(nb_docs, np.float32)
        s64 = np.zeros(nb_docs, np.float64)
        n = lib.orc_db_query_all(self.h, leaf.ctypes.data, leaf.size, K, s32.ctypes.data, s64.ctypes.data, nb_docs)
        return s32, s64, n


def ref_assign_test_inputs(nb_desc):
    """tests/KmeansAssignCenterTests.cpp:36-73 -> (tiles u32[t][32][32], centers u32[32][16], prev u8[n])."""
    nb_tiles = (nb_desc + 31) // 32
    tiles = np.zeros((nb_tiles, 32, 32), np.uint32)
    centers = np.zeros((32, 16), np.uint32)
    prev = np.zeros(nb_desc, np.uint8)
    lib.orc_ref_assign_test_inputs(nb_desc, tiles.ctypes.data, centers.ctypes.data, prev.ctypes.data)
    return tiles, centers, prev


def ref_reformat_test_inputs(nb_desc, nb_tiles, tile_size):
    desc = np.zeros((nb_desc, 128), np.uint8)
    idx = np.zeros((nb_tiles, tile_size), np.uint32)
    lib.orc_ref_reformat_test_inputs(nb_desc, nb_tiles, tile_size, desc.ctypes.data, idx.ctypes.data)
    return desc, idx


def ref_voc_test_inputs(which):
    n = 4096 if which == 0 else 32
    desc = np.zeros((n, 128), np.uint8)
    lib.orc_ref_voc_test_inputs(which, desc.ctypes.data)
    return desc


def tiles_to_rows(tiles):
    """[tiles][32 words][32 slots] u32 (DescTile, kmeans.cuh:53-60) -> [tiles*32][128] uint8 rows."""
    t = _c(tiles, np.uint32)
    return t.transpose(0, 2, 1).copy().view(np.uint8).reshape(-1, 128)


def centers_words_to_rows(centers_words):
    """[32 words][16 centers] u32 (Centers, kmeans.cuh:58) -> [16][128] uint8."""
    c = _c(centers_words, np.uint32)
    return c.T.copy().view(np.uint8).reshape(16, 128)
