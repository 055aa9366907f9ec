"""Python binding of the RapidBoW MI355X library (librapidbow_amd.so).

Mirrors the reference's public surface (RapidBoW.h:29-109): build a
vocabulary from device descriptors, look descriptors up into leaf ids,
serialise / deserialise, and the inverted-file database.  All compute runs in
the HIP library; importing this package fails loudly when it is not built.
"""
import ctypes
import os

import numpy as np

from . import hip
from .hip import DeviceArray, Stream, Event

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librapidbow_amd.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
lib = ctypes.CDLL(LIB_PATH)

_vp, _u32, _u64, _i32, _sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_size_t

# status codes (include/rapidbow.h)
OK, ERR_INVALID_ARG, ERR_UNSUPPORTED, ERR_INSUFFICIENT_DESC, ERR_NB_DOC, ERR_INVALID_BLOB, ERR_DUPLICATE_DOC, \
    ERR_HIP, ERR_OUT_OF_MEMORY, ERR_CAPACITY = range(10)
LOOKUP_AUTO, LOOKUP_PLAIN, LOOKUP_MFMA = 0, 1, 2


class DescriptorAttributes(ctypes.Structure):
    _fields_ = [("data_type", ctypes.c_int32), ("distance_type", ctypes.c_int32), ("nb_dims", ctypes.c_uint64)]


SIFT_ATTR = DescriptorAttributes(1, 2, 128)  # RapidBoW.h:50 SiftAttr


class BuildStats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("assign_ms", ctypes.c_double), ("update_ms", ctypes.c_double),
                ("point_passes", ctypes.c_uint64), ("assign_ops", ctypes.c_uint64), ("reduce_ms", ctypes.c_double),
                ("assign_launches", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("stream_ms", ctypes.c_double),
                ("stream_point_passes", ctypes.c_uint64), ("persist_ms", ctypes.c_double)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_uint32), ("nb_streams", ctypes.c_uint32), ("mem_budget", ctypes.c_uint64),
                ("max_results", ctypes.c_uint32), ("lookup_kernel", ctypes.c_int32),
                ("staging_bytes", ctypes.c_uint64), ("use_graphs", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("batch_max_desc", ctypes.c_uint64)]


class BuildOptions(ctypes.Structure):
    _fields_ = [("max_iters", ctypes.c_uint32), ("seed_mode", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("centroid_type", ctypes.c_uint32), ("kernel", ctypes.c_uint32),
                ("persist_max_tiles", ctypes.c_uint32), ("timing", ctypes.c_uint32)]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("rbow_last_error", ctypes.c_char_p, [])
_sig("rbow_build_options_default", None, [ctypes.POINTER(BuildOptions)])
_sig("rbow_build_vocabulary", _i32, [DescriptorAttributes, _u32, _vp, _u32, _u32, _u32, _vp, ctypes.POINTER(_vp)])
_sig("rbow_build_vocabulary_ex", _i32,
     [DescriptorAttributes, _u32, _vp, _u32, _u32, _u32, ctypes.POINTER(BuildOptions), _vp, ctypes.POINTER(_vp)])
_sig("rbow_voc_destroy", None, [_vp])
_sig("rbow_release_build_workspaces", _i32, [])
_sig("rbow_tune_build_options", _i32,
     [DescriptorAttributes, _u32, _vp, _u32, _u32, _u32, ctypes.POINTER(BuildOptions), ctypes.POINTER(ctypes.c_float), _vp])
_sig("rbow_voc_get_attributes", _i32, [_vp, ctypes.POINTER(DescriptorAttributes)])
_sig("rbow_voc_branch_factor", _u32, [_vp])
_sig("rbow_voc_nb_levels", _u32, [_vp])
_sig("rbow_voc_centroid_type", _u32, [_vp])
_sig("rbow_voc_lookup", _i32, [_vp, _vp, _u32, _vp, _vp])
_sig("rbow_voc_lookup_ex", _i32, [_vp, _vp, _u32, _vp, _i32, _vp, _vp])
_sig("rbow_voc_serialize", _i32, [_vp, _vp, _sz, ctypes.POINTER(_sz)])
_sig("rbow_voc_deserialize", _i32, [_vp, _sz, ctypes.POINTER(_vp)])
_sig("rbow_voc_level_weights", ctypes.POINTER(ctypes.c_float), [_vp, _u32])
_sig("rbow_voc_build_stats", _i32, [_vp, ctypes.POINTER(BuildStats)])
_sig("rbow_voc_build_passes", ctypes.POINTER(ctypes.c_uint32), [_vp])
_sig("rbow_voc_copy_centroids", _i32, [_vp, _u32, _vp, _sz])
_sig("rbow_kmeans_assign", _i32, [_vp, _vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _i32, _vp])
_sig("rbow_kmeans_update", _i32, [_vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp])
_sig("rbow_reformat_tiles", _i32, [_vp, _vp, _u32, _u32, _vp, _vp])
_sig("rbow_gather_rows", _i32, [_vp, _vp, _u64, _vp, _vp])
_sig("rbow_voc_tune_lookup", _i32, [_vp, _vp, _u32, ctypes.POINTER(_u32), _vp])
_sig("rbow_voc_set_lookup_crossover", _i32, [_vp, _u32])
_sig("rbow_voc_lookup_crossover", _u32, [_vp])
_sig("rbow_voc_set_lookup_fused_max", _i32, [_vp, _u32])
_sig("rbow_voc_lookup_fused_max", _u32, [_vp])
_sig("rbow_voc_set_lookup_leaf_item_rows", _i32, [_vp, _u32])
_sig("rbow_voc_lookup_leaf_item_rows", _u32, [_vp])
_sig("rbow_db_create", _i32, [_vp, ctypes.POINTER(_vp)])
_sig("rbow_db_destroy", None, [_vp])
_sig("rbow_db_add_doc", _i32, [_vp, _u32, _vp, _u32])
_sig("rbow_db_query", _i32, [_vp, _vp, _u32, _u32, _vp, _vp, ctypes.POINTER(_u32)])
_sig("rbow_db_query_and_add_doc", _i32, [_vp, _u32, _vp, _u32, _u32, _vp, ctypes.POINTER(_u32)])
_sig("rbow_db_add_docs_device", _i32, [_vp, _vp, _u32, _vp, _vp, _vp])
_sig("rbow_db_query_batch", _i32, [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _vp])
_sig("rbow_db_set_walk_steps", _i32, [_vp, _u32])
_sig("rbow_db_walk_steps", _u32, [_vp])
_sig("rbow_db_tune_scoring", _i32, [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp])
_sig("rbow_db_nb_docs", _u32, [_vp])
_sig("rbow_db_nb_postings", _u64, [_vp])
_sig("rbow_db_nb_delta_postings", _u64, [_vp])
_sig("rbow_db_compact", _i32, [_vp, _vp])
_sig("rbow_idf_weights", _i32, [_vp, _u64, _u32, _vp, _vp])
_sig("rbow_db_postings", _i32, [_vp, _u32, _vp, _u32, ctypes.POINTER(_u32)])
_sig("rbow_db_doc_rcp", _i32, [_vp, _u32, ctypes.POINTER(ctypes.c_float)])
_sig("rbow_server_options_default", None, [ctypes.POINTER(ServerOptions)])
_sig("rbow_server_create", _i32, [_vp, _vp, ctypes.POINTER(ServerOptions), ctypes.POINTER(_vp)])
_sig("rbow_server_destroy", None, [_vp])
_sig("rbow_server_submit", _i32, [_vp, _vp, _u32, ctypes.POINTER(_u64)])
_sig("rbow_server_wait", _i32, [_vp, _u64, _vp, _vp, ctypes.POINTER(_u32), ctypes.POINTER(ctypes.c_double)])
_sig("rbow_server_replay", _i32, [_vp, _vp, _vp, _u32, ctypes.c_double, _u64, _vp, _vp, _vp, _vp, _vp,
                                  ctypes.POINTER(ctypes.c_double)])
_sig("rbow_pipeline_create", _i32, [_vp, _vp, _u32, _u64, _u32, ctypes.POINTER(_vp)])
_sig("rbow_pipeline_destroy", None, [_vp])
_sig("rbow_pipeline_run", _i32, [_vp, _vp, _vp, _u32, _i32, _i32, _vp, _vp, _vp, _vp])
_sig("rbow_pipeline_graphs", _u32, [_vp])
_sig("rbow_synth_centres", None, [_u64, _u32, _vp])
_sig("rbow_synth_descriptors_device", _i32, [_vp, _vp, _u32, _u64, _u64, _u64, _vp, _vp, _vp])
_sig("rbow_synth_image_centres_device", _i32, [_vp, _u32, _u64, _u64, _u64, _u32, _u32, _vp, _u64, _vp])
_sig("rbow_synth_image_postings_device", _i32, [_vp, _u32, _u64, _u64, _u64, _u32, _u32, _vp, _u32, _vp, _u64, _vp])
_sig("rbow_synth_descriptors", None, [_u64, _u32, _u64, _u64, _u64, _vp, ctypes.c_float, _vp])


class RbowError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


def check(status):
    if status != OK:
        raise RbowError(status, lib.rbow_last_error().decode())


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    return int(x)


def _stream(s):
    return s.handle if isinstance(s, Stream) else s


class Vocabulary:
    """IVocabulary (RapidBoW.h:60-71)."""

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        if getattr(self, "handle", None):
            lib.rbow_voc_destroy(self.handle)
            self.handle = None

    @property
    def branch_factor(self):
        return lib.rbow_voc_branch_factor(self.handle)

    @property
    def nb_levels(self):
        return lib.rbow_voc_nb_levels(self.handle)

    @property
    def centroid_type(self):
        return lib.rbow_voc_centroid_type(self.handle)

    @property
    def nb_nodes(self):
        b, L = self.branch_factor, self.nb_levels
        return (b ** L - 1) // (b - 1)

    def attributes(self):
        a = DescriptorAttributes()
        check(lib.rbow_voc_get_attributes(self.handle, ctypes.byref(a)))
        return a

    def lookup(self, dev_desc, nb_desc, dev_leaf, kernel=LOOKUP_AUTO, dev_near_ties=None, stream=None):
        check(lib.rbow_voc_lookup_ex(self.handle, _ptr(dev_desc), nb_desc, _ptr(dev_leaf), kernel,
                                     _ptr(dev_near_ties), _stream(stream)))

    def lookup_host(self, desc, kernel=LOOKUP_AUTO, return_near_ties=False):
        """Convenience: host uint8[n][128] -> host uint32[n] leaf ids (synchronous)."""
        desc = np.ascontiguousarray(desc, dtype=np.uint8)
        n = desc.shape[0]
        d = DeviceArray.from_numpy(desc) if n else None
        out = DeviceArray((max(n, 1),), np.uint32)
        ties = DeviceArray.zeros((1,), np.uint32)
        self.lookup(d, n, out, kernel, ties)
        hip.synchronize()
        leaf = out.numpy(n)
        return (leaf, int(ties.numpy()[0])) if return_near_ties else leaf

    def tune_lookup(self, dev_desc, nb_desc, stream=None):
        """Autotune the AUTO lookup crossover on these descriptors; returns it."""
        c = _u32()
        check(lib.rbow_voc_tune_lookup(self.handle, _ptr(dev_desc), nb_desc, ctypes.byref(c), _stream(stream)))
        return c.value

    @property
    def lookup_crossover(self):
        return lib.rbow_voc_lookup_crossover(self.handle)

    @lookup_crossover.setter
    def lookup_crossover(self, n):
        check(lib.rbow_voc_set_lookup_crossover(self.handle, n))

    @property
    def lookup_fused_max(self):
        """uint8 b16 trees: MFMA-form batches up to this size run the fused single-launch descent."""
        return lib.rbow_voc_lookup_fused_max(self.handle)

    @lookup_fused_max.setter
    def lookup_fused_max(self, n):
        check(lib.rbow_voc_set_lookup_fused_max(self.handle, n))

    @property
    def lookup_leaf_item_rows(self):
        """Rows per work item at the leaf level of the bucketed MFMA descent (256 or 512)."""
        return lib.rbow_voc_lookup_leaf_item_rows(self.handle)

    @lookup_leaf_item_rows.setter
    def lookup_leaf_item_rows(self, n):
        check(lib.rbow_voc_set_lookup_leaf_item_rows(self.handle, n))

    def serialize(self):
        size = _sz(0)
        check(lib.rbow_voc_serialize(self.handle, None, 0, ctypes.byref(size)))
        buf = (ctypes.c_uint8 * size.value)()
        check(lib.rbow_voc_serialize(self.handle, buf, size.value, ctypes.byref(size)))
        return bytes(buf)

    def level_weights(self, level):
        p = lib.rbow_voc_level_weights(self.handle, level)
        if not p:
            raise IndexError(level)
        n = self.branch_factor ** (level + 1)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def weights(self):
        return np.concatenate([self.level_weights(l) for l in range(self.nb_levels)])

    def build_stats(self):
        st = BuildStats()
        check(lib.rbow_voc_build_stats(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in BuildStats._fields_}

    def build_passes(self):
        p = lib.rbow_voc_build_passes(self.handle)
        return None if not p else np.ctypeslib.as_array(p, shape=(self.nb_nodes,)).copy()

    def centroids(self, level):
        n = self.branch_factor ** (level + 1) * 128
        out = np.empty(n, np.float32)
        check(lib.rbow_voc_copy_centroids(self.handle, level, out.ctypes.data, n))
        return out.reshape(-1, self.branch_factor, 128)


def build_options(max_iters=1000, seed_mode=0, seed=5489, centroid_type=0, kernel=0, persist_max_tiles=None,
                  timing=False):
    o = BuildOptions()
    lib.rbow_build_options_default(ctypes.byref(o))
    o.max_iters, o.seed_mode, o.seed, o.centroid_type, o.kernel = max_iters, seed_mode, seed, centroid_type, kernel
    o.timing = int(bool(timing))
    if persist_max_tiles is not None:
        o.persist_max_tiles = persist_max_tiles
    return o


def build_vocabulary(dev_desc, nb_desc, nb_doc, nb_levels, branch_factor=16, stream=None, attr=SIFT_ATTR, **opts):
    """buildVocabulary (RapidBoW.h:82-88); keyword options -> rbow_build_options."""
    h = _vp()
    o = build_options(**opts)
    check(lib.rbow_build_vocabulary_ex(attr, branch_factor, _ptr(dev_desc), nb_desc, nb_doc, nb_levels,
                                       ctypes.byref(o), _stream(stream), ctypes.byref(h)))
    return Vocabulary(h.value)


def tune_build_options(dev_desc, nb_desc, nb_doc, nb_levels, branch_factor=16, stream=None, attr=SIFT_ATTR, **opts):
    """Build autotuner: the fastest persist_max_tiles for these data (see rbow_tune_build_options).
    Returns (tuned persist_max_tiles, {candidate: build ms})."""
    o = build_options(**opts)
    ms = (ctypes.c_float * 5)()
    check(lib.rbow_tune_build_options(attr, branch_factor, _ptr(dev_desc), nb_desc, nb_doc, nb_levels, ctypes.byref(o),
                                      ms, _stream(stream)))
    return o.persist_max_tiles, {c: float(m) for c, m in zip((0, 4, 8, 12, 16), ms)}


def release_build_workspaces():
    """Frees the work buffers that completed builds keep per device for the next build."""
    check(lib.rbow_release_build_workspaces())


def deserialize_vocabulary(blob):
    """deserializeVocabulary (RapidBoW.h:90)."""
    h = _vp()
    b = np.frombuffer(bytes(blob), np.uint8)
    check(lib.rbow_voc_deserialize(b.ctypes.data if b.size else None, b.size, ctypes.byref(h)))
    return Vocabulary(h.value)


def kmeans_assign(dev_points, task_cnt, b, centroid_type, dev_centers, dev_assign, dev_counts, dev_sums,
                  dev_changed, kernel=0, stream=None):
    tc = np.ascontiguousarray(task_cnt, dtype=np.uint32)
    check(lib.rbow_kmeans_assign(_ptr(dev_points), tc.ctypes.data, tc.size, b, centroid_type, _ptr(dev_centers),
                                 _ptr(dev_assign), _ptr(dev_counts), _ptr(dev_sums), _ptr(dev_changed), kernel,
                                 _stream(stream)))


def kmeans_update(task_cnt, b, centroid_type, dev_centers, dev_counts, dev_sums, dev_changed, stream=None):
    tc = np.ascontiguousarray(task_cnt, dtype=np.uint32)
    check(lib.rbow_kmeans_update(tc.ctypes.data, tc.size, b, centroid_type, _ptr(dev_centers), _ptr(dev_counts),
                                 _ptr(dev_sums), _ptr(dev_changed), _stream(stream)))


def reformat_tiles(dev_desc, dev_idx, nb_tiles, tile_size, dev_tiles, stream=None):
    check(lib.rbow_reformat_tiles(_ptr(dev_desc), _ptr(dev_idx), nb_tiles, tile_size, _ptr(dev_tiles),
                                  _stream(stream)))


def idf_weights(dev_occ, n, nb_doc, dev_weights, stream=None):
    check(lib.rbow_idf_weights(_ptr(dev_occ), n, nb_doc, _ptr(dev_weights), _stream(stream)))


def gather_rows(dev_desc, dev_idx, n, dev_out, stream=None):
    check(lib.rbow_gather_rows(_ptr(dev_desc), _ptr(dev_idx), n, _ptr(dev_out), _stream(stream)))


def synth_centres(centre_seed=0, nb_centres=2000):
    out = np.empty((nb_centres, 128), np.float32)
    lib.rbow_synth_centres(centre_seed, nb_centres, out.ctypes.data)
    return out


def synth_descriptors(n, seed, first=0, centre_seed=0, nb_centres=2000, centre_ids=None, sigma_scale=0.3):
    """Seeded SIFT-like uint8[n][128] (host)."""
    out = np.empty((n, 128), np.uint8)
    ids = None
    if centre_ids is not None:
        ids = np.ascontiguousarray(centre_ids, dtype=np.uint32)
        assert ids.size == n
    lib.rbow_synth_descriptors(centre_seed, nb_centres, seed, first, n, None if ids is None else ids.ctypes.data,
                               sigma_scale, out.ctypes.data)
    return out


class DeviceSynth:
    """SIFT-like data generated directly in HBM (same distribution as synth_descriptors)."""

    def __init__(self, centre_seed=0, nb_centres=2000, sigma_scale=0.3):
        c = synth_centres(centre_seed, nb_centres)
        sigma = (sigma_scale * np.sqrt((c.astype(np.float64) ** 2).sum(1)) / np.sqrt(128.0)).astype(np.float32)
        self.nb_centres = nb_centres
        self.centres = DeviceArray.from_numpy(c)
        self.sigma = DeviceArray.from_numpy(sigma)

    def descriptors(self, n, seed, first=0, dev_centre_ids=None, out=None, stream=None):
        out = out if out is not None else DeviceArray((max(n, 1), 128), np.uint8)
        check(lib.rbow_synth_descriptors_device(self.centres.ptr, self.sigma.ptr, self.nb_centres, seed, first, n,
                                                _ptr(dev_centre_ids), out.ptr, _stream(stream)))
        return out

    def image_centre_ids(self, dev_offsets, nb_images, first_image, image_seed, desc_seed, nb_desc, subset=64,
                         stream=None):
        ids = DeviceArray((max(nb_desc, 1),), np.uint32)
        check(lib.rbow_synth_image_centres_device(_ptr(dev_offsets), nb_images, first_image, image_seed, desc_seed,
                                                  self.nb_centres, subset, ids.ptr, nb_desc, _stream(stream)))
        return ids

    def image_postings(self, dev_offsets, nb_images, first_image, image_seed, post_seed, dev_cloud, cloud_size,
                       nb_postings, subset=64, out=None, stream=None):
        out = out if out is not None else DeviceArray((max(nb_postings, 1),), np.uint32)
        check(lib.rbow_synth_image_postings_device(_ptr(dev_offsets), nb_images, first_image, image_seed, post_seed,
                                                   self.nb_centres, subset, _ptr(dev_cloud), cloud_size, out.ptr,
                                                   nb_postings, _stream(stream)))
        return out


class DataBase:
    """IDataBase (RapidBoW.h:92-105) on the GPU; borrows `voc`."""

    def __init__(self, voc):
        h = _vp()
        check(lib.rbow_db_create(voc.handle, ctypes.byref(h)))
        self.handle, self.voc = h.value, voc

    def __del__(self):
        if getattr(self, "handle", None):
            lib.rbow_db_destroy(self.handle)
            self.handle = None

    def add_doc(self, doc, leaf):
        leaf = np.ascontiguousarray(leaf, np.uint32)
        check(lib.rbow_db_add_doc(self.handle, doc, leaf.ctypes.data if leaf.size else None, leaf.size))

    def query(self, leaf, K):
        leaf = np.ascontiguousarray(leaf, np.uint32)
        docs = np.zeros(max(K, 1), np.uint32)
        sc = np.zeros(max(K, 1), np.float32)
        n = ctypes.c_uint32()
        check(lib.rbow_db_query(self.handle, leaf.ctypes.data if leaf.size else None, leaf.size, K, docs.ctypes.data,
                                sc.ctypes.data, ctypes.byref(n)))
        return docs[:n.value], sc[:n.value]

    def query_and_add(self, doc, leaf, K):
        """queryAndAddDoc (RapidBoW.h:102-103): ranked doc ids of the query, then the insert."""
        leaf = np.ascontiguousarray(leaf, np.uint32)
        docs = np.zeros(max(K, 1), np.uint32)
        n = ctypes.c_uint32()
        check(lib.rbow_db_query_and_add_doc(self.handle, doc, leaf.ctypes.data if leaf.size else None, leaf.size, K,
                                            docs.ctypes.data, ctypes.byref(n)))
        return docs[:n.value]

    def add_docs_device(self, doc_ids, dev_leaf, offsets, stream=None):
        ids = np.ascontiguousarray(doc_ids, np.uint32)
        off = np.ascontiguousarray(offsets, np.uint64)
        check(lib.rbow_db_add_docs_device(self.handle, ids.ctypes.data, ids.size, _ptr(dev_leaf), off.ctypes.data,
                                          _stream(stream)))

    def query_batch(self, dev_leaf, offsets, K, dev_docs, dev_scores, dev_counts, stream=None):
        off = np.ascontiguousarray(offsets, np.uint64)
        check(lib.rbow_db_query_batch(self.handle, _ptr(dev_leaf), off.ctypes.data, off.size - 1, K, _ptr(dev_docs),
                                      _ptr(dev_scores), _ptr(dev_counts), _stream(stream)))

    @property
    def walk_steps(self):
        """Scoring tile shape: steps of 64 postings per wave and walk iteration (4 or 6)."""
        return lib.rbow_db_walk_steps(self.handle)

    @walk_steps.setter
    def walk_steps(self, n):
        check(lib.rbow_db_set_walk_steps(self.handle, n))

    def tune_scoring(self, dev_leaf, offsets, K, stream=None):
        """Times one query batch with each scoring tile shape and keeps the faster:
        returns (walk steps, {steps: median ms})."""
        off = np.ascontiguousarray(offsets, np.uint64)
        steps = ctypes.c_uint32()
        ms = (ctypes.c_float * 2)()
        check(lib.rbow_db_tune_scoring(self.handle, _ptr(dev_leaf), off.ctypes.data, off.size - 1, K,
                                       ctypes.byref(steps), ms, _stream(stream)))
        return steps.value, {4: ms[0], 6: ms[1]}

    @property
    def nb_docs(self):
        return lib.rbow_db_nb_docs(self.handle)

    @property
    def nb_postings(self):
        return lib.rbow_db_nb_postings(self.handle)

    @property
    def nb_delta_postings(self):
        return lib.rbow_db_nb_delta_postings(self.handle)

    def compact(self, stream=None):
        check(lib.rbow_db_compact(self.handle, _stream(stream)))


class QueryPipeline:
    """A batch of query images already in HBM through lookup, scoring and top-K, as one
    launch of a captured HIP graph (use_graph) or as direct launches."""

    def __init__(self, voc, db, max_queries, max_desc, max_results=50):
        h = _vp()
        check(lib.rbow_pipeline_create(voc.handle, db.handle, max_queries, max_desc, max_results, ctypes.byref(h)))
        self.handle, self.K, self._keep = h.value, max_results, (voc, db)

    def close(self):
        if getattr(self, "handle", None):
            lib.rbow_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def run(self, dev_desc, offsets, dev_docs, dev_scores, dev_counts, lookup_kernel=LOOKUP_AUTO, use_graph=True,
            stream=None):
        off = np.ascontiguousarray(offsets, np.uint64)
        check(lib.rbow_pipeline_run(self.handle, _ptr(dev_desc), off.ctypes.data, off.size - 1, lookup_kernel,
                                    int(bool(use_graph)), _ptr(dev_docs), _ptr(dev_scores), _ptr(dev_counts),
                                    _stream(stream)))

    @property
    def graphs(self):
        return lib.rbow_pipeline_graphs(self.handle)


class Server:
    """Continuous-batching query server (no counterpart in the reference)."""

    def __init__(self, voc, db, max_batch=64, nb_streams=2, mem_budget=4 << 30, max_results=50,
                 lookup_kernel=LOOKUP_AUTO, staging_bytes=None, use_graphs=True, batch_max_desc=0):
        o = ServerOptions()
        lib.rbow_server_options_default(ctypes.byref(o))
        o.max_batch, o.nb_streams, o.mem_budget, o.max_results, o.lookup_kernel = (
            max_batch, nb_streams, mem_budget, max_results, lookup_kernel)
        o.use_graphs, o.batch_max_desc = int(bool(use_graphs)), batch_max_desc
        if staging_bytes is not None:
            o.staging_bytes = staging_bytes
        h = _vp()
        check(lib.rbow_server_create(voc.handle, db.handle, ctypes.byref(o), ctypes.byref(h)))
        self.handle, self.K, self._keep = h.value, max_results, (voc, db)

    def close(self):
        if getattr(self, "handle", None):
            lib.rbow_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def submit(self, desc):
        desc = np.ascontiguousarray(desc, np.uint8)
        t = ctypes.c_uint64()
        check(lib.rbow_server_submit(self.handle, desc.ctypes.data if desc.size else None, desc.shape[0],
                                     ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        docs = np.zeros(self.K, np.uint32)
        sc = np.zeros(self.K, np.float32)
        n = ctypes.c_uint32()
        lat = ctypes.c_double()
        check(lib.rbow_server_wait(self.handle, ticket, docs.ctypes.data, sc.ctypes.data, ctypes.byref(n),
                                   ctypes.byref(lat)))
        return docs[:n.value], sc[:n.value], lat.value

    def replay(self, desc, offsets, qps, seed=4):
        desc = np.ascontiguousarray(desc, np.uint8)
        off = np.ascontiguousarray(offsets, np.uint64)
        nq = off.size - 1
        lat = np.zeros(nq, np.float64)
        docs = np.zeros((nq, self.K), np.uint32)
        scores = np.zeros((nq, self.K), np.float32)
        cnt = np.zeros(nq, np.uint32)
        bs = np.zeros(nq, np.uint32)
        el = ctypes.c_double()
        check(lib.rbow_server_replay(self.handle, desc.ctypes.data, off.ctypes.data, nq, qps, seed,
                                     lat.ctypes.data, docs.ctypes.data, scores.ctypes.data, cnt.ctypes.data,
                                     bs.ctypes.data, ctypes.byref(el)))
        return dict(latency_ms=lat, docs=docs, scores=scores, counts=cnt, batch_sizes=bs, elapsed_s=el.value)
