"""C-ABI library: loads, exports every declared symbol, host-side error paths (CPU only)."""
import ctypes
import glob
import os
import re
import struct

import numpy as np
import pytest

import rapidbow_amd as rb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "rapidbow*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rbow_[a-z0-9_]+)\s*\(", text))
    return names


def test_exports_every_declared_symbol():
    names = _declared()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(rb.lib, n)]
    assert not missing, missing


def test_facade_header_names_reference_surface():
    text = open(os.path.join(ROOT, "include", "RapidBoW.h")).read()
    for name in ("class IVocabulary", "class IDataBase", "buildVocabulary", "deserializeVocabulary",
                 "createDataBase", "SiftAttr", "getDescriptorBytes", "lookUp", "queryAndAddDoc"):
        assert name in text


def test_unsupported_attributes():
    h = ctypes.c_void_p()
    bad = rb.DescriptorAttributes(0, 0, 256)
    st = rb.lib.rbow_build_vocabulary(bad, 16, 0x1000, 100, 1, 2, None, ctypes.byref(h))
    assert st == rb.ERR_UNSUPPORTED and h.value is None
    st = rb.lib.rbow_build_vocabulary(rb.SIFT_ATTR, 8, 0x1000, 100, 1, 2, None, ctypes.byref(h))
    assert st == rb.ERR_UNSUPPORTED
    # branch factor 32 needs float32 centroids
    st = rb.lib.rbow_build_vocabulary(rb.SIFT_ATTR, 32, 0x1000, 100, 1, 2, None, ctypes.byref(h))
    assert st == rb.ERR_UNSUPPORTED


def test_build_option_timing_is_checked():
    o = rb.build_options(max_iters=6)
    o.timing = 2
    h = ctypes.c_void_p()
    st = rb.lib.rbow_build_vocabulary_ex(rb.SIFT_ATTR, 16, 0x1000, 100, 1, 2, ctypes.byref(o), None, ctypes.byref(h))
    assert st == rb.ERR_INVALID_ARG and h.value is None
    assert rb.lib.rbow_last_error() == b"timing must be 0 or 1"


def test_build_input_errors_match_reference_messages():
    with pytest.raises(rb.RbowError) as e:
        rb.build_vocabulary(0x1000, 15, 1, 2)
    assert e.value.status == rb.ERR_INSUFFICIENT_DESC and str(e.value) == "Insufficient descriptors"
    with pytest.raises(rb.RbowError) as e:
        rb.build_vocabulary(0x1000, 100, 0, 2)
    assert e.value.status == rb.ERR_NB_DOC and str(e.value) == "nbDoc must be positive"
    with pytest.raises(rb.RbowError) as e:
        rb.build_vocabulary(0x1001, 100, 1, 2)
    assert e.value.status == rb.ERR_INVALID_ARG


def test_invalid_blobs():
    for blob in (b"", b"\x01\x00", struct.pack("<I", 2) + b"\x00" * 10,
                 struct.pack("<I", 1) + b"\x00" * (2048 + 2 + 64 - 1), struct.pack("<I", 99)):
        with pytest.raises(rb.RbowError) as e:
            rb.deserialize_vocabulary(blob)
        assert e.value.status == rb.ERR_INVALID_BLOB and str(e.value) == "invalid blob"


def test_build_options_default_is_reference():
    o = rb.BuildOptions()
    rb.lib.rbow_build_options_default(ctypes.byref(o))
    assert (o.max_iters, o.seed_mode, o.seed, o.centroid_type, o.kernel) == (1000, 0, 5489, 0, 0)


def test_synth_is_deterministic_and_sliceable():
    a = rb.synth_descriptors(300, seed=5)
    b = rb.synth_descriptors(300, seed=5)
    np.testing.assert_array_equal(a, b)
    c = rb.synth_descriptors(100, seed=5, first=200)
    np.testing.assert_array_equal(a[200:], c)
    assert not np.array_equal(a, rb.synth_descriptors(300, seed=6))
    # SIFT-like: L2 norm ~512 before saturation, bytes clipped near 0.2*512
    n = np.sqrt((a.astype(np.float64) ** 2).sum(1))
    assert np.all(np.abs(n - 512) < 8)
