"""The allocation-free lookup (rbow_voc_lookup_noalloc) against the ordinary
one (rbow_voc_lookup_ex), through the C ABI: identical leaf ids and near-tie
counts on two calls in a row, the second smaller, on a uint8 and a float tree;
a batch too large for the static workspace is rejected before any launch."""
import ctypes

import numpy as np
import pytest

import rapidbow_amd as rb
from rapidbow_amd import hip

pytestmark = pytest.mark.gpu

_vp, _u32, _i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32
rb.lib.rbow_voc_lookup_noalloc.restype = _i32
rb.lib.rbow_voc_lookup_noalloc.argtypes = [_vp, _vp, _u32, _vp, _i32, _vp, _vp]


def _lookup(voc, fn, d, n, kernel):
    out = rb.DeviceArray((max(n, 1),), np.uint32)
    ties = rb.DeviceArray.zeros((1,), np.uint32)
    rb.check(fn(_vp(voc.handle), _vp(d.ptr), _u32(n), _vp(out.ptr), _i32(kernel), _vp(ties.ptr), None))
    hip.synchronize()
    return out.numpy(n), int(ties.numpy()[0])


@pytest.mark.parametrize("b,L,ct", [(16, 3, 0), (32, 3, 1)])
def test_lookup_noalloc_equals_lookup(b, L, ct):
    desc = rb.synth_descriptors(60_000, seed=41)
    voc = rb.build_vocabulary(rb.DeviceArray.from_numpy(desc), desc.shape[0], 100, L, branch_factor=b,
                              max_iters=6, seed_mode=1, centroid_type=ct)
    syn = rb.DeviceSynth()
    for n in (400_000, 70_001):  # the second call is smaller and reuses the array the first one filled
        q = syn.descriptors(n, seed=42 + n)
        for kernel in (rb.LOOKUP_MFMA, rb.LOOKUP_AUTO, rb.LOOKUP_PLAIN):
            ref, ref_ties = _lookup(voc, rb.lib.rbow_voc_lookup_ex, q, n, kernel)
            got, ties = _lookup(voc, rb.lib.rbow_voc_lookup_noalloc, q, n, kernel)
            np.testing.assert_array_equal(got, ref)
            assert ties == ref_ties
    # a batch whose workspace exceeds the static array: an error, nothing launched
    n = 12_000_000
    big = rb.DeviceArray((n, 128), np.uint8)
    out = rb.DeviceArray((n,), np.uint32)
    st = rb.lib.rbow_voc_lookup_noalloc(_vp(voc.handle), _vp(big.ptr), _u32(n), _vp(out.ptr), _i32(rb.LOOKUP_MFMA),
                                        None, None)
    assert st == rb.ERR_CAPACITY
    assert "scratch" in rb.lib.rbow_last_error().decode()
