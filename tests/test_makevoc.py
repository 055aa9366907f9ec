"""The command-line tool (the reference's main.cpp:25-49) built on the C++ facade."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import rapidbow_amd as rb

TOOL = os.path.join(os.path.dirname(rb.__file__), "makevoc")


def test_makevoc_usage_and_bad_input(tmp_path):
    r = subprocess.run([TOOL], capture_output=True, text=True)
    assert r.returncode != 0 and "usage: makevoc" in r.stderr
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\0" * 130)  # not a whole number of descriptors
    r = subprocess.run([TOOL, str(bad), str(tmp_path / "v.bin")], capture_output=True, text=True)
    assert r.returncode != 0 and "multiple of 128" in r.stderr


@pytest.mark.gpu
def test_makevoc_blob_equals_library_and_oracle(tmp_path):
    desc = rb.synth_descriptors(8000, seed=21)
    src = tmp_path / "desc.bin"
    src.write_bytes(desc.tobytes())
    out = tmp_path / "voc.bin"
    r = subprocess.run([TOOL, str(src), str(out), "--levels", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    blob = out.read_bytes()
    voc = rb.build_vocabulary(rb.DeviceArray.from_numpy(desc), 8000, 527, 3)  # reference seeding, cap 1000
    assert blob == voc.serialize()
    res = oracle.build(desc, 527, 3)
    n = res["nodes"].shape[0]
    assert len(blob) == 4 + 2114 * n  # Vocabulary.cpp:143,177-182
    np.testing.assert_array_equal(np.frombuffer(blob, np.uint8, n * 2048, 4).reshape(n, 2048), res["nodes"])
