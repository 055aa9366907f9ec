"""Work partition of the MFMA k-means assign (rapidbow_amd/csrc/kmpartition.hpp).

The partition is plain host C++; partition_check.cpp is compiled here with g++
and checks it on hand-picked and random levels (tasks of every size, tasks
whose Lloyd loop is not streamed between streamed ones, any resident wave
count) against the invariants the assign and update kernels index by.  It then
simulates the partial-sum bookkeeping of whole assign passes (stores of
unmerged chunks, LDS merge slots flushed for active tasks only, stale partials
of converged tasks, reduce items): the current one-range-per-wave layout loses
nothing, and a layout whose waves share a chunk's tiles is caught losing sums.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "rapidbow_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_partition_invariants(tmp_path):
    exe = tmp_path / "partition_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", CSRC,
                    os.path.join(HERE, "partition_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "0 failures" in r.stdout
    assert "per-wave ranges 0 wrong task totals" in r.stdout
