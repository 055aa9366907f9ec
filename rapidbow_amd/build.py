"""Builds the gfx950 shared library librapidbow_amd.so in-tree with hipcc.

Every source is compiled for gfx950 only (hipcc treats the host-only .cpp
files as HIP too, so they get the same target); packed-FP32 instruction
selection is disabled for all device code.  The oracle under oracle/ is built separately by its Makefile.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librapidbow_amd.so")
BUILD_DIR = os.path.join(HERE, "_build")

SOURCES = ["lookup.hip", "kmeans.hip", "database.hip", "vocabulary.cpp", "database.cpp", "capi.cpp", "server.cpp",
           "pipeline.cpp", "synth.cpp", "synth.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    procs = []
    for s in srcs:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD_DIR, s.replace(".", "_") + ".o")
        objs.append(obj)
        cmd = [HIPCC] + DEVICE + COMMON + ["-c", src, "-o", obj]  # host sources too: gfx950 is the only target
        if s == "synth.cpp":
            cmd += ["-fopenmp"]
        if s == "database.hip":  # binary32 denormals flushed in hardware: the reference's DAZ/FTZ scoring
            cmd += ["-fgpu-flush-denormals-to-zero"]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    errors = []
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops' is not a recognized" not in l)
        if p.returncode != 0:
            errors.append(f"{' '.join(cmd)}\n{out}")
        elif verbose and out.strip():
            print(out)
    if errors:
        raise RuntimeError("\n\n".join(errors))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-fopenmp", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    # the command-line tool (main.cpp of the reference), linked against the library
    tool = os.path.join(HERE, "makevoc")
    _run([HIPCC] + COMMON + ["-o", tool + ".tmp", os.path.join(HERE, "makevoc.cpp"), f"-L{HERE}", "-lrapidbow_amd",
                             "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tool + ".tmp", tool)
    return LIB


def build_oracle():
    d = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", d, "-B", "liboracle.so"])
    return os.path.join(d, "liboracle.so")


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_oracle())
    sys.exit(0)
