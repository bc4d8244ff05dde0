"""Build the HIP library libpsc.so in-tree with hipcc for gfx950, the
command-line driver psc_run (host C++ over the C ABI) and the copy probe."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsc.so")
SOURCES = ["psc_lib.hip", "psc_ic.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions off (v_pk_*_f32): plain scalar FP32 code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("hipcc failed building libpsc.so")
    os.replace(tmp, LIB)
    probe = os.path.join(HERE, "copy_probe")
    cmd = [hipcc, "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-o", probe + ".tmp",
           os.path.join(CSRC, "copy_probe.hip")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("hipcc failed building copy_probe")
    os.replace(probe + ".tmp", probe)
    run = os.path.join(HERE, "psc_run")
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-o", run + ".tmp", os.path.join(CSRC, "psc_run.cpp"),
           "-L" + HERE, "-lpsc", "-ldl", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("building psc_run failed")
    os.replace(run + ".tmp", run)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
