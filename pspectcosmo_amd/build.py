"""Build the HIP library in-tree with hipcc for gfx950, the command-line
driver psc_run (host C++ over the C ABI) and the copy probe.

The library's translation units (the C ABI, the initial-condition generator
and one unit per group of launch shapes) compile in parallel.  The potential
is a compile-time model header (the counterpart of the reference's model.h):
libpsc.so is built with models/preheat.hpp; build_lib(name, header) builds
libpsc_<name>.so with another header, e.g. a user's own model:

    python pspectcosmo_amd/build.py --model mymodel --header /path/to/mymodel.hpp
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsc.so")
UNITS = ["psc_lib.hip", "psc_ic.cpp", "psc_shapes_f32_small.hip", "psc_shapes_f32_256.hip",
         "psc_shapes_f32_512a.hip", "psc_shapes_f32_512b.hip", "psc_shapes_f64_small.hip",
         "psc_shapes_f64_256.hip", "psc_shapes_f64_512.hip", "psc_shapes_r3.hip", "psc_shapes_pad768.hip"]
# built-in models: name -> header (relative to csrc/)
MODELS = {"preheat": "models/preheat.hpp", "lambda_phi4": "models/lambda_phi4.hpp"}
DEFAULT_MODEL = "preheat"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions off (v_pk_*_f32): plain scalar FP32 code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Winfinite-recursion",
]


def lib_path(model=DEFAULT_MODEL):
    return LIB if model == DEFAULT_MODEL else os.path.join(HERE, f"libpsc_{model}.so")


def _hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _run(cmd, what):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"{what} failed")
    return r.stdout


def build_lib(model=DEFAULT_MODEL, header=None, default_shapes_only=False, jobs=None, verbose=False):
    """Compile every unit of the library with the model header and link
    libpsc[_model].so."""
    header = os.path.abspath(header) if header else os.path.join(CSRC, MODELS[model])
    objdir = os.path.join(HERE, "_build", model)
    os.makedirs(objdir, exist_ok=True)
    defs = [f'-DPSC_MODEL_HEADER="{header}"']
    if default_shapes_only:
        defs.append("-DPSC_DEFAULT_SHAPES_ONLY")
    jobs = jobs or min(len(UNITS), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    objs = []

    def compile_one(unit):
        obj = os.path.join(objdir, unit.rsplit(".", 1)[0] + ".o")
        cmd = [_hipcc()] + HIPCC_FLAGS + defs + ["-c", "-o", obj, os.path.join(CSRC, unit)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd, f"hipcc {unit} ({model})")
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, UNITS))
    out = lib_path(model)
    tmp = out + ".tmp"
    _run([_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, f"linking {out}")
    os.replace(tmp, out)
    return out


def build_tools(verbose=False):
    probe = os.path.join(HERE, "copy_probe")
    cmd = [_hipcc(), "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-o", probe + ".tmp",
           os.path.join(CSRC, "copy_probe.hip")]
    _run(cmd, "building copy_probe")
    os.replace(probe + ".tmp", probe)
    run = os.path.join(HERE, "psc_run")
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-o", run + ".tmp", os.path.join(CSRC, "psc_run.cpp"),
           "-L" + HERE, "-lpsc", "-ldl", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd, "building psc_run")
    os.replace(run + ".tmp", run)


def build(verbose=False):
    """The default library with every launch shape, the second built-in model
    with the default shapes, and the tools."""
    build_lib(DEFAULT_MODEL, verbose=verbose)
    for name in MODELS:
        if name != DEFAULT_MODEL:
            build_lib(name, default_shapes_only=True, verbose=verbose)
    build_tools(verbose=verbose)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--model", default=None, help="build only libpsc_<model>.so")
    ap.add_argument("--header", default=None, help="model header of --model")
    ap.add_argument("--default-shapes-only", action="store_true")
    a = ap.parse_args()
    if a.model:
        print(build_lib(a.model, a.header, a.default_shapes_only, verbose=True))
    else:
        print(build(verbose=True))
