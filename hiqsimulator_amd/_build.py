"""In-tree build of the native engine for gfx950.

Produces, next to this file:
  libhiqsim_amd.so          the C-ABI library (HIP kernels + C++ engine),
                            include/hiqsim_amd.h
  _cppsim_amd<ext>.so       the pybind11 module over that C ABI

Run as ``python hiqsimulator_amd/_build.py`` (it does not import the package) or through
``__graft_entry__.build()``.  Always recompiles.
"""
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
ARCH = os.environ.get("HIQSIM_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", "-I" + INCLUDE]
LIB_NAME = "libhiqsim_amd.so"


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def _run_all(cmds, verbose=False):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
             for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((c, text))
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + t for c, t in failed)
        raise RuntimeError("native build failed:\n" + msg)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in ("kernels.hip", "batch_kernels.hip", "noalloc.hip", "engine.cpp", "batch.cpp", "serve.cpp",
                "capi.cpp", "sched.cpp"):
        obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run_all(cmds, verbose)
    lib = os.path.join(PKG, LIB_NAME)
    tmp = lib + ".tmp"
    _run_all([[HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs])
    os.replace(tmp, lib)

    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    mod = os.path.join(PKG, "_cppsim_amd" + ext_suffix())
    tmp = mod + ".tmp"
    _run_all([["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
               "-I" + pybind11.get_include(), "-I" + py_inc, "-I" + INCLUDE,
               os.path.join(CSRC, "pybind_sim.cpp"), "-o", tmp, "-L" + PKG, "-lhiqsim_amd",
               "-Wl,-rpath,$ORIGIN"]])
    os.replace(tmp, mod)
    if verbose:
        print("built", lib, mod)
    return lib, mod


if __name__ == "__main__":
    build(verbose=True)
