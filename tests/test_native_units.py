"""Host-only C++ units of the native code, compiled with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer and run here:
* the request loop's pool allocator (first fit, alignment, coalescing,
  20000 random alloc/free steps against a byte ownership map);
* the schedulers, the tile planner (whole and one pass at a time, as the
  deferred clusters use it) and Batch::prepare() (planning, fusion, layout
  and the bounds check of every launch) on random workloads, with the plan
  invariants checked (tests/native/host_plans_test.cpp)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "hiqsimulator_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def run(exe):
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=ENV)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    return out.stdout


def test_range_alloc(tmp_path):
    exe = str(tmp_path / "range_alloc_test")
    src = os.path.join(HERE, "native", "range_alloc_test.cpp")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wall"] + SAN + [src, "-o", exe], check=True)
    assert "range_alloc ok" in run(exe)


def test_host_plans_under_sanitizers(tmp_path):
    hip_inc = "/opt/rocm/include"
    if not os.path.isdir(os.path.join(hip_inc, "hip")):
        pytest.skip("HIP headers not installed")
    flags = ["g++", "-O1", "-g", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + hip_inc] + SAN
    objs = []
    for src in (os.path.join(CSRC, "batch.cpp"), os.path.join(CSRC, "sched.cpp"),
                os.path.join(HERE, "native", "host_plans_test.cpp")):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run(flags + ["-c", src, "-o", obj], check=True)
        objs.append(obj)
    exe = str(tmp_path / "host_plans_test")
    subprocess.run(["g++"] + SAN + objs + ["-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", exe],
                   check=True)
    assert "host plans ok" in run(exe)
