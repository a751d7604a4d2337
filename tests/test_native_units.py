"""Host-only C++ units of the native code, compiled with g++ and run here:
the request loop's pool allocator (first fit, alignment, coalescing,
20000 random alloc/free steps against a byte ownership map)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_range_alloc(tmp_path):
    exe = str(tmp_path / "range_alloc_test")
    src = os.path.join(HERE, "native", "range_alloc_test.cpp")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "range_alloc ok" in out.stdout
