"""CPU test of the serving loop's arena allocator (first fit, coalescing frees): a randomised
run against a brute-force occupancy map, compiled with g++ from the library's header."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_range_alloc_against_brute_force():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "range_alloc_test")
        subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "msc-project_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "range_alloc_test.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
        assert out.strip() == "ok"
