"""CPU property test of the column-band arithmetic that lets frames and served iterations whose
per-sample state exceeds the budget run out of core (csrc/bands.h), compiled with g++."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bands_partition_columns_and_fit_their_room():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "bands_test")
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "msc-project_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "bands_test.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
        assert out.strip() == "ok"
