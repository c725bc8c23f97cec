"""Comparisons of a correctly rounded float root with a constant decided on
the root's argument (openabl_amd/csrc/abl_root_bounds.hpp, what generated
steps use for a distance they only compare): against the IEEE sqrtf of the
host on samples around every bound, random floats and special values for
constants as models write them (tests/host/root_bounds_check.cpp; its
"full" mode covers every non-negative float for three constants)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def test_bounds_equal_sqrtf_comparisons(tmp_path):
    exe = str(tmp_path / "rbc")
    src = os.path.join(HERE, "host", "root_bounds_check.cpp")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(REPO, "openabl_amd", "csrc"),
                        src, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0 and "not found" in r.stderr:
        pytest.skip("no g++")
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad 0" in r.stdout
