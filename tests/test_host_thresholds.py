"""Host-side near-test thresholds of the runtime (abl_host.cpp), checked by
brute force around every threshold against their definitions with glibc's
correctly rounded sqrtf (tests/native/thresholds_check.cpp):

  * sqrtf_le_threshold(r) = max float f with (double) sqrtf(f) <= r, so that
    dist <= r  <=>  (float) s <= T   (asset/c/libabl.h:156-168 dist_float2);
  * sqrtf_lt_threshold(r) = min float f with (double) sqrtf(f) >= r
    (strict separation tests, dist < r  <=>  (float) s < U);
  * double_le_threshold(t): (float) s <= t  <=>  s <= Td for every double s,
    which lets the FP64 kernels skip the rounding of s to float.

A planted error (dropping the ties-to-even rule of double_le_threshold) makes
the checker fail on ~20k of its cases."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_thresholds_brute_force(tmp_path):
    exe = str(tmp_path / "thresholds_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(REPO, "tests", "native", "thresholds_check.cpp"),
                    os.path.join(REPO, "openabl_amd", "csrc", "abl_host.cpp"), "-o", exe, "-lm"],
                   check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert out.stdout.startswith("ok ")
