"""The shared-reciprocal quotient the step programs use for quotients with a
common divisor (div_by_recip inside the div_ok range,
openabl_amd/csrc/abl_device.hpp), restated in C (tests/host/div_recip_check.c)
and compared with IEEE division bit for bit: counts 1..255 as divisors,
random divisors over the whole range, quotients next to rounding ties, the
circle model's force quotients, the range ends; FP64 and FP32.  A
restatement without the residual correction fails it (checked when the
test was written)."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_shared_reciprocal_quotient_is_ieee_division(tmp_path):
    exe = str(tmp_path / "divchk")
    subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off",
                    os.path.join(REPO, "tests", "host", "div_recip_check.c"), "-lm", "-o", exe],
                   check=True)
    r = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
