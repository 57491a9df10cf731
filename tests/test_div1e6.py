"""The device computes the ppm term of every tolerance bound, (m * tol) / 1e6, without a division instruction
sequence (div_1e6 in csrc/vh_common.h). Its result must be the IEEE quotient bit for bit, or the K1 counts and K2
windows would drift from the reference's (kernels.cu:36-38). The identity holds for the f32 operations alone, so
it is checked on the host over every f32 input."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _has_fma() -> bool:
    try:
        with open("/proc/cpuinfo") as fh:
            return " fma " in fh.read()
    except OSError:
        return False


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div1e6_equals_ieee_division_for_every_f32(tmp_path):
    exe = str(tmp_path / "div1e6_check")
    flags = ["-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math"] + (["-mfma"] if _has_fma() else [])
    subprocess.run(["gcc", *flags, "-o", exe, os.path.join(HERE, "div1e6_check.c"), "-lm"], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=600).stdout.split()
    bad, checked = int(out[0]), int(out[1])
    assert checked == 2 ** 32
    assert bad == 0, f"{bad} f32 inputs where the division-free form differs from x / 1e6"
