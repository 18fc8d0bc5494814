"""Division by a constant on the device (div_const, pcr_amd/csrc/pcr_device.h).

The bounce kernels divide by 5 (spectral table index), 150 (sigmoid-fit
wavelength) and pi (albedo / pi, cosine pdf) with a product and one
FMA-corrected step instead of a division.  Bit-exact parity with the CPU
oracle needs that to be the correctly rounded quotient: the check runs all
2^32 floats of the range the device takes the fast path for, and the header's
reciprocals must be the correctly rounded 1/c.
"""
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "pcr_amd", "csrc", "pcr_device.h")


def _hexf(name):
    m = re.search(r"constexpr float %s = ([0-9a-fx.p+-]+)f;" % name, open(HEADER).read())
    assert m, name
    return np.float32(float.fromhex(m.group(1)))


def test_reciprocals_are_correctly_rounded():
    pi = np.float32(3.14159274101257324)
    for name, c in (("kInvPiF", pi), ("kInv5F", np.float32(5)), ("kInv150F", np.float32(150))):
        # float64 1/c rounded to float32 is RN(1/c): 1/c is not a float32 midpoint for these c
        assert _hexf(name) == np.float32(1.0 / np.float64(c)), name


def test_div_const_is_correctly_rounded_everywhere(tmp_path):
    exe = str(tmp_path / "constdiv_check")
    subprocess.run(["gcc", "-O2", "-mfma", "-o", exe, os.path.join(ROOT, "tests", "csrc", "constdiv_check.c"),
                    "-lm", "-lpthread"], check=True, capture_output=True)
    for c in ("3.14159274101257324", "5", "150"):
        out = subprocess.run([exe, c], capture_output=True, text=True, timeout=600)
        assert out.returncode == 0 and "mismatches=0" in out.stdout, out.stdout
