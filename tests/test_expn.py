"""The FP64 exponential of the primitive arguments (qmc_expn.hpp: Cody-Waite
reduction, degree-13 Taylor polynomial, ldexp) against the C library's exp
on the host: within 1 ulp over [-745.2, 0], zero below the underflow
threshold and at -inf."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_exp_nonpos_within_one_ulp_of_libm():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "expn_check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(ROOT, "rust_qmc_amd", "csrc"), os.path.join(ROOT, "tests", "expn_check.cpp"),
                        "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    n, worst, worst_x, at_minus_inf, far_tail = int(out[0]), int(out[1]), float(out[2]), float(out[3]), float(out[4])
    assert n == 2000009
    assert worst <= 1, (worst, worst_x)
    assert at_minus_inf == 0.0 and far_tail == 0.0
