"""The FP64 exponentials of qmc_expn.hpp (Cody-Waite reduction, degree-13
Taylor polynomial, ldexp) against the C library's exp on the host:
exp_nonpos (primitive arguments) within 1 ulp over [-745.2, 0], zero below
the underflow threshold and at -inf; exp_clamped (acceptance exponents of
either sign) within 1 ulp over [-745.2, 709.78], infinite past the overflow
threshold and at +inf, zero at -inf and at NaN; sincos_turns (Box-Muller
angles) within 1 ulp of an extended-precision sin and cos of the exactly
reduced angle and within 1e-15 of the library's sin and cos of 2 pi u;
log_unit (Box-Muller radii) within 1 ulp of the library's log over (0, 1]."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fp64_exponentials_within_one_ulp_of_libm():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "expn_check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(ROOT, "rust_qmc_amd", "csrc"), os.path.join(ROOT, "tests", "expn_check.cpp"),
                        "-o", exe], check=True)
        lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()
    out = lines[0].split()
    n, worst, worst_x, at_minus_inf, far_tail = int(out[0]), int(out[1]), float(out[2]), float(out[3]), float(out[4])
    assert n == 2000009
    assert worst <= 1, (worst, worst_x)
    assert at_minus_inf == 0.0 and far_tail == 0.0
    out = lines[1].split()
    n, worst, worst_x = int(out[0]), int(out[1]), float(out[2])
    at_inf, past, at_nan, at_minus_inf = (float(v) for v in out[3:7])
    assert n == 2000011
    assert worst <= 1, (worst, worst_x)
    assert at_inf == float("inf") and past == float("inf") and at_nan == 0.0 and at_minus_inf == 0.0
    out = lines[2].split()
    n, worst, worst_x, worst_abs = int(out[0]), int(out[1]), float(out[2]), float(out[3])
    assert n == 3000009
    assert worst <= 1, (worst, worst_x)
    assert worst_abs < 1e-15
    out = lines[3].split()
    n, worst, worst_x, at_one = int(out[0]), int(out[1]), float(out[2]), float(out[3])
    assert n == 4000009
    assert worst <= 1, (worst, worst_x)
    assert at_one == 0.0
