"""CPU-side checks of the C ABI: the library loads, exports every function
include/qmc.h declares, and its host-only routines (statistics of
vmc.rs:141-188) agree with the oracle.  No GPU calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "qmc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qmc_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import rust_qmc_amd._lib as L
    lib = L.lib()
    decl = _declared_functions()
    assert len(decl) >= 30
    missing = [f for f in decl if not hasattr(lib, f)]
    assert not missing, missing
    assert set(decl) == set(L.EXPORTED)


def test_version_and_error_string():
    import rust_qmc_amd._lib as L
    assert b"gfx950" in L.lib().qmc_version()
    assert isinstance(L.lib().qmc_last_error(), bytes)


def test_statistics_match_oracle(oracle):
    import rust_qmc_amd as Q
    rng = np.random.default_rng(7)
    x = np.cumsum(rng.normal(size=3000)) * 0.01 + rng.normal(size=3000)
    tau_q = Q.autocorrelation_time(x)
    tau_o = oracle.autocorrelation_time(x)
    assert tau_q == pytest.approx(tau_o, rel=1e-12)
    assert Q.blocking_error(x, tau_q) == pytest.approx(oracle.blocking_error(x, tau_o), rel=1e-12)
    assert Q.autocorrelation_time(np.ones(10)) == 1.0


def test_invalid_arguments_rejected_without_gpu():
    import rust_qmc_amd._lib as L
    lib = L.lib()
    h = C.c_void_p()
    assert lib.qmc_system_create(None, 0, C.byref(h)) == L.QMC_ERR_INVALID
    assert b"null" in lib.qmc_last_error()
    assert lib.qmc_set_precision(None, 0) == L.QMC_ERR_INVALID
