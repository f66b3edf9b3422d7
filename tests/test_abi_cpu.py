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


def test_isdmc_block_error_matches_oracle_on_the_host(oracle):
    import rust_qmc_amd as Q
    rng = np.random.default_rng(5)
    x = np.cumsum(rng.normal(size=2000)) * 0.02 + rng.normal(size=2000)
    for n in (5, 19, 20, 400, 2000):
        assert Q.isdmc_block_error(x[:n]) == pytest.approx(oracle.isdmc_block_error(x[:n]), rel=1e-12, abs=0)


@pytest.mark.parametrize("doc,needle", [
    (b"{", b"JSON"),
    (b"[1, 2]", b"object"),
    (b'{"nuclei": [[0, 0, 0]], "charges": [1], "spins": [1]}', b"aos"),
    (b'{"nuclei": [[0, 0]], "charges": [1], "spins": [1], "aos": []}', b"nuclei"),
    (b'{"nuclei": [[0, 0, 0]], "charges": [1, 2], "spins": [1], "aos": [{}]}', b"length"),
    (b'{"nuclei": [[0, 0, 0]], "charges": [1], "spins": [1], "aos": [{"kind": "f", "center": 0}]}', b"kind"),
    (b'{"nuclei": [[0, 0, 0]], "charges": [1], "spins": [1], '
     b'"aos": [{"kind": "gto", "center": 3, "exp": [1.0], "coef": [1.0]}], "mo_up": [[1.0]]}', b"range"),
    (b'{"nuclei": [[0, 0, 0]], "charges": [1], "spins": [1], '
     b'"aos": [{"kind": "gto", "center": 0, "exp": [1.0, 2.0], "coef": [1.0]}], "mo_up": [[1.0]]}', b"same"),
    (b'{"nuclei": [[0, 0, 0]], "charges": [1], "spins": [1], '
     b'"aos": [{"kind": "gto", "center": 0, "exp": [1.0], "coef": [1.0]}], "mo_up": [[1.0, 2.0]]}', b"ragged"),
    (b'{"nuclei": [[0, 0, 0]], "charges": [1], "spins": [1], "flags": ["nope"], '
     b'"aos": [{"kind": "gto", "center": 0, "exp": [1.0], "coef": [1.0]}], "mo_up": [[1.0]]}', b"flag"),
    (b'{"system": "Argon"}', b"unknown system"),
    (b'{"system": "MethaneGTO", "basis": "cc-pVTZ"}', b"basis"),
    (b'{"orbital1": {"alpha": 1.3, "center": [0, 0, 0.7]}, "jastrow": {"cusp_param": 3}}', b"orbitals"),
    (b'{"orbital1": {"alpha": 1.3, "center": [0, 0]}, "orbital2": {"alpha": 1.3, "center": [0, 0, 1]}, '
     b'"jastrow": {"cusp_param": 3}}', b"center"),
    (b'{"a": [' * 70 + b']}' * 70, b"deep"),
    (b'{"a": 1} x', b"trailing"),
])
def test_json_descriptions_rejected_before_any_device_work(doc, needle):
    """Malformed or inconsistent descriptions fail in the host reader with a
    message naming the field, before a system (and device memory) exists."""
    import rust_qmc_amd._lib as L
    lib = L.lib()
    h = C.c_void_p()
    assert lib.qmc_system_from_json(doc, len(doc), 0, C.byref(h)) == L.QMC_ERR_INVALID
    assert needle in lib.qmc_last_error(), lib.qmc_last_error()
    assert not h.value
