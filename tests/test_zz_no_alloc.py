"""The allocation-free derivative entry point (qmc_derivatives_noalloc)
against qmc_derivatives, through the C ABI on device buffers.

Both run the same derivative kernels on the same configurations, so every
output (E_L, O_k, dE_L/dp_k, Hellmann-Feynman force, d ln psi / dR) must be
bit-identical.  Each case makes two calls in a row, the second with fewer
walkers: the scratch array keeps the first call's contents, which the second
call must not read.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _system(Q, name, precision):
    if name == "ch4_sto3g_series":
        g = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision)
        g.set_jastrow_series(6, 8)
        p = g.get_params()
        p[4:] = np.random.default_rng(7).normal(size=p.size - 4) * 0.05
        g.set_params(p)
        return g
    if name == "h2":
        return Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0, precision=precision)
    if name == "benzene":
        return Q.BenzeneGTO(precision=precision)
    raise KeyError(name)


def _positions(Q, g, W, seed):
    ens = Q.Ensemble(g, W, seed)
    ens.sweeps(2, 0.3)
    return ens.positions()


def _noalloc(Q, g, pos):
    W = pos.shape[0]
    P, M = g.num_params(), g.n_nuc
    sizes = dict(eloc=W, O=W * P, dE=W * P, F=W * M * 3, Om=W * M * 3)
    bufs = {k: Q.DeviceBuffer(8 * n) for k, n in sizes.items()}
    with Q.DeviceBuffer(pos.nbytes) as dpos:
        dpos.upload(pos)
        try:
            print(f"noalloc call: {W} walkers, P={P}, M={M}", flush=True)
            Q.check(Q.lib().qmc_derivatives_noalloc(g._h, W, dpos.ptr, bufs["eloc"].ptr, bufs["O"].ptr,
                                                    bufs["dE"].ptr, bufs["F"].ptr, bufs["Om"].ptr, None))
            assert Q._hip().hipDeviceSynchronize() == 0
            out = {k: bufs[k].download(np.float64, n) for k, n in sizes.items()}
        finally:
            for b in bufs.values():
                b.free()
    return dict(eloc=out["eloc"], O=out["O"].reshape(W, P), dE=out["dE"].reshape(W, P),
                F=out["F"].reshape(W, M, 3), Om=out["Om"].reshape(W, M, 3))


@pytest.mark.parametrize("name,precision,W1,W2", [("ch4_sto3g_series", "fp64", 4096, 1000),
                                                  ("ch4_sto3g_series", "mixed", 4096, 1000),
                                                  ("h2", "fp64", 3000, 77),
                                                  ("benzene", "fp64", 192, 70)])
def test_noalloc_derivatives_equal_the_allocating_entry_point(Q, name, precision, W1, W2):
    g = _system(Q, name, precision)
    for W, seed in ((W1, 3), (W2, 9)):
        pos = _positions(Q, g, W, seed)
        ref = g.derivatives_batch(pos)
        got = _noalloc(Q, g, pos)
        for k in ("eloc", "O", "dE", "F", "Om"):
            assert got[k].shape == ref[k].shape, k
            assert np.array_equal(got[k], ref[k]), (name, precision, W, k, np.max(np.abs(got[k] - ref[k])))


def test_noalloc_rejects_host_pointers_and_oversized_batches(Q):
    g = _system(Q, "h2", "fp64")
    pos = _positions(Q, g, 64, 1)
    el = np.zeros(64)
    rc = Q.lib().qmc_derivatives_noalloc(g._h, 64, pos.ctypes.data_as(C.c_void_p), el.ctypes.data_as(C.c_void_p),
                                         None, None, None, None, None)
    assert rc == Q._lib.QMC_ERR_INVALID
    assert "device pointers" in Q.lib().qmc_last_error().decode()
    with Q.DeviceBuffer(8 * 3 * 2) as dpos:
        rc = Q.lib().qmc_derivatives_noalloc(g._h, 1 << 40, dpos.ptr, None, None, None, None, None, None)
    assert rc == Q._lib.QMC_ERR_INVALID
    assert "exceed" in Q.lib().qmc_last_error().decode()
