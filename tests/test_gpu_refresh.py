"""The FP64 refresh of benzene-size ensembles on tiles of 16 walkers (Slater
rows on FP64 matrix cores, Gauss-Jordan with partial pivoting on register
rows) against the wave engine's Gauss-Jordan build (QMC_TILE_REFRESH=0).

The refresh runs at the start of a sweep call once the refresh interval has
passed; the call here then makes one sweep of zero step length, whose
proposals are the current positions: every ratio is 1 to the inverse's
accuracy, so the state after it is the refreshed state (ln|psi|, sign) and
its local energies are those of the refreshed inverses.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _refreshed(Q, g, W, seed, mode):
    old = os.environ.get("QMC_TILE_REFRESH")
    os.environ["QMC_TILE_REFRESH"] = mode
    try:
        ens = Q.Ensemble(g, W, seed)
        ens.sweeps(3, 0.2, refresh_interval=3)
        pos = ens.positions()
        es, e2, acc = ens.sweeps(1, 0.0, refresh_interval=3)
        return pos, ens.positions(), ens.state(), acc
    finally:
        if old is None:
            del os.environ["QMC_TILE_REFRESH"]
        else:
            os.environ["QMC_TILE_REFRESH"] = old


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("W", [100, 16])
def test_tile_refresh_matches_wave_build(Q, precision, W):
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision=precision)
    print(f"benzene {precision}, {W} walkers (tiles of 16, last one ragged when W % 16)", flush=True)
    p0t, pt, (lt, st, et), at = _refreshed(Q, g, W, 7, "1")
    p0w, pw, (lw, sw, ew), aw = _refreshed(Q, g, W, 7, "0")
    assert np.array_equal(p0t, p0w)  # the same walk up to the refresh
    assert np.array_equal(pt, p0t) and np.array_equal(pw, p0w)  # zero steps
    assert np.array_equal(st, sw)
    # zero-length moves: every ratio is 1 to the inverse's accuracy, so (bar
    # an FP32 ratio a few ulp under 1 meeting a uniform just below 1) all
    # are accepted
    n = W * 42
    assert float(np.sum(at)) >= n - (0 if precision == "fp64" else 3)
    assert float(np.sum(aw)) >= n - (0 if precision == "fp64" else 3)
    if precision == "fp64":
        # ln|D| and ln J of the two builds to 1e-12 relative; E_L from the
        # refreshed inverses to 1e-10
        assert np.allclose(lt, lw, rtol=1e-12, atol=0), np.abs(lt - lw).max()
        assert np.allclose(et, ew, rtol=1e-10, atol=0), np.abs(et - ew).max()
    else:
        # FP64 ln|D| plus 42 FP32 log-ratios of ~1; FP32 E_L
        assert np.allclose(lt, lw, rtol=0, atol=1e-4), np.abs(lt - lw).max()
        assert np.allclose(et, ew, rtol=2e-5, atol=2e-3), np.abs(et - ew).max()

