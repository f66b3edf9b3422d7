"""Heterogeneous batches (qmc_batch): benzene, CH4 STO-3G and H2 ensembles
sampled together, in one fused persistent launch or one launch per
ensemble, give exactly what each ensemble's own sweeps give."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _systems(Q):
    return [Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="mixed"),
            Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision="mixed"),
            Q.H2MoleculeMO(precision="mixed")]


# ragged walker counts: 3 tiles (last partial), 5 register blocks, 3 blocks
WALKERS = [37, 300, 130]
SEEDS = [11, 12, 13]
STEPS = [0.2, 0.3, 0.5]


def _individual(Q, n_steps, refresh, spl):
    out = []
    for s, W, seed, st in zip(_systems(Q), WALKERS, SEEDS, STEPS):
        e = Q.Ensemble(s, W, seed)
        es, e2, acc = e.sweeps(n_steps, st, refresh_interval=refresh, steps_per_launch=spl)
        out.append((es, acc, e.positions(), e.state()[2]))
    return out


def _batched(Q, n_steps, refresh, mode, spl):
    systems = _systems(Q)
    ens = [Q.Ensemble(s, W, seed) for s, W, seed in zip(systems, WALKERS, SEEDS)]
    b = Q.Batch(ens)
    b.set_config(mode=mode, blocks_per_cu=0, steps_per_launch=spl, order="cost")
    es, e2, acc = b.sweeps(n_steps, STEPS, refresh_interval=refresh)
    return b, [(es[i], acc[i], ens[i].positions(), ens[i].state()[2]) for i in range(3)], (systems, ens)


@pytest.mark.parametrize("mode", ["fused", "streams", "sequential"])
def test_batch_matches_individual_sweeps(Q, mode):
    ref = _individual(Q, 5, 2, 8)
    b, got, keep = _batched(Q, 5, 2, mode, 8)
    assert b.fusable
    ms, launches = b.timing()
    assert ms > 0 and launches == 3  # chunks of 2, 2, 1 sweeps
    # Same Philox counters and the same moves; the fused kernel compiles the
    # FP32 arithmetic apart from the ensembles' own kernels, so E_L may
    # differ by FP32 rounding (a few ulp of the kinetic and potential terms).
    for (es_r, acc_r, pos_r, el_r), (es_b, acc_b, pos_b, el_b) in zip(ref, got):
        assert np.array_equal(acc_r, acc_b)
        assert np.allclose(es_b, es_r, rtol=1e-7, atol=0)
        assert np.allclose(pos_b, pos_r, rtol=0, atol=1e-6)
        assert np.allclose(el_b, el_r, rtol=0, atol=1e-4)


def test_batch_autotuner_picks_a_working_configuration(Q):
    systems = _systems(Q)
    ens = [Q.Ensemble(s, W, seed) for s, W, seed in zip(systems, WALKERS, SEEDS)]
    b = Q.Batch(ens)
    cfg, ms = b.tune(2, STEPS, refresh_interval=4)
    assert ms > 0 and np.isfinite(ms)
    assert cfg["mode"] in ("fused", "streams", "sequential")
    es, e2, acc = b.sweeps(4, STEPS, refresh_interval=4)
    assert np.isfinite(es).all() and (acc > 0).all()


def test_batch_autotuner_refresh_interval_within_the_drift_guard(Q):
    """refresh_interval < 0 tunes the FP64 inverse refresh interval: the
    picked interval is one of the candidates, the sweeps then run with it,
    and after that many sweeps the local energies from the updated inverses
    stay within the guard's 0.01 mHa of a fresh inverse's (checked here on
    the ensembles themselves, mean over walkers)."""
    systems = _systems(Q)
    ens = [Q.Ensemble(s, W, seed) for s, W, seed in zip(systems, WALKERS, SEEDS)]
    b = Q.Batch(ens)
    cfg, ms = b.tune(2, STEPS, refresh_interval=-128)
    R = cfg["refresh_interval"]
    assert R in (16, 32, 64, 128)
    for e in ens:
        e.refresh()
    b.sweeps(R, STEPS, refresh_interval=0)
    for e in ens:
        cached = e.state()[2]
        e.refresh()
        fresh = e.state()[2]
        d = cached - fresh
        assert abs(d.mean()) < 1e-5 + 4 * d.std() / np.sqrt(d.size), (R, d.mean())
    es, e2, acc = b.sweeps(2 * R + 3, STEPS, refresh_interval=-1)
    assert np.isfinite(es).all()


def test_batch_with_an_fp64_ensemble_is_not_fused(Q):
    s64 = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision="fp64")
    s32 = Q.H2MoleculeMO(precision="mixed")
    ens = [Q.Ensemble(s64, 70, 1), Q.Ensemble(s32, 70, 2)]
    b = Q.Batch(ens)
    assert not b.fusable
    assert b.config()["mode"] == "streams"
    with pytest.raises(Q.QmcError):
        b.set_config(mode="fused")
    es, e2, acc = b.sweeps(3, [0.3, 0.5], refresh_interval=0)
    ref = Q.Ensemble(Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision="fp64"), 70, 1)
    es_r, _, acc_r = ref.sweeps(3, 0.3, refresh_interval=0, steps_per_launch=8)
    assert np.array_equal(acc[0], acc_r)
    assert np.allclose(es[0], es_r, rtol=1e-12, atol=0)
