"""Drift-diffusion VMC (DriftDiffusionVMC, vmc.rs:195-535) on the GPU against
the oracle's literal restatement (full recomputation of psi and of every
drift at each trial configuration)."""
import math

import numpy as np
import pytest

from test_gpu_parity import _close, _pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


@pytest.mark.parametrize("name", ["h2", "li_ref", "li_seeded", "ch4_sto3g", "ch4_631g"])
def test_dd_trajectory_parity_fp64(Q, oracle, name):
    """Same seed -> the Langevin single-electron moves (Sherman-Morrison
    ratio, drift from the inverse column) follow the oracle's chain step for
    step: identical decisions, positions and E_L."""
    g, o = _pair(Q, oracle, name)
    W, steps, seed, tau = 65, 4, 21, 0.05
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, tau, Q.QMC_MOVE_DRIFT, trace=True, refresh_interval=2, steps_per_launch=3,
                                 max_drift=1.0)
    pos_o, el_o, acc_o = o.dd_steps(p0, seed, 0, steps, tau, 1.0)
    assert acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr, el_o, 1e-9)


def test_dd_large_step_and_drift_cap(Q, oracle):
    """A time step large enough that tau F often exceeds max_drift (0.3) on
    CH4: the capped drift of vmc.rs:313-321 in both Green's functions."""
    g, o = _pair(Q, oracle, "ch4_sto3g")
    W, steps, seed, tau, cap = 40, 3, 4, 0.4, 0.3
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, tau, Q.QMC_MOVE_DRIFT, trace=True, refresh_interval=0, max_drift=cap)
    pos_o, el_o, acc_o = o.dd_steps(p0, seed, 0, steps, tau, cap)
    assert acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr, el_o, 1e-9)


def test_dd_generic_kernels_match(Q, oracle):
    g, o = _pair(Q, oracle, "ch4_631g")
    out = []
    for eng in ("auto", "generic"):
        g.set_engine(eng)
        ens = Q.Ensemble(g, 70, 2)
        es, e2, acc = ens.sweeps(3, 0.03, Q.QMC_MOVE_DRIFT, refresh_interval=2, max_drift=1.0)
        out.append((es, acc, ens.positions()))
    assert np.array_equal(out[0][1], out[1][1])
    assert np.allclose(out[0][0], out[1][0], rtol=1e-12, atol=0)
    assert np.allclose(out[0][2], out[1][2], rtol=0, atol=1e-12)


def test_dd_mixed_agrees_with_fp64(Q):
    res = {}
    for prec in ("fp64", "mixed"):
        s = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=prec)
        ens = Q.Ensemble(s, 4096, 3)
        ens.sweeps(40, 0.03, Q.QMC_MOVE_DRIFT, refresh_interval=16, max_drift=1.0)
        es, e2, acc = ens.sweeps(200, 0.03, Q.QMC_MOVE_DRIFT, refresh_interval=16, max_drift=1.0)
        tr = es / 4096
        assert np.isfinite(tr).all()
        tau = Q.autocorrelation_time(tr)
        res[prec] = (tr.mean(), max(Q.blocking_error(tr, tau), tr.std() / math.sqrt(len(tr))))
    (a, sa), (b, sb) = res["fp64"], res["mixed"]
    assert abs(a - b) < 3 * math.hypot(sa, sb) + 1e-4


def test_ddvmc_run_matches_metropolis_vmc(Q):
    """DriftDiffusionVMC::run and MCMCSimulation::run sample the same
    |psi|^2: H2 energies agree; the adapted time step stays in the
    reference's clamp [0.001, 0.05]."""
    h2 = Q.H2MoleculeMO()
    dd = Q.DriftDiffusionVMC(h2, Q.DDVMCParams(n_walkers=2048, n_steps=1500, n_burnin=300, time_step=0.02,
                                               adaptation_interval=100), seed=5)
    r = dd.run()
    assert 0.001 <= r.final_time_step <= 0.05
    assert 0.2 < r.acceptance_rate <= 1.0
    vmc = Q.MCMCSimulation(h2, Q.MCMCParams(n_walkers=2048, n_steps=1800, initial_step_size=0.5, max_step_size=0.5,
                                            min_step_size=0.5, n_burnin=300, move=Q.QMC_MOVE_SINGLE), seed=6)
    rv = vmc.run()
    assert abs(r.energy - rv.energy) < 3 * math.hypot(r.error, rv.error) + 2e-4


def test_dd_wave_benzene_trajectory_fp64(Q, oracle):
    """Benzene's drift-diffusion moves on the wave engine follow the
    oracle's chain in the log domain (the reference's 1e-30 guard on the
    linear psi, vmc.rs:363, rejects every benzene move: psi ~ e^-650)."""
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05)
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    W, steps, seed, tau = 3, 2, 7, 0.02
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, tau, Q.QMC_MOVE_DRIFT, trace=True, refresh_interval=1, max_drift=1.0)
    pos_o, el_o, acc_o = o.dd_steps(p0, seed, 0, steps, tau, 1.0, log_domain=True)
    assert acc_o > 0 and acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr, el_o, 1e-9)
    assert o.dd_steps(p0, seed, 0, 1, tau, 1.0)[2] == 0  # the reference's guard


@pytest.mark.parametrize("name", ["h2", "li_seeded", "ch4_sto3g"])
def test_dd_wave_engine_matches_register_engine(Q, oracle, name):
    out = []
    for eng in ("auto", "wave"):
        g, _ = _pair(Q, oracle, name)
        g.set_engine(eng)
        ens = Q.Ensemble(g, 48, 5)
        es, e2, acc = ens.sweeps(4, 0.05, Q.QMC_MOVE_DRIFT, refresh_interval=2, max_drift=1.0)
        out.append((es, acc, ens.positions()))
    assert np.array_equal(out[0][1], out[1][1])
    assert np.allclose(out[0][2], out[1][2], rtol=0, atol=1e-10)
    assert np.allclose(out[0][0], out[1][0], rtol=1e-10, atol=0)


def test_dd_benzene_mixed_and_ddvmc_run(Q):
    """DriftDiffusionVMC::run on benzene (drift-diffusion on the tile
    engine in both precisions) gives a finite energy with the time step
    inside the clamp, and the mixed chain's energy agrees with the FP64
    chain's."""
    res = {}
    for prec in ("fp64", "mixed"):
        s = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision=prec)
        dd = Q.DriftDiffusionVMC(s, Q.DDVMCParams(n_walkers=1024, n_steps=140, n_burnin=40, time_step=0.02,
                                                  adaptation_interval=20), seed=9)
        r = dd.run()
        assert np.isfinite(r.energy) and 0.001 <= r.final_time_step <= 0.05
        assert 0.3 < r.acceptance_rate <= 1.0
        res[prec] = r
    a, b = res["fp64"], res["mixed"]
    assert a.error > 0 and b.error > 0  # 100 production steps: the blocking error is defined
    assert abs(a.energy - b.energy) < 5 * math.hypot(a.error, b.error) + 0.01



def test_dd_tile_engine_benzene_fp64_vs_oracle(Q, oracle):
    """FP64 drift-diffusion on the tile engine (16 lanes per walker: the
    moved electron's AO weights staged in LDS, the quantum force at both
    ends of the move, the FP64 MFMA MO contraction for the Sherman-Morrison
    update) follows the oracle's log-domain chain: same decisions, positions
    and E_L to 1e-9.  17 walkers make a full tile and a ragged one."""
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="fp64")
    assert g.engine_name() == "tile"
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    W, steps, seed, tau = 17, 2, 7, 0.02
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, tau, Q.QMC_MOVE_DRIFT, trace=True, refresh_interval=1, max_drift=1.0)
    pos_o, el_o, acc_o = o.dd_steps(p0, seed, 0, steps, tau, 1.0, log_domain=True)
    assert acc_o > 0 and acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr, el_o, 1e-9)


def test_dd_tile_engine_fp64_matches_wave_engine(Q):
    """FP64 drift-diffusion sweeps of 37 walkers (three tiles, the last
    ragged) on the tile and the wave engine from the same seed, a refresh
    between launches, a drift cap that binds: the same decisions,
    positions, ln|psi| and E_L to 1e-10."""
    out = []
    for eng in ("tile", "wave"):
        g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="fp64")
        g.set_engine(eng)
        ens = Q.Ensemble(g, 37, 9)
        es, e2, acc, tr = ens.sweeps(3, 0.05, Q.QMC_MOVE_DRIFT, trace=True, refresh_interval=2, steps_per_launch=2,
                                     max_drift=0.3)
        out.append((ens.positions(), acc, tr, ens.state()))
    (p1, a1, t1, s1), (p2, a2, t2, s2) = out
    assert np.array_equal(a1, a2)
    assert np.allclose(p1, p2, rtol=0, atol=1e-10)
    assert np.allclose(t1, t2, rtol=1e-10, atol=0)
    assert np.allclose(s1[0], s2[0], rtol=1e-11, atol=0)
    assert np.array_equal(s1[1], s2[1])


def _walker_dev(p1, p2, t1, t2):
    """per-walker largest position and relative E_L deviation of two chains"""
    dp = np.abs(np.asarray(p1) - np.asarray(p2)).reshape(len(p1), -1).max(axis=1)
    de = (np.abs(np.asarray(t1) - np.asarray(t2)) / (np.abs(np.asarray(t2)) + 10.0)).max(axis=0)
    return dp, de


def test_dd_tile_engine_mixed_vs_wave_fp32_and_oracle(Q, oracle):
    """Mixed-precision benzene drift-diffusion on the tile engine against
    the wave engine's FP32 chain and the oracle's FP64 log-domain chain from
    the same start: the same decisions and, to FP32 rounding, the same
    positions and E_L walker by walker (17 walkers, 2 sweeps, 1,428 moves)."""
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    W, steps, seed, tau = 17, 2, 7, 0.02
    out = []
    for eng in ("tile", "wave"):
        g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="mixed")
        g.set_engine(eng)
        assert g.engine_name() == eng
        ens = Q.Ensemble(g, W, seed)
        p0 = ens.positions()
        es, e2, acc, tr = ens.sweeps(steps, tau, Q.QMC_MOVE_DRIFT, trace=True, refresh_interval=0, max_drift=1.0)
        out.append((p0, ens.positions(), acc, tr))
    (q1, p1, a1, t1), (q2, p2, a2, t2) = out
    assert np.array_equal(q1, q2)
    pos_o, el_o, acc_o = o.dd_steps(q1, seed, 0, steps, tau, 1.0, log_domain=True)
    assert a1.sum() == a2.sum() == acc_o
    for p, t in ((p2, t2), (pos_o, el_o)):
        dp, de = _walker_dev(p1, p, t1, t)
        assert dp.max() < 5e-4 and de.max() < 5e-4, (dp.max(), de.max())
