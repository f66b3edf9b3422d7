"""Importance-sampled DMC (ImportanceSampledDMC, is_dmc.rs:176-607) on the
GPU against the oracle's restatement (oracle/orc_dmc.c): the same Philox
counters give the same population step for step in FP64 -- survivors,
clones, drops at the capacity and re-initialisation after a collapse."""
import numpy as np
import pytest

from test_gpu_parity import _close, _pair

pytestmark = pytest.mark.gpu

H2_EXACT = -1.174475


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _run_both(Q, g, o, seed, W, n_steps, n_burnin, tau, damping=0.5, capacity=256, n_eq=6, n_fs=None,
              matched=False):
    n_fs = 3 * min(W, 50) if n_fs is None else n_fs
    P = Q.ISDMCParams(n_walkers=W, n_steps=n_steps, n_burnin=n_burnin, time_step=tau, max_drift=1.0,
                      e_ref_damping=damping, n_vmc_equilibrate=n_eq, n_vmc_force_samples=n_fs, capacity=capacity,
                      refresh_interval=0, matched_branching=matched)
    dmc = Q.ImportanceSampledDMC(g, P, seed=seed)
    r = dmc.run()
    ro = o.isdmc_run(seed, W, n_steps, n_burnin, tau, 1.0, damping, n_eq, n_fs, capacity=capacity,
                     matched_branching=matched)
    return dmc, r, ro


@pytest.mark.parametrize("name,W", [("h2", 64), ("li_ref", 48), ("ch4_sto3g", 40), ("ch4_631g", 33)])
def test_isdmc_population_parity_fp64(Q, oracle, name, W):
    g, o = _pair(Q, oracle, name)
    dmc, r, ro = _run_both(Q, g, o, 9, W, 12, 3, 0.01)
    assert np.array_equal(dmc.populations, ro["populations"])
    assert _close(dmc.energies, ro["energies"], 1e-9)
    assert r.energy == pytest.approx(ro["energy"], rel=1e-9)
    assert r.collapses == ro["collapses"] == 0
    assert r.acceptance_rate == pytest.approx(ro["acceptance_rate"], rel=1e-12)
    assert r.avg_population == pytest.approx(ro["avg_population"], rel=1e-12)
    scale = np.abs(ro["forces_vmc"]).max()
    assert np.allclose(r.forces_vmc, ro["forces_vmc"], rtol=1e-8, atol=1e-9 * scale)
    scale = np.abs(ro["forces_mixed"]).max()
    assert np.allclose(r.forces_mixed, ro["forces_mixed"], rtol=1e-8, atol=1e-9 * scale)
    assert np.allclose(r.forces_extrapolated, 2 * r.forces_mixed - r.forces_vmc, rtol=1e-12, atol=1e-12)
    assert r.capacity == 256


@pytest.mark.parametrize("name", ["h2", "ch4_sto3g"])
def test_isdmc_matched_branching_parity_fp64(Q, oracle, name):
    """The branching over 2 tau follows the oracle as the reference form does."""
    g, o = _pair(Q, oracle, name)
    dmc, r, ro = _run_both(Q, g, o, 5, 48, 10, 2, 0.01, matched=True)
    assert np.array_equal(dmc.populations, ro["populations"])
    assert _close(dmc.energies, ro["energies"], 1e-9)
    assert r.acceptance_rate == pytest.approx(ro["acceptance_rate"], rel=1e-12)


def test_isdmc_capacity_drops_match(Q, oracle):
    """Negative damping fills the capacity; the clones past it are dropped
    in source order exactly as the oracle drops them."""
    g, o = _pair(Q, oracle, "h2")
    dmc, r, ro = _run_both(Q, g, o, 2, 40, 30, 2, 0.01, damping=-2.0)
    assert dmc.populations.max() == 256
    assert np.array_equal(dmc.populations, ro["populations"])
    assert r.dropped == ro["dropped"] > 0
    assert _close(dmc.energies, ro["energies"], 1e-9)


def test_isdmc_collapse_reinitialises_like_the_oracle(Q, oracle):
    """A huge damping kills the population: the step's energy is the
    reference's 0/0, fresh walkers (new Philox walker ids) continue, and the
    remaining steps follow the oracle."""
    g, o = _pair(Q, oracle, "h2")
    dmc, r, ro = _run_both(Q, g, o, 2, 40, 30, 2, 0.01, damping=500.0)
    assert r.collapses == ro["collapses"] >= 1
    assert np.array_equal(dmc.populations, ro["populations"])
    assert np.array_equal(np.isnan(dmc.energies), np.isnan(ro["energies"]))
    ok = ~np.isnan(ro["energies"])
    assert _close(dmc.energies[ok], ro["energies"][ok], 1e-9)
    assert np.isnan(r.energy)


def test_isdmc_matched_branching_h2_exact_energy_mixed(Q):
    """Mixed precision, branching over the 2 tau the moves diffuse for: the
    nodeless H2 ground state energy -1.17447 Ha within the statistics."""
    g = Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0, precision="mixed")
    P = Q.ISDMCParams(n_walkers=4096, n_steps=1500, n_burnin=300, time_step=0.005, n_vmc_equilibrate=50,
                      n_vmc_force_samples=200, matched_branching=True)
    dmc = Q.ImportanceSampledDMC(g, P, seed=3)
    r = dmc.run()
    assert abs(r.energy - H2_EXACT) < max(3e-3, 4 * r.error)
    assert r.collapses == 0 and r.dropped == 0
    assert abs(r.avg_population - 4096) < 200


def test_isdmc_zero_variance_ion_mixed(Q):
    """E_L = -4.5 everywhere: weights are 1, the population never changes
    and the energy is exact to FP32 rounding of E_L."""
    s = Q.STOSlaterDet([([1.0], [0], [3.0])], [1], Z=3.0, flags=0, precision="mixed")
    P = Q.ISDMCParams(n_walkers=300, n_steps=40, n_burnin=5, time_step=0.01, n_vmc_equilibrate=5,
                      n_vmc_force_samples=50, refresh_interval=8)
    dmc = Q.ImportanceSampledDMC(s, P, seed=1)
    r = dmc.run()
    assert (dmc.populations == 300).all()
    assert abs(r.energy + 4.5) < 1e-5


def test_isdmc_mixed_agrees_with_fp64(Q):
    """Mixed vs FP64 populations: H2 energies agree within the combined
    statistical error; on CH4 STO-3G the first branchings pick the same
    walkers up to the few decisions that FP32 E_L rounding flips."""
    out = []
    for prec in ("fp64", "mixed"):
        g = Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0, precision=prec)
        P = Q.ISDMCParams(n_walkers=4096, n_steps=600, n_burnin=100, time_step=0.01, n_vmc_equilibrate=20,
                          n_vmc_force_samples=100, refresh_interval=16)
        out.append(Q.ImportanceSampledDMC(g, P, seed=4).run())
    a, b = out
    assert np.isfinite(a.energy) and np.isfinite(b.energy)
    assert abs(a.energy - b.energy) < 5 * np.hypot(a.error, b.error) + 1e-4
    pops = []
    for prec in ("fp64", "mixed"):
        g = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=prec)
        P = Q.ISDMCParams(n_walkers=4096, n_steps=8, n_burnin=0, time_step=0.005, n_vmc_equilibrate=2,
                          n_vmc_force_samples=50, refresh_interval=4)
        dmc = Q.ImportanceSampledDMC(g, P, seed=4)
        dmc.run()
        pops.append(dmc.populations)
    assert np.abs(pops[0] - pops[1]).max() <= 40


def test_isdmc_block_error_matches_oracle(Q, oracle):
    rng = np.random.default_rng(2)
    x = np.cumsum(rng.normal(size=3000)) * 0.01 + rng.normal(size=3000)
    for n in (10, 25, 3000):
        assert Q.isdmc_block_error(x[:n]) == pytest.approx(oracle.isdmc_block_error(x[:n]), rel=1e-12, abs=0)


def test_isdmc_rejects_bad_arguments(Q):
    g = Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0, precision="fp64")
    for bad in (dict(time_step=0.0), dict(capacity=10), dict(capacity=2_000_000), dict(n_steps=0)):
        kw = dict(n_walkers=64, n_steps=5, n_burnin=1, n_vmc_equilibrate=1, n_vmc_force_samples=10)
        kw.update(bad)
        with pytest.raises(Q.QmcError):
            Q.ImportanceSampledDMC(g, Q.ISDMCParams(**kw)).run()


def test_isdmc_benzene_wave_engine_parity_fp64(Q, oracle):
    """Benzene IS-DMC on the wave engine (drift-diffusion sweeps one
    wavefront per walker, all-electron VMC force phase) against the
    oracle's log-domain run: the reference's 1e-30 guard on the linear psi
    (is_dmc.rs:137, 234) would freeze every benzene walker."""
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05)
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    W, n_steps, n_burnin, tau, n_eq, n_fs = 6, 3, 1, 0.01, 2, 6
    P = Q.ISDMCParams(n_walkers=W, n_steps=n_steps, n_burnin=n_burnin, time_step=tau, max_drift=1.0,
                      e_ref_damping=0.5, n_vmc_equilibrate=n_eq, n_vmc_force_samples=n_fs, capacity=256,
                      refresh_interval=0)
    dmc = Q.ImportanceSampledDMC(g, P, seed=3)
    r = dmc.run()
    ro = o.isdmc_run(3, W, n_steps, n_burnin, tau, 1.0, 0.5, n_eq, n_fs, capacity=256, log_domain=True)
    assert np.array_equal(dmc.populations, ro["populations"])
    assert _close(dmc.energies, ro["energies"], 1e-9)
    assert r.acceptance_rate == pytest.approx(ro["acceptance_rate"], rel=1e-12)
    scale = np.abs(ro["forces_vmc"]).max()
    assert np.allclose(r.forces_vmc, ro["forces_vmc"], rtol=1e-8, atol=1e-9 * scale)
    scale = np.abs(ro["forces_mixed"]).max()
    assert np.allclose(r.forces_mixed, ro["forces_mixed"], rtol=1e-8, atol=1e-9 * scale)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_isdmc_wave_engine_matches_register_engine(Q, oracle, precision):
    """CH4 IS-DMC with the sweeps, builds and VMC force phase on the wave
    engine follows the register-engine run (same decisions, same
    population)."""
    out = []
    for eng in ("auto", "wave"):
        g, _ = _pair(Q, oracle, "ch4_sto3g", precision=precision)
        g.set_engine(eng)
        P = Q.ISDMCParams(n_walkers=64, n_steps=20, n_burnin=5, time_step=0.01, max_drift=1.0, e_ref_damping=0.5,
                          n_vmc_equilibrate=5, n_vmc_force_samples=100, capacity=256, refresh_interval=8)
        d = Q.ImportanceSampledDMC(g, P, seed=4)
        out.append((d.run(), d.populations, d.energies))
    (ra, pa, ea), (rb, pb, eb) = out
    assert np.array_equal(pa, pb)
    tol = 1e-9 if precision == "fp64" else 1e-5
    assert _close(ea, eb, tol)
    assert ra.acceptance_rate == pytest.approx(rb.acceptance_rate, rel=1e-12)
    assert np.allclose(ra.forces_vmc, rb.forces_vmc, rtol=tol, atol=tol * np.abs(rb.forces_vmc).max())


@pytest.mark.parametrize("name,precision", [("ch4_sto3g", "fp64"), ("ch4_sto3g", "mixed"), ("h2", "fp64")])
def test_isdmc_graph_replay_matches_individual_launches(Q, oracle, name, precision, monkeypatch):
    """Pairs of IS-DMC steps replay a captured graph whose kernels read the
    step, trace slot and force flag from the device; the same run with
    every step launched on its own (QMC_DMC_GRAPH=0) gives identical
    populations, energies and forces.  171 steps in chunks of 64 leave odd
    tails, the burn-in ends inside a pair, and the mixed run refreshes its
    inverses every 8 steps between replays."""
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("QMC_DMC_GRAPH", flag)
        g, _ = _pair(Q, oracle, name, precision=precision)
        P = Q.ISDMCParams(n_walkers=96, n_steps=150, n_burnin=21, time_step=0.01, max_drift=1.0, e_ref_damping=0.5,
                          n_vmc_equilibrate=5, n_vmc_force_samples=100, capacity=512, refresh_interval=8)
        d = Q.ImportanceSampledDMC(g, P, seed=6)
        r = d.run()
        out.append((r, d.populations.copy(), d.energies.copy()))
    (ra, pa, ea), (rb, pb, eb) = out
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ea, eb)
    np.testing.assert_array_equal(ra.forces_mixed, rb.forces_mixed)
    assert ra.energy == rb.energy and ra.acceptance_rate == rb.acceptance_rate
