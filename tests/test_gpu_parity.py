"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

FP64 mode must reproduce the oracle to 1e-10 relative on identical
configurations (log psi, grad ln psi = drift/2, lap psi / psi, E_L), and
follow the oracle's Metropolis trajectories step for step from the same
Philox seed.  Mixed mode is held to the FP32 tolerances stated per test.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL64 = 1e-10


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _pair(Q, oracle, name, precision="fp64"):
    """(GPU system, oracle system) built from the same reference constructor."""
    O = oracle.System
    if name == "h2":
        return Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0, precision=precision), O.h2mo(1.3, 0.7, 3.0)
    if name == "li_ref":
        return Q.Lithium(None, precision=precision), O.lithium_ref(exact=False)
    if name == "li_ref_exact":
        return Q.Lithium(None, flags=0, precision=precision), O.lithium_ref(exact=True)
    if name == "li_seeded":
        return Q.Lithium(0, precision=precision), O.lithium_config(0, exact=False)
    if name == "ch4_631g":
        return Q.MethaneGTO(1.5, 2.0, 2.05, "6-31G", precision=precision), O.methane_gto(1.5, 2.0, 2.05)
    if name == "ch4_sto3g":
        return (Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision),
                O.methane_sto3g(1.5, 2.0, 2.06, 0))
    raise KeyError(name)


SYSTEMS = ["h2", "li_ref", "li_ref_exact", "li_seeded", "ch4_631g", "ch4_sto3g"]


def _configs(oracle_sys, W, seed=3, equil=4):
    """Seeded configurations, partly equilibrated with oracle sweeps."""
    pos = oracle_sys.initialize_batch(seed, W)
    if equil:
        pos, _, _, _ = oracle_sys.single_electron_steps(pos, seed, 1000, equil, 0.4)
    return pos


def _close(a, b, rtol, scale=None):
    a = np.asarray(a)
    b = np.asarray(b)
    s = np.abs(b) if scale is None else scale
    return np.all(np.abs(a - b) <= rtol * np.maximum(s, 1e-300) + rtol * 1e-3)


@pytest.mark.parametrize("name", SYSTEMS)
def test_eval_parity_fp64(Q, oracle, name):
    g, o = _pair(Q, oracle, name)
    pos = _configs(o, 97)  # ragged: not a multiple of the 64-lane block
    ref = o.batch_eval(pos)
    out = g.evaluate_batch(pos)
    assert np.array_equal(out["sign"], ref["sign"])
    assert _close(out["logabs"], ref["logabs"], RTOL64)
    gs = np.abs(ref["grad"]).max(axis=(1, 2))[:, None, None]
    assert _close(out["grad"], ref["grad"], RTOL64, scale=gs)
    ls = np.abs(ref["lap"]).max(axis=1)[:, None]
    assert _close(out["lap"], ref["lap"], RTOL64, scale=ls)
    assert _close(out["eloc"], ref["eloc"], RTOL64)


@pytest.mark.parametrize("name", ["h2", "li_seeded", "ch4_631g", "ch4_sto3g"])
def test_eval_mixed_precision_error(Q, oracle, name):
    """FP32 orbitals/Jastrow with an FP64 inverse: per-configuration E_L
    within 2e-4 relative (plus 1e-3 Ha absolute), ln|psi| to 1e-12."""
    g, o = _pair(Q, oracle, name, precision="mixed")
    pos = _configs(o, 64)
    ref = o.batch_eval(pos)
    out = g.evaluate_batch(pos)
    assert _close(out["logabs"], ref["logabs"], 1e-12)
    err = np.abs(out["eloc"] - ref["eloc"])
    assert np.all(err <= 2e-4 * np.abs(ref["eloc"]) + 1e-3), err.max()


@pytest.mark.parametrize("name", SYSTEMS)
def test_single_electron_trajectory_parity(Q, oracle, name):
    """Same seed -> the GPU's Sherman-Morrison sweeps follow the oracle's
    brute-force single-electron Metropolis chain step for step."""
    g, o = _pair(Q, oracle, name)
    W, steps, seed, step_size = 70, 12, 11, 0.45
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, step_size, Q.QMC_MOVE_SINGLE, refresh_interval=5, steps_per_launch=3,
                                 trace=True)
    pos_o, lp_o, el_o, acc_o = o.single_electron_steps(p0, seed, 0, steps, step_size)
    assert acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr, el_o, 1e-9)
    la, sg, el = ens.state()
    assert _close(la, lp_o, 1e-9)
    assert np.allclose(es, el_o.sum(axis=1), rtol=1e-9)


@pytest.mark.parametrize("name", ["h2", "li_ref", "ch4_sto3g"])
def test_all_electron_trajectory_parity(Q, oracle, name):
    """MCMCSimulation::metropolis_step semantics (vmc.rs:95-119): all
    electrons moved, E_L recomputed on acceptance only."""
    g, o = _pair(Q, oracle, name)
    W, steps, seed, step_size = 66, 15, 5, 0.3
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    la0, _, el0 = ens.state()
    es, e2, acc, tr = ens.sweeps(steps, step_size, Q.QMC_MOVE_ALL, steps_per_launch=4, trace=True)
    ref = o.batch_eval(p0)
    assert _close(el0, ref["eloc"], 1e-10)
    pos_o, lp_o, el_o, acc_o = o.all_electron_steps(p0, ref["logabs"], ref["eloc"], seed, 0, steps, step_size)
    assert acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr[-1], el_o, 1e-9)


def test_vmc_run_reproduces_oracle_mcmc(Q, oracle):
    """qmc_vmc_run (MCMCSimulation::run) vs the oracle's literal run:
    identical per-step energy series, tau and blocking error."""
    g, o = _pair(Q, oracle, "h2")
    sim = Q.MCMCSimulation(g, Q.MCMCParams(64, 300, 0.5, 2.0, 0.05, 0.5, 100, 0, Q.QMC_MOVE_ALL), seed=2)
    res = sim.run()
    ores, oe = o.mcmc_run(64, 300, 0.5, seed=2, max_step=2.0, min_step=0.05, target=0.5, interval=100)
    assert np.allclose(sim.energies, oe, rtol=1e-9)
    assert res.energy == pytest.approx(ores.energy, rel=1e-10)
    assert res.autocorrelation_time == pytest.approx(ores.autocorrelation_time, rel=1e-6)
    assert res.acceptance == pytest.approx(ores.acceptance, rel=1e-12)
    assert res.final_step_size == pytest.approx(ores.final_step, rel=1e-12)


def test_sweeps_reproducible(Q, oracle):
    g, _ = _pair(Q, oracle, "ch4_sto3g", precision="mixed")
    runs = []
    for _ in range(2):
        ens = Q.Ensemble(g, 200, 9)
        runs.append(ens.sweeps(10, 0.3))
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_mixed_vs_fp64_ensemble_energy(Q, oracle):
    """Mixed precision samples the same |psi|^2: ensemble energies of FP64
    and mixed runs of CH4/STO-3G agree within 2.5 sigma."""
    res = {}
    for prec in ("fp64", "mixed"):
        g, _ = _pair(Q, oracle, "ch4_sto3g", precision=prec)
        ens = Q.Ensemble(g, 4096, 21)
        ens.sweeps(40, 0.3)
        es, e2, acc = ens.sweeps(160, 0.3, refresh_interval=64)
        tr = es / 4096
        tau = Q.autocorrelation_time(tr)
        res[prec] = (tr.mean(), Q.blocking_error(tr, tau))
    (a, sa), (b, sb) = res["fp64"], res["mixed"]
    assert abs(a - b) < 2.5 * math.hypot(sa, sb) + 1e-4


def test_device_pointer_inputs(Q, oracle):
    """Device-resident positions and outputs (buffers from the library's own
    allocator) give the same results as host buffers."""
    g, o = _pair(Q, oracle, "ch4_631g")
    pos = _configs(o, 33)
    out_h = g.evaluate_batch(pos)
    with Q.DeviceBuffer(pos.nbytes) as dpos, Q.DeviceBuffer(33 * 8) as del_:
        dpos.upload(pos)
        Q.check(Q.lib().qmc_local_energy(g._h, 33, dpos.ptr, del_.ptr, None))
        el = del_.download(np.float64, 33)
    assert np.array_equal(el, out_h["eloc"])


def test_singular_configuration_reported(Q, oracle):
    g, _ = _pair(Q, oracle, "ch4_631g")
    pos = _configs(oracle.System.methane_gto(), 4)
    # a spin-up electron 170 bohr out: every Gaussian underflows to exactly 0,
    # the Slater matrix has a zero row and its last pivot is exactly 0
    pos[1, 2] = [100.0, 100.0, 100.0]
    with pytest.raises(Q.QmcError) as ei:
        g.evaluate_batch(pos)
    assert ei.value.code == -4
    g.set_flags(Q.QMC_FLAGS_REFERENCE)  # methane.rs:829 identity fallback
    out = g.evaluate_batch(pos, ("logabs", "sign"))
    assert out["logabs"][1] == -np.inf and out["sign"][1] == 0.0
    assert np.isfinite(out["logabs"][[0, 2, 3]]).all()


def test_empty_and_single_walker(Q, oracle):
    g, o = _pair(Q, oracle, "h2")
    assert g.evaluate_batch(np.zeros((0, 2, 3)))["eloc"].shape == (0,)
    pos = _configs(o, 1)
    assert g.local_energy(pos[0]) == pytest.approx(o.local_energy(pos[0]), rel=RTOL64)


def test_hydrogenic_zero_variance(Q):
    """One electron in STO{c=1,p=0,zeta=3} about Z=3: E_L = -4.5 everywhere
    (lithium.rs:84-99)."""
    s = Q.STOSlaterDet([([1.0], [0], [3.0])], [1], Z=3.0, flags=0)
    rng = np.random.default_rng(0)
    pos = rng.normal(size=(50, 1, 3)) * 0.7
    el = s.evaluate_batch(pos, ("eloc",))["eloc"]
    assert np.allclose(el, -4.5, rtol=0, atol=1e-12)


def test_nuclei_and_params_roundtrip(Q, oracle):
    g, o = _pair(Q, oracle, "ch4_631g")
    assert g.nuclear_repulsion() == pytest.approx(13.499626641060864, rel=1e-14)
    R = g.get_nuclei()
    R2 = R + 0.01 * np.arange(R.size).reshape(R.shape)
    g.set_nuclei(R2)
    o.set_nuclei(R2)
    p = g.get_params()
    assert np.allclose(p, [1.5, 2.0, 0.5, 0.25])
    p2 = p * 1.1
    g.set_params(p2)
    o.set_params(p2)
    pos = _configs(o, 20)
    assert _close(g.evaluate_batch(pos)["eloc"], o.batch_eval(pos)["eloc"], RTOL64)


SPEC_NAMES = {"h2": "h2-slater1s", "li_ref": "li-init-li-sto", "li_ref_exact": "li-init-li-sto",
              "li_seeded": "li-sto3term", "ch4_631g": "ch4-631g", "ch4_sto3g": "ch4-sto3g"}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", SYSTEMS)
def test_compiled_shell_kernels_match_generic(Q, oracle, name, precision):
    """The kernels specialised to a system's shell structure are selected for
    the reference systems and agree with the run-time-loop kernels on
    evaluations and on a short Sherman-Morrison trajectory."""
    g, o = _pair(Q, oracle, name, precision=precision)
    assert g.engine_name() == "register/" + SPEC_NAMES[name]
    pos = _configs(o, 70)
    spec = g.evaluate_batch(pos)
    W, steps, seed, step = 65, 5, 11, 0.4
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    tr_spec = ens.sweeps(steps, step, trace=True, refresh_interval=3, steps_per_launch=2)[3]
    pos_spec = ens.positions()
    g.set_engine("generic")
    assert g.engine_name() == "register/generic"
    gen = g.evaluate_batch(pos)
    ens2 = Q.Ensemble(g, W, seed)
    assert np.array_equal(ens2.positions(), p0)
    tr_gen = ens2.sweeps(steps, step, trace=True, refresh_interval=3, steps_per_launch=2)[3]
    tol = 1e-12 if precision == "fp64" else 2e-5
    for k in ("logabs", "eloc"):
        assert _close(spec[k], gen[k], tol), k
    assert np.allclose(pos_spec, ens2.positions(), rtol=0, atol=1e-9 if precision == "fp64" else 1e-5)
    # FP32 E_L is a sum of terms of order 10 Ha that can cancel to ~0: compare
    # it to 1e-4 of max(|E_L|, 1 Ha)
    assert _close(tr_spec, tr_gen, 1e-10 if precision == "fp64" else 1e-4,
                  scale=None if precision == "fp64" else np.maximum(np.abs(tr_gen), 1.0))


@pytest.mark.parametrize("name", SYSTEMS)
def test_compiled_shell_kernels_all_electron(Q, oracle, name):
    """All-electron moves (vmc.rs:95-119) run the run-time-loop kernel under
    both engine settings (the compiled-shell all-electron kernels are not
    built; see test_compiled_shell_all_electron_kernel): the state built by
    the compiled-shell refresh equals the oracle's ln|psi|, and the steps
    taken after it equal those taken after the run-time-loop refresh."""
    g, o = _pair(Q, oracle, name)
    W, seed, step = 66, 5, 0.3
    res = []
    for eng in ("auto", "generic"):
        g.set_engine(eng)
        ens = Q.Ensemble(g, W, seed)
        st0 = ens.state()
        out = [st0]
        for _ in range(3):
            acc = ens.sweeps(1, step, Q.QMC_MOVE_ALL)[2]
            out.append(ens.state() + (acc,))
        res.append(out)
    ref0 = o.batch_eval(Q.Ensemble(g, W, seed).positions())
    assert _close(res[0][0][0], ref0["logabs"], 1e-10), "state ln|psi| after the build"
    for a, b in zip(res[0], res[1]):
        for x, y in zip(a, b):
            assert _close(x, y, 1e-10)


@pytest.mark.skip(reason="the all-electron kernel compiled for a shell structure is not built into the library: "
                         "its CH4 6-31G FP32 instance faulted the GPU (memory aperture violation) and its CH4 "
                         "STO-3G FP64 instance once gave wrong results; the host build of the same source agrees "
                         "with the run-time-loop kernel and the oracle, and no cause was found in the source")
def test_compiled_shell_all_electron_kernel():
    """Placeholder for the compiled-shell all-electron kernel, which stays out
    of the library until the cause of its fault is shown."""


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_compiled_shell_sweeps_match_generic_at_full_size(Q, oracle, precision):
    """The headline kernels (CH4 STO-3G sweeps and FP64 refresh compiled for
    the shell structure) against the run-time-loop kernels on the benchmark's
    65,536 walkers: the same state after the build, the same accepted moves
    and positions after four sweeps (FP64 to 1e-9 bohr; in mixed precision
    the two FP32 instruction orders may round a few acceptance tests
    differently, so at most 1e-4 of the walkers may part)."""
    g, _ = _pair(Q, oracle, "ch4_sto3g", precision=precision)
    W, seed = 65536, 21
    out = {}
    for eng in ("auto", "generic"):
        g.set_engine(eng)
        ens = Q.Ensemble(g, W, seed)
        st = ens.state()
        es, e2, acc = ens.sweeps(4, 0.3, refresh_interval=2, steps_per_launch=4)[:3]
        out[eng] = (st, acc, ens.positions(), ens.state())
    g.set_engine("auto")
    (st_a, acc_a, pos_a, end_a), (st_g, acc_g, pos_g, end_g) = out["auto"], out["generic"]
    assert _close(st_a[0], st_g[0], 1e-12), "ln|psi| after the build"
    assert np.array_equal(st_a[1], st_g[1]), "sign after the build"
    tol_el = 1e-10 if precision == "fp64" else 2e-4
    assert _close(st_a[2], st_g[2], tol_el, scale=np.maximum(np.abs(st_g[2]), 1.0)), "E_L after the build"
    moved = np.any(np.abs(pos_a - pos_g) > (1e-9 if precision == "fp64" else 1e-4), axis=(1, 2))
    if precision == "fp64":
        assert np.array_equal(acc_a, acc_g)
        assert not moved.any(), np.nonzero(moved)[0][:10]
        assert _close(end_a[0], end_g[0], 1e-10)
    else:
        assert moved.sum() <= W * 1e-4, moved.sum()
        assert np.all(np.abs(acc_a - acc_g) <= W * 1e-4 * 10)
