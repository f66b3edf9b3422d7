"""GPU parity of the SR-gradient and force path against the oracle.

O_k (OptimizableWfn::log_derivatives), Omega (wfn_nuclear_gradient) and the
Hellmann-Feynman force are compared with the oracle's analytic restatements
at 1e-10; d E_L / d p has no reference counterpart and is held to central
finite differences of the oracle's literal E_L (1e-5 relative); the ensemble
moments from the FP64 matrix-core Gram are compared with numpy sums of the
per-walker values; the SR solve and the ZV force combination with the
oracle's literal sr_update / zv_forces on the same samples.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _series_params(sys_, seed):
    p = sys_.get_params()
    rng = np.random.default_rng(seed)
    p[4:] = rng.normal(size=p.size - 4) * 0.05
    return p


def _pair(Q, oracle, name, series=False, precision="fp64"):
    if name == "ch4_sto3g":
        g = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision)
        o = oracle.System.methane_sto3g(1.5, 2.0, 2.06, 0)
    elif name == "ch4_631g":
        g = Q.MethaneGTO(2.0, 3.0, 2.05, "6-31G", precision=precision)
        o = oracle.System.methane_gto(2.0, 3.0, 2.05)
    elif name == "h2":
        g = Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0, precision=precision)
        o = oracle.System.h2mo(1.3, 0.7, 3.0)
    else:
        raise KeyError(name)
    if series:
        g.set_jastrow_series(6, 8)
        o.set_jastrow_series(6, 8)
        p = _series_params(g, 7)
        g.set_params(p)
        o.set_params(p)
    return g, o


def _configs(o, W, seed=4):
    pos = o.initialize_batch(seed, W)
    pos, _, _, _ = o.single_electron_steps(pos, seed, 500, 3, 0.4)
    return pos


@pytest.mark.parametrize("name,series", [("ch4_sto3g", False), ("ch4_sto3g", True), ("ch4_631g", False),
                                         ("ch4_631g", True), ("h2", False)])
def test_derivatives_parity(Q, oracle, name, series):
    g, o = _pair(Q, oracle, name, series)
    W = 70
    pos = _configs(o, W)
    out = g.derivatives_batch(pos)
    P = g.num_params()
    assert P == (32 if series else (4 if name != "h2" else 0))
    for w in range(W):
        r = pos[w]
        if P:
            Oo = o.log_derivatives(r)
            assert np.allclose(out["O"][w], Oo, rtol=1e-10, atol=1e-10 * np.abs(Oo).max())
        Fo = o.hf_force(r)
        assert np.allclose(out["F"][w], Fo, rtol=1e-10, atol=1e-10 * np.abs(Fo).max())
        Om = o.wfn_nuclear_gradient(r)
        # H2's reference Omega is a central difference with h = 1e-5 (hydrogen.rs:260-290)
        tol = 1e-7 if name == "h2" else 1e-10
        assert np.allclose(out["Om"][w], Om, rtol=tol, atol=tol * np.abs(Om).max())
        assert out["eloc"][w] == pytest.approx(o.local_energy(r), rel=1e-10)


@pytest.mark.parametrize("series", [False, True])
def test_dEL_dp_against_finite_differences(Q, oracle, series):
    g, o = _pair(Q, oracle, "ch4_sto3g", series)
    pos = _configs(o, 6)
    dE = g.derivatives_batch(pos)["dE"]
    p0 = o.get_params()
    for w in range(pos.shape[0]):
        for k in range(p0.size):
            h = 1e-5 * max(1.0, abs(p0[k]))
            pp = p0.copy(); pp[k] += h
            pm = p0.copy(); pm[k] -= h
            o.set_params(pp); ep = o.local_energy(pos[w])
            o.set_params(pm); em = o.local_energy(pos[w])
            o.set_params(p0)
            fd = (ep - em) / (2 * h)
            assert dE[w, k] == pytest.approx(fd, rel=1e-5, abs=1e-6), (w, k)


def test_hf_force_physical_sign(Q, oracle):
    """Without QMC_FLAG_HF_NN_REFSIGN the nucleus-nucleus force enters with
    the physical sign: F_exact = F_reference + 2 F_nn."""
    g, o = _pair(Q, oracle, "ch4_sto3g")
    pos = _configs(o, 5)
    F_ref = g.derivatives_batch(pos)["F"]
    g.set_flags(0)
    F_ex = g.derivatives_batch(pos)["F"]
    R, Z = o.nuclei()
    Fnn = np.zeros_like(R)
    for I in range(len(Z)):
        for J in range(len(Z)):
            if I != J:
                d = R[I] - R[J]
                Fnn[I] += Z[I] * Z[J] * d / np.linalg.norm(d) ** 3
    assert np.allclose(F_ex, F_ref + 2 * Fnn[None], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_moments_match_per_walker_sums(Q, oracle, precision):
    """The FP64 MFMA Gram of the feature matrix equals numpy sums of the
    per-walker derivatives of the same configurations."""
    g, o = _pair(Q, oracle, "ch4_sto3g", series=True, precision=precision)
    W = 300
    ens = Q.Ensemble(g, W, 13)
    ens.sweeps(5, 0.3)
    ens.measure(reset=True)
    m = ens.moments()
    pos = ens.positions()
    d = g.derivatives_batch(pos)
    E = d["eloc"]
    O = np.concatenate([d["O"], d["Om"].reshape(W, -1)], axis=1)
    om = O.mean(0)
    tol = 1e-9 if precision == "fp64" else 1e-9  # same precision path on both sides
    assert m["n_samples"] == W
    assert m["energy"] == pytest.approx(E.mean(), rel=tol)
    assert m["variance"] == pytest.approx(E.var(), rel=1e-7)
    assert np.allclose(m["o_mean"], om, rtol=tol, atol=tol)
    S = O.T @ O / W - np.outer(om, om)
    assert np.allclose(m["S"], S, rtol=1e-7, atol=1e-9 * np.abs(S).max())
    f = (E[:, None] * O).mean(0) - E.mean() * om
    assert np.allclose(m["f"], f, rtol=1e-7, atol=1e-9 * np.abs(f).max())
    assert np.allclose(m["de_mean"], d["dE"].mean(0), rtol=1e-9, atol=1e-9)
    assert np.allclose(m["f_hf"], d["F"].mean(0), rtol=1e-9, atol=1e-9)
    Om = d["Om"]
    cov = (E[:, None, None] * Om).mean(0) - E.mean() * Om.mean(0)
    assert np.allclose(m["f_zv"], d["F"].mean(0) - 2 * cov, rtol=1e-7, atol=1e-8)


def test_sr_step_and_zv_forces_match_oracle(Q, oracle):
    """One SR step (S + eps I) dp = -f and the ZV force from the ensemble
    moments equal the oracle's literal compute_sr_update / sample_forces
    combination on the same configurations."""
    g, o = _pair(Q, oracle, "ch4_631g")
    W = 400
    ens = Q.Ensemble(g, W, 3)
    ens.sweeps(20, 0.3)
    ens.measure(reset=True)
    m = ens.moments()
    pos = ens.positions()
    E = np.array([o.local_energy(r) for r in pos])
    Oj = np.array([o.log_derivatives(r) for r in pos])
    dp_o, em, var, sing = oracle.sr_update(E, Oj, 1e-3)
    P = 4
    dp, singular = Q.sr_solve(m["S"][:P, :P], m["f"][:P], 1e-3)
    assert not singular and sing == 0
    assert np.allclose(dp, dp_o, rtol=1e-7, atol=1e-9)
    assert m["energy"] == pytest.approx(em, rel=1e-10)
    F = np.array([o.hf_force(r) for r in pos])
    Om = np.array([o.wfn_nuclear_gradient(r) for r in pos])
    mf, me, err, fv = oracle.zv_forces(E, F, Om, True)
    assert np.allclose(m["f_zv"], mf, rtol=1e-7, atol=1e-8)
    mf0, _, _, fv0 = oracle.zv_forces(E, F, Om, False)
    assert np.allclose(m["f_hf"], mf0, rtol=1e-9, atol=1e-9)
    assert np.allclose(m["f_var"], fv0, rtol=1e-7, atol=1e-9)


def test_sr_step_47_columns_matches_oracle(Q, oracle):
    """The CH4 differentiation workload's 47-dimensional SR system (32
    Jastrow parameters + 15 nuclear coordinates, O = [O_J, Omega]): the step
    solved from the GPU moments equals the oracle's literal
    compute_sr_update (sr_optimize.rs:176-246) on the same 47-column
    samples, and so does its parameter block."""
    g, o = _pair(Q, oracle, "ch4_sto3g", series=True)
    W = 600
    ens = Q.Ensemble(g, W, 5)
    ens.sweeps(20, 0.3)
    ens.measure(reset=True)
    m = ens.moments()
    pos = ens.positions()
    E = np.array([o.local_energy(r) for r in pos])
    Oj = np.array([o.log_derivatives(r) for r in pos])
    Om = np.array([np.asarray(o.wfn_nuclear_gradient(r)).ravel() for r in pos])
    O = np.hstack([Oj, Om])
    assert O.shape == (W, 47) and m["S"].shape == (47, 47)
    dp_o, em, var, sing = oracle.sr_update(E, O, 1e-3)
    dp, singular = Q.sr_solve(m["S"], m["f"], 1e-3)
    assert not singular and sing == 0
    scale = np.abs(dp_o).max()
    assert np.allclose(dp, dp_o, rtol=1e-6, atol=1e-8 * scale), np.abs(dp - dp_o).max() / scale
    assert m["energy"] == pytest.approx(em, rel=1e-10)
    assert m["variance"] == pytest.approx(var, rel=1e-8)


def test_sr_optimize_reference_case(Q):
    """sr_optimize.rs:399-427: MethaneGTO from (b_ee, b_en) = (0.5, 1.0);
    the final energy lies in (-60, -30) Ha with a finite positive variance
    and the parameters move."""
    g = Q.MethaneGTO(0.5, 1.0, 2.05, "6-31G")
    opt = Q.SROptimizer(n_samples=3000 * 16, n_walkers=1024, n_equilibrate=200, max_iterations=15,
                        learning_rate=0.03, step_size=0.3)
    res = opt.optimize(g)
    assert -60.0 < res.final_energy < -30.0
    assert res.final_variance > 0 and math.isfinite(res.final_variance)
    assert np.abs(res.param_history[-1] - res.param_history[0]).max() > 1e-4


def test_sr_with_nuclear_block_runs(Q):
    """The 47-dimensional SR system of the CH4 differentiation workload
    (32 Jastrow parameters + 15 nuclear coordinates) solves and updates."""
    g = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0)
    g.set_jastrow_series(6, 8)
    opt = Q.SROptimizer(n_samples=4096, n_walkers=2048, n_equilibrate=50, max_iterations=3, step_size=0.3,
                        include_nuclear=True, move=Q.QMC_MOVE_SINGLE)
    res = opt.optimize(g)
    assert res.param_history.shape == (4, 32)
    assert np.isfinite(res.energy_history).all()


def test_sample_forces_bare_and_zero_variance(Q):
    g = Q.MethaneGTO(1.5, 2.0, 2.05, "6-31G")
    common = dict(n_samples=8192 * 4, n_walkers=8192, n_equilibrate=100, mc_step_size=0.3,
                  move=Q.QMC_MOVE_SINGLE, seed=1)
    Fb, eb, errb, fvb = Q.GeometryOptimizer(zero_variance=False, **common).sample_forces(g)
    Fz, ez, errz, fvz = Q.GeometryOptimizer(zero_variance=True, **common).sample_forces(g)
    assert eb == pytest.approx(ez, rel=1e-12)  # same seed, same walk
    assert np.isfinite(Fz).all() and np.isfinite(Fb).all()


@pytest.mark.parametrize("name", ["ch4_sto3g", "ch4_631g", "h2"])
def test_compiled_shell_derivatives_match_generic(Q, oracle, name):
    """Derivative kernels specialised to the shell structure agree with the
    run-time-loop kernels (O_k, dE_L/dp, Omega, HF force) with the Pade
    series switched on."""
    from test_gpu_parity import _configs, _pair
    g, o = _pair(Q, oracle, name)
    if name != "h2":
        g.set_jastrow_series(3, 2)
        p = g.get_params()
        p[4:] = np.linspace(-0.05, 0.05, p.size - 4)
        g.set_params(p)
    pos = _configs(o, 70)
    a = g.derivatives_batch(pos)
    g.set_engine("generic")
    b = g.derivatives_batch(pos)
    for k in a:
        if b[k].size:
            assert np.allclose(a[k], b[k], rtol=1e-11, atol=1e-11 * (1 + np.abs(b[k]).max())), k


@pytest.mark.parametrize("name", ["h2", "ch4_sto3g", "ch4_631g"])
def test_two_wave_mixed_derivatives_match_one_wave_kernel(Q, oracle, name):
    """Mixed precision with a compiled shell structure runs the derivative
    kernel with one wavefront per spin sector (each wave takes its own
    same-spin pairs and a share of the opposite-spin pairs); the run-time
    shell kernels (one lane per walker, every pair) give the same columns
    to FP32 rounding, so no pair is lost or counted twice."""
    from test_gpu_parity import _configs
    g, o = _pair(Q, oracle, name, series=name != "h2", precision="mixed")
    pos = _configs(o, 150)
    a = g.derivatives_batch(pos)
    g.set_engine("generic")
    b = g.derivatives_batch(pos)
    for k in a:
        if b[k].size:
            s = np.abs(b[k]).max()
            assert np.allclose(a[k], b[k], rtol=2e-5, atol=2e-5 * (1 + s)), (k, np.abs(a[k] - b[k]).max() / (1 + s))
    ref = g.derivatives_batch(pos)
    g64, _ = _pair(Q, oracle, name, series=name != "h2")
    c = g64.derivatives_batch(pos)
    assert np.allclose(a["eloc"], c["eloc"], rtol=1e-4, atol=1e-3)
    assert ref["O"].shape == c["O"].shape
