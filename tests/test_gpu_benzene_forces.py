"""Benzene's SR-gradient and force path on the GPU (the wave derivative
kernel: one wavefront per walker).

BenzeneGTO::set_nuclei rebuilds the C-H sigma MO coefficients from the
bond directions (benzene.rs:786-809), so the nuclear gradient Omega of
ln psi carries the coefficients' dependence on the geometry as well as the
AO centres and the e-n Jastrow.  The reference has no Omega for benzene
(the trait default panics); the oracle defines it by central differences
of ln|psi| through set_nuclei (h = 1e-5, the method of hydrogen.rs:260-290),
pinned there by the translation sum rule.  O_k, the Hellmann-Feynman force
and E_L are the oracle's analytic restatements, compared at 1e-10.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _bent(R0, seed=3, scale=0.12):
    return R0 + np.random.default_rng(seed).normal(size=R0.shape) * scale


def _pair(Q, oracle, precision="fp64", bent=True):
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision=precision)
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    if bent:
        R0, _ = o.nuclei()
        R = _bent(R0)
        o.set_nuclei(R)
        g.set_nuclei(R)
    return g, o


def _configs(o, W, seed=4):
    pos = o.initialize_batch(seed, W)
    pos, _, _, _ = o.single_electron_steps(pos, seed, 300, 2, 0.3)
    return pos


@pytest.mark.parametrize("bent", [False, True])
def test_benzene_derivatives_parity_fp64(Q, oracle, bent):
    g, o = _pair(Q, oracle, bent=bent)
    W = 6
    pos = _configs(o, W)
    out = g.derivatives_batch(pos)
    assert g.num_params() == 4
    for w in range(W):
        r = pos[w]
        Oo = o.log_derivatives(r)
        assert np.allclose(out["O"][w], Oo, rtol=1e-10, atol=1e-10 * np.abs(Oo).max()), w
        Fo = o.hf_force(r)
        assert np.allclose(out["F"][w], Fo, rtol=1e-10, atol=1e-10 * np.abs(Fo).max()), w
        assert out["eloc"][w] == pytest.approx(o.local_energy(r), rel=1e-10)
        Om = o.wfn_nuclear_gradient(r)  # central differences, h = 1e-5
        assert np.allclose(out["Om"][w], Om, rtol=1e-6, atol=1e-6 * np.abs(Om).max()), (w, np.abs(out["Om"][w] - Om).max())


def test_benzene_omega_translation_sum_rule(Q, oracle):
    """Moving every electron and nucleus together leaves every benzene MO
    (the C-H directions included) and so psi unchanged:
    sum_I Omega_I + sum_i grad_i ln psi = 0, with the GPU's own gradient."""
    g, o = _pair(Q, oracle)
    pos = _configs(o, 5, seed=9)
    Om = g.derivatives_batch(pos)["Om"].reshape(5, -1, 3)
    gr = g.evaluate_batch(pos, want=("grad",))["grad"]
    tot = Om.sum(axis=1) + gr.sum(axis=1)
    scale = np.abs(Om).sum(axis=(1, 2))[:, None]
    assert np.all(np.abs(tot) <= 1e-11 * scale + 1e-12), np.abs(tot).max()


def test_benzene_omega_without_the_mo_rule_misses_the_bond_term(Q, oracle):
    """The coefficient term is large enough for the parity test to see: the
    same molecule with its MO coefficients frozen at the bent geometry
    (mo_rule fixed: same psi, no dC/dR) misses the oracle's Omega by orders
    of magnitude more than the full kernel does."""
    g, o = _pair(Q, oracle, bent=True)
    pos = _configs(o, 2, seed=11)
    Om_gpu = g.derivatives_batch(pos)["Om"]
    fixed = Q.System.from_json(g.to_json().replace('"mo_rule": "benzene_ch"', '"mo_rule": "fixed"'))
    Om_fixed = fixed.derivatives_batch(pos)["Om"]
    Om_o = np.array([o.wfn_nuclear_gradient(r) for r in pos]).reshape(Om_gpu.shape)
    assert np.abs(Om_fixed - Om_o).max() > 1e3 * np.abs(Om_gpu - Om_o).max()


def test_benzene_derivatives_mixed_precision(Q, oracle):
    g64, o = _pair(Q, oracle)
    g32, _ = _pair(Q, oracle, precision="mixed")
    pos = _configs(o, 8, seed=5)
    a, b = g32.derivatives_batch(pos), g64.derivatives_batch(pos)
    for k, tol in (("O", 1e-4), ("F", 1e-4), ("Om", 1e-3), ("dE", 1e-3)):
        s = np.abs(b[k]).max()
        assert np.abs(a[k] - b[k]).max() <= tol * s, (k, np.abs(a[k] - b[k]).max() / s)
    assert np.all(np.abs(a["eloc"] - b["eloc"]) <= 2e-4 * np.abs(b["eloc"]) + 1e-2)


@pytest.mark.parametrize("name", ["h2", "ch4_sto3g"])
def test_wave_derivatives_match_register_kernels(Q, oracle, name):
    """The wave derivative kernel on systems the register kernels also run."""
    if name == "h2":
        mk = lambda: Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0)
        o = oracle.System.h2mo(1.3, 0.7, 3.0)
    else:
        mk = lambda: Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0)
        o = oracle.System.methane_sto3g(1.5, 2.0, 2.06, 0)
    reg, wav = mk(), mk()
    wav.set_engine("wave")
    pos = _configs(o, 40, seed=2)
    a, b = wav.derivatives_batch(pos), reg.derivatives_batch(pos)
    for k in ("eloc", "O", "dE", "F", "Om"):
        if b[k].size == 0:
            continue
        assert np.allclose(a[k], b[k], rtol=1e-10, atol=1e-10 * (1 + np.abs(b[k]).max())), k


def test_benzene_sample_forces_and_one_geometry_step(Q):
    """qmc_sample_forces and one iteration of the geometry loop on benzene
    (wave engine, single-electron moves): finite forces, and the first step
    is the charge-weighted projection of the sampled forces with each
    displacement capped at 0.02 bohr (geometry_opt.rs:283-303)."""
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05)
    Z = g.get_charges()
    R0 = g.get_nuclei()
    opt = Q.GeometryOptimizer(n_samples=256, n_walkers=128, n_equilibrate=20, mc_step_size=0.3,
                              zero_variance=True, max_iterations=1, step_size=0.01, force_threshold=0.0,
                              decorrelation=2, move=Q.QMC_MOVE_SINGLE, seed=5)
    F, e, err, fv = opt.sample_forces(g)
    assert np.all(np.isfinite(F)) and np.isfinite(e) and np.all(np.isfinite(fv))
    assert e < 0.0, e
    r = opt.optimize(g)  # its first sampling uses the same key as sample_forces above
    assert r.energy_history[0] == pytest.approx(e, rel=1e-12)
    com = (F * Z[:, None]).sum(0) / Z.sum()
    d = opt.step_size * (F - com)
    n = np.linalg.norm(d, axis=1, keepdims=True)
    d = np.where(n > 0.02, d * (0.02 / np.maximum(n, 1e-300)), d)
    R1 = g.get_nuclei()
    assert np.allclose(R1, R0 + d, rtol=0, atol=1e-12)
    assert np.allclose(r.final_nuclei, R1)
