"""GeometryOptimizer::optimize (geometry_opt.rs:233-347) on the GPU: the
steepest-descent loop over qmc_sample_forces with the charge-weighted
centre-of-mass projection and the 0.02 bohr displacement cap."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = 0x9E3779B97F4A7C15


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _h2(Q, half=0.9, flags=0):
    s = Q.H2MoleculeMO(1.3, (0, 0, half), (0, 0, -half), 3.0, precision="fp64")
    s.set_flags(flags)
    return s


def _opt(Q, **kw):
    base = dict(n_samples=2048, n_walkers=512, n_equilibrate=50, mc_step_size=0.5, zero_variance=True,
                max_iterations=4, step_size=0.05, force_threshold=0.0, move=Q.QMC_MOVE_SINGLE, seed=17)
    base.update(kw)
    return Q.GeometryOptimizer(**base)


def test_loop_matches_a_restatement_over_the_same_samplings(Q):
    """Each iteration samples forces with key seed + k * golden, removes the
    charge-weighted mean force, caps each displacement at 0.02 bohr and moves
    the nuclei (geometry_opt.rs:264-306)."""
    g = _opt(Q, step_size=0.3)
    a = _h2(Q)
    r = g.optimize(a)
    b = _h2(Q)
    Z = b.get_charges()
    R = b.get_nuclei()
    geo = [R.copy()]
    for k in range(g.max_iterations):
        gk = _opt(Q, step_size=0.3, seed=(g.seed + k * GOLD) % (1 << 64))
        F, e, err, fv = gk.sample_forces(b)
        assert e == pytest.approx(r.energy_history[k], rel=1e-12)
        assert np.linalg.norm(F, axis=1).max() == pytest.approx(r.force_history[k], rel=1e-12)
        com = (F * Z[:, None]).sum(0) / Z.sum()
        d = g.step_size * (F - com)
        n = np.linalg.norm(d, axis=1, keepdims=True)
        d = np.where(n > 0.02, d * (0.02 / np.maximum(n, 1e-300)), d)
        R = R + d
        b.set_nuclei(R)
        geo.append(R.copy())
    assert np.allclose(r.geometry_history, np.array(geo), rtol=0, atol=1e-12)
    assert np.allclose(r.final_nuclei, R, rtol=0, atol=1e-12)


def test_displacements_conserve_the_charge_weighted_centre(Q):
    """Below the cap the projected step moves no charge-weighted centre."""
    s = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision="mixed")
    s.set_flags(0)
    Z = s.get_charges()
    r = _opt(Q, step_size=1e-4, max_iterations=3).optimize(s)
    steps = np.diff(r.geometry_history, axis=0)
    assert (np.linalg.norm(steps, axis=2) < 0.02).all()
    assert np.allclose((steps * Z[None, :, None]).sum(axis=1), 0.0, atol=1e-13)


def test_displacement_cap(Q):
    r = _opt(Q, step_size=50.0, max_iterations=3).optimize(_h2(Q))
    n = np.linalg.norm(np.diff(r.geometry_history, axis=0), axis=2)
    assert (n <= 0.02 + 1e-12).all() and np.isclose(n.max(), 0.02, rtol=0, atol=1e-12)


def test_converged_at_the_first_sampling(Q):
    s = _h2(Q)
    R0 = s.get_nuclei()
    r = _opt(Q, force_threshold=1e9).optimize(s)
    assert r.converged and len(r.energy_history) == 1
    assert np.array_equal(s.get_nuclei(), R0) and r.geometry_history.shape[0] == 1
    assert np.isfinite(r.final_energy)


def test_h2_bond_relaxes_toward_equilibrium(Q):
    """Physical HF sign, zero-variance forces: from R = 1.8 bohr the bond of
    the LCAO H2 (alpha = 1.3) shortens toward its minimum near 1.4 bohr."""
    s = _h2(Q, half=0.9)
    r = _opt(Q, n_samples=32768, n_walkers=4096, step_size=0.5, max_iterations=8).optimize(s)
    bond = np.abs(r.geometry_history[:, 0, 2] - r.geometry_history[:, 1, 2])
    assert bond[0] == pytest.approx(1.8)
    assert bond[-1] < 1.8 - 0.1
    assert bond[-1] > 1.2
