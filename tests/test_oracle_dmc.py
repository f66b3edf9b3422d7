"""Pins of the IS-DMC oracle (oracle/orc_dmc.c; ImportanceSampledDMC,
src/sampling/is_dmc.rs:176-607).  Nothing here retypes the oracle's
formulas: the branching is checked against the expectation it must have,
the blocking error against closed cases and invariances, and the whole run
against the exact H2 ground-state energy, the zero-variance hydrogenic ion
and the population invariants."""
import math

import numpy as np
import pytest

H2_EXACT = -1.174475  # H2 at R = 1.4 bohr, Born-Oppenheimer (Kolos-Wolniewicz)


@pytest.fixture(scope="module")
def h2(oracle):
    return oracle.System.h2mo(1.3, 0.7, 3.0)


def test_branch_copies_have_the_weight_as_mean(oracle):
    """floor(W + xi) with xi uniform is an unbiased integer estimate of W
    below the clone cap (is_dmc.rs:284-289); W >= 3 gives exactly 3."""
    n, tau, E = 200_000, 0.01, -1.0
    el = np.full(n, E)
    for W in (0.3, 1.0, 1.7, 2.5):
        e_ref = E + math.log(W) / tau
        m = oracle.isdmc_branch(el, el, e_ref, tau, 7, 3)
        frac = W - math.floor(W)
        assert set(np.unique(m)) <= {math.floor(W), math.floor(W) + 1}
        sd = math.sqrt(max(frac * (1 - frac), 1e-12) / n)
        assert abs(m.mean() - W) < 5 * sd + 1e-12
    m = oracle.isdmc_branch(el, el, E + math.log(5.0) / tau, tau, 7, 3)
    assert (m == 3).all()
    # the weight uses the mean of the pre- and post-sweep energies
    m1 = oracle.isdmc_branch(np.full(n, -0.8), np.full(n, -1.2), E + math.log(1.7) / tau, tau, 7, 3)
    m2 = oracle.isdmc_branch(el, el, E + math.log(1.7) / tau, tau, 7, 3)
    assert np.array_equal(m1, m2)
    # a NaN energy gives no copies (the saturating float -> usize cast)
    m = oracle.isdmc_branch(np.array([np.nan, -1.0]), np.array([-1.0, -1.0]), -1.0, tau, 7, 3)
    assert m[0] == 0


def test_block_error_special_cases(oracle):
    be = oracle.isdmc_block_error
    assert be(np.arange(19.0)) == 0.0  # fewer than 20 samples
    assert be(np.full(100, 3.25)) == 0.0  # no variance
    # +1/-1 alternation: the first autocorrelation is negative, blocks of 2
    # all have mean 0, so the error is exactly 0
    assert be(np.tile([1.0, -1.0], 50)) == 0.0
    rng = np.random.default_rng(4)
    x = np.zeros(4000)
    for i in range(1, x.size):  # AR(1)
        x[i] = 0.8 * x[i - 1] + rng.normal()
    assert be(3.0 * x + 7.0) == pytest.approx(3.0 * be(x), rel=1e-9)
    # white noise: the blocked error matches sigma / sqrt(n)
    w = rng.normal(size=20_000)
    assert be(w) == pytest.approx(1.0 / math.sqrt(w.size), rel=0.15)
    # AR(1) with phi = 0.8: error -> sigma_x sqrt((1 + phi)/(1 - phi) / n)
    sx = 1.0 / math.sqrt(1 - 0.64)
    assert be(x) == pytest.approx(sx * math.sqrt(9.0 / x.size), rel=0.35)


def test_matched_branching_reaches_the_exact_h2_energy(h2):
    """H2's ground state is nodeless, so DMC is exact up to the time step:
    with the branching over the 2 tau the moves diffuse for, the energy
    reaches -1.17447 Ha."""
    r = h2.isdmc_run(5, 400, 3000, 300, 0.005, n_vmc_equilibrate=100, n_vmc_force_samples=200,
                     matched_branching=True)
    assert abs(r["energy"] - H2_EXACT) < max(4e-3, 3 * r["error"])
    assert r["collapses"] == 0 and r["dropped"] == 0
    assert abs(r["avg_population"] - 400) < 20


def test_reference_branching_is_biased_upward(h2):
    """is_dmc.rs branches over tau while its moves diffuse for 2 tau
    (r' = r + tau F + sqrt(2 tau) eta): the energy sits well above the
    exact one and above the matched run at the same tau."""
    r = h2.isdmc_run(5, 400, 1500, 300, 0.01, n_vmc_equilibrate=100, n_vmc_force_samples=200)
    m = h2.isdmc_run(5, 400, 1500, 300, 0.01, n_vmc_equilibrate=100, n_vmc_force_samples=200,
                     matched_branching=True)
    assert r["energy"] - H2_EXACT > 4e-3
    assert r["energy"] > m["energy"]


def test_zero_variance_ion_is_exact_and_keeps_its_population(oracle):
    """One electron in STO{c=1, p=0, zeta=3} about Z=3 is the exact 1s state
    (E_L = -4.5): every weight is 1, so no walker dies or clones and the
    DMC energy is exact with zero error."""
    s = oracle.System.sto_system([([1.0], [0], [3.0])], [1], Z=3)
    r = s.isdmc_run(1, 64, 50, 10, 0.01, n_vmc_equilibrate=10, n_vmc_force_samples=100)
    assert r["energy"] == pytest.approx(-4.5, abs=1e-12)
    assert r["error"] == 0.0
    assert (r["populations"] == 64).all()


def test_population_collapse_reinitialises(h2):
    """A huge E_ref damping kills the whole population: the step records the
    reference's 0/0 energy (NaN), and fresh walkers continue the run
    (is_dmc.rs:474-501)."""
    r = h2.isdmc_run(2, 40, 30, 2, 0.01, e_ref_damping=500.0, n_vmc_equilibrate=5, n_vmc_force_samples=40,
                     capacity=256)
    pops, es = r["populations"], r["energies"]
    assert r["collapses"] >= 1
    dead = np.flatnonzero(pops == 0)
    assert dead.size == r["collapses"]
    assert np.isnan(es[dead]).all() and np.isfinite(np.delete(es, dead)).all()
    assert np.isnan(r["energy"])


def test_capacity_drops_clones_in_order(h2):
    """Negative damping grows the population until the capacity holds it;
    the excess clones are counted (the reference's Vec has no bound)."""
    r = h2.isdmc_run(2, 40, 30, 2, 0.01, e_ref_damping=-2.0, n_vmc_equilibrate=5, n_vmc_force_samples=40,
                     capacity=256)
    assert r["populations"].max() == 256 and r["dropped"] > 0 and r["collapses"] == 0


def test_extrapolated_force_combines_mixed_and_vmc(h2):
    r = h2.isdmc_run(3, 64, 40, 10, 0.01, n_vmc_equilibrate=20, n_vmc_force_samples=100)
    assert np.allclose(r["forces_extrapolated"], 2 * r["forces_mixed"] - r["forces_vmc"], rtol=0, atol=1e-14)
    assert np.isfinite(r["forces_mixed"]).all() and np.isfinite(r["forces_vmc"]).all()
