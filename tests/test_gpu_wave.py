"""GPU parity of the wave engine (one wavefront per walker, state in LDS):
benzene (BenzeneGTO, benzene.rs) against the oracle, and the smaller
systems forced onto the wave engine against both the oracle and the
register engine."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    import rust_qmc_amd as Q
    Q.lib()
    return Q


def _close(a, b, rtol, scale=None):
    a = np.asarray(a)
    b = np.asarray(b)
    s = np.abs(b) if scale is None else scale
    return np.all(np.abs(a - b) <= rtol * np.maximum(s, 1e-300) + rtol * 1e-3)


@pytest.mark.parametrize("ch", [2.042, 2.05])
def test_benzene_eval_parity_fp64(Q, oracle, ch):
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, ch)
    o = oracle.System.benzene(1.5, 2.0, 2.64, ch)
    assert g.nuclear_repulsion() == pytest.approx(o.nuclear_repulsion(), rel=1e-14)
    pos = o.initialize_batch(2, 12)
    ref = o.batch_eval(pos)
    out = g.evaluate_batch(pos)
    assert np.array_equal(out["sign"], ref["sign"])
    assert _close(out["logabs"], ref["logabs"], 1e-10)
    gs = np.abs(ref["grad"]).max(axis=(1, 2))[:, None, None]
    assert _close(out["grad"], ref["grad"], 1e-10, scale=gs)
    ls = np.abs(ref["lap"]).max(axis=1)[:, None]
    assert _close(out["lap"], ref["lap"], 1e-10, scale=ls)
    assert _close(out["eloc"], ref["eloc"], 1e-10)


def test_benzene_mixed_precision_error(Q, oracle):
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="mixed")
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    pos = o.initialize_batch(5, 8)
    ref = o.batch_eval(pos)
    out = g.evaluate_batch(pos)
    assert _close(out["logabs"], ref["logabs"], 1e-12)
    err = np.abs(out["eloc"] - ref["eloc"])
    assert np.all(err <= 2e-4 * np.abs(ref["eloc"]) + 1e-2), err.max()


def test_benzene_trajectory_parity(Q, oracle):
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05)
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    W, steps, seed, step = 3, 2, 7, 0.2
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, step, trace=True, refresh_interval=1)
    pos_o, lp_o, el_o, acc_o = o.single_electron_steps(p0, seed, 0, steps, step)
    assert acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr, el_o, 1e-9)


@pytest.mark.parametrize("name", ["h2", "li_seeded", "ch4_sto3g"])
def test_wave_engine_matches_register_engine(Q, oracle, name):
    from test_gpu_parity import _configs, _pair
    g, o = _pair(Q, oracle, name)
    pos = _configs(o, 40)
    reg = g.evaluate_batch(pos)
    g.set_engine("wave")
    wav = g.evaluate_batch(pos)
    ref = o.batch_eval(pos)
    for k in ("logabs", "eloc"):
        assert _close(wav[k], ref[k], 1e-10), k
        assert _close(wav[k], reg[k], 1e-11), k
    W, steps, seed, step = 65, 6, 3, 0.4
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, step, trace=True, refresh_interval=4, steps_per_launch=3)
    pos_o, lp_o, el_o, acc_o = o.single_electron_steps(p0, seed, 0, steps, step)
    assert acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr, el_o, 1e-9)


def test_wave_engine_mixed_ensemble(Q, oracle):
    """Mixed-precision wave-engine sweeps of benzene stay finite and agree
    with FP64 sweeps of the same seed on the ensemble energy within 3 sigma."""
    res = {}
    for prec in ("fp64", "mixed"):
        g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision=prec)
        ens = Q.Ensemble(g, 512, 1)
        ens.sweeps(20, 0.2)
        es, e2, acc = ens.sweeps(40, 0.2, refresh_interval=20)
        tr = es / 512
        assert np.isfinite(tr).all()
        tau = Q.autocorrelation_time(tr)
        res[prec] = (tr.mean(), max(Q.blocking_error(tr, tau), tr.std() / np.sqrt(len(tr))))
    (a, sa), (b, sb) = res["fp64"], res["mixed"]
    assert abs(a - b) < 3 * np.hypot(sa, sb) + 1e-3


def test_tile_engine_selected_for_benzene(Q):
    assert Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="mixed").engine_name() == "tile"
    assert Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="fp64").engine_name() == "tile"
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="fp64")
    g.set_engine("wave")
    assert g.engine_name() == "wave"


def test_tile_engine_trajectory_vs_oracle(Q, oracle):
    """Tile-engine sweeps (FP32 MFMA MO contraction, register-resident
    inverse columns) follow the oracle's FP64 brute-force chain: same
    accept/reject decisions, positions to FP32 rounding.  3 walkers leave 13
    dead walkers in the tile."""
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="mixed")
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    W, steps, seed, step = 3, 2, 7, 0.2
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, step, trace=True, refresh_interval=0)
    pos_o, lp_o, el_o, acc_o = o.single_electron_steps(p0, seed, 0, steps, step)
    assert acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=2e-4)
    err = np.abs(tr - el_o)
    assert np.all(err <= 2e-4 * np.abs(el_o) + 1e-2), err.max()


def test_tile_engine_matches_wave_engine(Q):
    """Mixed-precision sweeps of 37 walkers (three tiles, the last ragged) on
    the tile and on the wave engine from the same seed: the same decisions
    and, to FP32 rounding, the same positions and energies."""
    out = []
    for eng in ("tile", "wave"):
        g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="mixed")
        g.set_engine(eng)
        ens = Q.Ensemble(g, 37, 9)
        es, e2, acc, tr = ens.sweeps(3, 0.2, trace=True, refresh_interval=2, steps_per_launch=2)
        out.append((ens.positions(), acc, tr, ens.state()))
    (p1, a1, t1, s1), (p2, a2, t2, s2) = out
    assert np.array_equal(a1, a2)
    assert np.allclose(p1, p2, rtol=0, atol=2e-4)
    assert _close(t1, t2, 2e-4, scale=np.abs(t2) + 10.0)
    assert _close(s1[0], s2[0], 1e-6, scale=np.abs(s2[0]))


def test_tile_engine_fp64_trajectory_vs_oracle(Q, oracle):
    """FP64 tile-engine sweeps (FP64 MFMA MO contraction v_mfma_f64_16x16x4f64,
    register-resident FP64 inverse columns) follow the oracle's FP64
    brute-force chain step for step: same decisions, positions and E_L to
    1e-9.  3 walkers leave 13 dead walkers in the tile."""
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="fp64")
    g.set_engine("tile")
    assert g.engine_name() == "tile"
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    W, steps, seed, step = 3, 2, 7, 0.2
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, step, trace=True, refresh_interval=0)
    pos_o, lp_o, el_o, acc_o = o.single_electron_steps(p0, seed, 0, steps, step)
    assert acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert np.allclose(tr, el_o, rtol=1e-9, atol=0), np.abs(tr - el_o).max()


def test_tile_engine_fp64_matches_wave_engine(Q):
    """FP64 sweeps of 37 walkers (three tiles, the last ragged) on the tile
    and the wave engine from the same seed, with a refresh between: the same
    decisions, positions, ln|psi| and E_L to 1e-10."""
    out = []
    for eng in ("tile", "wave"):
        g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="fp64")
        g.set_engine(eng)
        ens = Q.Ensemble(g, 37, 9)
        es, e2, acc, tr = ens.sweeps(3, 0.2, trace=True, refresh_interval=2, steps_per_launch=2)
        out.append((ens.positions(), acc, tr, ens.state()))
    (p1, a1, t1, s1), (p2, a2, t2, s2) = out
    assert np.array_equal(a1, a2)
    assert np.allclose(p1, p2, rtol=0, atol=1e-10)
    assert np.allclose(t1, t2, rtol=1e-10, atol=0)
    assert np.allclose(s1[0], s2[0], rtol=1e-11, atol=0)
    assert np.array_equal(s1[1], s2[1])


def test_wave_all_electron_benzene_parity_fp64(Q, oracle):
    """All-electron Metropolis (vmc.rs:95-119) on the wave engine: every
    electron moved, ln|psi'| from the FP64 rebuild, E_L on acceptance only."""
    g = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05)
    o = oracle.System.benzene(1.5, 2.0, 2.64, 2.05)
    W, steps, seed, step = 4, 4, 2, 0.05
    ens = Q.Ensemble(g, W, seed)
    p0 = ens.positions()
    es, e2, acc, tr = ens.sweeps(steps, step, Q.QMC_MOVE_ALL, trace=True)
    ref = o.batch_eval(p0)
    pos_o, lp_o, el_o, acc_o = o.all_electron_steps(p0, ref["logabs"], ref["eloc"], seed, 0, steps, step)
    assert acc_o > 0 and acc.sum() == acc_o
    assert np.allclose(ens.positions(), pos_o, rtol=0, atol=1e-9)
    assert _close(tr[-1], el_o, 1e-9)
    # the inverses are rebuilt before single-electron sweeps continue
    ens.sweeps(1, 0.2)
    la, sg, el = ens.state()
    ref = o.batch_eval(ens.positions())
    assert _close(la, ref["logabs"], 1e-10)


@pytest.mark.parametrize("name", ["h2", "li_seeded", "ch4_sto3g"])
def test_wave_all_electron_matches_register_engine(Q, oracle, name):
    from test_gpu_parity import _pair
    out = []
    for eng in ("auto", "wave"):
        g, _ = _pair(Q, oracle, name)
        g.set_engine(eng)
        ens = Q.Ensemble(g, 50, 3)
        es, e2, acc = ens.sweeps(6, 0.3, Q.QMC_MOVE_ALL)
        out.append((es, acc, ens.positions()))
    assert np.array_equal(out[0][1], out[1][1])
    assert np.allclose(out[0][2], out[1][2], rtol=0, atol=1e-12)
    assert np.allclose(out[0][0], out[1][0], rtol=1e-10, atol=0)
