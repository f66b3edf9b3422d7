"""Benchmark: VMC walker-moves/s and SR-grad samples/s on one MI355X.

Headline workload (CH4 Slater-Jastrow VMC): methane, C at the origin and
4 H at 2.06 bohr (tetrahedral), STO-3G contracted Gaussians (9 AOs, 3
primitives each), seeded orthonormal 5x9 MO coefficients shared by both
spins, Pade e-e/e-n Jastrow (b_ee=1.5, b_en=2.0), 65,536 walkers, Gaussian
single-electron moves of 0.3 bohr, Philox seed 0.

A step is one sweep of the whole ensemble: every electron of every walker
proposed once (Sherman-Morrison ratio/update), then one local-energy
measurement per walker, i.e. 65,536 walker-moves.  The headline value is
the mixed-precision mode (FP32 orbitals, Jastrow and Sherman-Morrison; FP64
inverse refresh every 64 sweeps, log-det and energy accumulation); the
timed window starts with a refresh, so it holds ceil(K / 64) of them.  The
FP64 mode is timed beside it.

The other workloads run at their own sizes beside it: H2 (4,096 walkers x
10^4 all-electron steps, FP64), the Li atom (16,384 walkers x 2x10^4 sweeps,
FP64 and mixed), the CH4 paired-chain bias (5x10^3 sweeps after 500 of
equilibration), CH4 SR (50 iterations of the 47-dimensional system over
65,536 walkers, with central finite differences at h = 1e-4), the
benzene + CH4 + H2 batch, benzene alone, and importance-sampled DMC.
Every mixed-precision result carries its difference from FP64.

Prints ONE JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walker-moves/s and SR-grad samples/s, mixed vs FP64 on 1 GPU at ≤0.1 mHa bias"
# Algorithmic flops per walker-move of CH4/STO-3G (9 AOs, 3 primitives, 5
# MOs per spin, 10 electrons, 5 nuclei): per electron move
# M*9 + A(8P+12) + 10nA + 2n + 2((N-1)+M)*13 + 0.5*4n^2, per sweep once
# 8 sum n^2 + 30(N(N-1)/2 + NM) + 10 N(N-1)/2 + 2NM + 10N + sum 2n^3/64
# (FMA = 2, exp/sqrt/div = 1).
FLOP_PER_WM = 16338
PEAK = {"fp64": 78.6, "mixed": 157.3}  # TFLOP/s, MI355X vector FP64 / FP32
# measured HBM traffic of one mixed-precision sweep launch at 65,536 walkers
# and 16 sweeps per launch (22,532 KiB read + 23,296 KiB written)
TRAFFIC_BYTES = int((2 * 11472.0 + 23296) * 1024)  # FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, KiB per launch
H2_README_HA = -1.15754  # README.md:22, -4.287 eV read as E - 2 E_H


def _stats(Q, tr):
    """Mean, blocking error with block size ceil(2 tau) and tau of a series
    (vmc.rs:141-188)."""
    import numpy as np

    tr = np.asarray(tr, dtype=float)
    tau = Q.autocorrelation_time(tr)
    err = Q.blocking_error(tr, tau)
    if not err > 0:
        err = float(np.std(tr) / math.sqrt(len(tr)))
    return float(tr.mean()), float(err), float(tau)


def _paired(Q, a, b):
    """Paired difference of two coupled per-sweep series: mean and blocking
    error of a - b, in mHa."""
    import numpy as np

    m, e, _ = _stats(Q, np.asarray(a) - np.asarray(b))
    return {"delta_mHa": m * 1e3, "err_mHa": e * 1e3}


def _aligned_warmup(warmup, refresh):
    """Untimed sweeps before the timed window: at least warmup, rounded up to
    a whole number of refresh intervals, so the window opens with a refresh."""
    if refresh <= 0:
        return warmup
    return max(1, -(-max(warmup, 1) // refresh)) * refresh


def run_mode(Q, precision, W, steps, warmup, spl, refresh):
    sys_ = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision)
    ens = Q.Ensemble(sys_, W, 0)
    nw = _aligned_warmup(warmup, refresh)
    ens.sweeps(nw, 0.3, refresh_interval=refresh, steps_per_launch=spl)
    ens.set_timing(True)
    # qmc_ensemble_sweeps returns after its stream has drained (it copies the
    # per-sweep sums back), so the wall clock brackets exactly the K sweeps
    # and the refreshes due inside them.
    t0 = time.perf_counter()
    es, e2, acc = ens.sweeps(steps, 0.3, refresh_interval=refresh, steps_per_launch=spl)
    wall = time.perf_counter() - t0
    kms, launches = ens.timing()
    return {"wall_s": wall, "kernel_ms": kms, "launches": launches, "warmup_sweeps": nw,
            "walker_moves_per_s": W * steps / wall, "acceptance": float(acc.sum() / (W * 10 * steps))}


def run_ch4_pairs(Q, W, n_equil, n_sweeps, refresh):
    """Config 3's bias: FP64 and mixed chains from the same seed (their
    accept decisions differ only when the uniform falls between the two
    precisions' ratios), n_equil sweeps of equilibration, then the per-sweep
    difference of the ensemble energies over n_sweeps; and the mixed
    sampler's cached E_L against the FP64 E_L of the same final positions."""
    import numpy as np

    out, tr, ens = {}, {}, {}
    for prec in ("fp64", "mixed"):
        s = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=prec)
        e = Q.Ensemble(s, W, 0)
        e.sweeps(n_equil, 0.3, refresh_interval=refresh)
        t0 = time.perf_counter()
        es, e2, acc = e.sweeps(n_sweeps, 0.3, refresh_interval=refresh)
        wall = time.perf_counter() - t0
        tr[prec] = es / W
        m, err, tau = _stats(Q, tr[prec])
        out[prec] = {"energy": m, "energy_err": err, "tau": tau, "acceptance": float(acc.sum() / (W * 10 * n_sweeps)),
                     "walker_moves_per_s": W * n_sweeps / wall}
        ens[prec] = (s, e)
    out["paired"] = _paired(Q, tr["mixed"], tr["fp64"])
    pm, pf = ens["mixed"][1].positions(), ens["fp64"][1].positions()
    out["decoupled_walkers"] = int(np.any(np.abs(pm - pf) > 1e-3, axis=(1, 2)).sum())
    el_mixed = ens["mixed"][1].state()[2]
    el_fp64 = ens["fp64"][0].evaluate_batch(pm, ("eloc",))["eloc"]
    d = el_mixed - el_fp64
    out["eval_bias_mHa"] = float(d.mean()) * 1e3
    out["eval_bias_err_mHa"] = float(d.std() / math.sqrt(len(d))) * 1e3
    out["eval_err_max_Ha"] = float(np.abs(d).max())
    # Resynchronised pairs: the two chains separate where a uniform falls
    # between the FP32 and FP64 acceptance ratios, after which the separated
    # walkers only add noise.  So every `window` sweeps the mixed walkers are
    # put back on the FP64 walkers' positions (with a fresh FP64 inverse) and
    # both continue with the same random streams; the window means of
    # (mixed - FP64) estimate the bias of the mixed sampler started from the
    # FP64 distribution, a window (2 tau) long.
    window = 64
    ef, em = ens["fp64"][1], ens["mixed"][1]
    diffs, dec = [], 0
    for _ in range(max(1, n_sweeps // window)):
        em.set_positions(ef.positions())
        a = em.sweeps(window, 0.3, refresh_interval=refresh)[0] / W
        b = ef.sweeps(window, 0.3, refresh_interval=refresh)[0] / W
        diffs.append(float((a - b).mean()))
        dec += int(np.any(np.abs(em.positions() - ef.positions()) > 1e-3, axis=(1, 2)).sum())
    diffs = np.array(diffs)
    out["resync_paired"] = {"delta_mHa": float(diffs.mean()) * 1e3,
                            "err_mHa": float(diffs.std(ddof=1) / math.sqrt(len(diffs))) * 1e3 if len(diffs) > 1 else None,
                            "window_sweeps": window, "windows": len(diffs),
                            "decoupled_walkers_per_window": dec / len(diffs)}
    out.update({"walkers": W, "equilibration_sweeps": n_equil, "sweeps": n_sweeps})
    return out


def run_h2(Q, W, n_steps, n_burn):
    """Config 1: H2MoleculeMO (LCAO Slater1s alpha = 1.3, protons at +-0.7
    bohr, Jastrow1 F = 3) in FP64 with the reference's all-electron
    Metropolis move (vmc.rs:95-119) at a fixed 0.5 bohr step; the energy
    over the steps after burn-in is compared with README.md:22."""
    g = Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0, precision="fp64")
    ens = Q.Ensemble(g, W, 0)
    ens.sweeps(n_burn, 0.5, move=Q.QMC_MOVE_ALL)
    ens.set_timing(True)
    t0 = time.perf_counter()
    es, e2, acc = ens.sweeps(n_steps - n_burn, 0.5, move=Q.QMC_MOVE_ALL, steps_per_launch=500)
    wall = time.perf_counter() - t0
    kms, _ = ens.timing()
    m, err, tau = _stats(Q, es / W)
    n = (n_steps - n_burn) * W
    return {"walkers": W, "steps": n_steps, "burn_in": n_burn, "move": "all-electron (vmc.rs:95-119)",
            "walker_moves_per_s": n / wall, "kernel_walker_moves_per_s": n / (kms / 1e3),
            "energy": m, "energy_err": err, "tau": tau, "acceptance": float(acc.sum() / n),
            "readme_energy": H2_README_HA, "deviation_sigma": abs(m - H2_README_HA) / err}


def run_li(Q, W, n_sweeps, n_equil):
    """Config 2: the Li atom with three seeded 3-term STO orbitals
    (p = 0, 1, 2; xi = 2.7, 0.65) and the Pade pair Jastrow b = 1, 16,384
    walkers, single-electron sweeps in FP64 and mixed precision from the
    same seed; the paired energy difference states the mixed error."""
    tr, out = {}, {"walkers": W, "sweeps": n_sweeps, "equilibration_sweeps": n_equil}
    for prec in ("fp64", "mixed"):
        g = Q.Lithium(0, precision=prec)
        ens = Q.Ensemble(g, W, 0)
        ens.sweeps(n_equil, 0.5)
        ens.set_timing(True)
        t0 = time.perf_counter()
        es, e2, acc = ens.sweeps(n_sweeps, 0.5, steps_per_launch=64)
        wall = time.perf_counter() - t0
        kms, _ = ens.timing()
        tr[prec] = es / W
        m, err, tau = _stats(Q, tr[prec])
        out[prec] = {"walker_moves_per_s": W * n_sweeps / wall, "kernel_walker_moves_per_s": W * n_sweeps / (kms / 1e3),
                     "energy": m, "energy_err": err, "tau": tau, "acceptance": float(acc.sum() / (W * 3 * n_sweeps)),
                     "engine": g.engine_name()}
    out["mixed_over_fp64"] = out["mixed"]["walker_moves_per_s"] / out["fp64"]["walker_moves_per_s"]
    out["paired"] = _paired(Q, tr["mixed"], tr["fp64"])
    return out


def _ch4_series(Q, precision):
    s = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision)
    s.set_jastrow_series(6, 8)
    return s


def run_sr_grad(Q, precision, W, n_meas, warmup):
    """SR-grad samples/s: per-walker E_L, O_k (32 Jastrow parameters + 15
    nuclear coordinates), dE_L/dp_k and the Hellmann-Feynman force of all W
    walkers, accumulated into the FP64 moment matrix (one measure call = W
    samples), on CH4/STO-3G with a 6-term e-e and 8-term e-n Pade series."""
    sys_ = _ch4_series(Q, precision)
    ens = Q.Ensemble(sys_, W, 0)
    ens.sweeps(warmup, 0.3, refresh_interval=64)
    ens.measure(reset=True)  # first call sizes the buffers
    t0 = time.perf_counter()
    for _ in range(n_meas):
        ens.measure()
    wall = time.perf_counter() - t0
    m = ens.moments()
    return {"samples_per_s": W * n_meas / wall, "params": sys_.num_params(), "nuclear": 3 * sys_.n_nuc,
            "ms_per_call": wall * 1e3 / n_meas, "energy": m["energy"], "ens": ens, "sys": sys_}


def sr_grad_errors(Q, sr):
    """Mixed-precision derivatives and moments against FP64 on identical
    configurations (the mixed ensemble's walkers)."""
    import numpy as np

    pos = sr["mixed"]["ens"].positions()[:8192]
    a = sr["mixed"]["sys"].derivatives_batch(pos)
    b = sr["fp64"]["sys"].derivatives_batch(pos)
    err = {}
    for k in ("eloc", "O", "dE", "F", "Om"):
        scale = max(1.0, float(np.abs(b[k]).max()))
        err[k] = float(np.abs(a[k] - b[k]).max() / scale)
    err["eloc_mean_mHa"] = float((a["eloc"] - b["eloc"]).mean()) * 1e3

    def sr_step(d):
        O = np.hstack([d["O"], d["Om"].reshape(len(pos), -1)])
        E = d["eloc"]
        Om = O.mean(0)
        S = O.T @ O / len(E) - np.outer(Om, Om)
        f = (E[:, None] * O).mean(0) - E.mean() * Om
        return Q.sr_solve(S, f, 1e-3)[0]

    da, db = sr_step(a), sr_step(b)
    err["sr_step_47"] = float(np.abs(da - db).max() / np.abs(db).max())
    return {"max_err_rel_to_max": err, "configurations": len(pos)}


def fd_check(Q, h=1e-4, W=32):
    """Central finite differences (h = 1e-4) of the FP64 wavefunction on W
    configurations: O_k = d ln|psi|/dp_k and dE_L/dp_k over the 32 Jastrow
    parameters, Omega = d ln|psi|/dR over the 15 nuclear coordinates."""
    import numpy as np

    g = _ch4_series(Q, "fp64")
    p0 = g.get_params()
    p0[4:] = np.random.default_rng(5).normal(size=p0.size - 4) * 0.05
    g.set_params(p0)
    ens = Q.Ensemble(g, W, 2)
    ens.sweeps(50, 0.3)
    pos = ens.positions()
    d = g.derivatives_batch(pos)
    P = p0.size
    fo, fe = np.zeros((W, P)), np.zeros((W, P))
    for k in range(P):
        vals = []
        for s in (1, -1):
            p = p0.copy()
            p[k] += s * h
            g.set_params(p)
            vals.append(g.evaluate_batch(pos, ("logabs", "eloc")))
        fo[:, k] = (vals[0]["logabs"] - vals[1]["logabs"]) / (2 * h)
        fe[:, k] = (vals[0]["eloc"] - vals[1]["eloc"]) / (2 * h)
    g.set_params(p0)
    R0 = g.get_nuclei()
    fom = np.zeros((W,) + R0.shape)
    for I in range(R0.shape[0]):
        for c in range(3):
            vals = []
            for s in (1, -1):
                R = R0.copy()
                R[I, c] += s * h
                g.set_nuclei(R)
                vals.append(g.evaluate_batch(pos, ("logabs",))["logabs"])
            fom[:, I, c] = (vals[0] - vals[1]) / (2 * h)
    g.set_nuclei(R0)

    def rel(a, b):
        return float(np.abs(a - b).max() / max(1.0, float(np.abs(b).max())))

    return {"h": h, "configurations": W, "O_k": rel(d["O"], fo), "dE_L_dp": rel(d["dE"], fe),
            "Omega": rel(d["Om"], fom), "measure": "max |analytic - FD| / max(1, max |FD|)"}


def run_sr_opt(Q, precision, W, iters):
    """Config 4: 50 SR iterations on the 47-dimensional system (32 Jastrow
    parameters + 15 nuclear coordinates, FP64 solve of (S + eps I) dp = -f),
    65,536 walkers per iteration (one measurement after 5 decorrelation
    sweeps)."""
    g = _ch4_series(Q, precision)
    opt = Q.SROptimizer(n_samples=W, n_walkers=W, n_equilibrate=200, max_iterations=iters, learning_rate=0.05,
                        tolerance=0.0, step_size=0.3, move=Q.QMC_MOVE_SINGLE, include_nuclear=True, seed=3)
    t0 = time.perf_counter()
    res = opt.optimize(g)
    wall = time.perf_counter() - t0
    n = len(res.energy_history)
    return {"iterations": n, "samples_per_iteration": W, "wall_s": wall,
            "samples_per_s": W * (n + 1) / wall, "initial_energy": float(res.energy_history[0]),
            "final_energy": res.final_energy, "final_variance": res.final_variance,
            "energy_history": [float(x) for x in res.energy_history]}


def run_batch(Q, sweeps, trial, precisions=("mixed", "fp64")):
    """Config 5: benzene (8192 walkers) + CH4 STO-3G (32768) + H2 (65536)
    sampled together; the autotuner picks the launch configuration, then the
    timed sweeps run with it.  The fused (one persistent launch) and
    sequential configurations are timed beside it.  The FP64 batch runs the
    same chains, so the per-molecule paired energy differences state the
    mixed error."""
    walkers = [8192, 32768, 65536]
    steps = [0.2, 0.3, 0.5]
    out, tr = {"walkers": walkers, "sweeps": sweeps}, {}
    for prec in precisions:
        systems = [Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision=prec),
                   Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=prec),
                   Q.H2MoleculeMO(precision=prec)]
        # the autotuner's trial sweeps advance its ensembles by a number of
        # sweeps that depends on the configurations it tries; the timed chains
        # start from fresh ensembles, so both precisions' chains stay paired
        tb = Q.Batch([Q.Ensemble(s, W, 100 + i) for i, (s, W) in enumerate(zip(systems, walkers))])
        tb.sweeps(16, steps, refresh_interval=64)
        cfg, tuned_ms = tb.tune(trial, steps, refresh_interval=-256)
        del tb
        ens = [Q.Ensemble(s, W, i) for i, (s, W) in enumerate(zip(systems, walkers))]
        b = Q.Batch(ens)
        b.set_config(**cfg)
        b.sweeps(64, steps, refresh_interval=-1)
        r = {"tuned_config": cfg, "engines": [s.engine_name() for s in systems]}
        t0 = time.perf_counter()
        es, e2, acc = b.sweeps(sweeps, steps, refresh_interval=-1)
        wall = time.perf_counter() - t0
        kms, nl = b.timing()
        tot = sum(walkers) * sweeps
        tr[prec] = [es[i] / walkers[i] for i in range(3)]
        r.update({"walker_moves_per_s": tot / wall, "kernel_walker_moves_per_s": tot / (kms / 1e3),
                  "ms_per_sweep": kms / sweeps, "launches": nl,
                  "energies": [float(t.mean()) for t in tr[prec]]})
        if prec == "mixed":
            for mode in ("fused", "sequential"):
                b.set_config(mode=mode, blocks_per_cu=0, steps_per_launch=cfg["steps_per_launch"], order="cost",
                             refresh_interval=cfg["refresh_interval"])
                b.sweeps(16, steps, refresh_interval=-1)
                r[f"ms_per_sweep_{mode}"] = b.timing()[0] / 16
        out[prec] = r
    if "fp64" in tr:
        out["paired_mixed_minus_fp64"] = {name: _paired(Q, tr["mixed"][i], tr["fp64"][i])
                                          for i, name in enumerate(("benzene", "ch4", "h2"))}
    return out


def run_benzene(Q, sweeps, W=8192, engine="auto"):
    """Benzene (BenzeneGTO, 42 electrons, 66 AOs, C-H 2.05 bohr), 8192
    walkers: Metropolis single-electron sweeps, drift-diffusion sweeps
    (tau = 0.01) and SR-grad samples (E_L, the 4 Jastrow O_k and dE_L/dp_k,
    HF force and Omega over 12 nuclei with the C-H MOs following the bonds);
    kernel time from HIP events.  Both precisions start from the same seed,
    so the per-sweep paired energy differences state the mixed error."""
    import numpy as np

    out, tr, keep = {"walkers": W, "sweeps": sweeps, "equilibration_sweeps": 64, "engine_option": engine}, {}, {}
    for prec in ("mixed", "fp64"):
        s = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision=prec)
        s.set_engine(engine)
        ens = Q.Ensemble(s, W, 3)
        ens.sweeps(64, 0.2, refresh_interval=64)
        keep[prec] = (s, ens)
        if prec == "mixed":
            # E_L of the mixed tile engine (FP32 orbitals and Sherman-Morrison
            # inverse) against FP64 E_L of the same configurations
            el_m = ens.state()[2]
            el_f = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="fp64").evaluate_batch(ens.positions(),
                                                                                    ("eloc",))["eloc"]
            d = el_m - el_f
            out["eval_bias_mHa"] = float(d.mean()) * 1e3
            out["eval_bias_err_mHa"] = float(d.std() / math.sqrt(len(d))) * 1e3
            out["eval_err_max_Ha"] = float(np.abs(d).max())
        r = {"engine": s.engine_name()}
        for name, move, step in (("metropolis", Q.QMC_MOVE_SINGLE, 0.2), ("drift_diffusion", Q.QMC_MOVE_DRIFT, 0.01)):
            ens.set_timing(True)
            es, e2, acc = ens.sweeps(sweeps, step, move=move, refresh_interval=64, steps_per_launch=8)
            kms, nl = ens.timing()
            ens.set_timing(False)
            tr[(prec, name)] = es / W
            r[name] = {"walker_moves_per_s": W * sweeps / (kms / 1e3), "ms_per_sweep": kms / sweeps,
                       "acceptance": float(acc.sum() / (W * 42 * sweeps)), "energy": float(es.mean() / W)}
        ens.measure(reset=True)
        t0 = time.perf_counter()
        for _ in range(4):
            ens.measure()
        wall = time.perf_counter() - t0
        r["sr_grad_samples_per_s"] = 4 * W / wall
        out[prec] = r
    out["paired_mixed_minus_fp64"] = {name: _paired(Q, tr[("mixed", name)], tr[("fp64", name)])
                                      for name in ("metropolis", "drift_diffusion")}
    return out


H2_EXACT_HA = -1.17447  # Kolos-Wolniewicz, R = 1.4 bohr


def run_isdmc(Q, precision, W, n_steps, system="h2"):
    """Importance-sampled DMC: one drift-diffusion sweep plus the branching,
    compaction and E_ref feedback per step, all on the device.  A
    walker-move is one step of one walker of the live population (every
    electron proposed once); the average live population is reported beside
    it.  H2 (Slater orbitals with their cusps) keeps a stable population and
    has a known exact energy; with the branching time matched to the 2 tau
    the moves diffuse for, DMC reaches it up to the time-step error.  CH4 in
    Gaussians has no nuclear cusp, E_L -> -inf at the nuclei, and the
    reference's branching (no cutoff) lets its population collapse, which is
    reproduced: its leg reports throughput and population only."""
    import numpy as np

    if system == "h2":
        g = Q.H2MoleculeMO(1.3, (0, 0, 0.7), (0, 0, -0.7), 3.0, precision=precision)
        P = Q.ISDMCParams(n_walkers=W, n_steps=n_steps, n_burnin=n_steps // 2, time_step=0.005,
                          n_vmc_equilibrate=200, n_vmc_force_samples=200, refresh_interval=64, matched_branching=True)
    else:
        g = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision)
        P = Q.ISDMCParams(n_walkers=W, n_steps=n_steps, n_burnin=n_steps // 4, time_step=0.005,
                          n_vmc_equilibrate=20, n_vmc_force_samples=100, refresh_interval=64)
    dmc = Q.ImportanceSampledDMC(g, P, seed=11)
    r = dmc.run()
    pops = dmc.populations
    swept = float(W + pops[:-1].sum())  # walkers entering each step's sweep
    out = {"system": system, "walker_moves_per_s": swept / (r.step_ms / 1e3), "ms_per_step": r.step_ms / n_steps,
           "avg_population": r.avg_population, "min_population": float(pops.min()), "target_population": W,
           "collapses": r.collapses, "dropped": r.dropped, "acceptance": r.acceptance_rate, "time_step": P.time_step}
    if system == "h2":
        out.update({"energy": r.energy, "energy_err": r.error, "exact": H2_EXACT_HA,
                    "production_energies": dmc.energies[P.n_burnin:]})
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _host_threads():
    """The host threads this process may use: its CPU affinity set, within
    the OMP_NUM_THREADS share the host grants it when that is set."""
    n = len(os.sched_getaffinity(0))
    lim = os.environ.get("OMP_NUM_THREADS")
    if lim and lim.isdigit() and int(lim) > 0:
        n = min(n, int(lim))
    return n


def cpu_baseline(seconds_target):
    """The reference algorithm (vmc.rs:95-119: all-electron proposal, full
    recomputation of psi, E_L with 8 sector LU per accepted move) on the CH4
    workload, timed on a bounded sample of walkers: one host thread (the
    reference is single-threaded on this path), and again with the walkers
    spread over every host thread this process may use."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.build()
    o = oracle.System.methane_sto3g(1.5, 2.0, 2.06, 0)
    # calibrate on a small sample, then size the timed sample to ~target s
    sec, _ = o.mcmc_time_steps(64, 1, 2, 0.3, 0)
    per_wm = sec / (64 * 2)
    W = 256
    steps = max(2, int(seconds_target / (per_wm * W)))
    sec, _ = o.mcmc_time_steps(W, 2, steps, 0.3, 0)
    cpu = _cpu_model()
    out = {
        "value": W * steps / sec,
        "unit": "walker-moves/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{W} walkers x {steps} steps (after 2 warm-up), CH4 STO-3G, oracle (gcc -O3 -march=x86-64-v3) "
                  f"in reference-algorithm mode, {cpu}, {sec:.1f} s",
    }
    T = _host_threads()
    if T > 1:
        Wm = 64 * T
        sm = max(2, int(seconds_target * T / (per_wm * Wm) / 2))
        secm, _ = o.mcmc_time_steps_mt(Wm, 2, sm, 0.3, 0, T)
        out["all_cores"] = {
            "value": Wm * sm / secm,
            "cores": T,
            "sample": f"{Wm} walkers x {sm} steps (after 2 warm-up) on {T} threads (all this process may use), "
                      f"{secm:.1f} s",
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--walkers", type=int, default=65536)
    ap.add_argument("--steps-per-launch", type=int, default=16)
    ap.add_argument("--refresh", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sr-calls", type=int, default=20, help="timed SR-grad measurement calls (W samples each)")
    ap.add_argument("--sr-iterations", type=int, default=50, help="SR iterations of the 47-dimensional system (0: skip)")
    ap.add_argument("--pair-sweeps", type=int, default=5000, help="CH4 paired-chain sweeps (0: skip)")
    ap.add_argument("--h2-steps", type=int, default=10000, help="H2 all-electron steps incl. 10%% burn-in (0: skip)")
    ap.add_argument("--li-sweeps", type=int, default=20000, help="Li sweeps per precision (0: skip)")
    ap.add_argument("--batch-sweeps", type=int, default=2000,
                    help="timed sweeps of the benzene + CH4 + H2 batch per precision (0: skip it)")
    ap.add_argument("--dmc-steps", type=int, default=400, help="IS-DMC steps per precision (0: skip)")
    ap.add_argument("--benzene-sweeps", type=int, default=128, help="timed benzene sweeps per leg (0: skip)")
    ap.add_argument("--benzene-engine", default="auto", choices=("auto", "wave"),
                    help="engine of the benzene leg: auto (tile engine) or the one-wave-per-walker engine")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 8 device-to-device copies of BYTES (HBM counter calibration) and exit")
    ap.add_argument("--only", default="", help="comma-separated legs to run (profiling): headline,pairs,h2,li,"
                                               "sr,fd,sropt,batch,dmc,benzene")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU; multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import rust_qmc_amd as Q

    if args.copy_calibration:
        n = args.copy_calibration
        with Q.DeviceBuffer(n) as a, Q.DeviceBuffer(n) as b:
            for _ in range(8):
                b.copy_from(a, n)
        print(json.dumps({"copy_bytes": n, "copies": 8}))
        return

    only = set(filter(None, args.only.split(",")))

    def want(leg):
        return not only or leg in only

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    W, K = args.walkers, args.steps
    out = {}
    if want("headline"):
        res = {m: run_mode(Q, m, W, K, args.warmup, args.steps_per_launch, args.refresh) for m in ("fp64", "mixed")}
        log("headline done")
    pairs = run_ch4_pairs(Q, W, 500, args.pair_sweeps, args.refresh) if args.pair_sweeps > 0 and want("pairs") else None
    log("pairs done")
    h2 = run_h2(Q, 4096, args.h2_steps, args.h2_steps // 10) if args.h2_steps > 0 and want("h2") else None
    li = run_li(Q, 16384, args.li_sweeps, 500) if args.li_sweeps > 0 and want("li") else None
    log("h2/li done")
    sr = {m: run_sr_grad(Q, m, W, args.sr_calls, 16) for m in ("fp64", "mixed")} if want("sr") else None
    sr_err = sr_grad_errors(Q, sr) if sr else None
    fd = fd_check(Q) if want("fd") else None
    sropt = ({m: run_sr_opt(Q, m, W, args.sr_iterations) for m in ("fp64", "mixed")}
             if args.sr_iterations > 0 and want("sropt") else None)
    log("sr done")
    batch = run_batch(Q, args.batch_sweeps, 4) if args.batch_sweeps > 0 and want("batch") else None
    log("batch done")
    dmc = None
    if args.dmc_steps > 0 and want("dmc"):
        dmc = {"h2": {m: run_isdmc(Q, m, W, args.dmc_steps, "h2") for m in ("mixed", "fp64")},
               "ch4": {m: run_isdmc(Q, m, W, args.dmc_steps // 4, "ch4") for m in ("mixed", "fp64")}}
        h = dmc["h2"]
        em, ef = h["mixed"].pop("production_energies"), h["fp64"].pop("production_energies")
        h["paired_mixed_minus_fp64"] = _paired(Q, em, ef)
    benzene = (run_benzene(Q, args.benzene_sweeps, engine=args.benzene_engine)
               if args.benzene_sweeps > 0 and want("benzene") else None)
    log("dmc/benzene done")
    if not want("headline"):
        print(json.dumps({"legs": sorted(only), "ch4_paired_chains": pairs, "h2": h2, "li": li,
                          "sr_errors": sr_err, "finite_differences": fd, "sr_optimization": sropt,
                          "heterogeneous_batch": batch, "isdmc": dmc, "benzene": benzene}, default=str))
        return
    mx, fp = res["mixed"], res["fp64"]
    # roofline of the dominant kernel (the mixed-precision sweep): HIP events
    # on the sweep's launch stream bracket each sweep launch
    avg_launch_s = mx["kernel_ms"] / 1e3 / mx["launches"]
    steps_per_launch = K / mx["launches"]
    achieved = FLOP_PER_WM * W * steps_per_launch / avg_launch_s / 1e12
    out = {
        "metric": METRIC,
        "value": mx["walker_moves_per_s"],
        "unit": "walker-moves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": mx["wall_s"] * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (Philox-seeded walkers and MO coefficients)",
        "config": {
            "workload": "CH4 STO-3G Slater-Jastrow VMC, 65536 walkers, single-electron moves 0.3 bohr, "
                        "mixed FP32/FP64 with FP64 inverse refresh every 64 sweeps",
            "walkers": W,
            "electrons": 10,
            "aos": 9,
            "refresh_interval": args.refresh,
            "refreshes_timed": -(-K // args.refresh) if args.refresh > 0 else 0,
            "warmup_sweeps_run": mx["warmup_sweeps"],
            "steps_per_launch": args.steps_per_launch,
        },
        "fp64": {"value": fp["walker_moves_per_s"], "ms_per_step": fp["wall_s"] * 1e3 / K,
                 "kernel_ms_per_launch": fp["kernel_ms"] / fp["launches"], "acceptance": fp["acceptance"]},
        "mixed": {"kernel_ms_per_launch": mx["kernel_ms"] / mx["launches"], "acceptance": mx["acceptance"]},
        "mixed_over_fp64": mx["walker_moves_per_s"] / fp["walker_moves_per_s"],
    }
    if pairs:
        rs = pairs["resync_paired"]
        out["delta_E_mHa"] = rs["delta_mHa"]
        out["delta_E_err_mHa"] = rs["err_mHa"]
        out["delta_E_estimator"] = (f"paired chains, same seed and streams, after {pairs['equilibration_sweeps']} + "
                                    f"{pairs['sweeps']} sweeps: the mixed walkers put back on the FP64 walkers every "
                                    f"{rs['window_sweeps']} sweeps, mean of (mixed - FP64) ensemble energy over "
                                    f"{rs['windows']} windows, standard error over windows")
        out["delta_E_eval_mHa"] = pairs["eval_bias_mHa"]
        out["delta_E_eval_err_mHa"] = pairs["eval_bias_err_mHa"]
        out["ch4_paired_chains"] = pairs
    out["sr_grad_samples_per_s"] = None if not sr else {
        "mixed": sr["mixed"]["samples_per_s"], "fp64": sr["fp64"]["samples_per_s"],
        "params": sr["mixed"]["params"], "nuclear_coords": sr["mixed"]["nuclear"],
        "ms_per_65536_samples": {"mixed": sr["mixed"]["ms_per_call"], "fp64": sr["fp64"]["ms_per_call"]},
        "mixed_vs_fp64": sr_err}
    out["sr_optimization"] = sropt
    out["finite_differences"] = fd
    out["h2"] = h2
    out["li"] = li
    out["heterogeneous_batch"] = batch
    out["isdmc"] = dmc
    out["benzene"] = benzene
    out["roofline"] = {
        "bound": "valu",
        "engine": "vector ALU at one wave per SIMD: the sweep has no GEMM-shaped work; the peak is MI355X's "
                  "FP32 vector rate",
        "achieved": achieved,
        "peak": PEAK["mixed"],
        "unit": "TFLOP/s",
        "frac": achieved / PEAK["mixed"],
        # HBM bytes per launch of the mixed sweep (16 sweeps x 65,536
        # walkers) from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the
        # factors from a copy kernel of known size
        "traffic": TRAFFIC_BYTES if (W == 65536 and args.steps_per_launch == 16) else None,
        "traffic_unit": "bytes per launch",
        "kernel": "sweep_kernel<5,5,CtBasisL<BdCH4Sto3g>,JsNone,float>",
        "flop_per_walker_move": FLOP_PER_WM,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
