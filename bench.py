"""Benchmark: CH4 Slater-Jastrow VMC walker-moves/s on one MI355X.

Workload (the CH4 configuration of the VMC hot path): methane, C at the
origin and 4 H at 2.06 bohr (tetrahedral), STO-3G contracted Gaussians (9
AOs, 3 primitives each), seeded orthonormal 5x9 MO coefficients shared by
both spins, Pade e-e/e-n Jastrow (b_ee=1.5, b_en=2.0), 65,536 walkers,
Gaussian single-electron moves of 0.3 bohr, Philox seed 0.

A step is one sweep of the whole ensemble: every electron of every walker
proposed once (Sherman-Morrison ratio/update), then one local-energy
measurement per walker, i.e. 65,536 walker-moves.  The headline value is
the mixed-precision mode (FP32 orbitals, Jastrow and Sherman-Morrison; FP64
inverse refresh every 64 sweeps, log-det and energy accumulation); the FP64
mode is timed beside it and the ensemble energies of both are compared.

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
TRAFFIC_BYTES = int((2 * 11485.5 + 23296) * 1024)  # FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, KiB per launch


def run_mode(Q, precision, W, steps, warmup, spl, refresh):
    import numpy as np

    sys_ = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision)
    ens = Q.Ensemble(sys_, W, 0)
    if warmup:
        ens.sweeps(warmup, 0.3, refresh_interval=refresh, steps_per_launch=spl)
    ens.set_timing(True)
    # qmc_ensemble_sweeps returns after its stream has drained (it copies the
    # per-sweep sums back), so the wall clock brackets exactly the K sweeps.
    t0 = time.perf_counter()
    es, e2, acc = ens.sweeps(steps, 0.3, refresh_interval=refresh, steps_per_launch=spl)
    wall = time.perf_counter() - t0
    kms, launches = ens.timing()
    tr = es / W
    tau = Q.autocorrelation_time(tr)
    err = Q.blocking_error(tr, tau)
    out = {
        "wall_s": wall,
        "kernel_ms": kms,
        "launches": launches,
        "walker_moves_per_s": W * steps / wall,
        "energy": float(tr.mean()),
        "energy_err": float(err) if err > 0 else float(np.std(tr) / math.sqrt(len(tr))),
        "acceptance": float(acc.sum() / (W * 10 * steps)),
        "trace": tr,
    }
    if precision == "mixed":
        # E_L bias of the mixed engine on identical configurations: the
        # sampler's cached E_L (FP32 orbitals, Sherman-Morrison-updated FP32
        # inverse) minus the FP64 E_L of the same positions
        ref = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision="fp64")
        el_mixed = ens.state()[2]
        el_fp64 = ref.evaluate_batch(ens.positions())["eloc"]
        d = el_mixed - el_fp64
        out["eval_bias"] = float(d.mean())
        out["eval_bias_sem"] = float(d.std() / math.sqrt(len(d)))
        out["eval_err_max"] = float(np.abs(d).max())
    return out


def run_sr(Q, precision, W, n_meas, warmup):
    """SR-grad samples/s: per-walker E_L, O_k (32 Jastrow parameters + 15
    nuclear coordinates), dE_L/dp_k and the Hellmann-Feynman force of all W
    walkers, accumulated into the FP64 moment matrix (one measure call = W
    samples), on CH4/STO-3G with a 6-term e-e and 8-term e-n Pade series."""
    sys_ = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision)
    sys_.set_jastrow_series(6, 8)
    ens = Q.Ensemble(sys_, W, 0)
    ens.sweeps(warmup, 0.3, refresh_interval=64)
    ens.measure(reset=True)  # first call sizes the buffers
    t0 = time.perf_counter()
    for _ in range(n_meas):
        ens.measure()
    wall = time.perf_counter() - t0
    m = ens.moments()
    return {"samples_per_s": W * n_meas / wall, "params": sys_.num_params(), "nuclear": 3 * sys_.n_nuc,
            "ms_per_call": wall * 1e3 / n_meas, "energy": m["energy"]}


def run_batch(Q, sweeps, trial):
    """Heterogeneous batch: benzene (8192 walkers) + CH4 STO-3G (32768) + H2
    (65536), mixed precision, sampled together; the autotuner picks the
    launch configuration, then the timed sweeps run with it.  The fused
    (one persistent launch) and sequential (one launch per ensemble)
    configurations are timed beside it."""
    systems = [Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision="mixed"),
               Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision="mixed"),
               Q.H2MoleculeMO(precision="mixed")]
    walkers = [8192, 32768, 65536]
    steps = [0.2, 0.3, 0.5]
    ens = [Q.Ensemble(s, W, i) for i, (s, W) in enumerate(zip(systems, walkers))]
    b = Q.Batch(ens)
    b.sweeps(8, steps, refresh_interval=64)
    cfg, tuned_ms = b.tune(trial, steps, refresh_interval=64)
    out = {"tuned_config": cfg, "engines": [s.engine_name() for s in systems], "walkers": walkers}
    t0 = time.perf_counter()
    es, e2, acc = b.sweeps(sweeps, steps, refresh_interval=64)
    wall = time.perf_counter() - t0
    kms, nl = b.timing()
    tot = sum(walkers) * sweeps
    out.update({"walker_moves_per_s": tot / wall, "kernel_walker_moves_per_s": tot / (kms / 1e3),
                "ms_per_sweep": kms / sweeps, "launches": nl,
                "energies": [float(es[i].mean() / walkers[i]) for i in range(3)]})
    for mode in ("fused", "sequential"):
        b.set_config(mode=mode, blocks_per_cu=0, steps_per_launch=cfg["steps_per_launch"], order="cost")
        b.sweeps(sweeps // 2, steps, refresh_interval=64)
        out[f"ms_per_sweep_{mode}"] = b.timing()[0] / (sweeps // 2)
    return out


def run_benzene(Q, sweeps, W=8192):
    """Benzene (BenzeneGTO, 42 electrons, 66 AOs, C-H 2.05 bohr), 8192
    walkers: Metropolis single-electron sweeps (mixed: tile engine, MFMA MO
    contraction; FP64: wave engine), drift-diffusion sweeps (wave engine,
    tau = 0.01) and SR-grad samples (wave derivative kernel: E_L, the 4
    Jastrow O_k and dE_L/dp_k, HF force and Omega over 12 nuclei with the
    C-H MOs following the bonds) -- kernel time from HIP events."""
    out = {"walkers": W}
    for prec in ("mixed", "fp64"):
        s = Q.BenzeneGTO(1.5, 2.0, 2.64, 2.05, precision=prec)
        ens = Q.Ensemble(s, W, 3)
        ens.sweeps(4, 0.2, refresh_interval=64)
        r = {"engine": s.engine_name()}
        for name, move, step in (("metropolis", Q.QMC_MOVE_SINGLE, 0.2), ("drift_diffusion", Q.QMC_MOVE_DRIFT, 0.01)):
            ens.set_timing(True)
            es, e2, acc = ens.sweeps(sweeps, step, move=move, refresh_interval=64, steps_per_launch=sweeps)
            kms, nl = ens.timing()
            ens.set_timing(False)
            r[name] = {"walker_moves_per_s": W * sweeps / (kms / 1e3), "ms_per_sweep": kms / sweeps,
                       "acceptance": float(acc.sum() / (W * 42 * sweeps)), "energy": float(es.mean() / W)}
        ens.measure(reset=True)
        t0 = time.perf_counter()
        for _ in range(4):
            ens.measure()
        wall = time.perf_counter() - t0
        r["sr_grad_samples_per_s"] = 4 * W / wall
        out[prec] = r
    return out


def run_isdmc(Q, precision, W, n_steps):
    """Importance-sampled DMC on CH4 STO-3G: one drift-diffusion sweep plus
    the branching, compaction and E_ref feedback per step, all on the
    device.  Throughput counts the walkers each step actually sweeps."""
    g = Q.MethaneGTO(1.5, 2.0, 2.06, "STO-3G", 0, precision=precision)
    P = Q.ISDMCParams(n_walkers=W, n_steps=n_steps, n_burnin=0, time_step=0.005, n_vmc_equilibrate=20,
                      n_vmc_force_samples=100, refresh_interval=64)
    dmc = Q.ImportanceSampledDMC(g, P, seed=11)
    r = dmc.run()
    pops = dmc.populations
    swept = float(W + pops[:-1].sum())  # walkers entering each step's sweep
    return {"walker_steps_per_s": swept / (r.step_ms / 1e3), "walker_moves_per_s": 10 * swept / (r.step_ms / 1e3),
            "ms_per_step": r.step_ms / n_steps, "avg_population": r.avg_population, "collapses": r.collapses,
            "dropped": r.dropped, "acceptance": r.acceptance_rate}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(seconds_target):
    """The reference algorithm (vmc.rs:95-119: all-electron proposal, full
    recomputation of psi, E_L with 8 sector LU per accepted move) on the CH4
    workload, timed on a bounded sample of walkers: one host thread (the
    reference is single-threaded on this path), and again with the walkers
    spread over the host threads this process may use (at most 16)."""
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
        "sample": f"{W} walkers x {steps} steps (after 2 warm-up), CH4 STO-3G, oracle in reference-algorithm "
                  f"mode, {cpu}, {sec:.1f} s",
    }
    T = min(16, len(os.sched_getaffinity(0)))
    if T > 1:
        Wm = 64 * T
        sm = max(2, int(seconds_target * T / (per_wm * Wm) / 2))
        secm, _ = o.mcmc_time_steps_mt(Wm, 2, sm, 0.3, 0, T)
        out["all_cores"] = {
            "value": Wm * sm / secm,
            "cores": T,
            "sample": f"{Wm} walkers x {sm} steps (after 2 warm-up) on {T} threads, {secm:.1f} s",
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
    ap.add_argument("--batch-sweeps", type=int, default=32,
                    help="timed sweeps of the benzene + CH4 + H2 batch (0: skip it)")
    ap.add_argument("--dmc-steps", type=int, default=100, help="IS-DMC steps per precision (0: skip)")
    ap.add_argument("--benzene-sweeps", type=int, default=8, help="timed benzene sweeps per leg (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 8 device-to-device copies of BYTES (HBM counter calibration) and exit")
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

    W, K = args.walkers, args.steps
    res = {m: run_mode(Q, m, W, K, args.warmup, args.steps_per_launch, args.refresh) for m in ("fp64", "mixed")}
    sr = {m: run_sr(Q, m, W, args.sr_calls, 16) for m in ("fp64", "mixed")}
    batch = run_batch(Q, args.batch_sweeps, 4) if args.batch_sweeps > 0 else None
    dmc = {m: run_isdmc(Q, m, W, args.dmc_steps) for m in ("mixed", "fp64")} if args.dmc_steps > 0 else None
    benzene = run_benzene(Q, args.benzene_sweeps) if args.benzene_sweeps > 0 else None
    mx, fp = res["mixed"], res["fp64"]
    # roofline of the dominant kernel (the mixed-precision sweep)
    avg_launch_s = mx["kernel_ms"] / 1e3 / mx["launches"]
    steps_per_launch = K / mx["launches"]
    achieved = FLOP_PER_WM * W * steps_per_launch / avg_launch_s / 1e12
    # Both modes start from the same Philox seed and run the same sweeps, so
    # their chains stay coupled (an accept decision differs only when the
    # uniform falls between the FP32 and FP64 ratios): the per-sweep
    # difference of ensemble energies is a paired estimator of the total
    # mixed-precision bias (sampling + E_L evaluation) with a small variance.
    d = mx["trace"] - fp["trace"]
    tau_d = Q.autocorrelation_time(d)
    dE = float(d.mean())
    sig = float(Q.blocking_error(d, tau_d))
    if not sig > 0:
        sig = float(d.std() / math.sqrt(len(d)))
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
            "steps_per_launch": args.steps_per_launch,
        },
        "fp64": {"value": fp["walker_moves_per_s"], "ms_per_step": fp["wall_s"] * 1e3 / K,
                 "kernel_ms_per_launch": fp["kernel_ms"] / fp["launches"], "energy": fp["energy"],
                 "energy_err": fp["energy_err"], "acceptance": fp["acceptance"]},
        "mixed": {"energy": mx["energy"], "energy_err": mx["energy_err"], "acceptance": mx["acceptance"],
                  "kernel_ms_per_launch": mx["kernel_ms"] / mx["launches"]},
        "mixed_over_fp64": mx["walker_moves_per_s"] / fp["walker_moves_per_s"],
        "delta_E_mHa": mx["eval_bias"] * 1e3,
        "delta_E_err_mHa": mx["eval_bias_sem"] * 1e3,
        "delta_E_estimator": "E_L bias on identical configurations: mixed sampler's cached E_L minus FP64 E_L "
                             "of the same 65536 final positions (mean, standard error)",
        "delta_E_max_abs_Ha": mx["eval_err_max"],
        "delta_E_chains_mHa": dE * 1e3,
        "delta_E_chains_err_mHa": sig * 1e3,
        "delta_E_chains_estimator": "paired chains, same seed: mean over sweeps of (mixed - FP64) ensemble energy",
        "sr_grad_samples_per_s": {"mixed": sr["mixed"]["samples_per_s"], "fp64": sr["fp64"]["samples_per_s"],
                                  "params": sr["mixed"]["params"], "nuclear_coords": sr["mixed"]["nuclear"],
                                  "ms_per_65536_samples": {"mixed": sr["mixed"]["ms_per_call"],
                                                           "fp64": sr["fp64"]["ms_per_call"]}},
        "heterogeneous_batch": batch,
        "isdmc": dmc,
        "benzene": benzene,
        "roofline": {
            "bound": "mfma",
            "engine": "vector ALU: the sweep has no GEMM-shaped work; 157.3 TFLOP/s is MI355X's FP32 peak "
                      "for both the vector and the matrix path",
            "achieved": achieved,
            "peak": PEAK["mixed"],
            "unit": "TFLOP/s",
            "frac": achieved / PEAK["mixed"],
            # HBM bytes per launch of the mixed sweep (16 sweeps x 65,536
            # walkers) from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the
            # factors from a copy kernel of known size
            "traffic": TRAFFIC_BYTES if (W == 65536 and args.steps_per_launch == 16) else None,
            "traffic_unit": "bytes per launch",
            "kernel": "sweep_kernel<5,5,float>",
            "flop_per_walker_move": FLOP_PER_WM,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
