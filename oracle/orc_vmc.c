/*
 * orc_vmc.c -- oracle samplers and estimators: MCMCSimulation
 * (vmc.rs:38-189), the single-electron Metropolis sweep used on the GPU,
 * SROptimizer::compute_sr_update (sr_optimize.rs:176-246) and the ZV force
 * combination of GeometryOptimizer::sample_forces (geometry_opt.rs:198-226).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static uint32_t philox_w0(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    uint32_t ctr[4] = {c0, c1, c2, c3}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    orc_philox4x32_10(ctr, key, o);
    return o[0];
}

/* vmc.rs:141-163 */
double orc_autocorrelation_time(const double* e, int64_t n) {
    double mean = 0.0;
    for (int64_t i = 0; i < n; ++i) mean += e[i];
    mean /= (double)n;
    double var = 0.0;
    for (int64_t i = 0; i < n; ++i) var += (e[i] - mean) * (e[i] - mean);
    var /= (double)n;
    if (var == 0.0) return 1.0;
    double ac = 1.0;
    for (int64_t t = 1; t < n / 2; ++t) {
        double s = 0.0;
        for (int64_t i = 0; i < n - t; ++i) s += (e[i] - mean) * (e[i + t] - mean);
        double at = s / ((double)(n - t) * var);
        if (at < 0.0) break;
        ac += 2.0 * at;
    }
    return ac;
}

/* vmc.rs:166-188 */
double orc_blocking_error(const double* e, int64_t n, double tau) {
    int64_t bs = (int64_t)ceil(2.0 * tau);
    int64_t nb = n / bs;
    if (nb < 2) return 0.0;
    double* bm = (double*)malloc(sizeof(double) * nb);
    for (int64_t b = 0; b < nb; ++b) {
        double s = 0.0;
        for (int64_t i = b * bs; i < b * bs + bs; ++i) s += e[i];
        bm[b] = s / (double)bs;
    }
    double mean = 0.0;
    for (int64_t b = 0; b < nb; ++b) mean += bm[b];
    mean /= (double)nb;
    double var = 0.0;
    for (int64_t b = 0; b < nb; ++b) var += (bm[b] - mean) * (bm[b] - mean);
    var /= (double)(nb - 1);
    free(bm);
    return sqrt(var / (double)nb);
}

typedef struct { double r[ORC_MAXE][3]; double psi; double energy; } mcmc_state;

/* vmc.rs:95-119 with Philox: electron e's normals from counter
 * (w, step, e, ALL), the acceptance uniform from (w, step, N, ALL). */
static int metropolis_step(const orc_system* s, mcmc_state* st, double step_size, uint64_t seed,
                           uint32_t w, uint32_t step) {
    int n = s->n_elec;
    double nr[ORC_MAXE][3];
    for (int e = 0; e < n; ++e) {
        double z[3];
        orc_normals3(seed, w, step, (uint32_t)e, ORC_RNG_ALL, z);
        for (int k = 0; k < 3; ++k) nr[e][k] = st->r[e][k] + step_size * z[k];
    }
    double npsi = orc_evaluate(s, (const double(*)[3])nr);
    double ratio = (npsi / st->psi) * (npsi / st->psi);
    double u = orc_u01(philox_w0(seed, w, step, (uint32_t)n, ORC_RNG_ALL));
    if (u < ratio) {
        memcpy(st->r, nr, sizeof(double) * 3 * n);
        st->psi = npsi;
        st->energy = orc_local_energy(s, (const double(*)[3])st->r);
        return 1;
    }
    return 0;
}

/* vmc.rs:69-138: no burn-in discard, adaptation every adaptation_interval */
void orc_mcmc_run(const orc_system* s, const orc_mcmc_params* p, uint64_t seed, double* energies,
                  orc_mcmc_results* res) {
    int64_t W = p->n_walkers;
    mcmc_state* st = (mcmc_state*)malloc(sizeof(mcmc_state) * W);
    for (int64_t w = 0; w < W; ++w) { /* vmc.rs:57-66 */
        orc_initialize(s, seed, (uint32_t)w, st[w].r);
        st[w].psi = orc_evaluate(s, (const double(*)[3])st[w].r);
        st[w].energy = orc_local_energy(s, (const double(*)[3])st[w].r);
    }
    double step_size = p->initial_step_size;
    int64_t acc = 0, acc_total = 0;
    for (int64_t step = 0; step < p->n_steps; ++step) {
        for (int64_t w = 0; w < W; ++w) acc += metropolis_step(s, &st[w], step_size, seed, (uint32_t)w, (uint32_t)step);
        double m = 0.0;
        for (int64_t w = 0; w < W; ++w) m += st[w].energy;
        energies[step] = m / (double)W;
        if ((step + 1) % p->adaptation_interval == 0) { /* vmc.rs:122-128 */
            double rate = (double)acc / (double)(W * p->adaptation_interval);
            double adj = sqrt(rate / p->target_acceptance);
            step_size = step_size * adj;
            if (step_size < p->min_step_size) step_size = p->min_step_size;
            if (step_size > p->max_step_size) step_size = p->max_step_size;
            acc_total += acc;
            acc = 0;
        }
    }
    acc_total += acc;
    double e = 0.0;
    for (int64_t i = 0; i < p->n_steps; ++i) e += energies[i];
    res->energy = e / (double)p->n_steps;
    res->autocorrelation_time = orc_autocorrelation_time(energies, p->n_steps);
    res->error = orc_blocking_error(energies, p->n_steps, res->autocorrelation_time);
    res->acceptance = (double)acc_total / (double)(W * p->n_steps);
    res->final_step = step_size;
    free(st);
}

/* Timing harness for the CPU baseline: the reference algorithm
 * (metropolis_step with full recomputation, E_L on accept) on W walkers,
 * `warm` untimed steps then `steps` timed steps, one thread. */
void orc_mcmc_time_steps(const orc_system* s, int64_t W, int64_t warm, int64_t steps, double step,
                         uint64_t seed, double* seconds, double* mean_e) {
    mcmc_state* st = (mcmc_state*)malloc(sizeof(mcmc_state) * W);
    for (int64_t w = 0; w < W; ++w) {
        orc_initialize(s, seed, (uint32_t)w, st[w].r);
        st[w].psi = orc_evaluate(s, (const double(*)[3])st[w].r);
        st[w].energy = orc_local_energy(s, (const double(*)[3])st[w].r);
    }
    for (int64_t k = 0; k < warm; ++k)
        for (int64_t w = 0; w < W; ++w) metropolis_step(s, &st[w], step, seed, (uint32_t)w, (uint32_t)k);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    double m = 0.0;
    for (int64_t k = 0; k < steps; ++k) {
        for (int64_t w = 0; w < W; ++w) metropolis_step(s, &st[w], step, seed, (uint32_t)w, (uint32_t)(warm + k));
        for (int64_t w = 0; w < W; ++w) m += st[w].energy;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    *mean_e = m / (double)(W * steps);
    free(st);
}

/* The same workload on `threads` host threads: the walkers' chains are
 * independent (every draw is keyed by walker and step), so each thread runs
 * whole chains of its share of the walkers, warm-up included; the timed
 * region is the chains after the warm-up, bounded by barriers. */
void orc_mcmc_time_steps_mt(const orc_system* s, int64_t W, int64_t warm, int64_t steps, double step,
                            uint64_t seed, int threads, double* seconds, double* mean_e) {
    mcmc_state* st = (mcmc_state*)malloc(sizeof(mcmc_state) * W);
    struct timespec t0, t1;
    double m = 0.0;
#pragma omp parallel num_threads(threads)
    {
#pragma omp for schedule(static)
        for (int64_t w = 0; w < W; ++w) {
            orc_initialize(s, seed, (uint32_t)w, st[w].r);
            st[w].psi = orc_evaluate(s, (const double(*)[3])st[w].r);
            st[w].energy = orc_local_energy(s, (const double(*)[3])st[w].r);
            for (int64_t k = 0; k < warm; ++k) metropolis_step(s, &st[w], step, seed, (uint32_t)w, (uint32_t)k);
        }
#pragma omp single
        clock_gettime(CLOCK_MONOTONIC, &t0);
#pragma omp for schedule(static) reduction(+ : m)
        for (int64_t w = 0; w < W; ++w)
            for (int64_t k = 0; k < steps; ++k) {
                metropolis_step(s, &st[w], step, seed, (uint32_t)w, (uint32_t)(warm + k));
                m += st[w].energy;
            }
#pragma omp single
        clock_gettime(CLOCK_MONOTONIC, &t1);
    }
    *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    *mean_e = m / (double)(W * steps);
    free(st);
}

/* Single-electron Metropolis sweeps, the move scheme of the GPU sampler:
 * for each electron e in sector order (spin-up electrons by increasing
 * index, then spin-down), r_e' = r_e + step * z with z from
 * counter (w, step, 2e, SINGLE), accepted when U < |psi'/psi|^2 with U
 * from (w, step, 2e+1, SINGLE) word 0.  The ratio is taken from full
 * recomputation of ln|psi| (brute force, no incremental update).  After
 * each sweep E_L is measured; eloc is [nsteps][W]. */
void orc_single_electron_steps(const orc_system* s, int64_t W, double* pos, double* logpsi,
                               uint64_t seed, uint32_t step0, int nsteps, double step,
                               double* eloc, int64_t* accepted) {
    int n = s->n_elec;
    int64_t acc = 0;
    int order[ORC_MAXE], q = 0;
    for (int e = 0; e < n; ++e)
        if (s->spins[e] == 1) order[q++] = e;
    for (int e = 0; e < n; ++e)
        if (s->spins[e] != 1) order[q++] = e;
    for (int64_t w = 0; w < W; ++w) {
        double(*r)[3] = (double(*)[3])(pos + w * n * 3);
        double lp, sg;
        orc_log_psi(s, (const double(*)[3])r, &lp, &sg);
        for (int k = 0; k < nsteps; ++k) {
            uint32_t st = step0 + (uint32_t)k;
            for (int qq = 0; qq < n; ++qq) {
                const int e = order[qq];
                double z[3], old[3], u;
                orc_move_draw(seed, (uint32_t)w, st, (uint32_t)e, ORC_RNG_SINGLE, z, &u);
                for (int q = 0; q < 3; ++q) { old[q] = r[e][q]; r[e][q] += step * z[q]; }
                double lpn, sgn;
                orc_log_psi(s, (const double(*)[3])r, &lpn, &sgn);
                double ratio = exp(2.0 * (lpn - lp));
                if (u < ratio) {
                    lp = lpn;
                    ++acc;
                } else {
                    for (int q = 0; q < 3; ++q) r[e][q] = old[q];
                }
            }
            if (eloc) eloc[(int64_t)k * W + w] = orc_local_energy(s, (const double(*)[3])r);
        }
        if (logpsi) logpsi[w] = lp;
    }
    if (accepted) *accepted = acc;
}

/* DriftDiffusionVMC (vmc.rs:195-438) literally: drift F_i = 2 grad_i psi / psi
 * (zero when |psi| < 1e-30, vmc.rs:255-261), the displacement tau F capped
 * at max_drift in norm (vmc.rs:313-321), Langevin proposal
 * r_i' = r_i + cap(tau F_i) + sqrt(2 tau) z, moves with |psi'| < 1e-30
 * skipped, acceptance |psi'/psi|^2 exp(G_b - G_f) with
 * G_f = -|r' - r - cap(tau F(r))|^2 / (4 tau),
 * G_b = -|r - r' - cap(tau F'(r'))|^2 / (4 tau) (vmc.rs:323-346, 375-431),
 * full recomputation of psi and of every drift at the trial configuration.
 * Electrons in the sector order of the GPU sampler; z from counter
 * (w, step, 2e, DRIFT), U from (w, step, 2e+1, DRIFT) word 0.  E_L after
 * each sweep into eloc [nsteps][W].  no_green != 0 drops the Green's
 * function factor (a deliberately wrong sampler for the tests). */
static void dd_drift(const orc_system* s, const double (*r)[3], double psi, int logd, double (*F)[3]) {
    int n = s->n_elec;
    if (logd ? psi == 0.0 : fabs(psi) < 1e-30) {
        for (int e = 0; e < n; ++e) F[e][0] = F[e][1] = F[e][2] = 0.0;
        return;
    }
    orc_derivative(s, r, F);
    for (int e = 0; e < n; ++e)
        for (int q = 0; q < 3; ++q) F[e][q] = 2.0 * F[e][q] / psi;
}

static void dd_cap(const double F[3], double tau, double max_drift, double d[3]) {
    double m = 0.0;
    for (int q = 0; q < 3; ++q) {
        d[q] = tau * F[q];
        m += d[q] * d[q];
    }
    m = sqrt(m);
    if (m > max_drift)
        for (int q = 0; q < 3; ++q) d[q] *= max_drift / m;
}

void orc_dd_steps(const orc_system* s, int64_t W, double* pos, uint64_t seed, uint32_t step0, int nsteps,
                  double tau, double max_drift, int flags, double* eloc, int64_t* accepted) {
    const int no_green = flags & ORC_DD_NO_GREEN, logd = flags & ORC_DD_LOG;
    int n = s->n_elec;
    int64_t acc = 0;
    int order[ORC_MAXE], q = 0;
    for (int e = 0; e < n; ++e)
        if (s->spins[e] == 1) order[q++] = e;
    for (int e = 0; e < n; ++e)
        if (s->spins[e] != 1) order[q++] = e;
    const double s2t = sqrt(2.0 * tau);
    /* ORC_DD_LOG: the ratio from ln|psi| and no 1e-30 guard (vmc.rs:363);
     * the drift 2 grad psi / psi stays a ratio of normal doubles for any
     * psi above ~1e-300, benzene's e^-650 included */
    for (int64_t w = 0; w < W; ++w) {
        double(*r)[3] = (double(*)[3])(pos + w * n * 3);
        double psi = orc_evaluate(s, (const double(*)[3])r), la = 0.0, sg = 0.0;
        if (logd) orc_log_psi(s, (const double(*)[3])r, &la, &sg);
        double F[ORC_MAXE][3], Fn[ORC_MAXE][3], t[ORC_MAXE][3];
        dd_drift(s, (const double(*)[3])r, psi, logd, F);
        for (int k = 0; k < nsteps; ++k) {
            uint32_t st = step0 + (uint32_t)k;
            for (int qq = 0; qq < n; ++qq) {
                const int e = order[qq];
                double z[3], d[3], u;
                orc_move_draw(seed, (uint32_t)w, st, (uint32_t)e, ORC_RNG_DRIFT, z, &u);
                dd_cap(F[e], tau, max_drift, d);
                memcpy(t, r, sizeof(double) * 3 * n);
                for (int c = 0; c < 3; ++c) t[e][c] = r[e][c] + d[c] + s2t * z[c];
                const double psin = orc_evaluate(s, (const double(*)[3])t);
                double lan = 0.0, sgn = 0.0;
                if (logd) {
                    orc_log_psi(s, (const double(*)[3])t, &lan, &sgn);
                    if (sgn == 0.0) continue;
                } else if (fabs(psin) < 1e-30) {
                    continue;
                }
                dd_drift(s, (const double(*)[3])t, psin, logd, Fn);
                double dn[3], gf = 0.0, gb = 0.0;
                dd_cap(Fn[e], tau, max_drift, dn);
                for (int c = 0; c < 3; ++c) {
                    const double a = t[e][c] - r[e][c] - d[c], b = r[e][c] - t[e][c] - dn[c];
                    gf -= a * a;
                    gb -= b * b;
                }
                gf /= 4.0 * tau;
                gb /= 4.0 * tau;
                double p;
                if (logd) {
                    p = exp(2.0 * (lan - la) + (no_green ? 0.0 : gb - gf));
                } else {
                    const double ratio = psin / psi;
                    p = ratio * ratio * (no_green ? 1.0 : exp(gb - gf));
                }
                if (p > 1.0) p = 1.0;
                if (u < p) {
                    memcpy(r, t, sizeof(double) * 3 * n);
                    psi = psin;
                    la = lan;
                    sg = sgn;
                    memcpy(F, Fn, sizeof(double) * 3 * n);
                    ++acc;
                }
            }
            if (eloc) eloc[(int64_t)k * W + w] = orc_local_energy(s, (const double(*)[3])r);
        }
        (void)sg;
    }
    if (accepted) *accepted = acc;
}

/* The all-electron move of vmc.rs:95-119 in the log domain, stepping a
 * caller-held ensemble: cached E_L kept on reject (vmc.rs:111-116).
 * eloc is [W] in/out (the current cached energies). */
void orc_all_electron_steps(const orc_system* s, int64_t W, double* pos, double* logpsi,
                            double* eloc, uint64_t seed, uint32_t step0, int nsteps, double step,
                            int64_t* accepted) {
    int n = s->n_elec;
    int64_t acc = 0;
    for (int64_t w = 0; w < W; ++w) {
        double(*r)[3] = (double(*)[3])(pos + w * n * 3);
        double lp = logpsi[w];
        for (int k = 0; k < nsteps; ++k) {
            uint32_t st = step0 + (uint32_t)k;
            double nr[ORC_MAXE][3];
            for (int e = 0; e < n; ++e) {
                double z[3];
                orc_normals3(seed, (uint32_t)w, st, (uint32_t)e, ORC_RNG_ALL, z);
                for (int q = 0; q < 3; ++q) nr[e][q] = r[e][q] + step * z[q];
            }
            double lpn, sgn;
            orc_log_psi(s, (const double(*)[3])nr, &lpn, &sgn);
            double ratio = exp(2.0 * (lpn - lp));
            double u = orc_u01(philox_w0(seed, (uint32_t)w, st, (uint32_t)n, ORC_RNG_ALL));
            if (u < ratio) {
                memcpy(r, nr, sizeof(double) * 3 * n);
                lp = lpn;
                eloc[w] = orc_local_energy(s, (const double(*)[3])r);
                ++acc;
            }
        }
        logpsi[w] = lp;
    }
    if (accepted) *accepted = acc;
}

/* sr_optimize.rs:176-246.  O is [ns][P].  Returns 1 if S was singular
 * (dp then falls back to -f, as at :239-242). */
int orc_sr_update(int64_t ns, int P, const double* E, const double* O, double eps, double* dp,
                  double* e_mean, double* variance) {
    double n = (double)ns;
    double em = 0.0;
    for (int64_t k = 0; k < ns; ++k) em += E[k];
    em /= n;
    double var = 0.0;
    for (int64_t k = 0; k < ns; ++k) var += (E[k] - em) * (E[k] - em);
    var /= n;
    double om[ORC_MAXP];
    for (int i = 0; i < P; ++i) om[i] = 0.0;
    for (int64_t k = 0; k < ns; ++k)
        for (int i = 0; i < P; ++i) om[i] += O[k * P + i];
    for (int i = 0; i < P; ++i) om[i] /= n;
    double* S = (double*)calloc((size_t)P * P, sizeof(double));
    for (int64_t k = 0; k < ns; ++k)
        for (int i = 0; i < P; ++i)
            for (int j = 0; j < P; ++j) S[i * P + j] += O[k * P + i] * O[k * P + j];
    for (int i = 0; i < P * P; ++i) S[i] /= n;
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < P; ++j) S[i * P + j] -= om[i] * om[j];
    for (int i = 0; i < P; ++i) S[i * P + i] += eps;
    double f[ORC_MAXP], nf[ORC_MAXP] = {0};
    for (int i = 0; i < P; ++i) f[i] = 0.0;
    for (int64_t k = 0; k < ns; ++k)
        for (int i = 0; i < P; ++i) f[i] += E[k] * O[k * P + i];
    for (int i = 0; i < P; ++i) { f[i] /= n; f[i] -= em * om[i]; nf[i] = -f[i]; }
    int singular = orc_lu_solve(P, S, nf, dp);
    if (singular)
        for (int i = 0; i < P; ++i) dp[i] = nf[i];
    free(S);
    *e_mean = em;
    *variance = var;
    return singular;
}

/* geometry_opt.rs:174-226 accumulation and combination.  F and Om are
 * [ns][M][3]; mean_force is [M][3]; fvar is [M]. */
void orc_zv_forces(int64_t ns, int M, const double* E, const double* F, const double* Om,
                   int use_zv, double* mean_force, double* mean_e, double* err, double* fvar) {
    double fa[ORC_MAXN][3] = {{0}}, fsq[ORC_MAXN] = {0}, elo[ORC_MAXN][3] = {{0}}, oa[ORC_MAXN][3] = {{0}};
    double ea = 0.0, esq = 0.0;
    for (int64_t k = 0; k < ns; ++k) {
        for (int I = 0; I < M; ++I) {
            const double* f = F + (k * M + I) * 3;
            for (int q = 0; q < 3; ++q) fa[I][q] += f[q];
            fsq[I] += f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        }
        ea += E[k];
        esq += E[k] * E[k];
        if (use_zv)
            for (int I = 0; I < M; ++I) {
                const double* o = Om + (k * M + I) * 3;
                for (int q = 0; q < 3; ++q) { elo[I][q] += E[k] * o[q]; oa[I][q] += o[q]; }
            }
    }
    double n = (double)ns;
    double me = ea / n;
    double ev = esq / n - me * me;
    *mean_e = me;
    *err = sqrt(ev / n);
    for (int I = 0; I < M; ++I) {
        for (int q = 0; q < 3; ++q) {
            double mf = fa[I][q] / n;
            if (use_zv) {
                double cov = elo[I][q] / n - me * (oa[I][q] / n);
                mf = mf - 2.0 * cov;
            }
            mean_force[I * 3 + q] = mf;
        }
        double msq = fsq[I] / n;
        double sq = (fa[I][0] * fa[I][0] + fa[I][1] * fa[I][1] + fa[I][2] * fa[I][2]) / (n * n);
        double v = msq - sq;
        fvar[I] = v > 0.0 ? v : 0.0;
    }
}
