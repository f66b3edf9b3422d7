/*
 * orc_dmc.c -- oracle restatement of ImportanceSampledDMC
 * (is_dmc.rs:176-607) with the GPU's Philox counters: the drift-diffusion
 * sweep of orc_dd_steps (is_dmc.rs:198-265), branching with at most three
 * copies (is_dmc.rs:272-300), the E_ref feedback (is_dmc.rs:411-422,
 * 485-489), re-initialisation of a dead population (is_dmc.rs:424-433,
 * 492-501), VMC forces from all-electron walkers (is_dmc.rs:305-370) and
 * the blocking error (is_dmc.rs:555-607).  The population keeps the GPU's
 * order: survivors in source order with clones adjacent, and at most
 * `capacity` walkers (clones past it are dropped and counted).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ISDMC_MAX_CLONES 3               /* is_dmc.rs:43 */
#define ISDMC_VMC_KEY 0xD1B54A32D192ED03ull

static uint32_t philox_x(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    uint32_t ctr[4] = {c0, c1, c2, c3}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    orc_philox4x32_10(ctr, key, o);
    return o[0];
}

/* copies of each walker: m = min(floor(W + xi), 3) with
 * W = exp(tau (E_ref - (E_L' + E_L) / 2)) (is_dmc.rs:284-289); the
 * saturating float->usize cast maps NaN to 0. */
void orc_isdmc_branch(int64_t n, const double* el, const double* el_prev, double e_ref, double tau, uint64_t seed,
                      uint32_t step, int* copies) {
    for (int64_t w = 0; w < n; ++w) {
        const double wgt = exp(tau * (e_ref - 0.5 * (el[w] + el_prev[w])));
        const double xi = orc_u01(philox_x(seed, (uint32_t)w, step, 0u, ORC_RNG_BRANCH));
        const double v = wgt + xi;
        copies[w] = v >= ISDMC_MAX_CLONES ? ISDMC_MAX_CLONES : (v >= 0.0 ? (int)floor(v) : 0);
    }
}

/* is_dmc.rs:555-607 */
double orc_isdmc_block_error(const double* e, int64_t n) {
    if (n < 20) return 0.0;
    double mean = 0.0;
    for (int64_t i = 0; i < n; ++i) mean += e[i];
    mean /= (double)n;
    double var = 0.0;
    for (int64_t i = 0; i < n; ++i) var += (e[i] - mean) * (e[i] - mean);
    var /= (double)n;
    if (var < 1e-30) return 0.0;
    double ac = 1.0;
    for (int64_t t = 1; t < n / 2; ++t) {
        double s = 0.0;
        for (int64_t i = 0; i < n - t; ++i) s += (e[i] - mean) * (e[i + t] - mean);
        const double at = s / ((double)(n - t) * var);
        if (at < 0.0) break;
        ac += 2.0 * at;
    }
    const int64_t bs = (int64_t)ceil(2.0 * ac), nb = n / bs;
    if (nb < 2) return sqrt(var * ac / (double)n);
    double bmean = 0.0, bvar = 0.0;
    double* bm = (double*)malloc(sizeof(double) * (size_t)nb);
    for (int64_t b = 0; b < nb; ++b) {
        double s = 0.0;
        for (int64_t i = b * bs; i < b * bs + bs; ++i) s += e[i];
        bm[b] = s / (double)bs;
        bmean += bm[b];
    }
    bmean /= (double)nb;
    for (int64_t b = 0; b < nb; ++b) bvar += (bm[b] - bmean) * (bm[b] - bmean);
    bvar /= (double)(nb - 1);
    free(bm);
    return sqrt(bvar / (double)nb);
}

/* n_walkers fresh walkers from ids w0 + w under key seed */
static double isdmc_init(const orc_system* s, int64_t nw, uint64_t seed, int64_t w0, double* pos, double* el,
                         double* el_prev) {
    const int n = s->n_elec;
    double sum = 0.0;
    for (int64_t w = 0; w < nw; ++w) {
        double(*r)[3] = (double(*)[3])(pos + w * n * 3);
        orc_initialize(s, seed, (uint32_t)(w0 + w), r);
        el[w] = orc_local_energy(s, (const double(*)[3])r);
        el_prev[w] = el[w];
        sum += el[w];
    }
    return sum / (double)nw;
}

static void add_hf(const orc_system* s, const double* pos, int64_t nw, double* acc) {
    const int n = s->n_elec, M = s->n_nuc;
    double F[ORC_MAXN][3];
    for (int64_t w = 0; w < nw; ++w) {
        orc_hf_force(s, (const double(*)[3])(pos + w * n * 3), F);
        for (int I = 0; I < M; ++I)
            for (int c = 0; c < 3; ++c) acc[3 * I + c] += F[I][c];
    }
}

void orc_isdmc_run(const orc_system* s, const orc_isdmc_params* p, uint64_t seed, orc_isdmc_results* res,
                   double* f_mixed, double* f_vmc, double* f_extrap, double* e_trace, double* pop_trace) {
    const int n = s->n_elec, M = s->n_nuc;
    const int64_t cap = p->capacity, N0 = p->n_walkers, T = p->n_burnin + p->n_steps;
    /* the moves diffuse for 2 tau (is_dmc.rs:215-222); the reference branches over tau */
    const double tau_b = p->matched_branching ? 2.0 * p->time_step : p->time_step;
    /* phase 1: VMC forces with all-electron moves of step 0.5 */
    {
        const int64_t nv = N0 < 50 ? N0 : 50;
        const uint64_t kv = seed ^ ISDMC_VMC_KEY;
        double* pos = (double*)malloc(sizeof(double) * (size_t)(nv * n * 3));
        double* lp = (double*)malloc(sizeof(double) * (size_t)nv);
        double* el = (double*)malloc(sizeof(double) * (size_t)nv);
        for (int64_t w = 0; w < nv; ++w) {
            double(*r)[3] = (double(*)[3])(pos + w * n * 3);
            double sg;
            orc_initialize(s, kv, (uint32_t)w, r);
            orc_log_psi(s, (const double(*)[3])r, &lp[w], &sg);
            el[w] = orc_local_energy(s, (const double(*)[3])r);
        }
        int64_t acc = 0;
        uint32_t step = 0;
        orc_all_electron_steps(s, nv, pos, lp, el, kv, step, (int)p->n_vmc_equilibrate, 0.5, &acc);
        step += (uint32_t)p->n_vmc_equilibrate;
        int64_t nfb = p->n_vmc_force_samples / nv;
        if (nfb < 1) nfb = 1;
        double F[3 * ORC_MAXN] = {0};
        for (int64_t it = 0; it < nfb; ++it) {
            orc_all_electron_steps(s, nv, pos, lp, el, kv, step, 5, 0.5, &acc);
            step += 5;
            add_hf(s, pos, nv, F);
        }
        for (int q = 0; q < 3 * M; ++q) f_vmc[q] = F[q] / (double)(nv * nfb);
        free(pos);
        free(lp);
        free(el);
    }
    /* phases 2-4 */
    double* pos = (double*)malloc(sizeof(double) * (size_t)(cap * n * 3));
    double* npos = (double*)malloc(sizeof(double) * (size_t)(cap * n * 3));
    double* el_prev = (double*)malloc(sizeof(double) * (size_t)cap);
    double* eln = (double*)malloc(sizeof(double) * (size_t)cap);
    double* nel = (double*)malloc(sizeof(double) * (size_t)cap);
    double* el = (double*)malloc(sizeof(double) * (size_t)cap);
    int* copies = (int*)malloc(sizeof(int) * (size_t)cap);
    int64_t pop = N0, collapses = 0, dropped = 0, n_fs = 0;
    double e_ref = isdmc_init(s, N0, seed, 0, pos, el, el_prev);
    double Facc[3 * ORC_MAXN] = {0};
    double acc_p = 0.0, moves_p = 0.0, pop_sum = 0.0, e_sum = 0.0;
    for (int64_t t = 0; t < T; ++t) {
        int64_t acc = 0;
        orc_dd_steps(s, pop, pos, seed, (uint32_t)t, 1, p->time_step, p->max_drift, p->log_domain ? ORC_DD_LOG : 0,
                     eln, &acc);
        const int prod = t >= p->n_burnin;
        if (prod) {
            acc_p += (double)acc;
            moves_p += (double)pop * n;
        }
        orc_isdmc_branch(pop, eln, el_prev, e_ref, tau_b, seed, (uint32_t)t, copies);
        int64_t nn = 0;
        for (int64_t w = 0; w < pop; ++w)
            for (int c = 0; c < copies[w]; ++c) {
                if (nn < cap) {
                    memcpy(npos + nn * n * 3, pos + w * n * 3, sizeof(double) * n * 3);
                    nel[nn] = eln[w];
                    ++nn;
                } else {
                    ++dropped;
                }
            }
        double* tp = pos;
        pos = npos;
        npos = tp;
        double es = 0.0;
        for (int64_t w = 0; w < nn; ++w) {
            el_prev[w] = nel[w];
            es += nel[w];
        }
        const double e_mean = es / (double)nn;
        if (prod) {
            e_trace[t - p->n_burnin] = e_mean;
            pop_trace[t - p->n_burnin] = (double)nn;
            pop_sum += (double)nn;
            e_sum += e_mean;
            add_hf(s, pos, nn, Facc);
            n_fs += nn;
        }
        e_ref = e_mean - p->e_ref_damping * log((double)nn / (double)N0) / tau_b;
        pop = nn;
        if (nn == 0) {
            ++collapses;
            e_ref = isdmc_init(s, N0, seed, collapses * N0, pos, el, el_prev);
            pop = N0;
        }
    }
    res->energy = e_sum / (double)p->n_steps;
    res->error = orc_isdmc_block_error(e_trace, p->n_steps);
    res->acceptance_rate = moves_p > 0.0 ? acc_p / moves_p : 0.0;
    res->avg_population = pop_sum / (double)p->n_steps;
    res->collapses = collapses;
    res->dropped = dropped;
    for (int q = 0; q < 3 * M; ++q) {
        f_mixed[q] = Facc[q] / (double)n_fs;
        f_extrap[q] = 2.0 * f_mixed[q] - f_vmc[q];
    }
    free(pos);
    free(npos);
    free(el_prev);
    free(eln);
    free(nel);
    free(el);
    free(copies);
}
