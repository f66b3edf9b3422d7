/*
 * orc_core.c -- oracle building blocks: Philox4x32-10, LU, orbitals, Jastrow.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Citations are rust_qmc paths.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* ------------------------------------------------------------------------
 * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11.  Multipliers and Weyl key increments are the
 * published constants of the method.  Replaces rand::thread_rng
 * (vmc.rs:51) with a counter-based, seedable stream.
 * ---------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* uniform on the open interval (0, 1): never 0, so log() is safe */
double orc_u01(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

/* Three N(0,1) draws by the Box-Muller transform from one Philox block:
 * (w0, w1) -> (z0, z1), (w2, w3) -> z2 (cosine branch). */
void orc_normals3(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, double z[3]) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const double twopi = 6.283185307179586;
    double u1 = orc_u01(o[0]), u2 = orc_u01(o[1]), u3 = orc_u01(o[2]), u4 = orc_u01(o[3]);
    double rad1 = sqrt(-2.0 * log(u1));
    double rad2 = sqrt(-2.0 * log(u3));
    z[0] = rad1 * cos(twopi * u2);
    z[1] = rad1 * sin(twopi * u2);
    z[2] = rad2 * cos(twopi * u4);
}

/* One single-electron move from one Philox block (the GPU's Rng::move):
 * Box-Muller normals from the top 24 bits of the four words, the acceptance
 * uniform from their low bytes. */
void orc_move_draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, double z[3], double* u) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const double twopi = 6.283185307179586;
    double v[4];
    for (int i = 0; i < 4; ++i) v[i] = ((double)(o[i] >> 8) + 0.5) * 5.9604644775390625e-08;
    double rad1 = sqrt(-2.0 * log(v[0]));
    double rad2 = sqrt(-2.0 * log(v[2]));
    z[0] = rad1 * cos(twopi * v[1]);
    z[1] = rad1 * sin(twopi * v[1]);
    z[2] = rad2 * cos(twopi * v[3]);
    const uint32_t low = (o[0] & 0xFFu) | ((o[1] & 0xFFu) << 8) | ((o[2] & 0xFFu) << 16) | ((o[3] & 0xFFu) << 24);
    *u = orc_u01(low);
}

/* ------------------------------------------------------------------------
 * LU with partial pivoting (row-major a[n*n]).  Stands in for nalgebra's
 * DMatrix::determinant and try_inverse (methane.rs:828-829,
 * slater.rs:217,233) including the silent identity fallback on a singular
 * matrix.  Returns 0 on success, 1 if singular.
 * ---------------------------------------------------------------------- */
static int lu_factor(int n, double* lu, int* piv, double* sgn) {
    *sgn = 1.0;
    for (int i = 0; i < n; ++i) piv[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(lu[k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            double v = fabs(lu[i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        if (best == 0.0) return 1;
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                double t = lu[k * n + j];
                lu[k * n + j] = lu[p * n + j];
                lu[p * n + j] = t;
            }
            int t = piv[k]; piv[k] = piv[p]; piv[p] = t;
            *sgn = -*sgn;
        }
        double inv_piv = 1.0 / lu[k * n + k];
        for (int i = k + 1; i < n; ++i) {
            double f = lu[i * n + k] * inv_piv;
            lu[i * n + k] = f;
            for (int j = k + 1; j < n; ++j) lu[i * n + j] -= f * lu[k * n + j];
        }
    }
    return 0;
}

static void lu_solve_inplace(int n, const double* lu, const int* piv, const double* b, double* x) {
    double y[ORC_MAXAO];
    for (int i = 0; i < n; ++i) {
        double s = b[piv[i]];
        for (int j = 0; j < i; ++j) s -= lu[i * n + j] * y[j];
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < n; ++j) s -= lu[i * n + j] * x[j];
        x[i] = s / lu[i * n + i];
    }
}

int orc_lu_det_inv(int n, const double* a, double* det, double* inv, double* logabs, double* sgn) {
    double lu[ORC_MAXE * ORC_MAXE];
    int piv[ORC_MAXE];
    double s;
    memcpy(lu, a, sizeof(double) * n * n);
    int singular = lu_factor(n, lu, piv, &s);
    if (singular) {
        if (det) *det = 0.0;
        if (logabs) *logabs = -INFINITY;
        if (sgn) *sgn = 0.0;
        if (inv) {
            for (int i = 0; i < n * n; ++i) inv[i] = 0.0;
            for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
        }
        return 1;
    }
    double d = s, la = 0.0, sg = s;
    for (int i = 0; i < n; ++i) {
        double v = lu[i * n + i];
        d *= v;
        la += log(fabs(v));
        if (v < 0) sg = -sg;
    }
    if (det) *det = d;
    if (logabs) *logabs = la;
    if (sgn) *sgn = sg;
    if (inv) {
        double e[ORC_MAXE], x[ORC_MAXE];
        for (int c = 0; c < n; ++c) {
            for (int i = 0; i < n; ++i) e[i] = (i == c) ? 1.0 : 0.0;
            lu_solve_inplace(n, lu, piv, e, x);
            for (int i = 0; i < n; ++i) inv[i * n + c] = x[i];
        }
    }
    return 0;
}

/* nalgebra lu().solve (sr_optimize.rs:238); returns 1 when singular */
int orc_lu_solve(int n, const double* a, const double* b, double* x) {
    double lu[ORC_MAXP * ORC_MAXP];
    int piv[ORC_MAXP];
    double s;
    memcpy(lu, a, sizeof(double) * n * n);
    if (lu_factor(n, lu, piv, &s)) return 1;
    lu_solve_inplace(n, lu, piv, b, x);
    return 0;
}

/* ------------------------------------------------------------------------
 * Orbitals
 * ---------------------------------------------------------------------- */
static double norm3(const double d[3]) { return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]); }

/* slater.rs:17-21 */
double orc_slater1s_eval(const orc_slater1s* o, const double r[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    return exp(-o->alpha * norm3(d));
}

/* slater.rs:23-31 */
void orc_slater1s_grad(const orc_slater1s* o, const double r[3], double g[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    double rn = norm3(d);
    if (rn == 0.0) { g[0] = g[1] = g[2] = 0.0; return; }
    double scalar = -o->alpha / rn * exp(-o->alpha * rn);
    for (int k = 0; k < 3; ++k) g[k] = d[k] * scalar;
}

/* slater.rs:33-41 */
double orc_slater1s_lap(const orc_slater1s* o, const double r[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    double rn = norm3(d);
    if (rn == 0.0) return -INFINITY;
    double e = exp(-o->alpha * rn);
    return (o->alpha * o->alpha - 2.0 * o->alpha / rn) * e;
}

static double powi(double s, int p) {
    double v = 1.0;
    for (int i = 0; i < p; ++i) v *= s;
    return v;
}

/* slater.rs:92-99 */
double orc_sto_eval(const orc_sto* o, const double r[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    double s = norm3(d);
    double v = 0.0;
    for (int k = 0; k < o->n; ++k) v += o->coef[k] * powi(s, o->pw[k]) * exp(-o->zeta[k] * s);
    return v;
}

/* slater.rs:101-121 */
void orc_sto_grad(const orc_sto* o, const double r[3], double g[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    double s = norm3(d);
    if (s == 0.0) { g[0] = g[1] = g[2] = 0.0; return; }
    double rd = 0.0;
    for (int k = 0; k < o->n; ++k) {
        double pf = (double)o->pw[k];
        double ef = exp(-o->zeta[k] * s);
        double sp = powi(s, o->pw[k]);
        double spm1 = sp / s;
        rd += o->coef[k] * ef * (-o->zeta[k] * sp + pf * spm1);
    }
    for (int k = 0; k < 3; ++k) g[k] = d[k] / s * rd;
}

/* slater.rs:123-144 */
double orc_sto_lap(const orc_sto* o, const double r[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    double s = norm3(d);
    if (s == 0.0) return 0.0;
    double v = 0.0;
    for (int k = 0; k < o->n; ++k) {
        double pf = (double)o->pw[k], z = o->zeta[k];
        double ef = exp(-z * s);
        double sp = powi(s, o->pw[k]);
        double spm1 = sp / s;
        double spm2 = spm1 / s;
        double d2 = z * z * sp - 2.0 * z * pf * spm1 + pf * (pf - 1.0) * spm2;
        double d1 = 2.0 * (-z * sp + pf * spm1) / s;
        v += o->coef[k] * ef * (d2 + d1);
    }
    return v;
}

/* methane.rs:498-519 */
double orc_cgto_eval(const orc_cgto* o, const double r[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double ang = 1.0;
    if (o->l == 1) ang = d[o->m < 2 ? o->m : 2];
    double rad = 0.0;
    for (int k = 0; k < o->nprim; ++k) rad += o->coef[k] * exp(-o->expo[k] * r2);
    return ang * rad;
}

/* methane.rs:526-556 */
void orc_cgto_grad(const orc_cgto* o, const double r[3], double g[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double rad = 0.0, rada = 0.0;
    for (int k = 0; k < o->nprim; ++k) rad += o->coef[k] * exp(-o->expo[k] * r2);
    for (int k = 0; k < o->nprim; ++k) rada += o->coef[k] * o->expo[k] * exp(-o->expo[k] * r2);
    double gr[3] = {-2.0 * rada * d[0], -2.0 * rada * d[1], -2.0 * rada * d[2]};
    if (o->l == 0) {
        g[0] = gr[0]; g[1] = gr[1]; g[2] = gr[2];
    } else if (o->l == 1) {
        int mi = o->m;
        double xm = d[mi];
        for (int k = 0; k < 3; ++k) g[k] = xm * gr[k];
        g[mi] += rad;
    } else {
        g[0] = g[1] = g[2] = 0.0;
    }
}

/* methane.rs:563-590 */
double orc_cgto_lap(const orc_cgto* o, const double r[3]) {
    double d[3] = {r[0] - o->center[0], r[1] - o->center[1], r[2] - o->center[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double lr = 0.0;
    for (int k = 0; k < o->nprim; ++k) {
        double e = exp(-o->expo[k] * r2);
        lr += o->coef[k] * (-6.0 * o->expo[k] + 4.0 * o->expo[k] * o->expo[k] * r2) * e;
    }
    if (o->l == 0) return lr;
    if (o->l == 1) {
        double xm = d[o->m];
        double rada = 0.0;
        for (int k = 0; k < o->nprim; ++k) rada += o->coef[k] * o->expo[k] * exp(-o->expo[k] * r2);
        double drdxm = -2.0 * xm * rada;
        return xm * lr + 2.0 * drdxm;
    }
    return 0.0;
}

/* ------------------------------------------------------------------------
 * Jastrow factors
 * ---------------------------------------------------------------------- */
static double dist(const double a[3], const double b[3]) {
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return norm3(d);
}

/* jastrow.rs:30-33 */
double orc_j1_eval(const orc_jastrow1* j, const double (*r)[3]) {
    double r12 = dist(r[0], r[1]);
    return exp(-j->F / (2.0 * (1.0 + r12 / j->F)));
}

/* jastrow.rs:35-42 */
void orc_j1_deriv(const orc_jastrow1* j, const double (*r)[3], double (*g)[3]) {
    double d[3] = {r[0][0] - r[1][0], r[0][1] - r[1][1], r[0][2] - r[1][2]};
    double rn = norm3(d);
    double psi = orc_j1_eval(j, r);
    double f = 1.0 + rn / j->F;
    double denom = 2.0 * f * f * rn;
    double gf = psi / denom;
    for (int k = 0; k < 3; ++k) { g[0][k] = gf * d[k]; g[1][k] = -gf * d[k]; }
}

/* jastrow.rs:44-53 */
void orc_j1_lap(const orc_jastrow1* j, const double (*r)[3], double* l) {
    double rn = dist(r[0], r[1]);
    double f = 1.0 + rn / j->F;
    double psi = orc_j1_eval(j, r);
    double denom = 2.0 * f * f;
    double gsq = 1.0 / (denom * denom);
    double lf = 1.0 / (rn * f * f * f);
    double comp = (gsq + lf) * psi;
    l[0] = comp;
    l[1] = comp;
}

/* jastrow.rs:85-93 */
double orc_j2_eval(const orc_jastrow2* j, const double (*r)[3]) {
    double sum = 0.0;
    for (int a = 0; a < j->n; ++a)
        for (int b = a + 1; b < j->n; ++b) sum += -j->F / (2.0 * (1.0 + dist(r[a], r[b]) / j->F));
    return exp(sum);
}

/* jastrow.rs:95-108 (exact=0), or the exact gradient J * sum_j u'(r_ij) rhat_ij */
void orc_j2_deriv(const orc_jastrow2* j, const double (*r)[3], double (*g)[3]) {
    double psi = orc_j2_eval(j, r);
    for (int a = 0; a < j->n; ++a) g[a][0] = g[a][1] = g[a][2] = 0.0;
    for (int a = 0; a < j->n; ++a)
        for (int b = a + 1; b < j->n; ++b) {
            double d[3] = {r[a][0] - r[b][0], r[a][1] - r[b][1], r[a][2] - r[b][2]};
            double rn = norm3(d);
            double f = 1.0 + rn / j->F;
            double factor = j->exact ? 1.0 / (2.0 * f * f * rn) * psi
                                     : j->F / (2.0 * f * f * rn) * psi;
            for (int k = 0; k < 3; ++k) { g[a][k] += factor * d[k]; g[b][k] -= factor * d[k]; }
        }
}

/* jastrow.rs:110-125 (exact=0: per-pair sum, no cross terms), or the exact
 * J (|sum_j grad u_ij|^2 + sum_j lap u_ij) */
void orc_j2_lap(const orc_jastrow2* j, const double (*r)[3], double* l) {
    double psi = orc_j2_eval(j, r);
    for (int a = 0; a < j->n; ++a) l[a] = 0.0;
    if (!j->exact) {
        for (int a = 0; a < j->n; ++a)
            for (int b = a + 1; b < j->n; ++b) {
                double rn = dist(r[a], r[b]);
                double f = 1.0 + rn / j->F;
                double denom = 2.0 * f * f;
                double gsq = 1.0 / (denom * denom);
                double lf = 1.0 / (rn * f * f * f);
                double comp = (gsq + lf) * psi;
                l[a] += comp;
                l[b] += comp;
            }
        return;
    }
    for (int a = 0; a < j->n; ++a) {
        double gu[3] = {0, 0, 0}, lu = 0.0;
        for (int b = 0; b < j->n; ++b) {
            if (b == a) continue;
            double d[3] = {r[a][0] - r[b][0], r[a][1] - r[b][1], r[a][2] - r[b][2]};
            double rn = norm3(d);
            double f = 1.0 + rn / j->F;
            double up = 1.0 / (2.0 * f * f);
            for (int k = 0; k < 3; ++k) gu[k] += up * d[k] / rn;
            lu += 1.0 / (rn * f * f * f);
        }
        l[a] = psi * (gu[0] * gu[0] + gu[1] * gu[1] + gu[2] * gu[2] + lu);
    }
}

/* Pade power series s(x) = sum_{k=0}^{K-1} c_k x^{k+2} with its first two
 * x-derivatives (series extension; not in the reference). */
static void series(int K, const double* c, double x, double* s, double* s1, double* s2) {
    double v = 0.0, v1 = 0.0, v2 = 0.0;
    double xp = 1.0; /* x^k */
    for (int k = 0; k < K; ++k) {
        double n = (double)(k + 2);
        v += c[k] * xp * x * x;
        v1 += c[k] * n * xp * x;
        v2 += c[k] * n * (n - 1.0) * xp;
        xp *= x;
    }
    *s = v;
    *s1 = v1;
    *s2 = v2;
}

/* jastrow.rs:218-233 (+ series) */
static double j3_u_ee(const orc_jastrow3* j, const double (*r)[3]) {
    double sum = 0.0;
    for (int a = 0; a < j->n; ++a)
        for (int b = a + 1; b < j->n; ++b) {
            double rij = dist(r[a], r[b]);
            int same = j->spins[a] == j->spins[b];
            double aa = same ? j->a_para : j->a_anti;
            sum += aa * rij / (1.0 + j->b_ee * rij);
            if (j->k_ee > 0) {
                double s, s1, s2;
                series(j->k_ee, same ? j->c_para : j->c_anti, rij / (1.0 + j->b_ee * rij), &s, &s1, &s2);
                sum += s;
            }
        }
    return sum;
}

/* jastrow.rs:236-247 (+ series) */
static double j3_u_en(const orc_jastrow3* j, const double (*r)[3]) {
    double sum = 0.0;
    for (int a = 0; a < j->n; ++a)
        for (int I = 0; I < j->n_nuc; ++I) {
            double rin = dist(r[a], j->nuc[I]);
            double aen = j->Z[I];
            sum -= aen * rin / (1.0 + j->b_en * rin);
            if (j->k_en > 0) {
                double s, s1, s2;
                series(j->k_en, j->d_en[j->species[I]], rin / (1.0 + j->b_en * rin), &s, &s1, &s2);
                sum += s;
            }
        }
    return sum;
}

double orc_j3_u(const orc_jastrow3* j, const double (*r)[3]) { return j3_u_ee(j, r) + j3_u_en(j, r); }

/* jastrow.rs:250-269 (+ series) */
static void j3_grad_u_ee(const orc_jastrow3* j, const double (*r)[3], int i, double g[3]) {
    g[0] = g[1] = g[2] = 0.0;
    for (int b = 0; b < j->n; ++b) {
        if (b == i) continue;
        double d[3] = {r[i][0] - r[b][0], r[i][1] - r[b][1], r[i][2] - r[b][2]};
        double rij = norm3(d);
        if (rij < 1e-10) continue;
        int same = j->spins[i] == j->spins[b];
        double aa = same ? j->a_para : j->a_anti;
        double denom = 1.0 + j->b_ee * rij;
        double factor = aa / (denom * denom * rij);
        if (j->k_ee > 0) {
            double s, s1, s2;
            series(j->k_ee, same ? j->c_para : j->c_anti, rij / denom, &s, &s1, &s2);
            factor += s1 / (denom * denom * rij);
        }
        for (int k = 0; k < 3; ++k) g[k] += factor * d[k];
    }
}

/* jastrow.rs:272-286 (+ series) */
static void j3_grad_u_en(const orc_jastrow3* j, const double (*r)[3], int i, double g[3]) {
    g[0] = g[1] = g[2] = 0.0;
    for (int I = 0; I < j->n_nuc; ++I) {
        double d[3] = {r[i][0] - j->nuc[I][0], r[i][1] - j->nuc[I][1], r[i][2] - j->nuc[I][2]};
        double rin = norm3(d);
        if (rin < 1e-10) continue;
        double aen = j->Z[I];
        double denom = 1.0 + j->b_en * rin;
        double factor = -aen / (denom * denom * rin);
        if (j->k_en > 0) {
            double s, s1, s2;
            series(j->k_en, j->d_en[j->species[I]], rin / denom, &s, &s1, &s2);
            factor += s1 / (denom * denom * rin);
        }
        for (int k = 0; k < 3; ++k) g[k] += factor * d[k];
    }
}

/* second radial derivative plus 2/r first derivative of the series term:
 * d2s/dr2 = s''(x)/D^4 - 2 b s'(x)/D^3, ds/dr = s'(x)/D^2 */
static double series_lap(double s1, double s2, double b, double r, double denom) {
    double d2 = denom * denom;
    return s2 / (d2 * d2) - 2.0 * b * s1 / (d2 * denom) + 2.0 * s1 / (d2 * r);
}

/* jastrow.rs:289-306 (+ series) */
static double j3_lap_u_ee(const orc_jastrow3* j, const double (*r)[3], int i) {
    double lap = 0.0;
    for (int b = 0; b < j->n; ++b) {
        if (b == i) continue;
        double rij = dist(r[i], r[b]);
        if (rij < 1e-10) continue;
        int same = j->spins[i] == j->spins[b];
        double aa = same ? j->a_para : j->a_anti;
        double denom = 1.0 + j->b_ee * rij;
        lap += 2.0 * aa / (rij * denom * denom) - 2.0 * aa * j->b_ee / (denom * denom * denom);
        if (j->k_ee > 0) {
            double s, s1, s2;
            series(j->k_ee, same ? j->c_para : j->c_anti, rij / denom, &s, &s1, &s2);
            lap += series_lap(s1, s2, j->b_ee, rij, denom);
        }
    }
    return lap;
}

/* jastrow.rs:309-321 (+ series) */
static double j3_lap_u_en(const orc_jastrow3* j, const double (*r)[3], int i) {
    double lap = 0.0;
    for (int I = 0; I < j->n_nuc; ++I) {
        double rin = dist(r[i], j->nuc[I]);
        if (rin < 1e-10) continue;
        double aen = j->Z[I];
        double denom = 1.0 + j->b_en * rin;
        lap += -2.0 * aen / (rin * denom * denom) + 2.0 * aen * j->b_en / (denom * denom * denom);
        if (j->k_en > 0) {
            double s, s1, s2;
            series(j->k_en, j->d_en[j->species[I]], rin / denom, &s, &s1, &s2);
            lap += series_lap(s1, s2, j->b_en, rin, denom);
        }
    }
    return lap;
}

/* jastrow.rs:428-430 */
double orc_j3_eval(const orc_jastrow3* j, const double (*r)[3]) { return exp(j3_u_ee(j, r) + j3_u_en(j, r)); }

/* jastrow.rs:432-437 */
void orc_j3_deriv(const orc_jastrow3* j, const double (*r)[3], double (*g)[3]) {
    double J = orc_j3_eval(j, r);
    for (int i = 0; i < j->n; ++i) {
        double ge[3], gn[3];
        j3_grad_u_ee(j, r, i, ge);
        j3_grad_u_en(j, r, i, gn);
        for (int k = 0; k < 3; ++k) g[i][k] = J * (ge[k] + gn[k]);
    }
}

/* jastrow.rs:439-449 */
void orc_j3_lap(const orc_jastrow3* j, const double (*r)[3], double* l) {
    double J = orc_j3_eval(j, r);
    for (int i = 0; i < j->n; ++i) {
        double ge[3], gn[3];
        j3_grad_u_ee(j, r, i, ge);
        j3_grad_u_en(j, r, i, gn);
        double gu[3] = {ge[0] + gn[0], ge[1] + gn[1], ge[2] + gn[2]};
        double lu = j3_lap_u_ee(j, r, i) + j3_lap_u_en(j, r, i);
        l[i] = J * (lu + gu[0] * gu[0] + gu[1] * gu[1] + gu[2] * gu[2]);
    }
}

int orc_j3_num_params(const orc_jastrow3* j) { return 4 + 2 * j->k_ee + j->n_species * j->k_en; }

/* methane.rs:1191-1198 ordering [b_ee, b_en, a_anti, a_para], then series */
void orc_j3_get_params(const orc_jastrow3* j, double* p) {
    p[0] = j->b_ee;
    p[1] = j->b_en;
    p[2] = j->a_anti;
    p[3] = j->a_para;
    int q = 4;
    for (int k = 0; k < j->k_ee; ++k) p[q++] = j->c_anti[k];
    for (int k = 0; k < j->k_ee; ++k) p[q++] = j->c_para[k];
    for (int s = 0; s < j->n_species; ++s)
        for (int k = 0; k < j->k_en; ++k) p[q++] = j->d_en[s][k];
}

/* methane.rs:1200-1206 */
void orc_j3_set_params(orc_jastrow3* j, const double* p) {
    j->b_ee = p[0];
    j->b_en = p[1];
    j->a_anti = p[2];
    j->a_para = p[3];
    int q = 4;
    for (int k = 0; k < j->k_ee; ++k) j->c_anti[k] = p[q++];
    for (int k = 0; k < j->k_ee; ++k) j->c_para[k] = p[q++];
    for (int s = 0; s < j->n_species; ++s)
        for (int k = 0; k < j->k_en; ++k) j->d_en[s][k] = p[q++];
}

/* jastrow.rs:326-391 (methane.rs:1212-1219 ordering), plus the series
 * parameters: dx/db = -x^2 so d/db [a x + s(x)] = -(a + s'(x)) x^2. */
void orc_j3_log_derivs(const orc_jastrow3* j, const double (*r)[3], double* O) {
    int P = orc_j3_num_params(j);
    for (int q = 0; q < P; ++q) O[q] = 0.0;
    /* d_ln_j_d_bee, jastrow.rs:326-342 */
    for (int a = 0; a < j->n; ++a)
        for (int b = a + 1; b < j->n; ++b) {
            double rij = dist(r[a], r[b]);
            int same = j->spins[a] == j->spins[b];
            double aa = same ? j->a_para : j->a_anti;
            double denom = 1.0 + j->b_ee * rij;
            O[0] += -aa * rij * rij / (denom * denom);
            double x = rij / denom;
            if (same) O[3] += x; else O[2] += x;  /* jastrow.rs:364-391 */
            if (j->k_ee > 0) {
                double s, s1, s2;
                const double* c = same ? j->c_para : j->c_anti;
                series(j->k_ee, c, x, &s, &s1, &s2);
                O[0] += -s1 * x * x;
                double xp = x * x;
                int base = 4 + (same ? j->k_ee : 0);
                for (int k = 0; k < j->k_ee; ++k) { O[base + k] += xp; xp *= x; }
            }
        }
    /* d_ln_j_d_ben, jastrow.rs:347-359 */
    for (int a = 0; a < j->n; ++a)
        for (int I = 0; I < j->n_nuc; ++I) {
            double rin = dist(r[a], j->nuc[I]);
            double aen = j->Z[I];
            double denom = 1.0 + j->b_en * rin;
            O[1] += aen * rin * rin / (denom * denom);
            if (j->k_en > 0) {
                double y = rin / denom, s, s1, s2;
                series(j->k_en, j->d_en[j->species[I]], y, &s, &s1, &s2);
                O[1] += -s1 * y * y;
                double yp = y * y;
                int base = 4 + 2 * j->k_ee + j->species[I] * j->k_en;
                for (int k = 0; k < j->k_en; ++k) { O[base + k] += yp; yp *= y; }
            }
        }
}
