/*
 * orc_systems.c -- oracle restatement of the reference's physical systems:
 * H2MoleculeMO (hydrogen.rs), Lithium (lithium.rs + STOSlaterDet in
 * slater.rs), and MethaneGTO / BenzeneGTO generalised to any number of AOs
 * and MOs (methane.rs:475-1277, benzene.rs).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double norm3(const double d[3]) { return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]); }
static double dist(const double a[3], const double b[3]) {
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return norm3(d);
}
static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static orc_system* sys_alloc(int kind) {
    orc_system* s = (orc_system*)calloc(1, sizeof(orc_system));
    s->kind = kind;
    return s;
}

void orc_system_free(orc_system* s) { free(s); }
int orc_system_n_elec(const orc_system* s) { return s->n_elec; }
int orc_system_n_nuc(const orc_system* s) { return s->n_nuc; }
int orc_system_n_ao(const orc_system* s) { return s->n_ao; }
int orc_system_n_mo(const orc_system* s) { return s->n_mo; }

void orc_system_get_nuclei(const orc_system* s, double* R, double* Z) {
    for (int I = 0; I < s->n_nuc; ++I) {
        for (int k = 0; k < 3; ++k) R[3 * I + k] = s->nuc[I][k];
        if (Z) Z[I] = s->Z[I];
    }
}

void orc_system_get_mo(const orc_system* s, int spin, double* mo) {
    for (int m = 0; m < s->n_mo; ++m)
        for (int a = 0; a < s->n_ao; ++a) mo[m * s->n_ao + a] = s->mo[spin][m][a];
}

/* ======================================================================
 * Basis-set data
 * ==================================================================== */

/* methane.rs:701-708, benzene.rs:161-168 (6-31G carbon 1s) */
static const double C631_1S_E[6] = {3047.5249, 457.36951, 103.94869, 29.210155, 9.2866630, 3.1639270};
static const double C631_1S_C[6] = {0.0018347, 0.0140373, 0.0688426, 0.2321844, 0.4679413, 0.3623120};
/* methane.rs:712-714 (6-31G carbon SP inner) */
static const double C631_SP_E[3] = {7.8682724, 1.8812885, 0.5442493};
static const double C631_SP_S[3] = {-0.1193324, -0.1608542, 1.1434564};
static const double C631_SP_P[3] = {0.0689991, 0.3164240, 0.7443083};
/* methane.rs:733 */
static const double C631_OUT_E = 0.1687145;
/* methane.rs:752-756 */
static const double H631_IN_E[3] = {18.7311370, 2.8253937, 0.6401217};
static const double H631_IN_C[3] = {0.0334946, 0.2347270, 0.8137573};
/* methane.rs:764 */
static const double H631_OUT_E = 0.1612778;

/* STO-3G (Hehre, Stewart, Pople, J. Chem. Phys. 51, 2657 (1969)); the
 * reference carries no STO-3G data, so this is the published table. */
static const double H_STO3G_E[3] = {3.42525091, 0.62391373, 0.16885540};
static const double STO3G_1S_C[3] = {0.15432897, 0.53532814, 0.44463454};
static const double C_STO3G_1S_E[3] = {71.6168370, 13.0450960, 3.5305122};
static const double C_STO3G_SP_E[3] = {2.9412494, 0.6834831, 0.2222899};
static const double C_STO3G_SP_S[3] = {-0.09996723, 0.39951283, 0.70011547};
static const double C_STO3G_SP_P[3] = {0.15591627, 0.60768372, 0.39195739};

static void add_cgto(orc_system* s, int nuc, int l, int m, int np, const double* e, const double* c) {
    orc_cgto* o = &s->ao[s->n_ao];
    for (int k = 0; k < 3; ++k) o->center[k] = s->nuc[nuc][k];
    o->nprim = np;
    for (int k = 0; k < np; ++k) { o->expo[k] = e[k]; o->coef[k] = c[k]; }
    o->l = l;
    o->m = m;
    s->ao_nuc[s->n_ao] = nuc;
    s->n_ao++;
}

/* methane.rs:697-776: 9 AOs on a carbon */
static void add_c631(orc_system* s, int nuc) {
    double one = 1.0;
    add_cgto(s, nuc, 0, 0, 6, C631_1S_E, C631_1S_C);
    add_cgto(s, nuc, 0, 0, 3, C631_SP_E, C631_SP_S);
    for (int m = 0; m < 3; ++m) add_cgto(s, nuc, 1, m, 3, C631_SP_E, C631_SP_P);
    add_cgto(s, nuc, 0, 0, 1, &C631_OUT_E, &one);
    for (int m = 0; m < 3; ++m) add_cgto(s, nuc, 1, m, 1, &C631_OUT_E, &one);
}

static void jastrow3_init(orc_system* s, double b_ee, double b_en) {
    /* jastrow.rs:165-185 (new_general) and :188-215 (new_ch4) */
    orc_jastrow3* j = &s->j3;
    memset(j, 0, sizeof(*j));
    j->n = s->n_elec;
    j->a_anti = 0.5;
    j->a_para = 0.25;
    j->b_ee = b_ee;
    j->b_en = b_en;
    j->n_nuc = s->n_nuc;
    for (int I = 0; I < s->n_nuc; ++I) {
        for (int k = 0; k < 3; ++k) j->nuc[I][k] = s->nuc[I][k];
        j->Z[I] = s->Z[I];
    }
    for (int e = 0; e < s->n_elec; ++e) j->spins[e] = s->spins[e];
    /* species for the series extension: distinct charges in order of first appearance */
    j->n_species = 0;
    double seen[ORC_MAXSPEC];
    for (int I = 0; I < s->n_nuc; ++I) {
        int sp = -1;
        for (int q = 0; q < j->n_species; ++q)
            if (seen[q] == s->Z[I]) sp = q;
        if (sp < 0 && j->n_species < ORC_MAXSPEC) { seen[j->n_species] = s->Z[I]; sp = j->n_species++; }
        j->species[I] = sp < 0 ? 0 : sp;
    }
}

void orc_set_jastrow_series(orc_system* s, int k_ee, int k_en) {
    if (s->kind != ORC_SYS_GTO) return;
    s->j3.k_ee = k_ee;
    s->j3.k_en = k_en;
}

/* ======================================================================
 * Constructors
 * ==================================================================== */

/* hydrogen.rs:28-35; nuclei at +-half_bond on z (conf/h2.yaml:1-8) */
orc_system* orc_h2mo_new_general(const double c1[3], const double c2[3], double alpha, double F,
                                 double init_sigma) {
    orc_system* s = sys_alloc(ORC_SYS_H2MO);
    s->n_elec = 2;
    s->spins[0] = 1;
    s->spins[1] = -1;
    s->n_nuc = 2;
    for (int k = 0; k < 3; ++k) { s->nuc[0][k] = c1[k]; s->nuc[1][k] = c2[k]; }
    s->Z[0] = s->Z[1] = 1.0;
    s->orb1.alpha = alpha;
    s->orb2.alpha = alpha;
    for (int k = 0; k < 3; ++k) { s->orb1.center[k] = c1[k]; s->orb2.center[k] = c2[k]; }
    s->j1.F = F;
    s->init_nuc[0] = 0;
    s->init_nuc[1] = 1;
    s->init_sigma = init_sigma;
    return s;
}

orc_system* orc_h2mo_new(double alpha, double half_bond, double F, double init_sigma) {
    double c1[3] = {0, 0, half_bond}, c2[3] = {0, 0, -half_bond};
    return orc_h2mo_new_general(c1, c2, alpha, F, init_sigma);
}

/* lithium.rs:9-19 with STOSlaterDet::new (slater.rs:162-172) */
orc_system* orc_lithium_new(int n, const orc_sto* orbitals, const int* spins, double F, int Z,
                            int exact_j2) {
    orc_system* s = sys_alloc(ORC_SYS_LITHIUM);
    s->n_elec = n;
    s->n_orb = n;
    for (int e = 0; e < n; ++e) {
        s->spins[e] = spins[e];
        s->orbitals[e] = orbitals[e];
        s->init_nuc[e] = 0;
    }
    s->n_nuc = 1;
    s->nuc[0][0] = s->nuc[0][1] = s->nuc[0][2] = 0.0;
    s->Z[0] = (double)Z;
    s->li_Z = Z;
    s->j2.F = F;
    s->j2.n = n;
    s->j2.exact = exact_j2;
    s->init_sigma = 1.0; /* slater.rs:198-204: N(0,1) about the origin */
    return s;
}

/* slater.rs:68-89: the 7-primitive init_li_sto used by main.rs:37-48 */
static void init_li_sto(orc_sto* o, int n) {
    static const double ex[7] = {0.72089388, 2.61691643, 0.69257443, 1.37137558, 3.97864549, 13.52900016, 19.30801440};
    static const int pw[7] = {1, 1, 2, 2, 2, 2, 3};
    static const double c1[7] = {-0.12220686, 1.11273225, 0.04125378, 0.09306499, -0.10260021, -0.00034191, 0.00021963};
    static const double c2[7] = {0.47750469, 0.11140449, -1.25954273, -0.18475003, -0.02736293, -0.00025064, 0.00057962};
    memset(o, 0, sizeof(*o));
    o->n = 7;
    for (int k = 0; k < 7; ++k) {
        o->zeta[k] = ex[k];
        o->pw[k] = pw[k];
        o->coef[k] = (n == 1) ? c1[k] : (n == 2 ? c2[k] : 0.0);
    }
}

/* src/main.rs:36-49: 1s, 1s, 2s with spins [+1,-1,+1], Jastrow2 F=1 */
orc_system* orc_lithium_ref_new(int exact_j2) {
    orc_sto orb[3];
    init_li_sto(&orb[0], 1);
    init_li_sto(&orb[1], 1);
    init_li_sto(&orb[2], 2);
    int spins[3] = {1, -1, 1};
    return orc_lithium_new(3, orb, spins, 1.0, 3, exact_j2);
}

static double factorial(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

/* Seeded Li STO orbitals: psi_n = A sum_{nu} c_nu r^{p_nu} exp(-xi_n r),
 * p = 0, 1, 2, xi = 2.7 (1s) and 0.65 (2s), c_nu ~ U(-1,1) from Philox
 * (counter (3 n + nu, 1, 0, SYS)), A normalising int psi^2 d^3r = 1. */
orc_system* orc_lithium_config_new(uint64_t seed, int exact_j2) {
    const double xi[2] = {2.7, 0.65};
    orc_sto base[2];
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int o = 0; o < 2; ++o) {
        memset(&base[o], 0, sizeof(orc_sto));
        base[o].n = 3;
        double c[3];
        for (int nu = 0; nu < 3; ++nu) {
            uint32_t ctr[4] = {(uint32_t)(3 * o + nu), 1u, 0u, ORC_RNG_SYS}, w[4];
            orc_philox4x32_10(ctr, key, w);
            c[nu] = 2.0 * orc_u01(w[0]) - 1.0;
        }
        double norm = 0.0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                int q = a + b;
                norm += c[a] * c[b] * 4.0 * M_PI * factorial(q + 2) / pow(2.0 * xi[o], q + 3);
            }
        double A = 1.0 / sqrt(norm);
        for (int nu = 0; nu < 3; ++nu) {
            base[o].coef[nu] = A * c[nu];
            base[o].pw[nu] = nu;
            base[o].zeta[nu] = xi[o];
        }
    }
    orc_sto orb[3] = {base[0], base[0], base[1]};
    int spins[3] = {1, -1, 1};
    return orc_lithium_new(3, orb, spins, 1.0, 3, exact_j2);
}

static void methane_geometry(orc_system* s, double ch) {
    /* methane.rs:625-632 */
    double d = ch / sqrt(3.0);
    double H[4][3] = {{d, d, d}, {d, -d, -d}, {-d, d, -d}, {-d, -d, d}};
    s->n_nuc = 5;
    s->nuc[0][0] = s->nuc[0][1] = s->nuc[0][2] = 0.0;
    s->Z[0] = 6.0;
    for (int h = 0; h < 4; ++h) {
        for (int k = 0; k < 3; ++k) s->nuc[1 + h][k] = H[h][k];
        s->Z[1 + h] = 1.0;
    }
    s->n_elec = 10;
    for (int e = 0; e < 10; ++e) {
        s->spins[e] = e < 5 ? 1 : -1;          /* methane.rs:640 */
        s->init_nuc[e] = e < 6 ? 0 : e - 5;    /* methane.rs:1110-1119 */
    }
    s->init_sigma = 0.5;
}

/* methane.rs:623-685 (17-AO 6-31G, fixed MO coefficients) */
orc_system* orc_methane_gto_new(double b_ee, double b_en, double ch_bond) {
    static const double MO[5][17] = {
        {0.9943, 0.0234, 0.0, 0.0, 0.0, 0.0020, 0.0, 0.0, 0.0, 0.0018, 0.0018, 0.0018, 0.0018, 0.0015, 0.0015, 0.0015, 0.0015},
        {-0.0237, 0.2223, 0.0, 0.0, 0.0, 0.8340, 0.0, 0.0, 0.0, 0.0673, 0.0673, 0.0673, 0.0673, 0.0437, 0.0437, 0.0437, 0.0437},
        {0.0, 0.0, 0.2090, 0.0, 0.0, 0.0, 0.5730, 0.0, 0.0, 0.1020, 0.1020, -0.1020, -0.1020, 0.0670, 0.0670, -0.0670, -0.0670},
        {0.0, 0.0, 0.0, 0.2090, 0.0, 0.0, 0.0, 0.5730, 0.0, 0.1020, -0.1020, 0.1020, -0.1020, 0.0670, -0.0670, 0.0670, -0.0670},
        {0.0, 0.0, 0.0, 0.0, 0.2090, 0.0, 0.0, 0.0, 0.5730, 0.1020, -0.1020, -0.1020, 0.1020, 0.0670, -0.0670, -0.0670, 0.0670},
    };
    orc_system* s = sys_alloc(ORC_SYS_GTO);
    methane_geometry(s, ch_bond);
    add_c631(s, 0);
    double one = 1.0;
    for (int h = 0; h < 4; ++h) add_cgto(s, 1 + h, 0, 0, 3, H631_IN_E, H631_IN_C);
    for (int h = 0; h < 4; ++h) add_cgto(s, 1 + h, 0, 0, 1, &H631_OUT_E, &one);
    s->n_mo = 5;
    for (int sp = 0; sp < 2; ++sp)
        for (int m = 0; m < 5; ++m)
            for (int a = 0; a < 17; ++a) s->mo[sp][m][a] = MO[m][a];
    jastrow3_init(s, b_ee, b_en);
    return s;
}

/* Seeded orthonormal MO rows: C_{mu nu} = 2 U - 1 with U from Philox
 * counter (mu*n_ao + nu, 0, 0, SYS), then modified Gram-Schmidt over rows
 * in the Euclidean metric. */
void orc_random_mo(uint64_t seed, int n_mo, int n_ao, double* mo) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int m = 0; m < n_mo; ++m)
        for (int a = 0; a < n_ao; ++a) {
            uint32_t ctr[4] = {(uint32_t)(m * n_ao + a), 0u, 0u, ORC_RNG_SYS}, w[4];
            orc_philox4x32_10(ctr, key, w);
            mo[m * n_ao + a] = 2.0 * orc_u01(w[0]) - 1.0;
        }
    for (int m = 0; m < n_mo; ++m) {
        for (int q = 0; q < m; ++q) {
            double p = 0.0;
            for (int a = 0; a < n_ao; ++a) p += mo[m * n_ao + a] * mo[q * n_ao + a];
            for (int a = 0; a < n_ao; ++a) mo[m * n_ao + a] -= p * mo[q * n_ao + a];
        }
        double nn = 0.0;
        for (int a = 0; a < n_ao; ++a) nn += mo[m * n_ao + a] * mo[m * n_ao + a];
        double inv = 1.0 / sqrt(nn);
        for (int a = 0; a < n_ao; ++a) mo[m * n_ao + a] *= inv;
    }
}

/* CH4 with the STO-3G basis (9 AOs: C 1s, 2s, 2p x3; H 1s x4) and seeded
 * orthonormal 5x9 MO coefficients shared by both spins. */
orc_system* orc_methane_sto3g_new(double b_ee, double b_en, double ch_bond, uint64_t mo_seed) {
    orc_system* s = sys_alloc(ORC_SYS_GTO);
    methane_geometry(s, ch_bond);
    add_cgto(s, 0, 0, 0, 3, C_STO3G_1S_E, STO3G_1S_C);
    add_cgto(s, 0, 0, 0, 3, C_STO3G_SP_E, C_STO3G_SP_S);
    for (int m = 0; m < 3; ++m) add_cgto(s, 0, 1, m, 3, C_STO3G_SP_E, C_STO3G_SP_P);
    for (int h = 0; h < 4; ++h) add_cgto(s, 1 + h, 0, 0, 3, H_STO3G_E, STO3G_1S_C);
    s->n_mo = 5;
    double mo[5 * 9];
    orc_random_mo(mo_seed, 5, 9, mo);
    for (int sp = 0; sp < 2; ++sp)
        for (int m = 0; m < 5; ++m)
            for (int a = 0; a < 9; ++a) s->mo[sp][m][a] = mo[m * 9 + a];
    jastrow3_init(s, b_ee, b_en);
    return s;
}

/* benzene.rs:345-381: MOs 9-14, C-H sigma bonds; the C 2p part points from
 * carbon n to hydrogen n, so it is rebuilt whenever the nuclei move
 * (BenzeneGTO::set_nuclei, benzene.rs:786-809). */
static void benzene_ch_mos(double mo[21][66], const orc_system* s) {
    for (int n = 0; n < 6; ++n) {
        int m = 9 + n;
        double dir[3] = {s->nuc[6 + n][0] - s->nuc[n][0], s->nuc[6 + n][1] - s->nuc[n][1],
                         s->nuc[6 + n][2] - s->nuc[n][2]};
        double dn = norm3(dir);
        for (int k = 0; k < 3; ++k) dir[k] /= dn;
        mo[m][9 * n + 1] = 0.15;
        mo[m][9 * n + 5] = 0.30;
        for (int k = 0; k < 3; ++k) {
            mo[m][9 * n + 2 + k] = 0.20 * dir[k];
            mo[m][9 * n + 6 + k] = 0.55 * dir[k];
        }
        mo[m][54 + 2 * n] = 0.25;
        mo[m][54 + 2 * n + 1] = 0.15;
    }
}

/* benzene.rs:71-112, geometry :119-140, basis :152-245, MOs :262-470 */
orc_system* orc_benzene_new(double b_ee, double b_en, double cc_bond, double ch_bond) {
    orc_system* s = sys_alloc(ORC_SYS_GTO);
    s->n_nuc = 12;
    for (int i = 0; i < 6; ++i) {
        double ang = M_PI / 3.0 * (double)i;
        double rh = cc_bond + ch_bond;
        s->nuc[i][0] = cc_bond * cos(ang);
        s->nuc[i][1] = cc_bond * sin(ang);
        s->nuc[i][2] = 0.0;
        s->Z[i] = 6.0;
        s->nuc[6 + i][0] = rh * cos(ang);
        s->nuc[6 + i][1] = rh * sin(ang);
        s->nuc[6 + i][2] = 0.0;
        s->Z[6 + i] = 1.0;
    }
    s->n_elec = 42;
    for (int e = 0; e < 42; ++e) {
        s->spins[e] = e < 21 ? 1 : -1;                 /* benzene.rs:82-84 */
        s->init_nuc[e] = e < 36 ? e / 6 : 6 + (e - 36); /* benzene.rs:642-655 */
    }
    s->init_sigma = 0.5;
    for (int c = 0; c < 6; ++c) add_c631(s, c);
    double one = 1.0;
    for (int h = 0; h < 6; ++h) {
        add_cgto(s, 6 + h, 0, 0, 3, H631_IN_E, H631_IN_C);
        add_cgto(s, 6 + h, 0, 0, 1, &H631_OUT_E, &one);
    }
    /* MO coefficients, benzene.rs:262-470 */
    double mo[21][66];
    memset(mo, 0, sizeof(mo));
    int m = 0;
#define CO(n) (9 * (n))
#define HO(n) (54 + 2 * (n))
    for (int n = 0; n < 6; ++n, ++m) {
        mo[m][CO(n)] = 0.9943;
        mo[m][CO(n) + 1] = 0.0234;
        mo[m][CO(n) + 5] = 0.0020;
    }
    {
        double cs = 1.0 / sqrt(6.0);
        for (int n = 0; n < 6; ++n) { mo[m][CO(n) + 1] = 0.20 * cs; mo[m][CO(n) + 5] = 0.80 * cs; }
        ++m;
    }
    for (int n = 0; n < 6; ++n) {
        double ang = M_PI / 3.0 * (double)n;
        double w = cos(ang) / sqrt(3.0);
        mo[m][CO(n) + 1] = 0.20 * w;
        mo[m][CO(n) + 5] = 0.80 * w;
        mo[m][CO(n) + 2] = 0.50 * w;
        mo[m][CO(n) + 6] = 0.55 * w;
    }
    ++m;
    for (int n = 0; n < 6; ++n) {
        double ang = M_PI / 3.0 * (double)n;
        double w = sin(ang) / sqrt(3.0);
        mo[m][CO(n) + 1] = 0.20 * w;
        mo[m][CO(n) + 5] = 0.80 * w;
        mo[m][CO(n) + 3] = 0.50 * w;
        mo[m][CO(n) + 7] = 0.55 * w;
    }
    ++m;
    benzene_ch_mos(mo, s);
    m += 6;
    {
        double cpz = 1.0 / sqrt(6.0);
        for (int n = 0; n < 6; ++n) { mo[m][CO(n) + 4] = 0.30 * cpz; mo[m][CO(n) + 8] = 0.70 * cpz; }
        ++m;
    }
    for (int n = 0; n < 6; ++n) {
        double ph = cos(M_PI / 3.0 * (double)n) / sqrt(3.0);
        mo[m][CO(n) + 4] = 0.30 * ph;
        mo[m][CO(n) + 8] = 0.70 * ph;
    }
    ++m;
    for (int n = 0; n < 6; ++n) {
        double ph = sin(M_PI / 3.0 * (double)n) / sqrt(3.0);
        mo[m][CO(n) + 4] = 0.30 * ph;
        mo[m][CO(n) + 8] = 0.70 * ph;
    }
    ++m;
    {
        double nrm = 1.0 / sqrt(6.0);
        for (int n = 0; n < 6; ++n) {
            double ang = M_PI / 3.0 * (double)n;
            double tx = -sin(ang), ty = cos(ang);
            mo[m][CO(n) + 2] = 0.25 * tx * nrm;
            mo[m][CO(n) + 3] = 0.25 * ty * nrm;
            mo[m][CO(n) + 6] = 0.60 * tx * nrm;
            mo[m][CO(n) + 7] = 0.60 * ty * nrm;
        }
        ++m;
    }
    for (int pass = 0; pass < 2; ++pass, ++m) {
        for (int n = 0; n < 6; ++n) {
            double ang = M_PI / 3.0 * (double)n;
            double ph = (pass == 0 ? cos(2.0 * ang) : sin(2.0 * ang)) / sqrt(3.0);
            double tx = -sin(ang), ty = cos(ang);
            mo[m][CO(n) + 2] = 0.25 * tx * ph;
            mo[m][CO(n) + 3] = 0.25 * ty * ph;
            mo[m][CO(n) + 6] = 0.60 * tx * ph;
            mo[m][CO(n) + 7] = 0.60 * ty * ph;
        }
    }
#undef CO
#undef HO
    s->n_mo = 21;
    for (int sp = 0; sp < 2; ++sp)
        for (int q = 0; q < 21; ++q)
            for (int a = 0; a < 66; ++a) s->mo[sp][q][a] = mo[q][a];
    jastrow3_init(s, b_ee, b_en);
    s->mo_rule = ORC_MO_BENZENE_CH;
    return s;
}

/* ======================================================================
 * Nuclear geometry (ForceCalculator::set_nuclei, methane.rs:965-978,
 * benzene.rs:786-806, hydrogen.rs:221-225)
 * ==================================================================== */
void orc_system_set_nuclei(orc_system* s, const double* R) {
    for (int I = 0; I < s->n_nuc; ++I)
        for (int k = 0; k < 3; ++k) s->nuc[I][k] = R[3 * I + k];
    if (s->kind == ORC_SYS_H2MO) {
        for (int k = 0; k < 3; ++k) { s->orb1.center[k] = R[k]; s->orb2.center[k] = R[3 + k]; }
    } else if (s->kind == ORC_SYS_LITHIUM) {
        for (int o = 0; o < s->n_orb; ++o)
            for (int k = 0; k < 3; ++k) s->orbitals[o].center[k] = R[k];
    } else {
        for (int a = 0; a < s->n_ao; ++a)
            for (int k = 0; k < 3; ++k) s->ao[a].center[k] = R[3 * s->ao_nuc[a] + k];
        for (int I = 0; I < s->n_nuc; ++I)
            for (int k = 0; k < 3; ++k) s->j3.nuc[I][k] = R[3 * I + k];
        if (s->mo_rule == ORC_MO_BENZENE_CH) {
            double mo[21][66];
            for (int q = 0; q < 21; ++q)
                for (int a = 0; a < 66; ++a) mo[q][a] = s->mo[0][q][a];
            benzene_ch_mos(mo, s);
            for (int sp = 0; sp < 2; ++sp)
                for (int q = 9; q < 15; ++q)
                    for (int a = 0; a < 66; ++a) s->mo[sp][q][a] = mo[q][a];
        }
    }
}

/* methane.rs:908-924 / benzene.rs:601-628 (all-pairs form) */
double orc_nuclear_repulsion(const orc_system* s) {
    double v = 0.0;
    for (int I = 0; I < s->n_nuc; ++I)
        for (int J = I + 1; J < s->n_nuc; ++J) v += s->Z[I] * s->Z[J] / dist(s->nuc[I], s->nuc[J]);
    return v;
}

/* ======================================================================
 * GTO Slater-Jastrow (MethaneGTO generalised)
 * ==================================================================== */

/* methane.rs:781-787 */
static double mo_evaluate(const orc_system* s, int sp, int mo, const double r[3]) {
    double v = 0.0;
    for (int a = 0; a < s->n_ao; ++a) v += s->mo[sp][mo][a] * orc_cgto_eval(&s->ao[a], r);
    return v;
}

/* methane.rs:792-798 */
static void mo_derivative(const orc_system* s, int sp, int mo, const double r[3], double g[3]) {
    g[0] = g[1] = g[2] = 0.0;
    for (int a = 0; a < s->n_ao; ++a) {
        double ga[3];
        orc_cgto_grad(&s->ao[a], r, ga);
        for (int k = 0; k < 3; ++k) g[k] += s->mo[sp][mo][a] * ga[k];
    }
}

/* methane.rs:803-809 */
static double mo_laplacian(const orc_system* s, int sp, int mo, const double r[3]) {
    double v = 0.0;
    for (int a = 0; a < s->n_ao; ++a) v += s->mo[sp][mo][a] * orc_cgto_lap(&s->ao[a], r);
    return v;
}

static int sector_indices(const orc_system* s, int spin, int* idx) {
    int n = 0;
    for (int e = 0; e < s->n_elec; ++e)
        if (s->spins[e] == spin) idx[n++] = e;
    return n;
}

/* methane.rs:812-831: (det, inverse) of one spin sector; inv may be NULL */
static double slater_sector(const orc_system* s, const double (*r)[3], int spin, double* inv,
                            double* logabs, double* sgn) {
    int idx[ORC_MAXE];
    int n = sector_indices(s, spin, idx);
    int sp = spin == 1 ? 0 : 1;
    double a[ORC_MAXMO * ORC_MAXMO];
    for (int row = 0; row < n; ++row)
        for (int col = 0; col < n; ++col) a[row * n + col] = mo_evaluate(s, sp, col, r[idx[row]]);
    double det;
    orc_lu_det_inv(n, a, &det, inv, logabs, sgn);
    return det;
}

/* methane.rs:833-837 */
static double gto_slater_evaluate(const orc_system* s, const double (*r)[3]) {
    double du = slater_sector(s, r, 1, NULL, NULL, NULL);
    double dd = slater_sector(s, r, -1, NULL, NULL, NULL);
    return du * dd;
}

/* methane.rs:839-871 */
static void gto_slater_derivative(const orc_system* s, const double (*r)[3], double (*g)[3]) {
    double inv_u[ORC_MAXMO * ORC_MAXMO], inv_d[ORC_MAXMO * ORC_MAXMO];
    double du = slater_sector(s, r, 1, inv_u, NULL, NULL);
    double dd = slater_sector(s, r, -1, inv_d, NULL, NULL);
    for (int e = 0; e < s->n_elec; ++e) g[e][0] = g[e][1] = g[e][2] = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        int spin = pass == 0 ? 1 : -1, sp = pass;
        int idx[ORC_MAXE];
        int n = sector_indices(s, spin, idx);
        const double* inv = pass == 0 ? inv_u : inv_d;
        for (int row = 0; row < n; ++row) {
            double gl[3] = {0, 0, 0};
            for (int col = 0; col < n; ++col) {
                double gm[3];
                mo_derivative(s, sp, col, r[idx[row]], gm);
                for (int k = 0; k < 3; ++k) gl[k] += inv[col * n + row] * gm[k];
            }
            for (int k = 0; k < 3; ++k) g[idx[row]][k] = du * dd * gl[k];
        }
    }
}

/* methane.rs:873-905 */
static void gto_slater_laplacian(const orc_system* s, const double (*r)[3], double* l) {
    double inv_u[ORC_MAXMO * ORC_MAXMO], inv_d[ORC_MAXMO * ORC_MAXMO];
    double du = slater_sector(s, r, 1, inv_u, NULL, NULL);
    double dd = slater_sector(s, r, -1, inv_d, NULL, NULL);
    for (int e = 0; e < s->n_elec; ++e) l[e] = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        int spin = pass == 0 ? 1 : -1, sp = pass;
        int idx[ORC_MAXE];
        int n = sector_indices(s, spin, idx);
        const double* inv = pass == 0 ? inv_u : inv_d;
        for (int row = 0; row < n; ++row) {
            double ll = 0.0;
            for (int col = 0; col < n; ++col) ll += inv[col * n + row] * mo_laplacian(s, sp, col, r[idx[row]]);
            l[idx[row]] = du * dd * ll;
        }
    }
}

/* methane.rs:1007-1063 generalised: d ln|D| / dR_I */
static void gto_slater_nuclear_gradient(const orc_system* s, const double (*r)[3], int I, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    int any = 0;
    for (int a = 0; a < s->n_ao; ++a)
        if (s->ao_nuc[a] == I) any = 1;
    if (!any) return;
    for (int pass = 0; pass < 2; ++pass) {
        int spin = pass == 0 ? 1 : -1, sp = pass;
        double inv[ORC_MAXMO * ORC_MAXMO];
        slater_sector(s, r, spin, inv, NULL, NULL);
        int idx[ORC_MAXE];
        int n = sector_indices(s, spin, idx);
        for (int row = 0; row < n; ++row)
            for (int col = 0; col < n; ++col) {
                double dphi[3] = {0, 0, 0};
                for (int a = 0; a < s->n_ao; ++a) {
                    if (s->ao_nuc[a] != I) continue;
                    double ga[3];
                    orc_cgto_grad(&s->ao[a], r[idx[row]], ga);
                    for (int k = 0; k < 3; ++k) dphi[k] -= s->mo[sp][col][a] * ga[k];
                }
                for (int k = 0; k < 3; ++k) out[k] += inv[col * n + row] * dphi[k];
            }
    }
}

/* methane.rs:1071-1090 (+ series): du_en / dR_I */
static void gto_jastrow_nuclear_gradient(const orc_system* s, const double (*r)[3], int I, double out[3]) {
    const orc_jastrow3* j = &s->j3;
    double zi = s->Z[I], b = j->b_en;
    out[0] = out[1] = out[2] = 0.0;
    for (int e = 0; e < s->n_elec; ++e) {
        double d[3] = {r[e][0] - s->nuc[I][0], r[e][1] - s->nuc[I][1], r[e][2] - s->nuc[I][2]};
        double rin = norm3(d);
        if (rin < 1e-10) continue;
        double denom = 1.0 + b * rin;
        double f = zi / (rin * denom * denom);
        if (j->k_en > 0) {
            /* -d/dr of the series term, times the unit vector */
            double y = rin / denom, v = 0.0, yp = y;
            for (int k = 0; k < j->k_en; ++k) { v += j->d_en[j->species[I]][k] * (double)(k + 2) * yp; yp *= y; }
            f -= v / (rin * denom * denom);
        }
        for (int k = 0; k < 3; ++k) out[k] += f * d[k];
    }
}

/* ======================================================================
 * STOSlaterDet (slater.rs:197-270), N x N spin-blocked matrix
 * ==================================================================== */
static void stodet_matrix(const orc_system* s, const double (*r)[3], double* m) {
    int n = s->n_elec;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            m[i * n + j] = (s->spins[i] == s->spins[j]) ? orc_sto_eval(&s->orbitals[j], r[i]) : 0.0;
}

/* slater.rs:197-209 */
static double stodet_evaluate(const orc_system* s, const double (*r)[3]) {
    double m[ORC_MAXE * ORC_MAXE], det;
    stodet_matrix(s, r, m);
    orc_lu_det_inv(s->n_elec, m, &det, NULL, NULL, NULL);
    return det;
}

/* slater.rs:211-235 */
static void stodet_derivative(const orc_system* s, const double (*r)[3], double (*g)[3]) {
    int n = s->n_elec;
    double psi = stodet_evaluate(s, r);
    double m[ORC_MAXE * ORC_MAXE], inv[ORC_MAXE * ORC_MAXE], det;
    stodet_matrix(s, r, m);
    orc_lu_det_inv(n, m, &det, inv, NULL, NULL);
    for (int i = 0; i < n; ++i) {
        double gs[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j) {
            if (s->spins[i] != s->spins[j]) continue;
            double go[3];
            orc_sto_grad(&s->orbitals[j], r[i], go);
            for (int k = 0; k < 3; ++k) gs[k] += inv[j * n + i] * go[k];
        }
        for (int k = 0; k < 3; ++k) g[i][k] = psi * gs[k];
    }
}

/* slater.rs:237-261 */
static void stodet_laplacian(const orc_system* s, const double (*r)[3], double* l) {
    int n = s->n_elec;
    double psi = stodet_evaluate(s, r);
    double m[ORC_MAXE * ORC_MAXE], inv[ORC_MAXE * ORC_MAXE], det;
    stodet_matrix(s, r, m);
    orc_lu_det_inv(n, m, &det, inv, NULL, NULL);
    for (int i = 0; i < n; ++i) {
        double ls = 0.0;
        for (int j = 0; j < n; ++j) {
            if (s->spins[i] != s->spins[j]) continue;
            ls += inv[j * n + i] * orc_sto_lap(&s->orbitals[j], r[i]);
        }
        l[i] = psi * ls;
    }
}

/* ======================================================================
 * Trait dispatch
 * ==================================================================== */

/* MultiWfn::initialize with a Philox stream: electron e of walker w is
 * nuc[init_nuc[e]] + init_sigma * z, z from counter (w, 0, e, INIT).
 * Electron-to-centre rules follow methane.rs:1110-1119,
 * benzene.rs:642-655, slater.rs:198-204; H2 uses one electron per proton. */
void orc_initialize(const orc_system* s, uint64_t seed, uint32_t walker, double (*r)[3]) {
    for (int e = 0; e < s->n_elec; ++e) {
        double z[3];
        orc_normals3(seed, walker, 0u, (uint32_t)e, ORC_RNG_INIT, z);
        const double* c = s->nuc[s->init_nuc[e]];
        for (int k = 0; k < 3; ++k) r[e][k] = c[k] + s->init_sigma * z[k];
    }
}

double orc_evaluate(const orc_system* s, const double (*r)[3]) {
    switch (s->kind) {
    case ORC_SYS_H2MO: { /* hydrogen.rs:120-124 */
        double p1 = orc_slater1s_eval(&s->orb1, r[0]) + orc_slater1s_eval(&s->orb2, r[0]);
        double p2 = orc_slater1s_eval(&s->orb1, r[1]) + orc_slater1s_eval(&s->orb2, r[1]);
        return p1 * p2 * orc_j1_eval(&s->j1, r);
    }
    case ORC_SYS_LITHIUM: /* lithium.rs:24-26 */
        return stodet_evaluate(s, r) * orc_j2_eval(&s->j2, r);
    default: /* methane.rs:1122-1124 */
        return gto_slater_evaluate(s, r) * orc_j3_eval(&s->j3, r);
    }
}

void orc_derivative(const orc_system* s, const double (*r)[3], double (*g)[3]) {
    int n = s->n_elec;
    switch (s->kind) {
    case ORC_SYS_H2MO: { /* hydrogen.rs:126-139 */
        double ps1 = orc_slater1s_eval(&s->orb1, r[0]) + orc_slater1s_eval(&s->orb2, r[0]);
        double ps2 = orc_slater1s_eval(&s->orb1, r[1]) + orc_slater1s_eval(&s->orb2, r[1]);
        double ga[3], gb[3], gs1[3], gs2[3];
        orc_slater1s_grad(&s->orb1, r[0], ga);
        orc_slater1s_grad(&s->orb2, r[0], gb);
        for (int k = 0; k < 3; ++k) gs1[k] = ga[k] + gb[k];
        orc_slater1s_grad(&s->orb1, r[1], ga);
        orc_slater1s_grad(&s->orb2, r[1], gb);
        for (int k = 0; k < 3; ++k) gs2[k] = ga[k] + gb[k];
        double j = orc_j1_eval(&s->j1, r);
        double gj[2][3];
        orc_j1_deriv(&s->j1, r, gj);
        for (int k = 0; k < 3; ++k) {
            g[0][k] = gs1[k] * ps2 * j + ps1 * ps2 * gj[0][k];
            g[1][k] = gs2[k] * ps1 * j + ps1 * ps2 * gj[1][k];
        }
        return;
    }
    case ORC_SYS_LITHIUM: { /* lithium.rs:28-38 */
        double psi = stodet_evaluate(s, r);
        double j = orc_j2_eval(&s->j2, r);
        double gp[ORC_MAXE][3], gj[ORC_MAXE][3];
        stodet_derivative(s, r, gp);
        orc_j2_deriv(&s->j2, r, gj);
        for (int e = 0; e < n; ++e)
            for (int k = 0; k < 3; ++k) g[e][k] = j * gp[e][k] + psi * gj[e][k];
        return;
    }
    default: { /* methane.rs:1126-1136 */
        double sv = gto_slater_evaluate(s, r);
        double j = orc_j3_eval(&s->j3, r);
        double gs[ORC_MAXE][3], gj[ORC_MAXE][3];
        gto_slater_derivative(s, r, gs);
        orc_j3_deriv(&s->j3, r, gj);
        for (int e = 0; e < n; ++e)
            for (int k = 0; k < 3; ++k) g[e][k] = gs[e][k] * j + sv * gj[e][k];
        return;
    }
    }
}

void orc_laplacian(const orc_system* s, const double (*r)[3], double* l) {
    int n = s->n_elec;
    switch (s->kind) {
    case ORC_SYS_H2MO: { /* hydrogen.rs:141-161 */
        double ps1 = orc_slater1s_eval(&s->orb1, r[0]) + orc_slater1s_eval(&s->orb2, r[0]);
        double ps2 = orc_slater1s_eval(&s->orb1, r[1]) + orc_slater1s_eval(&s->orb2, r[1]);
        double ga[3], gb[3], gs1[3], gs2[3];
        orc_slater1s_grad(&s->orb1, r[0], ga);
        orc_slater1s_grad(&s->orb2, r[0], gb);
        for (int k = 0; k < 3; ++k) gs1[k] = ga[k] + gb[k];
        orc_slater1s_grad(&s->orb1, r[1], ga);
        orc_slater1s_grad(&s->orb2, r[1], gb);
        for (int k = 0; k < 3; ++k) gs2[k] = ga[k] + gb[k];
        double ls1 = orc_slater1s_lap(&s->orb1, r[0]) + orc_slater1s_lap(&s->orb2, r[0]);
        double ls2 = orc_slater1s_lap(&s->orb1, r[1]) + orc_slater1s_lap(&s->orb2, r[1]);
        double j = orc_j1_eval(&s->j1, r);
        double gj[2][3], lj[2];
        orc_j1_deriv(&s->j1, r, gj);
        orc_j1_lap(&s->j1, r, lj);
        l[0] = ls1 * ps2 * j + 2.0 * dot3(gs1, gj[0]) * ps2 + ps1 * ps2 * lj[0];
        l[1] = ls2 * ps1 * j + 2.0 * dot3(gs2, gj[1]) * ps1 + ps1 * ps2 * lj[1];
        return;
    }
    case ORC_SYS_LITHIUM: { /* lithium.rs:40-57 */
        double psi = stodet_evaluate(s, r);
        double j = orc_j2_eval(&s->j2, r);
        double lp[ORC_MAXE], lj[ORC_MAXE], gp[ORC_MAXE][3], gj[ORC_MAXE][3];
        stodet_laplacian(s, r, lp);
        orc_j2_lap(&s->j2, r, lj);
        stodet_derivative(s, r, gp);
        orc_j2_deriv(&s->j2, r, gj);
        for (int e = 0; e < n; ++e) l[e] = j * lp[e] + psi * lj[e] + 2.0 * dot3(gp[e], gj[e]);
        return;
    }
    default: { /* methane.rs:1138-1151 */
        double sv = gto_slater_evaluate(s, r);
        double j = orc_j3_eval(&s->j3, r);
        double ls[ORC_MAXE], lj[ORC_MAXE], gs[ORC_MAXE][3], gj[ORC_MAXE][3];
        gto_slater_laplacian(s, r, ls);
        orc_j3_lap(&s->j3, r, lj);
        gto_slater_derivative(s, r, gs);
        orc_j3_deriv(&s->j3, r, gj);
        for (int e = 0; e < n; ++e) l[e] = ls[e] * j + 2.0 * dot3(gs[e], gj[e]) + sv * lj[e];
        return;
    }
    }
}

double orc_local_energy(const orc_system* s, const double (*r)[3]) {
    int n = s->n_elec;
    double psi = orc_evaluate(s, r);
    double lap[ORC_MAXE];
    orc_laplacian(s, r, lap);
    double lsum = 0.0;
    for (int e = 0; e < n; ++e) lsum += lap[e];
    double kinetic = -0.5 * lsum / psi;
    switch (s->kind) {
    case ORC_SYS_H2MO: { /* hydrogen.rs:183-200 */
        double r12 = dist(r[0], r[1]);
        double pot = 1.0 / r12 - 1.0 / dist(r[0], s->orb1.center) - 1.0 / dist(r[1], s->orb2.center) -
                     1.0 / dist(r[0], s->orb2.center) - 1.0 / dist(r[1], s->orb1.center) +
                     1.0 / dist(s->orb1.center, s->orb2.center);
        return kinetic + pot;
    }
    case ORC_SYS_LITHIUM: { /* lithium.rs:60-82 */
        double ven = 0.0, vee = 0.0;
        for (int e = 0; e < n; ++e) ven += -(double)s->li_Z / norm3(r[e]);
        for (int a = 0; a < n; ++a)
            for (int b = a + 1; b < n; ++b) vee += 1.0 / dist(r[a], r[b]);
        return kinetic + ven + vee;
    }
    default: { /* methane.rs:1154-1182 */
        double ven = 0.0, vee = 0.0;
        for (int e = 0; e < n; ++e)
            for (int I = 0; I < s->n_nuc; ++I) ven += -s->Z[I] / dist(r[e], s->nuc[I]);
        for (int a = 0; a < n; ++a)
            for (int b = a + 1; b < n; ++b) vee += 1.0 / dist(r[a], r[b]);
        return kinetic + ven + vee + orc_nuclear_repulsion(s);
    }
    }
}

/* ln|evaluate| and its sign, accumulated in the log domain (LU pivots,
 * Jastrow exponent) so that it stays finite where the linear value would
 * underflow. */
void orc_log_psi(const orc_system* s, const double (*r)[3], double* logabs, double* sgn) {
    switch (s->kind) {
    case ORC_SYS_H2MO: {
        double p1 = orc_slater1s_eval(&s->orb1, r[0]) + orc_slater1s_eval(&s->orb2, r[0]);
        double p2 = orc_slater1s_eval(&s->orb1, r[1]) + orc_slater1s_eval(&s->orb2, r[1]);
        double r12 = dist(r[0], r[1]);
        double u = -s->j1.F / (2.0 * (1.0 + r12 / s->j1.F));
        *logabs = log(fabs(p1)) + log(fabs(p2)) + u;
        *sgn = (p1 * p2 < 0) ? -1.0 : 1.0;
        return;
    }
    case ORC_SYS_LITHIUM: {
        double m[ORC_MAXE * ORC_MAXE], la, sg;
        stodet_matrix(s, r, m);
        orc_lu_det_inv(s->n_elec, m, NULL, NULL, &la, &sg);
        double u = 0.0;
        for (int a = 0; a < s->n_elec; ++a)
            for (int b = a + 1; b < s->n_elec; ++b) u += -s->j2.F / (2.0 * (1.0 + dist(r[a], r[b]) / s->j2.F));
        *logabs = la + u;
        *sgn = sg;
        return;
    }
    default: {
        double lu_, su, ld, sd;
        slater_sector(s, r, 1, NULL, &lu_, &su);
        slater_sector(s, r, -1, NULL, &ld, &sd);
        *logabs = lu_ + ld + orc_j3_u(&s->j3, r);
        *sgn = su * sd;
        return;
    }
    }
}

int orc_num_params(const orc_system* s) { return s->kind == ORC_SYS_GTO ? orc_j3_num_params(&s->j3) : 0; }
void orc_get_params(const orc_system* s, double* p) { if (s->kind == ORC_SYS_GTO) orc_j3_get_params(&s->j3, p); }
void orc_set_params(orc_system* s, const double* p) { if (s->kind == ORC_SYS_GTO) orc_j3_set_params(&s->j3, p); }

/* methane.rs:1212-1219 */
void orc_log_derivatives(const orc_system* s, const double (*r)[3], double* O) {
    if (s->kind == ORC_SYS_GTO) orc_j3_log_derivs(&s->j3, r, O);
}

/* methane.rs:1241-1272 (identical in hydrogen.rs:227-258, benzene.rs:828-859) */
void orc_hf_force(const orc_system* s, const double (*r)[3], double (*F)[3]) {
    int M = s->n_nuc;
    for (int I = 0; I < M; ++I) F[I][0] = F[I][1] = F[I][2] = 0.0;
    for (int I = 0; I < M; ++I) {
        double zi = s->Z[I];
        for (int e = 0; e < s->n_elec; ++e) {
            double d[3] = {r[e][0] - s->nuc[I][0], r[e][1] - s->nuc[I][1], r[e][2] - s->nuc[I][2]};
            double dd = norm3(d);
            if (dd > 1e-10)
                for (int k = 0; k < 3; ++k) F[I][k] += zi * d[k] / (dd * dd * dd);
        }
    }
    for (int I = 0; I < M; ++I)
        for (int J = 0; J < M; ++J) {
            if (I == J) continue;
            double d[3] = {s->nuc[I][0] - s->nuc[J][0], s->nuc[I][1] - s->nuc[J][1], s->nuc[I][2] - s->nuc[J][2]};
            double dd = norm3(d);
            if (dd > 1e-10)
                for (int k = 0; k < 3; ++k) F[I][k] -= s->Z[I] * s->Z[J] * d[k] / (dd * dd * dd);
        }
}

void orc_wfn_nuclear_gradient(const orc_system* s, const double (*r)[3], double (*Om)[3]) {
    if (s->kind == ORC_SYS_GTO && s->mo_rule == ORC_MO_FIXED) { /* methane.rs:1096-1101 */
        for (int I = 0; I < s->n_nuc; ++I) {
            double a[3], b[3];
            gto_slater_nuclear_gradient(s, r, I, a);
            gto_jastrow_nuclear_gradient(s, r, I, b);
            for (int k = 0; k < 3; ++k) Om[I][k] = a[k] + b[k];
        }
        return;
    }
    /* hydrogen.rs:260-290: central differences of ln|psi|, h = 1e-5.  Also the
     * definition used for BenzeneGTO, whose MOs follow the geometry and for
     * which the reference has no nuclear gradient (the trait default panics). */
    const double h = 1e-5;
    orc_system* w = (orc_system*)malloc(sizeof(orc_system));
    double R0[3 * ORC_MAXN];
    for (int I = 0; I < s->n_nuc; ++I)
        for (int k = 0; k < 3; ++k) R0[3 * I + k] = s->nuc[I][k];
    for (int I = 0; I < s->n_nuc; ++I)
        for (int k = 0; k < 3; ++k) {
            double R[3 * ORC_MAXN];
            memcpy(w, s, sizeof(orc_system));
            memcpy(R, R0, sizeof(R0));
            R[3 * I + k] += h;
            orc_system_set_nuclei(w, R);
            double lf, sf, lb, sb;
            orc_log_psi(w, r, &lf, &sf); /* ln|psi|: benzene's psi underflows a double */
            memcpy(w, s, sizeof(orc_system));
            memcpy(R, R0, sizeof(R0));
            R[3 * I + k] -= h;
            orc_system_set_nuclei(w, R);
            orc_log_psi(w, r, &lb, &sb);
            Om[I][k] = (sf != 0.0 && sb != 0.0 && lf > -1e300 && lb > -1e300) ? (lf - lb) / (2.0 * h) : 0.0;
        }
    free(w);
}

void orc_batch_eval(const orc_system* s, int64_t W, const double* pos, double* logabs, double* sgn,
                    double* grad, double* lap, double* el) {
    int n = s->n_elec;
    for (int64_t w = 0; w < W; ++w) {
        const double(*r)[3] = (const double(*)[3])(pos + w * n * 3);
        double la, sg;
        orc_log_psi(s, r, &la, &sg);
        if (logabs) logabs[w] = la;
        if (sgn) sgn[w] = sg;
        if (grad || lap) {
            double psi = orc_evaluate(s, r);
            if (grad) {
                double g[ORC_MAXE][3];
                orc_derivative(s, r, g);
                for (int e = 0; e < n; ++e)
                    for (int k = 0; k < 3; ++k) grad[(w * n + e) * 3 + k] = g[e][k] / psi;
            }
            if (lap) {
                double l[ORC_MAXE];
                orc_laplacian(s, r, l);
                for (int e = 0; e < n; ++e) lap[w * n + e] = l[e] / psi;
            }
        }
        if (el) el[w] = orc_local_energy(s, r);
    }
}

/* Flat-array front end to orc_lithium_new for callers without the orc_sto
 * struct: orbital o has nprim[o] terms starting at offset sum(nprim[:o]). */
orc_system* orc_sto_system_new(int n, const int* nprim, const double* coef, const int* pw,
                               const double* zeta, const int* spins, double F, int Z, int exact_j2) {
    orc_sto orb[ORC_MAXE];
    int off = 0;
    for (int o = 0; o < n; ++o) {
        memset(&orb[o], 0, sizeof(orc_sto));
        orb[o].n = nprim[o];
        for (int k = 0; k < nprim[o]; ++k) {
            orb[o].coef[k] = coef[off + k];
            orb[o].pw[k] = pw[off + k];
            orb[o].zeta[k] = zeta[off + k];
        }
        off += nprim[o];
    }
    return orc_lithium_new(n, orb, spins, F, Z, exact_j2);
}

/* The Slater matrix the determinant is taken of: for GTO systems the
 * spin-sector matrix S[row][col] = phi_col(r_row) (methane.rs:812-826,
 * spin = +1/-1); for STOSlaterDet the full N x N spin-blocked matrix
 * (slater.rs:205-216, spin ignored).  Returns the matrix order. */
int orc_slater_matrix(const orc_system* s, const double (*r)[3], int spin, double* out) {
    if (s->kind == ORC_SYS_LITHIUM) {
        stodet_matrix(s, r, out);
        return s->n_elec;
    }
    if (s->kind == ORC_SYS_H2MO) {
        int e = spin == 1 ? 0 : 1;
        out[0] = orc_slater1s_eval(&s->orb1, r[e]) + orc_slater1s_eval(&s->orb2, r[e]);
        return 1;
    }
    int idx[ORC_MAXE];
    int n = sector_indices(s, spin, idx);
    int sp = spin == 1 ? 0 : 1;
    for (int row = 0; row < n; ++row)
        for (int col = 0; col < n; ++col) out[row * n + col] = mo_evaluate(s, sp, col, r[idx[row]]);
    return n;
}

/* Jastrow exponent u = ln J of the system's Jastrow factor */
double orc_jastrow_u(const orc_system* s, const double (*r)[3]) {
    if (s->kind == ORC_SYS_H2MO) return log(orc_j1_eval(&s->j1, r));
    if (s->kind == ORC_SYS_LITHIUM) return log(orc_j2_eval(&s->j2, r));
    return orc_j3_u(&s->j3, r);
}
