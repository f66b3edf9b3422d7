/*
 * oracle.h -- CPU FP64 restatement of rust_qmc's VMC walker-ensemble hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (rust_qmc_amd/libqmc_hip.so) never links,
 * loads or calls it.
 *
 * Every routine restates one reference function literally, in plain C99,
 * and cites the reference file:line it follows (paths relative to the
 * rust_qmc crate root).  The reference's own numerical quirks are kept
 * (Jastrow2 gradient factor F and missing Laplacian cross terms,
 * unnormalised Gaussian primitives, identity fallback for singular
 * matrices); an "exact" switch exists only where the GPU offers an exact
 * alternative.
 *
 * Things the reference does not define and that this restatement adds,
 * each pinned by a check that does not retype the formula:
 *   - Philox4x32-10 (Salmon et al., SC'11) replaces rand::thread_rng;
 *     pinned by the published Random123 known-answer vectors and an
 *     independent Python restatement.
 *   - LU with partial pivoting replaces nalgebra's determinant /
 *     try_inverse / lu().solve; pinned by numpy/LAPACK slogdet, inv, solve.
 *   - The 32-parameter Pade-series Jastrow extension and dE_L/dp; pinned by
 *     central finite differences, and by reduction to the reference's
 *     4-parameter Jastrow3 when the series coefficients are zero.
 *   - Single-electron Metropolis sweeps (the GPU's move scheme); pinned by
 *     detailed balance against the all-electron sampler (ensemble energies
 *     agree within 2 sigma) and by brute-force full recomputation.
 */
#ifndef QMC_ORACLE_H
#define QMC_ORACLE_H

#include <stdint.h>

#define ORC_MAXE 64      /* electrons */
#define ORC_MAXN 16      /* nuclei */
#define ORC_MAXAO 128    /* atomic orbitals */
#define ORC_MAXPRIM 8    /* primitives per orbital */
#define ORC_MAXMO 32     /* molecular orbitals per spin */
#define ORC_MAXK 8       /* Pade power-series terms */
#define ORC_MAXSPEC 4    /* nuclear species in the e-n series */
#define ORC_MAXP 64      /* optimisable parameters */

/* Philox streams (4th counter word); shared with include/qmc.h */
#define ORC_RNG_INIT 1u
#define ORC_RNG_SINGLE 2u
#define ORC_RNG_ALL 3u
#define ORC_RNG_SYS 4u
#define ORC_RNG_DRIFT 5u
#define ORC_RNG_BRANCH 6u

/* ---- slater.rs:10-42 ---- */
typedef struct { double alpha; double center[3]; } orc_slater1s;
/* ---- slater.rs:48-59 ---- */
typedef struct {
    int n;
    double coef[ORC_MAXPRIM];
    int pw[ORC_MAXPRIM];
    double zeta[ORC_MAXPRIM];
    double center[3];
} orc_sto;
/* ---- methane.rs:481-495 ---- */
typedef struct {
    double center[3];
    int nprim;
    double expo[ORC_MAXPRIM];
    double coef[ORC_MAXPRIM];
    int l, m;
} orc_cgto;

/* ---- jastrow.rs:15-18 ---- */
typedef struct { double F; } orc_jastrow1;
/* ---- jastrow.rs:60-65; exact=1 replaces the quirky gradient/Laplacian ---- */
typedef struct { double F; int n; int exact; } orc_jastrow2;
/* ---- jastrow.rs:138-155, plus the Pade-series extension ---- */
typedef struct {
    int n;
    double a_anti, a_para, b_ee, b_en;
    int n_nuc;
    double nuc[ORC_MAXN][3];
    double Z[ORC_MAXN];
    int spins[ORC_MAXE];
    /* extension: u_ee += sum_{k=2}^{k_ee+1} c_{s,k} x^k,  x = r/(1+b_ee r)
     *            u_en += sum_{k=2}^{k_en+1} d_{sp(I),k} y^k, y = r/(1+b_en r) */
    int k_ee;
    double c_anti[ORC_MAXK], c_para[ORC_MAXK];
    int k_en;
    int n_species;
    int species[ORC_MAXN];
    double d_en[ORC_MAXSPEC][ORC_MAXK];
} orc_jastrow3;

enum { ORC_SYS_H2MO = 1, ORC_SYS_LITHIUM = 2, ORC_SYS_GTO = 3 };

typedef struct {
    int kind;
    int n_elec;
    int spins[ORC_MAXE];
    int n_nuc;
    double nuc[ORC_MAXN][3];
    double Z[ORC_MAXN];
    /* initial-position rule: electron e ~ nuc[init_nuc[e]] + N(0, init_sigma) */
    int init_nuc[ORC_MAXE];
    double init_sigma;
    /* H2MoleculeMO (hydrogen.rs:28-35) */
    orc_slater1s orb1, orb2;
    orc_jastrow1 j1;
    /* Lithium = STOSlaterDet x Jastrow2 (lithium.rs:9-12, slater.rs:149-160) */
    int n_orb;
    orc_sto orbitals[ORC_MAXE];
    orc_jastrow2 j2;
    int li_Z;            /* nuclear charge at the origin (lithium.rs:60-63) */
    /* GTO Slater-Jastrow: MethaneGTO / BenzeneGTO generalised to any A, n */
    int n_ao;
    orc_cgto ao[ORC_MAXAO];
    int ao_nuc[ORC_MAXAO];
    int n_mo;
    double mo[2][ORC_MAXMO][ORC_MAXAO];  /* [spin: 0 up, 1 down][mo][ao] */
    orc_jastrow3 j3;
    /* MO coefficients that follow the geometry on set_nuclei:
     * ORC_MO_BENZENE_CH = the C-H sigma MOs of BenzeneGTO (benzene.rs:350, :797) */
    int mo_rule;
} orc_system;

enum { ORC_MO_FIXED = 0, ORC_MO_BENZENE_CH = 1 };

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 and normals */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u01(uint32_t w);
void orc_normals3(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, double z[3]);
void orc_move_draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, double z[3], double* u);

/* linear algebra */
int orc_lu_det_inv(int n, const double* a, double* det, double* inv, double* logabs, double* sgn);
int orc_lu_solve(int n, const double* a, const double* b, double* x);

/* orbitals */
double orc_slater1s_eval(const orc_slater1s* o, const double r[3]);
void orc_slater1s_grad(const orc_slater1s* o, const double r[3], double g[3]);
double orc_slater1s_lap(const orc_slater1s* o, const double r[3]);
double orc_sto_eval(const orc_sto* o, const double r[3]);
void orc_sto_grad(const orc_sto* o, const double r[3], double g[3]);
double orc_sto_lap(const orc_sto* o, const double r[3]);
double orc_cgto_eval(const orc_cgto* o, const double r[3]);
void orc_cgto_grad(const orc_cgto* o, const double r[3], double g[3]);
double orc_cgto_lap(const orc_cgto* o, const double r[3]);

/* Jastrow factors (values are J, derivatives are of J, as in the reference) */
double orc_j1_eval(const orc_jastrow1* j, const double (*r)[3]);
void orc_j1_deriv(const orc_jastrow1* j, const double (*r)[3], double (*g)[3]);
void orc_j1_lap(const orc_jastrow1* j, const double (*r)[3], double* l);
double orc_j2_eval(const orc_jastrow2* j, const double (*r)[3]);
void orc_j2_deriv(const orc_jastrow2* j, const double (*r)[3], double (*g)[3]);
void orc_j2_lap(const orc_jastrow2* j, const double (*r)[3], double* l);
double orc_j3_u(const orc_jastrow3* j, const double (*r)[3]);
double orc_j3_eval(const orc_jastrow3* j, const double (*r)[3]);
void orc_j3_deriv(const orc_jastrow3* j, const double (*r)[3], double (*g)[3]);
void orc_j3_lap(const orc_jastrow3* j, const double (*r)[3], double* l);
int orc_j3_num_params(const orc_jastrow3* j);
void orc_j3_get_params(const orc_jastrow3* j, double* p);
void orc_j3_set_params(orc_jastrow3* j, const double* p);
void orc_j3_log_derivs(const orc_jastrow3* j, const double (*r)[3], double* O);

/* system constructors (return heap objects; free with orc_system_free) */
orc_system* orc_h2mo_new(double alpha, double half_bond, double F, double init_sigma);
orc_system* orc_h2mo_new_general(const double c1[3], const double c2[3], double alpha, double F,
                                 double init_sigma);
orc_system* orc_lithium_new(int n, const orc_sto* orbitals, const int* spins, double F, int Z,
                            int exact_j2);
orc_system* orc_lithium_ref_new(int exact_j2);
orc_system* orc_lithium_config_new(uint64_t seed, int exact_j2);
orc_system* orc_sto_system_new(int n, const int* nprim, const double* coef, const int* pw,
                               const double* zeta, const int* spins, double F, int Z, int exact_j2);
orc_system* orc_methane_gto_new(double b_ee, double b_en, double ch_bond);
orc_system* orc_methane_sto3g_new(double b_ee, double b_en, double ch_bond, uint64_t mo_seed);
orc_system* orc_benzene_new(double b_ee, double b_en, double cc_bond, double ch_bond);
void orc_system_free(orc_system* s);
int orc_system_n_elec(const orc_system* s);
int orc_system_n_nuc(const orc_system* s);
void orc_system_get_nuclei(const orc_system* s, double* R, double* Z);
void orc_system_set_nuclei(orc_system* s, const double* R);
int orc_system_n_ao(const orc_system* s);
int orc_system_n_mo(const orc_system* s);
void orc_system_get_mo(const orc_system* s, int spin, double* mo);
int orc_num_params(const orc_system* s);
void orc_get_params(const orc_system* s, double* p);
void orc_set_params(orc_system* s, const double* p);
void orc_set_jastrow_series(orc_system* s, int k_ee, int k_en);
double orc_nuclear_repulsion(const orc_system* s);

/* MultiWfn / EnergyCalculator / OptimizableWfn / ForceCalculator */
void orc_initialize(const orc_system* s, uint64_t seed, uint32_t walker, double (*r)[3]);
double orc_evaluate(const orc_system* s, const double (*r)[3]);
void orc_derivative(const orc_system* s, const double (*r)[3], double (*g)[3]);
void orc_laplacian(const orc_system* s, const double (*r)[3], double* l);
double orc_local_energy(const orc_system* s, const double (*r)[3]);
void orc_log_psi(const orc_system* s, const double (*r)[3], double* logabs, double* sgn);
void orc_log_derivatives(const orc_system* s, const double (*r)[3], double* O);
void orc_hf_force(const orc_system* s, const double (*r)[3], double (*F)[3]);
void orc_wfn_nuclear_gradient(const orc_system* s, const double (*r)[3], double (*Om)[3]);
int orc_slater_matrix(const orc_system* s, const double (*r)[3], int spin, double* out);
double orc_jastrow_u(const orc_system* s, const double (*r)[3]);

/* batched helpers over W configurations laid out [W][N][3] */
void orc_batch_eval(const orc_system* s, int64_t W, const double* pos, double* logabs,
                    double* sgn, double* grad, double* lap, double* el);

/* samplers */
typedef struct {
    int64_t n_walkers, n_steps;
    double initial_step_size, max_step_size, min_step_size, target_acceptance;
    int64_t adaptation_interval;
} orc_mcmc_params;
typedef struct { double energy, error, autocorrelation_time, acceptance, final_step; } orc_mcmc_results;
double orc_autocorrelation_time(const double* e, int64_t n);
double orc_blocking_error(const double* e, int64_t n, double tau);
void orc_mcmc_run(const orc_system* s, const orc_mcmc_params* p, uint64_t seed, double* energies,
                  orc_mcmc_results* res);
void orc_mcmc_time_steps(const orc_system* s, int64_t W, int64_t warm, int64_t steps, double step,
                         uint64_t seed, double* seconds, double* mean_e);
void orc_mcmc_time_steps_mt(const orc_system* s, int64_t W, int64_t warm, int64_t steps, double step,
                            uint64_t seed, int threads, double* seconds, double* mean_e);
void orc_single_electron_steps(const orc_system* s, int64_t W, double* pos, double* logpsi,
                               uint64_t seed, uint32_t step0, int nsteps, double step,
                               double* eloc, int64_t* accepted);
/* flags: ORC_DD_NO_GREEN drops the Green's function ratio; ORC_DD_LOG takes
 * the acceptance ratio from ln|psi| without the reference's 1e-30 guard */
enum { ORC_DD_NO_GREEN = 1, ORC_DD_LOG = 2 };
void orc_dd_steps(const orc_system* s, int64_t W, double* pos, uint64_t seed, uint32_t step0, int nsteps,
                  double tau, double max_drift, int flags, double* eloc, int64_t* accepted);
void orc_all_electron_steps(const orc_system* s, int64_t W, double* pos, double* logpsi,
                            double* eloc, uint64_t seed, uint32_t step0, int nsteps, double step,
                            int64_t* accepted);

/* ImportanceSampledDMC (is_dmc.rs:176-607), orc_dmc.c */
typedef struct {
    int64_t n_walkers, n_steps, n_burnin;
    double time_step, max_drift, e_ref_damping;
    int64_t n_vmc_equilibrate, n_vmc_force_samples, capacity;
    int32_t matched_branching; /* 1: weights and E_ref feedback over 2 tau */
    int32_t log_domain;        /* 1: drift-diffusion acceptance from ln|psi| (ORC_DD_LOG) */
} orc_isdmc_params;
typedef struct {
    double energy, error, acceptance_rate, avg_population;
    int64_t collapses, dropped;
} orc_isdmc_results;
void orc_isdmc_branch(int64_t n, const double* el, const double* el_prev, double e_ref, double tau, uint64_t seed,
                      uint32_t step, int* copies);
double orc_isdmc_block_error(const double* e, int64_t n);
void orc_isdmc_run(const orc_system* s, const orc_isdmc_params* p, uint64_t seed, orc_isdmc_results* res,
                   double* f_mixed, double* f_vmc, double* f_extrap, double* e_trace, double* pop_trace);

/* SR (sr_optimize.rs:176-246) and ZV forces (geometry_opt.rs:198-226) */
int orc_sr_update(int64_t ns, int P, const double* E, const double* O, double eps, double* dp,
                  double* e_mean, double* variance);
void orc_zv_forces(int64_t ns, int M, const double* E, const double* F, const double* Om,
                   int use_zv, double* mean_force, double* mean_e, double* err, double* fvar);

/* seeded system data */
void orc_random_mo(uint64_t seed, int n_mo, int n_ao, double* mo);

#ifdef __cplusplus
}
#endif
#endif
