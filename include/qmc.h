/*
 * qmc.h -- C ABI of rust_qmc_amd, the MI355X (gfx950) walker-ensemble
 * engine for variational Monte Carlo of Slater-Jastrow wavefunctions.
 *
 * The ABI mirrors the rust_qmc trait surface (paths relative to the rust_qmc
 * crate root):
 *   MultiWfn          src/wavefunction/traits.rs:48-59
 *   OptimizableWfn    src/wavefunction/traits.rs:98-115
 *   EnergyCalculator  src/sampling/traits.rs:6-8
 *   ForceCalculator   src/sampling/traits.rs:20-56
 *   MCMCSimulation    src/sampling/vmc.rs:38-189
 * Each entry point below names the reference item it replaces.
 *
 * Conventions
 *   - Every function returns QMC_OK (0) or a negative QMC_ERR_* code;
 *     qmc_last_error() returns a human-readable message for the last
 *     failure on the calling thread.
 *   - Positions are double [W][N][3] in the reference's electron order
 *     (Vec<Vector3<f64>> per walker).  Buffers may be host or device
 *     memory; the library detects which and stages host memory itself.
 *   - `stream` is a hipStream_t or NULL for the default stream; calls are
 *     synchronous with respect to the host when they return results in
 *     host memory.
 *   - The reference returns linear psi, grad psi and lap psi.  This ABI
 *     returns ln|psi| with its sign, grad ln|psi| and lap(psi)/psi, which
 *     stay finite where the linear values under- or overflow (benzene:
 *     ln|psi| ~ -650).  evaluate = sign*exp(logabs),
 *     derivative = evaluate*grad, laplacian = evaluate*lap.
 *   - A system handle may be shared read-only between threads; calls that
 *     mutate it (set_params, set_nuclei, set_precision, set_flags) need
 *     external synchronisation.
 */
#ifndef RUST_QMC_AMD_QMC_H
#define RUST_QMC_AMD_QMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define QMC_OK 0
#define QMC_ERR_INVALID (-1)     /* bad argument (null pointer, size mismatch) */
#define QMC_ERR_UNSUPPORTED (-2) /* shape not compiled into the engine */
#define QMC_ERR_HIP (-3)         /* HIP runtime failure */
#define QMC_ERR_SINGULAR (-4)    /* a Slater matrix was exactly singular */
#define QMC_ERR_NOMEM (-5)

/* ---- precision modes ---- */
#define QMC_FP64 0  /* everything in FP64 */
#define QMC_MIXED 1 /* FP32 orbitals, Jastrow and Sherman-Morrison; FP64
                       inverse refresh, log-det, energy accumulators */

/* ---- behaviour flags (qmc_system_desc.flags, qmc_set_flags) ----
 * The reference's own quirks, reproduced when the flag is set. */
#define QMC_FLAG_J2_GRAD_QUIRK (1u << 0)     /* Jastrow2 gradient times F, jastrow.rs:102 */
#define QMC_FLAG_J2_LAP_QUIRK (1u << 1)      /* Jastrow2 Laplacian w/o cross terms, jastrow.rs:110-125 */
#define QMC_FLAG_HF_NN_REFSIGN (1u << 2)     /* HF force subtracts the nucleus-nucleus force, methane.rs:1259-1269 */
#define QMC_FLAG_SINGULAR_IDENTITY (1u << 3) /* singular Slater matrix -> identity inverse, methane.rs:829 */
#define QMC_FLAGS_REFERENCE \
  (QMC_FLAG_J2_GRAD_QUIRK | QMC_FLAG_J2_LAP_QUIRK | QMC_FLAG_HF_NN_REFSIGN | QMC_FLAG_SINGULAR_IDENTITY)

/* ---- MO coefficients that follow the geometry (qmc_set_nuclei) ----
 * QMC_MO_BENZENE_CH: MOs 9-14 of BenzeneGTO are the C-H sigma bonds, whose
 * C 2p part points from carbon n to hydrogen n (benzene.rs:345-381); they
 * are rebuilt whenever the nuclei move, as BenzeneGTO::set_nuclei does
 * (benzene.rs:786-809).  The nuclear gradient Omega includes the
 * coefficients' dependence on the geometry. */
#define QMC_MO_FIXED 0
#define QMC_MO_BENZENE_CH 1

/* ---- system description ---- */
#define QMC_AO_GTO 0      /* CGTO, methane.rs:488-591 (unnormalised primitives) */
#define QMC_AO_STO 1      /* STO radial sum c r^p exp(-zeta r), slater.rs:48-145 */
#define QMC_AO_SLATER1S 2 /* Slater1s exp(-alpha r), slater.rs:10-42 */

typedef struct {
  int32_t kind;   /* QMC_AO_* */
  int32_t center; /* nucleus index */
  int32_t l, m;   /* GTO angular part: l = 0 (s) or 1 (p, m = 0,1,2 -> x,y,z) */
  int32_t nprim;
  int32_t prim0; /* offset into the prim_* arrays */
} qmc_ao;

#define QMC_JAS_NONE 0
#define QMC_JAS_PAIR 1 /* Jastrow1 / Jastrow2: u = sum_{i<j} -F/(2(1+r_ij/F)), jastrow.rs:15-126 */
#define QMC_JAS_PADE 2 /* Jastrow3 (jastrow.rs:138-450) plus an optional Pade power series */

typedef struct {
  int32_t n_elec;
  const int32_t* spins; /* [n_elec], +1 or -1 */
  int32_t n_nuc;
  const double* nuc_pos;    /* [n_nuc][3] */
  const double* nuc_charge; /* [n_nuc] */
  int32_t n_ao;
  const qmc_ao* aos;        /* [n_ao] */
  const double* prim_exp;   /* GTO alpha or STO zeta */
  const double* prim_coef;  /* contraction coefficients */
  const int32_t* prim_pow;  /* STO powers (unused by GTO; may be NULL without STO) */
  int32_t n_mo;             /* rows of each MO matrix, >= electrons of each spin */
  const double* mo_up;      /* [n_mo][n_ao]; spin-up determinant uses rows 0..n_up-1 */
  const double* mo_dn;      /* [n_mo][n_ao]; may equal mo_up (restricted orbitals) */
  int32_t jastrow_kind;     /* QMC_JAS_* */
  double jas_F;             /* QMC_JAS_PAIR cusp parameter F */
  double b_ee, b_en, a_anti, a_para; /* QMC_JAS_PADE (Jastrow3 fields) */
  int32_t k_ee, k_en;       /* Pade series lengths; 0, 0 is the reference Jastrow3 */
  const double* c_anti;     /* [k_ee] coefficients of x^2..x^(k_ee+1), antiparallel pairs */
  const double* c_para;     /* [k_ee] parallel pairs */
  int32_t n_species;        /* nuclear species of the e-n series */
  const int32_t* species;   /* [n_nuc] species index of each nucleus (NULL if k_en = 0) */
  const double* d_en;       /* [n_species][k_en] coefficients of y^2..y^(k_en+1) */
  const int32_t* init_nuc;  /* [n_elec] nucleus each electron is initialised around */
  double init_sigma;        /* width of the initial Gaussian cloud */
  uint32_t flags;           /* QMC_FLAG_* */
  int32_t mo_rule;          /* QMC_MO_*: rows the rule owns are recomputed from nuc_pos */
} qmc_system_desc;

typedef struct qmc_system qmc_system;

/* Build a system from a description (all arrays are copied). */
int qmc_system_create(const qmc_system_desc* desc, int32_t precision, qmc_system** out);
int qmc_system_destroy(qmc_system* sys);

/* Constructors of the reference's systems.  flags select reference quirks. */
/* H2MoleculeMO {Slater1s(alpha, c1), Slater1s(alpha, c2), Jastrow1{F}}, hydrogen.rs:28-35 */
int qmc_h2_molecule_mo_create(double alpha, const double c1[3], const double c2[3], double F,
                              double init_sigma, int32_t precision, qmc_system** out);
/* Lithium{STOSlaterDet{orbitals, spins}, Jastrow2{F, n}}, lithium.rs:9-19.
 * Orbital o has nprim[o] terms (coef, pow, zeta) starting at sum(nprim[:o]);
 * all centred at the origin with nuclear charge Z (lithium.rs:60-63). */
int qmc_sto_slater_det_create(int32_t n, const int32_t* nprim, const double* coef, const int32_t* pow,
                              const double* zeta, const int32_t* spins, double F, double Z,
                              uint32_t flags, int32_t precision, qmc_system** out);
/* Li atom with seeded STO orbitals (the Li benchmark workload): electrons
 * [up, down, up] in orbitals [1s, 1s, 2s] as src/main.rs:36-49; each
 * orbital A sum_{nu=0..2} c_nu r^nu exp(-xi r) with xi = 2.7 (1s), 0.65
 * (2s), c_nu = 2U-1 from Philox counter (3o+nu, 1, 0, 4) and A normalising
 * the orbital; Jastrow2 with F = 1. */
int qmc_lithium_seeded_create(uint64_t seed, uint32_t flags, int32_t precision, qmc_system** out);
/* MethaneGTO::new(b_ee, b_en), methane.rs:623-685; basis 0 = the reference's
 * 17-AO 6-31G with its fixed MO coefficients, 1 = STO-3G (9 AOs) with
 * seeded orthonormal 5x9 MO coefficients; ch_bond = C-H distance (bohr). */
#define QMC_BASIS_631G 0
#define QMC_BASIS_STO3G 1
int qmc_methane_gto_create(double b_ee, double b_en, double ch_bond, int32_t basis, uint64_t mo_seed,
                           int32_t precision, qmc_system** out);
/* BenzeneGTO::new(b_ee, b_en), benzene.rs:71-112, with C-C / C-H lengths */
int qmc_benzene_gto_create(double b_ee, double b_en, double cc_bond, double ch_bond, int32_t precision,
                           qmc_system** out);

/* System descriptions as JSON text, replacing the YAML loaders of
 * io/config.rs:20-35 (which read H2 only).  The document is one object in
 * one of three forms:
 *  - the serde layout of H2MoleculeMO (hydrogen.rs:27-35):
 *      {"orbital1": {"alpha": a, "center": [x, y, z]}, "orbital2": {...},
 *       "jastrow": {"cusp_param": F}}
 *    (conf/h2.yaml's keys H1/H2 {alpha, R} and J {F} are read the same way);
 *  - a named constructor: {"system": "MethaneGTO", "b_ee", "b_en",
 *    "ch_bond", "basis": "6-31G" | "STO-3G", "mo_seed"}, {"system":
 *    "BenzeneGTO", "b_ee", "b_en", "cc_bond", "ch_bond"}, {"system":
 *    "LithiumSeeded", "seed"} (missing numbers take the constructors'
 *    reference values);
 *  - any molecule, the fields of qmc_system_desc: "nuclei" [[x,y,z]...],
 *    "charges", "spins" (+1/-1), "aos" [{"kind": "gto"|"sto"|"slater1s",
 *    "center", "l", "m", "exp": [...], "coef": [...], "pow": [...]}],
 *    "mo_up" [[C_ma]...] (rows = MOs), optional "mo_dn", "jastrow" {"kind":
 *    "none"|"pair"|"pade", "F", "b_ee", "b_en", "a_anti", "a_para",
 *    "c_anti", "c_para", "species", "d_en" [[...] per species]},
 *    "init_nuc", "init_sigma", "mo_rule": "fixed"|"benzene_ch" (the MO rows
 *    the rule owns are recomputed from "nuclei").
 * "flags" is an integer or a list of "j2_grad_quirk", "j2_lap_quirk",
 * "hf_nn_refsign", "singular_identity", "reference".  len = 0 means a
 * NUL-terminated text.  Errors name the field or the byte offset. */
int qmc_system_from_json(const char* text, int64_t len, int32_t precision, qmc_system** out);
/* The system in the generic form above (exact: %.17g numbers).  Call with
 * buf = NULL, len = 0 to get the size (with the NUL) in *needed. */
int qmc_system_to_json(const qmc_system* sys, char* buf, int64_t len, int64_t* needed);

int qmc_set_precision(qmc_system* sys, int32_t precision);
/* Engine selection.  The register engine keeps one walker per lane in
 * VGPRs (compiled electron counts up to 5 + 5); the wave engine keeps one
 * walker per wavefront in LDS (any count up to 32 per spin); the tile
 * engine moves 16 walkers per workgroup together with an MFMA MO
 * contraction (mixed-precision sweeps of 21 + 21-electron systems such as
 * benzene; state build, refresh and evaluation stay on the wave engine).
 * AUTO picks the register engine when it exists for the system, else the
 * tile engine for mixed-precision sweeps when it fits, else the wave
 * engine.  The register engine uses kernels specialised to the system's
 * shell structure when one is compiled (H2, both Li atoms, CH4 STO-3G and
 * 6-31G); GENERIC forces its kernels with run-time shell loops. */
#define QMC_ENGINE_AUTO 0
#define QMC_ENGINE_REGISTER 1
#define QMC_ENGINE_WAVE 2
#define QMC_ENGINE_GENERIC 3
#define QMC_ENGINE_TILE 4
int qmc_set_engine(qmc_system* sys, int32_t engine);
/* Name of the kernels a system's sweeps run on: "tile", "wave",
 * "register/generic" or "register/<shell structure>" (no reference
 * counterpart; diagnostics). */
int qmc_engine_name(const qmc_system* sys, char* buf, int32_t len);
int qmc_set_flags(qmc_system* sys, uint32_t flags);
int qmc_system_info(const qmc_system* sys, int32_t* n_elec, int32_t* n_up, int32_t* n_nuc, int32_t* n_ao);
/* MO matrix of one spin (0 up, 1 down) as [n_mo][n_ao] */
int qmc_get_mo(const qmc_system* sys, int32_t spin, double* mo);
/* MethaneGTO::nuclear_repulsion, methane.rs:908-924 */
int qmc_nuclear_repulsion(const qmc_system* sys, double* vnn);

/* OptimizableWfn (traits.rs:98-115); parameter order [b_ee, b_en, a_anti,
 * a_para] as methane.rs:1191-1198, then the Pade series coefficients */
int qmc_num_params(const qmc_system* sys, int32_t* n);
int qmc_get_params(const qmc_system* sys, double* p);
int qmc_set_params(qmc_system* sys, const double* p, int32_t n);
/* switch the Pade series on (k_ee, k_en > 0) or off; coefficients start at 0 */
int qmc_set_jastrow_series(qmc_system* sys, int32_t k_ee, int32_t k_en);

/* ForceCalculator geometry (sampling/traits.rs:20-35) */
int qmc_get_nuclei(const qmc_system* sys, double* R, double* Z);
int qmc_set_nuclei(qmc_system* sys, const double* R, int32_t n_nuc);

/* ---- batched per-configuration trait methods ---- */
/* MultiWfn::evaluate (traits.rs:53) as ln|psi| and sign */
int qmc_eval_log_psi(const qmc_system* sys, int64_t W, const double* pos, double* logabs, double* sign,
                     void* stream);
/* MultiWfn::derivative / psi (traits.rs:56); drift = 2 g (vmc.rs:255-261) */
int qmc_grad_log_psi(const qmc_system* sys, int64_t W, const double* pos, double* grad, void* stream);
/* MultiWfn::laplacian / psi (traits.rs:59), per electron [W][N] */
int qmc_lap_over_psi(const qmc_system* sys, int64_t W, const double* pos, double* lap, void* stream);
/* EnergyCalculator::local_energy (sampling/traits.rs:7) */
int qmc_local_energy(const qmc_system* sys, int64_t W, const double* pos, double* eloc, void* stream);
/* all of the above in one pass; any output may be NULL */
int qmc_evaluate(const qmc_system* sys, int64_t W, const double* pos, double* logabs, double* sign,
                 double* grad, double* lap, double* eloc, void* stream);

/* ---- per-configuration derivatives (any output may be NULL) ----
 * O   [W][P]    OptimizableWfn::log_derivatives (traits.rs:114; methane.rs:1212-1219)
 * dE  [W][P]    d E_L / d p_k (no reference counterpart; pinned by finite differences)
 * F   [W][M][3] ForceCalculator::hellmann_feynman_force (sampling/traits.rs:38; methane.rs:1241-1272)
 * Om  [W][M][3] ForceCalculator::wfn_nuclear_gradient (sampling/traits.rs:50; methane.rs:1096-1101)
 * P = qmc_num_params, M = number of nuclei. */
int qmc_derivatives(const qmc_system* sys, int64_t W, const double* pos, double* eloc, double* O, double* dE,
                    double* F, double* Om, void* stream);
int qmc_log_derivs(const qmc_system* sys, int64_t W, const double* pos, double* O, void* stream);
int qmc_dEL_dp(const qmc_system* sys, int64_t W, const double* pos, double* dE, void* stream);
int qmc_hf_force(const qmc_system* sys, int64_t W, const double* pos, double* F, void* stream);
int qmc_nuc_grad_log(const qmc_system* sys, int64_t W, const double* pos, double* Om, void* stream);

/* qmc_derivatives without any allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * pos and every non-NULL output must be device memory (host pointers are
 * rejected with QMC_ERR_INVALID; nothing is staged), and the call is
 * asynchronous on `stream`: the outputs are complete when the stream is.
 * The per-walker feature matrix lives in a fixed 64 MiB device array that
 * every GPU running the library holds until the process ends, so W is
 * limited to 8 Mi / Qp walkers, Qp = 2P + 6M + 2 rounded up to 16 (87,381
 * walkers at Qp = 96: CH4 with 32 Jastrow parameters, benzene with 4);
 * larger W gives QMC_ERR_INVALID.  Calls that overlap on one GPU (two
 * streams, two threads) share that array and race; serialise them.
 * Singular Slater matrices take the identity inverse, as in qmc_derivatives. */
int qmc_derivatives_noalloc(const qmc_system* sys, int64_t W, const double* pos, double* eloc, double* O,
                            double* dE, double* F, double* Om, void* stream);

/* (S + eps I) dp = -f by LU in FP64, falling back to dp = -f when the
 * matrix is singular (SROptimizer::compute_sr_update, sr_optimize.rs:219-242).
 * Returns QMC_OK, or QMC_ERR_SINGULAR when the fallback was taken. */
int qmc_sr_solve(int32_t n, const double* S, const double* f, double eps, double* dp);

/* ---- walker ensembles resident in HBM ---- */
typedef struct qmc_ensemble qmc_ensemble;

/* W walkers initialised from Philox counter (walker, 0, electron, 1):
 * electron e at nucleus init_nuc[e] + init_sigma * N(0,1)^3.  Replaces
 * MCMCSimulation::initialize (vmc.rs:57-66). */
int qmc_ensemble_create(const qmc_system* sys, int64_t W, uint64_t seed, qmc_ensemble** out);
int qmc_ensemble_destroy(qmc_ensemble* ens);
int qmc_ensemble_get_positions(const qmc_ensemble* ens, double* pos);
int qmc_ensemble_set_positions(qmc_ensemble* ens, const double* pos);
/* current ln|psi|, sign and cached E_L per walker */
int qmc_ensemble_get_state(const qmc_ensemble* ens, double* logabs, double* sign, double* eloc);
/* recompute inverses, ln|D|, ln J and E_L from the positions (FP64 refresh) */
int qmc_ensemble_refresh(qmc_ensemble* ens);

#define QMC_MOVE_SINGLE 0 /* single-electron Gaussian moves, Sherman-Morrison */
#define QMC_MOVE_ALL 1    /* all-electron moves, MCMCSimulation::metropolis_step (vmc.rs:95-119) */
#define QMC_MOVE_DRIFT 2  /* drift-diffusion single-electron moves, DriftDiffusionVMC::sweep
                             (vmc.rs:382-438); register engine */

typedef struct {
  int32_t move;             /* QMC_MOVE_* */
  int32_t n_steps;          /* sweeps to run */
  double step_size;         /* Gaussian proposal width (bohr), or the time step tau of QMC_MOVE_DRIFT */
  int32_t refresh_interval; /* FP64 inverse refresh every this many sweeps (0: never) */
  int32_t steps_per_launch; /* sweeps per persistent launch (0: library default) */
  double max_drift;         /* QMC_MOVE_DRIFT: cap of |tau F| per electron (vmc.rs:313-321) */
} qmc_sweep_params;

/* Run sweeps.  Step counters continue across calls.  Outputs (any may be
 * NULL, host or device): per-sweep ensemble sums of E_L and E_L^2 and the
 * number of accepted moves, each [n_steps]; eloc_trace [n_steps][W]. */
int qmc_ensemble_sweeps(qmc_ensemble* ens, const qmc_sweep_params* p, double* e_sum, double* e2_sum,
                        double* accepted, double* eloc_trace);

/* Kernel timing: when enabled, every sweep launch of qmc_ensemble_sweeps is
 * bracketed by HIP events on the launch stream; qmc_ensemble_timing
 * returns the summed sweep-kernel milliseconds and launch count of the
 * last qmc_ensemble_sweeps call. */
int qmc_ensemble_set_timing(qmc_ensemble* ens, int32_t on);
int qmc_ensemble_timing(const qmc_ensemble* ens, double* sweep_kernel_ms, int64_t* launches);

/* ---- SR-grad samples and force moments over an ensemble ----
 * qmc_ensemble_measure evaluates, for every walker of the ensemble, E_L,
 * O = [O_J (P), Omega (3M)], dE_L/dp (P) and the HF force (3M), and adds
 * their first and second moments (one FP64 matrix-core Gram product) to
 * the ensemble's accumulator; reset != 0 clears it first.
 * qmc_ensemble_moments derives from the accumulator: */
typedef struct {
  int64_t n_samples;
  int32_t P, M;      /* Jastrow parameters, nuclei; O has Q = P + 3M entries */
  double energy;     /* <E_L> */
  double variance;   /* <E_L^2> - <E_L>^2 */
  double* o_mean;    /* [Q]   <O>          (caller buffers; NULL to skip) */
  double* S;         /* [Q*Q] <O O^T> - <O><O>^T   (sr_optimize.rs:203-217) */
  double* f;         /* [Q]   <E_L O> - <E_L><O>    (sr_optimize.rs:224-234) */
  double* de_mean;   /* [P]   <dE_L/dp> */
  double* f_hf;      /* [M*3] <F_HF> with the nucleus-nucleus term (sign per QMC_FLAG_HF_NN_REFSIGN) */
  double* f_zv;      /* [M*3] <F_HF> - 2(<E_L Omega> - <E_L><Omega>)  (geometry_opt.rs:204-213) */
  double* f_var;     /* [M]   <|F_HF|^2> - |<F_HF>|^2  (geometry_opt.rs:220-224) */
} qmc_moments;
int qmc_ensemble_measure(qmc_ensemble* ens, int32_t reset);
int qmc_ensemble_moments(const qmc_ensemble* ens, qmc_moments* out);

/* SROptimizer (sr_optimize.rs:21-40) */
typedef struct {
  int64_t n_samples;       /* samples per iteration (walkers x measurement rounds) */
  int64_t n_walkers;
  int64_t n_equilibrate;   /* sweeps before the first iteration */
  int64_t max_iterations;
  double learning_rate;
  double sr_epsilon;
  double tolerance;
  double step_size;        /* Metropolis step (bohr) */
  int32_t decorrelation;   /* sweeps between measurements (reference: 5, sr_optimize.rs:145) */
  int32_t move;            /* QMC_MOVE_ALL as the reference, or QMC_MOVE_SINGLE */
  int32_t include_nuclear; /* 1: solve the (P+3M) system with Omega appended to O (the
                            * 47 x 47 S of the CH4 differentiation workload) and apply its
                            * first P entries: the parameter step with the parameter-
                            * geometry covariances of S taken into account.  The 3M
                            * nuclear entries are not applied (nuclei move only through
                            * qmc_set_nuclei / qmc_geometry_optimize).  0: the reference's
                            * P x P update (sr_optimize.rs:238). */
  int32_t refresh_interval;
} qmc_sr_params;
/* SRResult (sr_optimize.rs:351-366) */
typedef struct {
  double final_energy;
  double final_variance;
  int32_t iterations;
} qmc_sr_results;
/* SROptimizer::optimize (sr_optimize.rs:252-348); histories (host, may be
 * NULL) hold max_iterations entries, param_history (max_iterations+1)*P. */
int qmc_sr_optimize(qmc_system* sys, const qmc_sr_params* p, uint64_t seed, qmc_sr_results* res,
                    double* energy_history, double* variance_history, double* param_history);

/* GeometryOptimizer::sample_forces (geometry_opt.rs:105-227): forces [M*3]
 * (bare HF or zero-variance), mean energy, its error, force variance [M]. */
typedef struct {
  int64_t n_samples, n_walkers, n_equilibrate;
  double mc_step_size;
  int32_t zero_variance; /* ForceEstimator::ZeroVariance (force_variance.rs:27-39) */
  int32_t decorrelation;
  int32_t move;
  int32_t refresh_interval;
} qmc_force_params;
int qmc_sample_forces(const qmc_system* sys, const qmc_force_params* p, uint64_t seed, double* forces,
                      double* mean_energy, double* error, double* force_variance);

/* GeometryOptimizer::optimize (geometry_opt.rs:233-347): steepest descent
 * R_I += alpha (F_I - F_com) with F_com the charge-weighted mean force and
 * each displacement capped at max_displacement (0.02 bohr in the reference,
 * geometry_opt.rs:296-303), until max_I |F_I| < force_threshold or
 * max_iterations force samplings; then one more sampling for the final
 * energy.  Sampling k uses Philox key seed + k * 0x9E3779B97F4A7C15 (the
 * reference draws fresh walkers each call).  The system's nuclei are moved
 * (qmc_set_nuclei).  Histories (optional, host): energy and max |F| per
 * sampling [max_iterations], geometries [max_iterations + 1][M][3], of
 * which res->iterations (+1) entries are filled. */
typedef struct {
  double step_size;        /* alpha, bohr^2 / Ha */
  int64_t max_iterations;
  double force_threshold;  /* Ha / bohr */
  double max_displacement; /* bohr per nucleus and step */
} qmc_geomopt_params;
typedef struct {
  double final_energy;
  double final_max_force;
  int64_t iterations;  /* force samplings before the stop */
  int32_t converged;
} qmc_geomopt_results;
int qmc_geometry_optimize(qmc_system* sys, const qmc_force_params* fp, const qmc_geomopt_params* gp, uint64_t seed,
                          qmc_geomopt_results* res, double* energy_history, double* force_history,
                          double* geometry_history);

/* ---- heterogeneous batches: several ensembles in one persistent launch ----
 * No reference counterpart (the reference runs one system at a time); the
 * batch serves ensembles of different molecules sampled together.  In
 * FUSED mode one kernel launch runs every ensemble: workgroups take tiles
 * of walkers from a queue (most expensive first) until it is empty.  The
 * fused kernel is compiled for mixed-precision CH4/STO-3G and H2 register
 * ensembles and benzene-size tile ensembles, without Pade series; other
 * ensembles run in STREAMS (one launch per ensemble on concurrent streams)
 * or SEQUENTIAL mode.  Results per ensemble are those of
 * qmc_ensemble_sweeps with the same arguments. */
#define QMC_BATCH_FUSED 0
#define QMC_BATCH_STREAMS 1
#define QMC_BATCH_SEQUENTIAL 2
#define QMC_BATCH_ORDER_COST 0  /* queue holds the most expensive tiles first */
#define QMC_BATCH_ORDER_GIVEN 1 /* queue follows the ensemble order */
typedef struct {
  int32_t mode;             /* QMC_BATCH_* */
  int32_t blocks_per_cu;    /* FUSED: persistent workgroups per CU (0: as many as fit) */
  int32_t steps_per_launch; /* sweeps per launch */
  int32_t order;            /* QMC_BATCH_ORDER_* */
  int32_t refresh_interval; /* FP64 inverse refresh interval the tuner picked (0: none picked) */
} qmc_batch_config;
typedef struct qmc_batch qmc_batch;
/* ens: n distinct ensembles (kept by the caller, n <= 8) */
int qmc_batch_create(qmc_ensemble* const* ens, int32_t n, qmc_batch** out);
int qmc_batch_destroy(qmc_batch* b);
/* 1 when every ensemble can run in the fused kernel */
int qmc_batch_fusable(const qmc_batch* b, int32_t* fusable);
int qmc_batch_set_config(qmc_batch* b, const qmc_batch_config* cfg);
int qmc_batch_get_config(const qmc_batch* b, qmc_batch_config* cfg);
/* n_steps single-electron sweeps of every ensemble; step_size[n]; outputs
 * (host, may be NULL) [n][n_steps] per-sweep ensemble sums as
 * qmc_ensemble_sweeps; refresh_interval < 0 takes the configuration's tuned
 * interval (64 when none was tuned) */
int qmc_batch_sweeps(qmc_batch* b, int32_t n_steps, const double* step_size, int32_t refresh_interval, double* e_sum,
                     double* e2_sum, double* accepted);
/* kernel milliseconds (first launch start to last launch end, summed over
 * chunks) and launch count of the last qmc_batch_sweeps call */
int qmc_batch_timing(const qmc_batch* b, double* kernel_ms, int64_t* launches);
/* Autotuner: times every candidate configuration (modes, workgroups per
 * CU, sweeps per launch, queue order) on trial_steps sweeps of the
 * ensembles themselves (which advance), keeps the fastest and returns it
 * with its kernel milliseconds per sweep.  refresh_interval > 0 is the
 * FP64 inverse refresh interval of the trials; refresh_interval < 0 also
 * tunes it: of 16, 32, 64, 128 and 256 sweeps (at most -refresh_interval)
 * it keeps the longest after which, in every ensemble, the local energies
 * the sweeps computed from the Sherman-Morrison-updated inverses differ from
 * those of a fresh FP64 inverse by a mean below 0.01 mHa (two standard
 * errors included), and stores it in the configuration.  The tuned interval
 * is the one the trials then use. */
int qmc_batch_tune(qmc_batch* b, int32_t trial_steps, const double* step_size, int32_t refresh_interval,
                   qmc_batch_config* best, double* best_ms_per_sweep);

/* MCMCParams (vmc.rs:12-21) plus the options the reference hard-wires */
typedef struct {
  int64_t n_walkers;
  int64_t n_steps;
  double initial_step_size;
  double max_step_size;
  double min_step_size;
  double target_acceptance;
  int64_t adaptation_interval;
  int64_t n_burnin;         /* leading steps left out of the statistics (reference: 0) */
  int32_t move;             /* QMC_MOVE_ALL reproduces metropolis_step */
  int32_t refresh_interval;
} qmc_mcmc_params;

/* MCMCResults (vmc.rs:30-35) plus acceptance and the final step size */
typedef struct {
  double energy;
  double error;
  double autocorrelation_time;
  double acceptance;
  double final_step_size;
} qmc_vmc_results;

/* MCMCSimulation::run (vmc.rs:69-92) with compute_results (vmc.rs:131-188).
 * energy_trace (host) receives the per-step mean energies [n_steps]. */
int qmc_vmc_run(const qmc_system* sys, const qmc_mcmc_params* p, uint64_t seed, qmc_vmc_results* res,
                double* energy_trace);

/* DDVMCParams (vmc.rs:196-212) */
typedef struct {
  int64_t n_walkers;
  int64_t n_steps;  /* production sweeps */
  int64_t n_burnin; /* burn-in sweeps (discarded) */
  double time_step;
  double max_drift;
  double target_acceptance;
  int64_t adaptation_interval;
  int32_t refresh_interval;
} qmc_ddvmc_params;
/* DDVMCResults (vmc.rs:228-239) */
typedef struct {
  double energy;
  double error;
  double autocorrelation_time;
  double acceptance_rate;
  double final_time_step;
} qmc_ddvmc_results;
/* DriftDiffusionVMC::run (vmc.rs:441-499): burn-in with time-step
 * adaptation, production with windowed adaptation (vmc.rs:373-380), mean
 * E_L of every production sweep; energy_trace (host) receives [n_steps]. */
int qmc_ddvmc_run(const qmc_system* sys, const qmc_ddvmc_params* p, uint64_t seed, qmc_ddvmc_results* res,
                  double* energy_trace);

/* ISDMCParams (is_dmc.rs:49-84); print_interval has no counterpart (the
 * library prints nothing).  capacity bounds each population buffer (0: four
 * times n_walkers, rounded up to 256; at most 1,048,576): clones past it are
 * dropped and counted, where the reference's Vec grows without bound.
 * refresh_interval rebuilds the FP64 inverses of a mixed-precision population
 * every that many steps (0: never; FP64 systems ignore it).
 * The reference's moves r' = r + tau F + sqrt(2 tau) eta with F = 2 grad psi/psi
 * diffuse for an imaginary time 2 tau while its branching weight and E_ref
 * feedback use tau (is_dmc.rs:215-222, 286, 487-489), which biases the energy
 * (H2 at R = 1.4: -1.1646(17) against the exact -1.1745).  matched_branching = 1
 * uses 2 tau in the weight and the feedback; 0 reproduces the reference. */
typedef struct {
  int64_t n_walkers;
  int64_t n_steps;
  int64_t n_burnin;
  double time_step;
  double max_drift;
  double e_ref_damping;
  int64_t n_vmc_equilibrate;
  int64_t n_vmc_force_samples;
  int64_t capacity;
  int32_t refresh_interval;
  int32_t matched_branching;
} qmc_isdmc_params;
/* ISDMCResults (is_dmc.rs:90-108) plus the population bookkeeping */
typedef struct {
  double energy;
  double error;
  double acceptance_rate;
  double avg_population;
  int64_t collapses;  /* re-initialisations after the population died out (is_dmc.rs:424-433) */
  int64_t dropped;    /* clones dropped at the capacity */
  int64_t capacity;   /* walkers per population buffer */
  double step_ms;     /* device time of the burn-in and production steps */
} qmc_isdmc_results;
/* ImportanceSampledDMC::run (is_dmc.rs:376-546) on the register engine:
 * VMC forces from min(n_walkers, 50) all-electron walkers (step 0.5,
 * is_dmc.rs:305-370; Philox key seed ^ 0xD1B54A32D192ED03), then n_burnin +
 * n_steps steps of one drift-diffusion sweep (is_dmc.rs:198-265), branching
 * with at most 3 copies (is_dmc.rs:272-300) and the E_ref feedback.  Walkers
 * start from Philox counters (w, 0, e, INIT) under key seed, re-initialised
 * populations from walker ids collapses * n_walkers + w; the sweep of step t
 * uses counters (slot, t, ...) and the branching (slot, t, 0, 6).
 * forces_* are [M*3]; energy_trace and population_trace (host, optional)
 * receive the [n_steps] production series. */
int qmc_isdmc_run(const qmc_system* sys, const qmc_isdmc_params* p, uint64_t seed, qmc_isdmc_results* res,
                  double* forces_mixed, double* forces_vmc, double* forces_extrapolated, double* energy_trace,
                  double* population_trace);
/* ImportanceSampledDMC::block_error (is_dmc.rs:555-607) on a host series */
double qmc_isdmc_block_error(const double* e, int64_t n);

/* vmc.rs:141-163 and vmc.rs:166-188 on a host series */
double qmc_autocorrelation_time(const double* e, int64_t n);
double qmc_blocking_error(const double* e, int64_t n, double tau);

const char* qmc_last_error(void);
const char* qmc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* RUST_QMC_AMD_QMC_H */
