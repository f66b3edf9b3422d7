// qmc_sys.hpp -- the system description every kernel reads (DevSys) and
// the fixed capacities of the engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmc {

constexpr int MAXN = 16;      // nuclei
constexpr int MAXSH = 96;     // shells
constexpr int MAXAO = 128;    // atomic orbitals
constexpr int MAXPRIM = 512;  // primitive slots over all shells
constexpr int MAXMO = 32;     // molecular orbitals per spin
constexpr int MAXK = 8;       // Pade series terms
constexpr int MAXSPEC = 4;    // nuclear species of the e-n series
constexpr int MAXE = 64;      // electrons

// shell kinds
enum : int { SH_S = 0, SH_P = 1, SH_SP = 2, SH_P1 = 3, SH_STO = 4, SH_SLATER1S = 5 };
// Jastrow kinds
enum : int { JAS_NONE = 0, JAS_PAIR = 1, JAS_PADE = 2 };
// Philox streams (fourth counter word)
enum : uint32_t { RNG_INIT = 1u, RNG_SINGLE = 2u, RNG_ALL = 3u, RNG_SYS = 4u, RNG_DRIFT = 5u, RNG_BRANCH = 6u };

// flags (mirrors include/qmc.h)
enum : uint32_t {
  F_J2_GRAD_QUIRK = 1u << 0,
  F_J2_LAP_QUIRK = 1u << 1,
  F_HF_NN_REFSIGN = 1u << 2,
  F_SINGULAR_IDENTITY = 1u << 3,
};

// Everything a kernel needs to know about the system.  All fields are read
// with wave-uniform addresses, so they come through the scalar cache.
template <class T>
struct DevSys {
  int n_elec, n_up, n_dn, n_nuc, n_shell, n_ao, n_mo;
  int jas_kind, en_on, k_ee, k_en, n_species;
  uint32_t flags;
  T b_ee, b_en, a_anti, a_para, cst_ee, grad_scale;
  T c_anti[MAXK], c_para[MAXK];
  T d_en[MAXSPEC][MAXK];
  int species[MAXN];
  T nuc[MAXN][3];
  T Z[MAXN];
  double vnn;
  int sh_kind[MAXSH], sh_center[MAXSH], sh_nprim[MAXSH], sh_prim0[MAXSH], sh_ao0[MAXSH], sh_m[MAXSH];
  T prim_exp[MAXPRIM], coef_s[MAXPRIM], coef_p[MAXPRIM];
  T prim_nl2e[MAXPRIM];  // -prim_exp log2(e), rounded once: the FP32 primitives' 2^y argument factor
  int prim_pow[MAXPRIM];
  int sector_elec[2][MAXE];   // global index of the k-th electron of each spin sector
  int elec_spin[MAXE];        // 0 up, 1 down
  int elec_slot[MAXE];        // row of the electron in its sector
  int init_nuc[MAXE];
  T init_sigma;
  T moT[2][MAXAO][MAXMO];     // [spin][ao][mo], mo fastest
};

template <int N>
struct Cap {
  static constexpr int v = N > 0 ? N : 1;
};

}  // namespace qmc
