// qmc_host.cpp -- host side of rust_qmc_amd: the C ABI of include/qmc.h.
// Builds systems (the reference's constructors and a generic description),
// uploads them as DevSys parameter blocks, owns walker ensembles in HBM and
// drives the kernels of qmc_kernels.hip.
#include "qmc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "qmc_derivs.hpp"
#include "qmc_device.hpp"
#include "qmc_dmc.hpp"
#include "qmc_kernels.hpp"
#include "qmc_noalloc.hpp"
#include "qmc_refresh.hpp"
#include "qmc_batch.hpp"
#include "qmc_tile.hpp"
#include "qmc_wave.hpp"

using namespace qmc;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg);
}  // namespace

// error hook of the JSON reader (qmc_io.cpp)
__attribute__((visibility("hidden"))) int qmc_io_fail(int code, const char* msg) { return fail(code, msg); }

namespace {

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) return fail(QMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// ---------------------------------------------------------------------------
// Philox4x32-10 on the host (for seeded system data); same stream as the
// device and the published method (Salmon et al., SC'11).
// ---------------------------------------------------------------------------
void philox_host(const uint32_t ctr[4], uint64_t seed, uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
double u01(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

// ---------------------------------------------------------------------------
// Host copy of a system description
// ---------------------------------------------------------------------------
struct HostSys {
  std::vector<int32_t> spins;
  std::vector<double> nuc;  // [M][3]
  std::vector<double> Z;
  std::vector<qmc_ao> aos;
  std::vector<double> pexp, pcoef;
  std::vector<int32_t> ppow;
  int32_t n_mo = 0;
  std::vector<double> mo_up, mo_dn;  // [n_mo][n_ao]
  int32_t jas_kind = QMC_JAS_NONE;
  double F = 1.0, b_ee = 0, b_en = 0, a_anti = 0, a_para = 0;
  int32_t k_ee = 0, k_en = 0, n_species = 0;
  std::vector<double> c_anti, c_para, d_en;  // d_en [n_species][k_en]
  std::vector<int32_t> species;
  std::vector<int32_t> init_nuc;
  double init_sigma = 1.0;
  uint32_t flags = 0;
  int32_t mo_rule = QMC_MO_FIXED;  // MO coefficients that follow the geometry

  int n_elec() const { return (int)spins.size(); }
  int n_nuc() const { return (int)Z.size(); }
  int n_ao() const { return (int)aos.size(); }
  int n_up() const { return (int)std::count(spins.begin(), spins.end(), 1); }
  int n_dn() const { return n_elec() - n_up(); }
  int n_params() const { return jas_kind == QMC_JAS_PADE ? 4 + 2 * k_ee + n_species * k_en : 0; }
};

}  // namespace

struct qmc_system {
  HostSys h;
  int32_t precision = QMC_FP64;
  DevSys<double>* d64 = nullptr;  // device parameter blocks
  DevSys<float>* d32 = nullptr;
  const EngineOps* ops = nullptr;      // register-engine kernels in use
  const EngineOps* ops_spec = nullptr; // compiled shell structure matching the system, if any
  const EngineOps* ops_gen = nullptr;  // generic kernels of the electron-count shape
  const DerivOps* dops = nullptr;      // derivative kernels in use
  const DerivOps* dops_spec = nullptr; // specialised to the shell structure (may equal dops_gen)
  const DerivOps* dops_gen = nullptr;
  int32_t engine = QMC_ENGINE_AUTO;
  void pick_ops() {
    ops = (engine == QMC_ENGINE_GENERIC || !ops_spec) ? ops_gen : ops_spec;
    dops = (engine == QMC_ENGINE_GENERIC || !dops_spec) ? dops_gen : dops_spec;
  }
  bool dirty = true;
  uint64_t version = 0;  // bumped on every upload; ensembles refresh when it moves
  TilePlan tile;  // tile-engine plan (benzene-size sectors, mixed precision)
  // the wave engine (and, for mixed-precision sweeps, the tile engine) runs
  // when selected, or when no register engine was compiled for this
  // electron count
  bool wave() const {
    return engine == QMC_ENGINE_WAVE || engine == QMC_ENGINE_TILE || (engine == QMC_ENGINE_AUTO && !ops);
  }
  // single-electron sweeps on the tile engine (FP32 or FP64 state) unless
  // the wave engine is selected
  bool tile_sweeps(int32_t prec) const {
    (void)prec;
    return wave() && tile.ok && engine != QMC_ENGINE_WAVE;
  }
  // drift-diffusion sweeps of a plain ensemble on the tile engine, the same rule
  bool tile_dd(int32_t prec) const { return tile_sweeps(prec); }
  WaveShape shape() const { return WaveShape{h.n_elec(), h.n_up(), h.n_dn(), h.n_ao()}; }
  // the Jastrow has no Pade series terms (the kernels compiled without them apply)
  bool no_series() const { return h.k_ee == 0 && h.k_en == 0; }
  // dC/dR of MO coefficients that follow the geometry (device tables for
  // the wave derivative kernel; empty for fixed coefficients)
  MoDerivPair* mdp = nullptr;
  MoDerivEntry* mde = nullptr;
  int mdp0[3] = {0, 0, 0};
  ~qmc_system() {
    if (d64) (void)hipFree(d64);
    if (d32) (void)hipFree(d32);
    if (mdp) (void)hipFree(mdp);
    if (mde) (void)hipFree(mde);
  }
};

namespace {

int validate(const HostSys& h) {
  if (h.n_elec() < 1 || h.n_elec() > MAXE) return fail(QMC_ERR_INVALID, "n_elec out of range");
  for (int s : h.spins)
    if (s != 1 && s != -1) return fail(QMC_ERR_INVALID, "spins must be +1 or -1");
  if (h.n_nuc() < 1 || h.n_nuc() > MAXN) return fail(QMC_ERR_INVALID, "n_nuc out of range");
  if (h.n_ao() < 1 || h.n_ao() > MAXAO) return fail(QMC_ERR_INVALID, "n_ao out of range");
  if (h.n_mo < std::max(h.n_up(), h.n_dn()) || h.n_mo > MAXMO)
    return fail(QMC_ERR_INVALID, "n_mo must cover each spin sector and be <= 32");
  for (const auto& a : h.aos) {
    if (a.center < 0 || a.center >= h.n_nuc()) return fail(QMC_ERR_INVALID, "AO centre out of range");
    if (a.nprim < 1 || a.prim0 < 0 || a.prim0 + a.nprim > (int)h.pexp.size())
      return fail(QMC_ERR_INVALID, "AO primitive range out of bounds");
    if (a.kind == QMC_AO_GTO && (a.l < 0 || a.l > 1 || (a.l == 1 && (a.m < 0 || a.m > 2))))
      return fail(QMC_ERR_INVALID, "GTO angular momentum must be s or p");
    if (a.kind != QMC_AO_GTO && a.kind != QMC_AO_STO && a.kind != QMC_AO_SLATER1S)
      return fail(QMC_ERR_INVALID, "unknown AO kind");
  }
  if (h.k_ee < 0 || h.k_ee > MAXK || h.k_en < 0 || h.k_en > MAXK) return fail(QMC_ERR_INVALID, "series length > 8");
  if (h.k_en > 0 && (h.n_species < 1 || h.n_species > MAXSPEC)) return fail(QMC_ERR_INVALID, "bad n_species");
  if ((int)h.init_nuc.size() != h.n_elec()) return fail(QMC_ERR_INVALID, "init_nuc size");
  for (int c : h.init_nuc)
    if (c < 0 || c >= h.n_nuc()) return fail(QMC_ERR_INVALID, "init_nuc out of range");
  return QMC_OK;
}

bool same_prims(const HostSys& h, const qmc_ao& a, const qmc_ao& b, bool coefs) {
  if (a.nprim != b.nprim || a.center != b.center) return false;
  for (int k = 0; k < a.nprim; ++k) {
    if (h.pexp[a.prim0 + k] != h.pexp[b.prim0 + k]) return false;
    if (coefs && h.pcoef[a.prim0 + k] != h.pcoef[b.prim0 + k]) return false;
  }
  return true;
}

bool is_p(const qmc_ao& a, int m) { return a.kind == QMC_AO_GTO && a.l == 1 && a.m == m; }

// benzene.rs:345-381: MOs 9-14 are the C-H sigma bonds; the C 2p part of
// MO 9+n points from carbon n to hydrogen n (inner 0.20, outer 0.55 along
// the bond), so BenzeneGTO::set_nuclei (benzene.rs:786-809) rebuilds them.
// AO layout: carbon n owns AOs 9n..9n+8 (1s, 2s in, 2p in x/y/z, 2s out,
// 2p out x/y/z), hydrogen n owns 54+2n (inner) and 55+2n (outer).
static bool benzene_layout(const HostSys& h) {
  return h.n_nuc() == 12 && h.n_ao() == 66 && h.n_mo == 21 && h.n_up() == 21 && h.n_dn() == 21;
}

static void benzene_ch_dir(const HostSys& h, int n, double dir[3], double& len) {
  for (int c = 0; c < 3; ++c) dir[c] = h.nuc[3 * (6 + n) + c] - h.nuc[3 * n + c];
  len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  for (int c = 0; c < 3; ++c) dir[c] /= len;
}

static bool benzene_ch_ok(const HostSys& h) {
  if (!benzene_layout(h) || (int)h.mo_up.size() != 21 * 66 || (int)h.mo_dn.size() != 21 * 66) return false;
  for (int n = 0; n < 6; ++n)
    for (int c = 0; c < 3; ++c)
      if (!is_p(h.aos[9 * n + 2 + c], c) || !is_p(h.aos[9 * n + 6 + c], c) || h.aos[9 * n + 2 + c].center != n ||
          h.aos[9 * n + 6 + c].center != n)
        return false;
  return true;
}

int apply_mo_rule(HostSys& h) {
  if (h.mo_rule == QMC_MO_FIXED) return QMC_OK;
  if (h.mo_rule != QMC_MO_BENZENE_CH) return fail(QMC_ERR_INVALID, "unknown MO rule");
  if (!benzene_ch_ok(h)) return fail(QMC_ERR_INVALID, "the benzene C-H MO rule needs the BenzeneGTO AO/MO layout");
  for (int n = 0; n < 6; ++n) {
    double dir[3], len;
    benzene_ch_dir(h, n, dir, len);
    if (!(len > 0.0) || !std::isfinite(len)) return fail(QMC_ERR_INVALID, "a C-H bond of zero length");
  }
  const int NA = 66;
  for (std::vector<double>* mo : {&h.mo_up, &h.mo_dn})
    for (int n = 0; n < 6; ++n) {
      double* row = mo->data() + (size_t)(9 + n) * NA;
      double dir[3], len;
      benzene_ch_dir(h, n, dir, len);
      std::fill(row, row + NA, 0.0);
      row[9 * n + 1] = 0.15;
      row[9 * n + 5] = 0.30;
      for (int c = 0; c < 3; ++c) {
        row[9 * n + 2 + c] = 0.20 * dir[c];
        row[9 * n + 6 + c] = 0.55 * dir[c];
      }
      row[54 + 2 * n] = 0.25;
      row[54 + 2 * n + 1] = 0.15;
    }
  return QMC_OK;
}

template <class T>
int fill_devsys(const HostSys& h, DevSys<T>& D) {
  std::memset(&D, 0, sizeof(D));
  D.n_elec = h.n_elec();
  D.n_up = h.n_up();
  D.n_dn = h.n_dn();
  D.n_nuc = h.n_nuc();
  D.n_ao = h.n_ao();
  D.n_mo = h.n_mo;
  D.flags = h.flags;
  for (int I = 0; I < h.n_nuc(); ++I) {
    for (int c = 0; c < 3; ++c) D.nuc[I][c] = (T)h.nuc[3 * I + c];
    D.Z[I] = (T)h.Z[I];
  }
  double vnn = 0.0;
  for (int I = 0; I < h.n_nuc(); ++I)
    for (int J = I + 1; J < h.n_nuc(); ++J) {
      double dx = h.nuc[3 * I] - h.nuc[3 * J], dy = h.nuc[3 * I + 1] - h.nuc[3 * J + 1],
             dz = h.nuc[3 * I + 2] - h.nuc[3 * J + 2];
      vnn += h.Z[I] * h.Z[J] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  D.vnn = vnn;
  // Jastrow
  D.jas_kind = h.jas_kind;
  if (h.jas_kind == QMC_JAS_PAIR) {
    D.a_anti = D.a_para = (T)0.5;
    D.b_ee = (T)(1.0 / h.F);
    D.cst_ee = (T)(-0.5 * h.F);
    D.en_on = 0;
    D.grad_scale = (T)((h.flags & QMC_FLAG_J2_GRAD_QUIRK) ? h.F : 1.0);
  } else if (h.jas_kind == QMC_JAS_PADE) {
    D.a_anti = (T)h.a_anti;
    D.a_para = (T)h.a_para;
    D.b_ee = (T)h.b_ee;
    D.b_en = (T)h.b_en;
    D.cst_ee = 0;
    D.en_on = 1;
    D.grad_scale = 1;
    D.k_ee = h.k_ee;
    D.k_en = h.k_en;
    D.n_species = h.n_species;
    for (int k = 0; k < h.k_ee; ++k) {
      D.c_anti[k] = (T)h.c_anti[k];
      D.c_para[k] = (T)h.c_para[k];
    }
    for (int s = 0; s < h.n_species; ++s)
      for (int k = 0; k < h.k_en; ++k) D.d_en[s][k] = (T)h.d_en[s * h.k_en + k];
    for (int I = 0; I < h.n_nuc(); ++I) D.species[I] = h.species.empty() ? 0 : h.species[I];
  } else {
    D.grad_scale = 1;
  }
  // the quirky Laplacian only exists for the pair Jastrow
  if (h.jas_kind != QMC_JAS_PAIR) D.flags &= ~(uint32_t)QMC_FLAG_J2_LAP_QUIRK;
  // shells: consecutive AOs sharing a centre and primitives are evaluated once
  int nsh = 0, np = 0;
  const int A = h.n_ao();
  for (int i = 0; i < A;) {
    if (nsh >= MAXSH) return fail(QMC_ERR_UNSUPPORTED, "too many shells");
    const qmc_ao& a = h.aos[i];
    int kind, nao, m = 0;
    const qmc_ao* pa = nullptr;
    if (a.kind != QMC_AO_GTO) {
      kind = a.kind == QMC_AO_SLATER1S ? SH_SLATER1S : SH_STO;
      nao = 1;
    } else if (a.l == 0 && i + 3 < A && is_p(h.aos[i + 1], 0) && is_p(h.aos[i + 2], 1) && is_p(h.aos[i + 3], 2) &&
               same_prims(h, a, h.aos[i + 1], false) && same_prims(h, h.aos[i + 1], h.aos[i + 2], true) &&
               same_prims(h, h.aos[i + 1], h.aos[i + 3], true)) {
      kind = SH_SP;
      nao = 4;
      pa = &h.aos[i + 1];
    } else if (is_p(a, 0) && i + 2 < A && is_p(h.aos[i + 1], 1) && is_p(h.aos[i + 2], 2) &&
               same_prims(h, a, h.aos[i + 1], true) && same_prims(h, a, h.aos[i + 2], true)) {
      kind = SH_P;
      nao = 3;
      pa = &a;
    } else if (a.l == 0) {
      kind = SH_S;
      nao = 1;
    } else {
      kind = SH_P1;
      nao = 1;
      m = a.m;
      pa = &a;
    }
    if (np + a.nprim > MAXPRIM) return fail(QMC_ERR_UNSUPPORTED, "too many primitives");
    D.sh_kind[nsh] = kind;
    D.sh_center[nsh] = a.center;
    D.sh_nprim[nsh] = a.nprim;
    D.sh_prim0[nsh] = np;
    D.sh_ao0[nsh] = i;
    D.sh_m[nsh] = m;
    for (int k = 0; k < a.nprim; ++k) {
      D.prim_exp[np + k] = (T)h.pexp[a.prim0 + k];
      D.prim_nl2e[np + k] = (T)(-h.pexp[a.prim0 + k] * 1.4426950408889634);  // log2(e)
      const bool s_part = (kind == SH_S || kind == SH_SP || kind == SH_STO || kind == SH_SLATER1S);
      D.coef_s[np + k] = s_part ? (T)h.pcoef[a.prim0 + k] : (T)0;
      D.coef_p[np + k] = pa ? (T)h.pcoef[pa->prim0 + k] : (T)0;
      D.prim_pow[np + k] = (a.kind == QMC_AO_GTO || h.ppow.empty()) ? 0 : h.ppow[a.prim0 + k];
    }
    np += a.nprim;
    ++nsh;
    i += nao;
  }
  D.n_shell = nsh;
  for (int mo = 0; mo < h.n_mo; ++mo)
    for (int ao = 0; ao < A; ++ao) {
      D.moT[0][ao][mo] = (T)h.mo_up[mo * A + ao];
      D.moT[1][ao][mo] = (T)h.mo_dn[mo * A + ao];
    }
  int ku = 0, kd = 0;
  for (int e = 0; e < h.n_elec(); ++e) {
    if (h.spins[e] == 1) {
      D.elec_spin[e] = 0;
      D.elec_slot[e] = ku;
      D.sector_elec[0][ku++] = e;
    } else {
      D.elec_spin[e] = 1;
      D.elec_slot[e] = kd;
      D.sector_elec[1][kd++] = e;
    }
    D.init_nuc[e] = h.init_nuc[e];
  }
  D.init_sigma = (T)h.init_sigma;
  return QMC_OK;
}

// dC[spin][m][ao] / dR_{I,c} of the MO rows a rule owns, as the pair and
// entry tables of the wave derivative kernel.  Benzene's C-H sigma MO n
// (benzene.rs:350-381) has C 2p coefficients 0.20 u and 0.55 u with
// u = (R_H - R_C) / |R_H - R_C|, so d u_c / d R_H,c' = (delta_cc' - u_c u_c')
// / |R_H - R_C| and d u / d R_C is its negative.
int mo_deriv_tables(const HostSys& h, const DevSys<double>& D, std::vector<MoDerivPair>& pairs,
                    std::vector<MoDerivEntry>& ents, int p0[3]) {
  pairs.clear();
  ents.clear();
  p0[0] = p0[1] = p0[2] = 0;
  if (h.mo_rule != QMC_MO_BENZENE_CH) return QMC_OK;
  auto locate = [&](int ao, int& sh, int& q) {
    for (sh = 0; sh < D.n_shell; ++sh) {
      const int k = D.sh_kind[sh], nao = k == SH_SP ? 4 : (k == SH_P ? 3 : 1);
      if (ao >= D.sh_ao0[sh] && ao < D.sh_ao0[sh] + nao) {
        q = ao - D.sh_ao0[sh];
        return true;
      }
    }
    return false;
  };
  for (int spin = 0; spin < 2; ++spin) {
    p0[spin] = (int)pairs.size();
    for (int n = 0; n < 6; ++n) {
      double u[3], len;
      benzene_ch_dir(h, n, u, len);
      for (int part = 0; part < 2; ++part) {
        const double coef = part ? 0.55 : 0.20;
        for (int c = 0; c < 3; ++c) {
          MoDerivPair P{spin, 9 + n, 0, 0, (int)ents.size(), 0};
          if (!locate(9 * n + (part ? 6 : 2) + c, P.sh, P.q))
            return fail(QMC_ERR_INVALID, "the benzene C-H MO rule found no shell for a 2p AO");
          for (int c2 = 0; c2 < 3; ++c2) {
            const double d = coef * ((c == c2 ? 1.0 : 0.0) - u[c] * u[c2]) / len;
            ents.push_back({6 + n, c2, d});
            ents.push_back({n, c2, -d});
          }
          P.e1 = (int)ents.size();
          pairs.push_back(P);
        }
      }
    }
  }
  p0[2] = (int)pairs.size();
  return QMC_OK;
}

int upload(qmc_system* s) {
  if (!s->dirty) return QMC_OK;
  int rc = validate(s->h);
  if (rc) return rc;
  auto d64 = std::make_unique<DevSys<double>>();
  auto d32 = std::make_unique<DevSys<float>>();
  if ((rc = fill_devsys(s->h, *d64))) return rc;
  if ((rc = fill_devsys(s->h, *d32))) return rc;
  if (!s->d64) HIP_TRY(hipMalloc(&s->d64, sizeof(DevSys<double>)));
  if (!s->d32) HIP_TRY(hipMalloc(&s->d32, sizeof(DevSys<float>)));
  HIP_TRY(hipMemcpy(s->d64, d64.get(), sizeof(DevSys<double>), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d32, d32.get(), sizeof(DevSys<float>), hipMemcpyHostToDevice));
  {
    std::vector<MoDerivPair> pairs;
    std::vector<MoDerivEntry> ents;
    if ((rc = mo_deriv_tables(s->h, *d64, pairs, ents, s->mdp0))) return rc;
    if (s->mdp) HIP_TRY(hipFree(s->mdp));
    if (s->mde) HIP_TRY(hipFree(s->mde));
    s->mdp = nullptr;
    s->mde = nullptr;
    if (!pairs.empty()) {
      HIP_TRY(hipMalloc(&s->mdp, pairs.size() * sizeof(MoDerivPair)));
      HIP_TRY(hipMalloc(&s->mde, ents.size() * sizeof(MoDerivEntry)));
      HIP_TRY(hipMemcpy(s->mdp, pairs.data(), pairs.size() * sizeof(MoDerivPair), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(s->mde, ents.data(), ents.size() * sizeof(MoDerivEntry), hipMemcpyHostToDevice));
    }
  }
  s->ops_spec = find_engine(*d64, false);
  if (s->ops_spec && !s->ops_spec->match) s->ops_spec = nullptr;
  s->ops_gen = find_engine(*d64, true);
  s->dops_spec = find_derivs(*d64, false);
  s->dops_gen = find_derivs(*d64, true);
  s->pick_ops();
  tile_plan(*d32, s->tile);
  s->dirty = false;
  s->version++;
  return QMC_OK;
}

// feature layout of the derivative matrix Y (mirrors DerivLayout in qmc_derivs.hip)
struct Cols {
  int P, M, Qp;
  explicit Cols(const HostSys& h) : P(h.n_params()), M(h.n_nuc()) { Qp = (2 * P + 6 * M + 2 + 15) / 16 * 16; }
  int e() const { return P + 3 * M; }
  int one() const { return P + 3 * M + 1; }
  int de(int k) const { return P + 3 * M + 2 + k; }
  int f(int I, int c) const { return 2 * P + 3 * M + 2 + 3 * I + c; }
};

// nucleus-nucleus part of the Hellmann-Feynman force: the physical
// repulsion is +Z_I Z_J (R_I - R_J)/|R_I - R_J|^3; the reference subtracts
// it (methane.rs:1259-1269), reproduced under QMC_FLAG_HF_NN_REFSIGN.
// F must hold 3 n_nuc doubles; written in place (no allocation)
void hf_nn_fill(const HostSys& h, double* F) {
  const int M = h.n_nuc();
  for (int q = 0; q < 3 * M; ++q) F[q] = 0.0;
  const double sg = (h.flags & QMC_FLAG_HF_NN_REFSIGN) ? -1.0 : 1.0;
  for (int I = 0; I < M; ++I)
    for (int J = 0; J < M; ++J) {
      if (I == J) continue;
      double d[3], r2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        d[c] = h.nuc[3 * I + c] - h.nuc[3 * J + c];
        r2 += d[c] * d[c];
      }
      const double r = std::sqrt(r2);
      if (r > 1e-10)
        for (int c = 0; c < 3; ++c) F[3 * I + c] += sg * h.Z[I] * h.Z[J] * d[c] / (r2 * r);
    }
}
std::vector<double> hf_nn(const HostSys& h) {
  std::vector<double> F(3 * h.n_nuc(), 0.0);
  hf_nn_fill(h, F.data());
  return F;
}

// copy a host vector to a caller buffer that may be host or device memory
int put(double* dst, const double* src, size_t n);

// LU with partial pivoting on a dense FP64 system (host)
bool lu_solve(int n, std::vector<double> a, std::vector<double> b, std::vector<double>& x) {
  std::vector<int> piv(n);
  for (int i = 0; i < n; ++i) piv[i] = i;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(a[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > best) {
        best = std::fabs(a[i * n + k]);
        p = i;
      }
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = a[i * n + k] / a[k * n + k];
      a[i * n + k] = f;
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      b[i] -= f * b[k];
    }
  }
  x.assign(n, 0.0);
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
    x[i] = s / a[i * n + i];
  }
  return true;
}

int need_engine(const qmc_system* s) {
  if (s->wave()) {
    if (std::max(s->h.n_up(), s->h.n_dn()) > MAXMO)
      return fail(QMC_ERR_UNSUPPORTED, "more than 32 electrons of one spin");
    return QMC_OK;
  }
  if (!s->ops)
    return fail(QMC_ERR_UNSUPPORTED, "no register engine compiled for " + std::to_string(s->h.n_up()) + " up / " +
                                         std::to_string(s->h.n_dn()) + " down electrons");
  return QMC_OK;
}

// The derivative matrix Y of A.W walkers in the arithmetic type prec: the
// wave derivative kernel when the wave or tile engine runs the system (or
// its MO coefficients follow the geometry), else the register kernels.
int launch_derivs(const qmc_system* s, DerivArgs A, int32_t prec, hipStream_t st) {
  if (s->wave() || !s->dops || s->h.mo_rule != QMC_MO_FIXED) {
    if (int rc = need_engine(s)) return rc;
    if (s->h.n_elec() > 64) return fail(QMC_ERR_UNSUPPORTED, "the wave derivative kernel takes up to 64 electrons");
    A.mdp = s->mdp;
    A.mde = s->mde;
    for (int k = 0; k < 3; ++k) A.mdp0[k] = s->mdp0[k];
    if (prec == QMC_FP64)
      HIP_TRY(wave_derivs<double>(s->d64, s->d64, A, s->shape(), st));
    else
      HIP_TRY(wave_derivs<float>(s->d64, s->d32, A, s->shape(), st));
    return QMC_OK;
  }
  if (prec == QMC_FP64)
    HIP_TRY(s->dops->derivs64(s->d64, s->d64, A, st));
  else
    HIP_TRY(s->dops->derivs32(s->d64, s->d32, A, st));
  return QMC_OK;
}

int make_system(HostSys&& h, int32_t precision, qmc_system** out) {
  if (!out) return fail(QMC_ERR_INVALID, "out is null");
  if (precision != QMC_FP64 && precision != QMC_MIXED) return fail(QMC_ERR_INVALID, "bad precision");
  auto s = std::make_unique<qmc_system>();
  s->h = std::move(h);
  s->precision = precision;
  int rc = upload(s.get());
  if (rc) return rc;
  *out = s.release();
  return QMC_OK;
}

// --- basis data -------------------------------------------------------------
// 6-31G as carried by the reference (methane.rs:701-770, benzene.rs:161-226)
const double C631_1S_E[6] = {3047.5249, 457.36951, 103.94869, 29.210155, 9.2866630, 3.1639270};
const double C631_1S_C[6] = {0.0018347, 0.0140373, 0.0688426, 0.2321844, 0.4679413, 0.3623120};
const double C631_SP_E[3] = {7.8682724, 1.8812885, 0.5442493};
const double C631_SP_S[3] = {-0.1193324, -0.1608542, 1.1434564};
const double C631_SP_P[3] = {0.0689991, 0.3164240, 0.7443083};
const double C631_OUT_E = 0.1687145;
const double H631_IN_E[3] = {18.7311370, 2.8253937, 0.6401217};
const double H631_IN_C[3] = {0.0334946, 0.2347270, 0.8137573};
const double H631_OUT_E = 0.1612778;
// STO-3G, Hehre, Stewart and Pople, J. Chem. Phys. 51, 2657 (1969)
const double STO3G_1S_C[3] = {0.15432897, 0.53532814, 0.44463454};
const double H_STO3G_E[3] = {3.42525091, 0.62391373, 0.16885540};
const double C_STO3G_1S_E[3] = {71.6168370, 13.0450960, 3.5305122};
const double C_STO3G_SP_E[3] = {2.9412494, 0.6834831, 0.2222899};
const double C_STO3G_SP_S[3] = {-0.09996723, 0.39951283, 0.70011547};
const double C_STO3G_SP_P[3] = {0.15591627, 0.60768372, 0.39195739};

void add_ao(HostSys& h, int kind, int center, int l, int m, int np, const double* e, const double* c,
            const int32_t* pw = nullptr) {
  qmc_ao a;
  a.kind = kind;
  a.center = center;
  a.l = l;
  a.m = m;
  a.nprim = np;
  a.prim0 = (int32_t)h.pexp.size();
  for (int k = 0; k < np; ++k) {
    h.pexp.push_back(e[k]);
    h.pcoef.push_back(c[k]);
    h.ppow.push_back(pw ? pw[k] : 0);
  }
  h.aos.push_back(a);
}

void add_c631(HostSys& h, int nuc) {
  const double one = 1.0;
  add_ao(h, QMC_AO_GTO, nuc, 0, 0, 6, C631_1S_E, C631_1S_C);
  add_ao(h, QMC_AO_GTO, nuc, 0, 0, 3, C631_SP_E, C631_SP_S);
  for (int m = 0; m < 3; ++m) add_ao(h, QMC_AO_GTO, nuc, 1, m, 3, C631_SP_E, C631_SP_P);
  add_ao(h, QMC_AO_GTO, nuc, 0, 0, 1, &C631_OUT_E, &one);
  for (int m = 0; m < 3; ++m) add_ao(h, QMC_AO_GTO, nuc, 1, m, 1, &C631_OUT_E, &one);
}

// Jastrow3 defaults a_anti = 1/2, a_para = 1/4 (jastrow.rs:177-178); the e-n
// series species are the distinct charges in order of first appearance.
void jastrow3(HostSys& h, double b_ee, double b_en) {
  h.jas_kind = QMC_JAS_PADE;
  h.a_anti = 0.5;
  h.a_para = 0.25;
  h.b_ee = b_ee;
  h.b_en = b_en;
  std::vector<double> seen;
  h.species.assign(h.n_nuc(), 0);
  for (int I = 0; I < h.n_nuc(); ++I) {
    auto it = std::find(seen.begin(), seen.end(), h.Z[I]);
    if (it == seen.end()) {
      if ((int)seen.size() < MAXSPEC) seen.push_back(h.Z[I]);
      it = seen.end() - 1;
    }
    h.species[I] = (int32_t)(it - seen.begin());
  }
  h.n_species = (int32_t)seen.size();
}

void methane_geometry(HostSys& h, double ch) {
  const double d = ch / std::sqrt(3.0);  // methane.rs:625-632
  const double H[4][3] = {{d, d, d}, {d, -d, -d}, {-d, d, -d}, {-d, -d, d}};
  h.nuc = {0.0, 0.0, 0.0};
  h.Z = {6.0};
  for (int k = 0; k < 4; ++k) {
    h.nuc.insert(h.nuc.end(), H[k], H[k] + 3);
    h.Z.push_back(1.0);
  }
  h.spins = {1, 1, 1, 1, 1, -1, -1, -1, -1, -1};  // methane.rs:640
  h.init_nuc.resize(10);
  for (int e = 0; e < 10; ++e) h.init_nuc[e] = e < 6 ? 0 : e - 5;  // methane.rs:1110-1119
  h.init_sigma = 0.5;
}

// Seeded orthonormal MO rows: C = 2U-1 from Philox counter (mo*n_ao+ao, 0,
// 0, SYS), then modified Gram-Schmidt over rows (Euclidean metric).
void random_mo(uint64_t seed, int n_mo, int n_ao, std::vector<double>& mo) {
  mo.assign((size_t)n_mo * n_ao, 0.0);
  for (int m = 0; m < n_mo; ++m)
    for (int a = 0; a < n_ao; ++a) {
      uint32_t ctr[4] = {(uint32_t)(m * n_ao + a), 0u, 0u, RNG_SYS}, w[4];
      philox_host(ctr, seed, w);
      mo[m * n_ao + a] = 2.0 * u01(w[0]) - 1.0;
    }
  for (int m = 0; m < n_mo; ++m) {
    for (int q = 0; q < m; ++q) {
      double p = 0.0;
      for (int a = 0; a < n_ao; ++a) p += mo[m * n_ao + a] * mo[q * n_ao + a];
      for (int a = 0; a < n_ao; ++a) mo[m * n_ao + a] -= p * mo[q * n_ao + a];
    }
    double nn = 0.0;
    for (int a = 0; a < n_ao; ++a) nn += mo[m * n_ao + a] * mo[m * n_ao + a];
    const double inv = 1.0 / std::sqrt(nn);
    for (int a = 0; a < n_ao; ++a) mo[m * n_ao + a] *= inv;
  }
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int put(double* dst, const double* src, size_t n) {
  if (!dst || n == 0) return QMC_OK;
  if (is_device_ptr(dst))
    HIP_TRY(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
  else
    std::copy(src, src + n, dst);
  return QMC_OK;
}

// Device view of a caller buffer: the buffer itself when it is device
// memory, otherwise a staging allocation copied in (in=true) and/or out.
struct Staged {
  void* user = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
  bool owned = false;
  hipStream_t st = nullptr;
  ~Staged() {
    if (owned && dev) (void)hipFree(dev);
  }
  int init(const void* p, size_t nbytes, bool copy_in, hipStream_t s) {
    user = const_cast<void*>(p);
    bytes = nbytes;
    st = s;
    if (!p) return QMC_OK;
    if (is_device_ptr(p)) {
      dev = user;
      return QMC_OK;
    }
    HIP_TRY(hipMalloc(&dev, std::max<size_t>(nbytes, 8)));
    owned = true;
    if (copy_in) HIP_TRY(hipMemcpyAsync(dev, p, nbytes, hipMemcpyHostToDevice, s));
    return QMC_OK;
  }
  int copy_out() {
    if (owned && user) HIP_TRY(hipMemcpyAsync(user, dev, bytes, hipMemcpyDeviceToHost, st));
    return QMC_OK;
  }
};

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* qmc_last_error(void) { return g_err.c_str(); }
const char* qmc_version(void) { return "rust_qmc_amd 0.1 (gfx950)"; }

int qmc_system_create(const qmc_system_desc* d, int32_t precision, qmc_system** out) {
  if (!d || !d->spins || !d->nuc_pos || !d->nuc_charge || !d->aos || !d->prim_exp || !d->prim_coef || !d->mo_up)
    return fail(QMC_ERR_INVALID, "null field in qmc_system_desc");
  if (d->n_elec < 1 || d->n_elec > MAXE || d->n_nuc < 1 || d->n_nuc > MAXN || d->n_ao < 1 || d->n_ao > MAXAO ||
      d->n_mo < 1 || d->n_mo > MAXMO)
    return fail(QMC_ERR_INVALID, "size out of range in qmc_system_desc");
  HostSys h;
  h.spins.assign(d->spins, d->spins + d->n_elec);
  h.nuc.assign(d->nuc_pos, d->nuc_pos + 3 * d->n_nuc);
  h.Z.assign(d->nuc_charge, d->nuc_charge + d->n_nuc);
  h.aos.assign(d->aos, d->aos + d->n_ao);
  int nprim = 0;
  for (int i = 0; i < d->n_ao; ++i) nprim = std::max(nprim, d->aos[i].prim0 + d->aos[i].nprim);
  if (nprim > MAXPRIM) return fail(QMC_ERR_INVALID, "too many primitives");
  h.pexp.assign(d->prim_exp, d->prim_exp + nprim);
  h.pcoef.assign(d->prim_coef, d->prim_coef + nprim);
  if (d->prim_pow)
    h.ppow.assign(d->prim_pow, d->prim_pow + nprim);
  else
    h.ppow.assign(nprim, 0);
  h.n_mo = d->n_mo;
  h.mo_up.assign(d->mo_up, d->mo_up + (size_t)d->n_mo * d->n_ao);
  const double* dn = d->mo_dn ? d->mo_dn : d->mo_up;
  h.mo_dn.assign(dn, dn + (size_t)d->n_mo * d->n_ao);
  h.jas_kind = d->jastrow_kind;
  h.F = d->jas_F;
  h.b_ee = d->b_ee;
  h.b_en = d->b_en;
  h.a_anti = d->a_anti;
  h.a_para = d->a_para;
  h.k_ee = d->k_ee;
  h.k_en = d->k_en;
  h.n_species = d->n_species;
  if (h.k_ee > 0) {
    if (!d->c_anti || !d->c_para) return fail(QMC_ERR_INVALID, "series coefficients missing");
    h.c_anti.assign(d->c_anti, d->c_anti + h.k_ee);
    h.c_para.assign(d->c_para, d->c_para + h.k_ee);
  }
  if (h.k_en > 0) {
    if (!d->d_en || !d->species) return fail(QMC_ERR_INVALID, "e-n series coefficients missing");
    h.d_en.assign(d->d_en, d->d_en + h.n_species * h.k_en);
  }
  if (d->species) h.species.assign(d->species, d->species + d->n_nuc);
  if (d->init_nuc)
    h.init_nuc.assign(d->init_nuc, d->init_nuc + d->n_elec);
  else
    h.init_nuc.assign(d->n_elec, 0);
  h.init_sigma = d->init_sigma > 0 ? d->init_sigma : 1.0;
  h.flags = d->flags;
  h.mo_rule = d->mo_rule;
  if (int rc = apply_mo_rule(h)) return rc;
  return make_system(std::move(h), precision, out);
}

int qmc_system_destroy(qmc_system* sys) {
  delete sys;
  return QMC_OK;
}

int qmc_h2_molecule_mo_create(double alpha, const double c1[3], const double c2[3], double F, double init_sigma,
                              int32_t precision, qmc_system** out) {
  if (!c1 || !c2) return fail(QMC_ERR_INVALID, "null centre");
  HostSys h;
  h.spins = {1, -1};
  h.nuc = {c1[0], c1[1], c1[2], c2[0], c2[1], c2[2]};
  h.Z = {1.0, 1.0};
  const double one = 1.0;
  add_ao(h, QMC_AO_SLATER1S, 0, 0, 0, 1, &alpha, &one);
  add_ao(h, QMC_AO_SLATER1S, 1, 0, 0, 1, &alpha, &one);
  // psi = (phi1 + phi2)(r1) (phi1 + phi2)(r2) J (hydrogen.rs:120-124)
  h.n_mo = 1;
  h.mo_up = {1.0, 1.0};
  h.mo_dn = {1.0, 1.0};
  h.jas_kind = QMC_JAS_PAIR;
  h.F = F;
  h.init_nuc = {0, 1};
  h.init_sigma = init_sigma;
  // Jastrow1 is exact (jastrow.rs:35-53); the HF force keeps the
  // reference's nucleus-nucleus sign (hydrogen.rs:245-255)
  h.flags = QMC_FLAG_HF_NN_REFSIGN;
  return make_system(std::move(h), precision, out);
}

int qmc_sto_slater_det_create(int32_t n, const int32_t* nprim, const double* coef, const int32_t* pw,
                              const double* zeta, const int32_t* spins, double F, double Z, uint32_t flags,
                              int32_t precision, qmc_system** out) {
  if (n < 1 || n > MAXE || !nprim || !coef || !pw || !zeta || !spins) return fail(QMC_ERR_INVALID, "bad arguments");
  HostSys h;
  h.spins.assign(spins, spins + n);
  h.nuc = {0.0, 0.0, 0.0};
  h.Z = {Z};
  // STOSlaterDet (slater.rs:205-216): column j holds orbital j and couples
  // only to electrons of spin_j, so the spin-up determinant uses the
  // orbitals of the spin-up columns in order, likewise spin down.
  int off = 0;
  std::vector<int> up_orb, dn_orb;
  for (int o = 0; o < n; ++o) {
    add_ao(h, QMC_AO_STO, 0, 0, 0, nprim[o], zeta + off, coef + off, pw + off);
    off += nprim[o];
    (spins[o] == 1 ? up_orb : dn_orb).push_back(o);
  }
  h.n_mo = std::max((int)up_orb.size(), (int)dn_orb.size());
  h.mo_up.assign((size_t)h.n_mo * n, 0.0);
  h.mo_dn.assign((size_t)h.n_mo * n, 0.0);
  for (size_t k = 0; k < up_orb.size(); ++k) h.mo_up[k * n + up_orb[k]] = 1.0;
  for (size_t k = 0; k < dn_orb.size(); ++k) h.mo_dn[k * n + dn_orb[k]] = 1.0;
  h.jas_kind = QMC_JAS_PAIR;
  h.F = F;
  h.init_nuc.assign(n, 0);
  h.init_sigma = 1.0;  // slater.rs:198-204
  h.flags = flags;
  return make_system(std::move(h), precision, out);
}

int qmc_lithium_seeded_create(uint64_t seed, uint32_t flags, int32_t precision, qmc_system** out) {
  const double xi[2] = {2.7, 0.65};
  double coef[2][3];
  for (int o = 0; o < 2; ++o) {
    double c[3];
    for (int nu = 0; nu < 3; ++nu) {
      uint32_t ctr[4] = {(uint32_t)(3 * o + nu), 1u, 0u, RNG_SYS}, w[4];
      philox_host(ctr, seed, w);
      c[nu] = 2.0 * u01(w[0]) - 1.0;
    }
    // int |psi|^2 d^3r = sum_ab c_a c_b 4 pi (a+b+2)! / (2 xi)^(a+b+3)
    double norm = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double f = 1.0;
        for (int i = 2; i <= a + b + 2; ++i) f *= i;
        norm += c[a] * c[b] * 4.0 * M_PI * f / std::pow(2.0 * xi[o], a + b + 3);
      }
    const double A = 1.0 / std::sqrt(norm);
    for (int nu = 0; nu < 3; ++nu) coef[o][nu] = A * c[nu];
  }
  const int32_t nprim[3] = {3, 3, 3};
  const double cf[9] = {coef[0][0], coef[0][1], coef[0][2], coef[0][0], coef[0][1], coef[0][2],
                        coef[1][0], coef[1][1], coef[1][2]};
  const int32_t pw[9] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
  const double zt[9] = {xi[0], xi[0], xi[0], xi[0], xi[0], xi[0], xi[1], xi[1], xi[1]};
  const int32_t spins[3] = {1, -1, 1};
  return qmc_sto_slater_det_create(3, nprim, cf, pw, zt, spins, 1.0, 3.0, flags, precision, out);
}

int qmc_methane_gto_create(double b_ee, double b_en, double ch_bond, int32_t basis, uint64_t mo_seed,
                           int32_t precision, qmc_system** out) {
  HostSys h;
  methane_geometry(h, ch_bond);
  const double one = 1.0;
  if (basis == QMC_BASIS_631G) {
    // methane.rs:697-776 and the MO coefficients of methane.rs:662-674
    add_c631(h, 0);
    for (int k = 0; k < 4; ++k) add_ao(h, QMC_AO_GTO, 1 + k, 0, 0, 3, H631_IN_E, H631_IN_C);
    for (int k = 0; k < 4; ++k) add_ao(h, QMC_AO_GTO, 1 + k, 0, 0, 1, &H631_OUT_E, &one);
    static const double MO[5][17] = {
        {0.9943, 0.0234, 0.0, 0.0, 0.0, 0.0020, 0.0, 0.0, 0.0, 0.0018, 0.0018, 0.0018, 0.0018, 0.0015, 0.0015, 0.0015, 0.0015},
        {-0.0237, 0.2223, 0.0, 0.0, 0.0, 0.8340, 0.0, 0.0, 0.0, 0.0673, 0.0673, 0.0673, 0.0673, 0.0437, 0.0437, 0.0437, 0.0437},
        {0.0, 0.0, 0.2090, 0.0, 0.0, 0.0, 0.5730, 0.0, 0.0, 0.1020, 0.1020, -0.1020, -0.1020, 0.0670, 0.0670, -0.0670, -0.0670},
        {0.0, 0.0, 0.0, 0.2090, 0.0, 0.0, 0.0, 0.5730, 0.0, 0.1020, -0.1020, 0.1020, -0.1020, 0.0670, -0.0670, 0.0670, -0.0670},
        {0.0, 0.0, 0.0, 0.0, 0.2090, 0.0, 0.0, 0.0, 0.5730, 0.1020, -0.1020, -0.1020, 0.1020, 0.0670, -0.0670, -0.0670, 0.0670},
    };
    h.n_mo = 5;
    h.mo_up.assign(&MO[0][0], &MO[0][0] + 5 * 17);
  } else if (basis == QMC_BASIS_STO3G) {
    add_ao(h, QMC_AO_GTO, 0, 0, 0, 3, C_STO3G_1S_E, STO3G_1S_C);
    add_ao(h, QMC_AO_GTO, 0, 0, 0, 3, C_STO3G_SP_E, C_STO3G_SP_S);
    for (int m = 0; m < 3; ++m) add_ao(h, QMC_AO_GTO, 0, 1, m, 3, C_STO3G_SP_E, C_STO3G_SP_P);
    for (int k = 0; k < 4; ++k) add_ao(h, QMC_AO_GTO, 1 + k, 0, 0, 3, H_STO3G_E, STO3G_1S_C);
    h.n_mo = 5;
    random_mo(mo_seed, 5, 9, h.mo_up);
  } else {
    return fail(QMC_ERR_INVALID, "unknown basis");
  }
  h.mo_dn = h.mo_up;  // both spins share the MOs (methane.rs:824)
  jastrow3(h, b_ee, b_en);
  h.flags = QMC_FLAG_HF_NN_REFSIGN;  // methane.rs:1259-1269
  return make_system(std::move(h), precision, out);
}

int qmc_benzene_gto_create(double b_ee, double b_en, double cc, double ch, int32_t precision, qmc_system** out) {
  HostSys h;
  // geometry, benzene.rs:119-140: C at (cc cos, cc sin, 0), H radially beyond
  for (int i = 0; i < 6; ++i) {
    const double ang = M_PI / 3.0 * i;
    h.nuc.insert(h.nuc.end(), {cc * std::cos(ang), cc * std::sin(ang), 0.0});
    h.Z.push_back(6.0);
  }
  for (int i = 0; i < 6; ++i) {
    const double ang = M_PI / 3.0 * i, rh = cc + ch;
    h.nuc.insert(h.nuc.end(), {rh * std::cos(ang), rh * std::sin(ang), 0.0});
    h.Z.push_back(1.0);
  }
  h.spins.assign(42, 1);
  for (int e = 21; e < 42; ++e) h.spins[e] = -1;
  h.init_nuc.resize(42);
  for (int e = 0; e < 42; ++e) h.init_nuc[e] = e < 36 ? e / 6 : 6 + (e - 36);  // benzene.rs:642-655
  h.init_sigma = 0.5;
  for (int c = 0; c < 6; ++c) add_c631(h, c);
  const double one = 1.0;
  for (int k = 0; k < 6; ++k) {
    add_ao(h, QMC_AO_GTO, 6 + k, 0, 0, 3, H631_IN_E, H631_IN_C);
    add_ao(h, QMC_AO_GTO, 6 + k, 0, 0, 1, &H631_OUT_E, &one);
  }
  // MO coefficients, benzene.rs:262-470
  const int NA = 66;
  std::vector<double> mo(21 * NA, 0.0);
  auto C = [&](int m, int n, int k) -> double& { return mo[m * NA + 9 * n + k]; };
  int m = 0;
  for (int n = 0; n < 6; ++n, ++m) {
    C(m, n, 0) = 0.9943;
    C(m, n, 1) = 0.0234;
    C(m, n, 5) = 0.0020;
  }
  const double s6 = 1.0 / std::sqrt(6.0), s3 = std::sqrt(3.0);
  for (int n = 0; n < 6; ++n) {
    C(m, n, 1) = 0.20 * s6;
    C(m, n, 5) = 0.80 * s6;
  }
  ++m;
  for (int pass = 0; pass < 2; ++pass, ++m)
    for (int n = 0; n < 6; ++n) {
      const double ang = M_PI / 3.0 * n;
      const double w = (pass == 0 ? std::cos(ang) : std::sin(ang)) / s3;
      C(m, n, 1) = 0.20 * w;
      C(m, n, 5) = 0.80 * w;
      C(m, n, pass == 0 ? 2 : 3) = 0.50 * w;
      C(m, n, pass == 0 ? 6 : 7) = 0.55 * w;
    }
  m += 6;  // MOs 9-14: the C-H sigma bonds, from the geometry (benzene_ch_mos)
  for (int n = 0; n < 6; ++n) {
    C(m, n, 4) = 0.30 * s6;
    C(m, n, 8) = 0.70 * s6;
  }
  ++m;
  for (int pass = 0; pass < 2; ++pass, ++m)
    for (int n = 0; n < 6; ++n) {
      const double ang = M_PI / 3.0 * n;
      const double ph = (pass == 0 ? std::cos(ang) : std::sin(ang)) / s3;
      C(m, n, 4) = 0.30 * ph;
      C(m, n, 8) = 0.70 * ph;
    }
  for (int n = 0; n < 6; ++n) {
    const double ang = M_PI / 3.0 * n, tx = -std::sin(ang), ty = std::cos(ang);
    C(m, n, 2) = 0.25 * tx * s6;
    C(m, n, 3) = 0.25 * ty * s6;
    C(m, n, 6) = 0.60 * tx * s6;
    C(m, n, 7) = 0.60 * ty * s6;
  }
  ++m;
  for (int pass = 0; pass < 2; ++pass, ++m)
    for (int n = 0; n < 6; ++n) {
      const double ang = M_PI / 3.0 * n, tx = -std::sin(ang), ty = std::cos(ang);
      const double ph = (pass == 0 ? std::cos(2.0 * ang) : std::sin(2.0 * ang)) / s3;
      C(m, n, 2) = 0.25 * tx * ph;
      C(m, n, 3) = 0.25 * ty * ph;
      C(m, n, 6) = 0.60 * tx * ph;
      C(m, n, 7) = 0.60 * ty * ph;
    }
  h.n_mo = 21;
  h.mo_up = mo;
  h.mo_dn = mo;
  h.mo_rule = QMC_MO_BENZENE_CH;
  if (int rc = apply_mo_rule(h)) return rc;
  jastrow3(h, b_ee, b_en);
  h.flags = QMC_FLAG_HF_NN_REFSIGN;  // benzene.rs:846-856
  return make_system(std::move(h), precision, out);
}

int qmc_set_precision(qmc_system* s, int32_t precision) {
  if (!s || (precision != QMC_FP64 && precision != QMC_MIXED)) return fail(QMC_ERR_INVALID, "bad precision");
  s->precision = precision;
  return QMC_OK;
}

int qmc_set_engine(qmc_system* s, int32_t engine) {
  if (!s || engine < QMC_ENGINE_AUTO || engine > QMC_ENGINE_TILE) return fail(QMC_ERR_INVALID, "bad engine");
  if (engine == QMC_ENGINE_TILE && !s->tile.ok) return fail(QMC_ERR_UNSUPPORTED, "no tile kernel for this system");
  if ((engine == QMC_ENGINE_REGISTER || engine == QMC_ENGINE_GENERIC) && !s->ops_gen)
    return fail(QMC_ERR_UNSUPPORTED, "no register engine for this electron count");
  s->engine = engine;
  s->pick_ops();
  s->version++;  // ensembles rebuild their state with the new engine
  return QMC_OK;
}

int qmc_engine_name(const qmc_system* s, char* buf, int32_t len) {
  if (!s || !buf || len < 1) return fail(QMC_ERR_INVALID, "bad arguments");
  std::string n = s->tile_sweeps(s->precision) ? std::string("tile")
                  : s->wave()                    ? std::string("wave")
                                                 : std::string("register/") + (s->ops ? s->ops->name : "none");
  std::snprintf(buf, (size_t)len, "%s", n.c_str());
  return QMC_OK;
}

int qmc_set_flags(qmc_system* s, uint32_t flags) {
  if (!s) return fail(QMC_ERR_INVALID, "null system");
  s->h.flags = flags;
  s->dirty = true;
  return upload(s);
}

int qmc_system_info(const qmc_system* s, int32_t* n_elec, int32_t* n_up, int32_t* n_nuc, int32_t* n_ao) {
  if (!s) return fail(QMC_ERR_INVALID, "null system");
  if (n_elec) *n_elec = s->h.n_elec();
  if (n_up) *n_up = s->h.n_up();
  if (n_nuc) *n_nuc = s->h.n_nuc();
  if (n_ao) *n_ao = s->h.n_ao();
  return QMC_OK;
}

int qmc_get_mo(const qmc_system* s, int32_t spin, double* mo) {
  if (!s || !mo || (spin != 0 && spin != 1)) return fail(QMC_ERR_INVALID, "bad arguments");
  const auto& v = spin == 0 ? s->h.mo_up : s->h.mo_dn;
  std::copy(v.begin(), v.end(), mo);
  return QMC_OK;
}

int qmc_nuclear_repulsion(const qmc_system* s, double* vnn) {
  if (!s || !vnn) return fail(QMC_ERR_INVALID, "bad arguments");
  DevSys<double>* D = new DevSys<double>();
  int rc = fill_devsys(s->h, *D);
  *vnn = D->vnn;
  delete D;
  return rc;
}

int qmc_num_params(const qmc_system* s, int32_t* n) {
  if (!s || !n) return fail(QMC_ERR_INVALID, "bad arguments");
  *n = s->h.n_params();
  return QMC_OK;
}

int qmc_get_params(const qmc_system* s, double* p) {
  if (!s || !p) return fail(QMC_ERR_INVALID, "bad arguments");
  const HostSys& h = s->h;
  if (h.jas_kind != QMC_JAS_PADE) return QMC_OK;
  int q = 0;
  p[q++] = h.b_ee;
  p[q++] = h.b_en;
  p[q++] = h.a_anti;
  p[q++] = h.a_para;
  for (int k = 0; k < h.k_ee; ++k) p[q++] = h.c_anti[k];
  for (int k = 0; k < h.k_ee; ++k) p[q++] = h.c_para[k];
  for (int k = 0; k < h.n_species * h.k_en; ++k) p[q++] = h.d_en[k];
  return QMC_OK;
}

int qmc_set_params(qmc_system* s, const double* p, int32_t n) {
  if (!s || !p) return fail(QMC_ERR_INVALID, "bad arguments");
  HostSys& h = s->h;
  // methane.rs:1200-1206 asserts the length; here it is an error code
  if (n != h.n_params()) return fail(QMC_ERR_INVALID, "parameter count mismatch");
  int q = 0;
  h.b_ee = p[q++];
  h.b_en = p[q++];
  h.a_anti = p[q++];
  h.a_para = p[q++];
  for (int k = 0; k < h.k_ee; ++k) h.c_anti[k] = p[q++];
  for (int k = 0; k < h.k_ee; ++k) h.c_para[k] = p[q++];
  for (int k = 0; k < h.n_species * h.k_en; ++k) h.d_en[k] = p[q++];
  s->dirty = true;
  return upload(s);
}

int qmc_set_jastrow_series(qmc_system* s, int32_t k_ee, int32_t k_en) {
  if (!s || k_ee < 0 || k_ee > MAXK || k_en < 0 || k_en > MAXK) return fail(QMC_ERR_INVALID, "bad arguments");
  HostSys& h = s->h;
  if (h.jas_kind != QMC_JAS_PADE) return fail(QMC_ERR_INVALID, "series needs the Pade Jastrow");
  h.k_ee = k_ee;
  h.k_en = k_en;
  h.c_anti.assign(k_ee, 0.0);
  h.c_para.assign(k_ee, 0.0);
  h.d_en.assign((size_t)h.n_species * k_en, 0.0);
  s->dirty = true;
  return upload(s);
}

int qmc_get_nuclei(const qmc_system* s, double* R, double* Z) {
  if (!s) return fail(QMC_ERR_INVALID, "null system");
  if (R) std::copy(s->h.nuc.begin(), s->h.nuc.end(), R);
  if (Z) std::copy(s->h.Z.begin(), s->h.Z.end(), Z);
  return QMC_OK;
}

int qmc_set_nuclei(qmc_system* s, const double* R, int32_t n) {
  if (!s || !R) return fail(QMC_ERR_INVALID, "bad arguments");
  // methane.rs:965-978 asserts the count and rebuilds basis centres; AO
  // centres here are nucleus indices, so only the coordinates change.
  if (n != s->h.n_nuc()) return fail(QMC_ERR_INVALID, "nucleus count mismatch");
  HostSys h = s->h;
  h.nuc.assign(R, R + 3 * n);
  if (int rc = apply_mo_rule(h)) return rc;  // benzene.rs:797: the C-H sigma MOs follow the bonds
  s->h = std::move(h);
  s->dirty = true;
  return upload(s);
}

int qmc_evaluate(const qmc_system* sys, int64_t W, const double* pos, double* logabs, double* sign, double* grad,
                 double* lap, double* eloc, void* stream) {
  if (!sys || !pos || W < 0) return fail(QMC_ERR_INVALID, "bad arguments");
  if (W == 0) return QMC_OK;
  int rc = need_engine(sys);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int N = sys->h.n_elec();
  Staged P, L, S, G, Lp, E;
  if ((rc = P.init(pos, sizeof(double) * W * N * 3, true, st))) return rc;
  if ((rc = L.init(logabs, sizeof(double) * W, false, st))) return rc;
  if ((rc = S.init(sign, sizeof(double) * W, false, st))) return rc;
  if ((rc = G.init(grad, sizeof(double) * W * N * 3, false, st))) return rc;
  if ((rc = Lp.init(lap, sizeof(double) * W * N, false, st))) return rc;
  if ((rc = E.init(eloc, sizeof(double) * W, false, st))) return rc;
  int* sing = nullptr;
  HIP_TRY(hipMalloc(&sing, sizeof(int)));
  std::unique_ptr<int, void (*)(int*)> sing_guard(sing, [](int* p) { (void)hipFree(p); });
  HIP_TRY(hipMemsetAsync(sing, 0, sizeof(int), st));
  EvalOut O{(double*)L.dev, (double*)S.dev, (double*)G.dev, (double*)Lp.dev, (double*)E.dev};
  if (sys->wave()) {
    if (sys->precision == QMC_FP64)
      HIP_TRY(wave_eval<double>(sys->d64, sys->d64, W, (const double*)P.dev, sys->shape(), O, sing, st));
    else
      HIP_TRY(wave_eval<float>(sys->d64, sys->d32, W, (const double*)P.dev, sys->shape(), O, sing, st));
  } else if (sys->precision == QMC_FP64) {
    HIP_TRY(sys->ops->eval64(sys->d64, sys->d64, W, (const double*)P.dev, O, sing, st));
  } else {
    HIP_TRY(sys->ops->eval32(sys->d64, sys->d32, W, (const double*)P.dev, O, sing, st));
  }
  for (Staged* s : {&L, &S, &G, &Lp, &E})
    if ((rc = s->copy_out())) return rc;
  int nsing = 0;
  HIP_TRY(hipMemcpyAsync(&nsing, sing, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (nsing > 0 && !(sys->h.flags & QMC_FLAG_SINGULAR_IDENTITY))
    return fail(QMC_ERR_SINGULAR, std::to_string(nsing) + " configurations with a singular Slater matrix");
  return QMC_OK;
}

int qmc_eval_log_psi(const qmc_system* s, int64_t W, const double* pos, double* logabs, double* sign, void* st) {
  return qmc_evaluate(s, W, pos, logabs, sign, nullptr, nullptr, nullptr, st);
}
int qmc_grad_log_psi(const qmc_system* s, int64_t W, const double* pos, double* g, void* st) {
  return qmc_evaluate(s, W, pos, nullptr, nullptr, g, nullptr, nullptr, st);
}
int qmc_lap_over_psi(const qmc_system* s, int64_t W, const double* pos, double* l, void* st) {
  return qmc_evaluate(s, W, pos, nullptr, nullptr, nullptr, l, nullptr, st);
}
int qmc_local_energy(const qmc_system* s, int64_t W, const double* pos, double* el, void* st) {
  return qmc_evaluate(s, W, pos, nullptr, nullptr, nullptr, nullptr, el, st);
}

int qmc_derivatives(const qmc_system* sys, int64_t W, const double* pos, double* eloc, double* O, double* dE,
                    double* F, double* Om, void* stream) {
  if (!sys || !pos || W < 0) return fail(QMC_ERR_INVALID, "bad arguments");
  if (W == 0) return QMC_OK;
  hipStream_t st = (hipStream_t)stream;
  const int N = sys->h.n_elec();
  const Cols C(sys->h);
  int rc;
  Staged P;
  if ((rc = P.init(pos, sizeof(double) * W * N * 3, true, st))) return rc;
  double* Y = nullptr;
  HIP_TRY(hipMalloc(&Y, sizeof(double) * (size_t)C.Qp * W));
  std::unique_ptr<double, void (*)(double*)> yg(Y, [](double* p) { (void)hipFree(p); });
  DerivArgs A{W, C.P, C.M, C.Qp, (const double*)P.dev, nullptr, nullptr, 0, Y, W};
  if ((rc = launch_derivs(sys, A, sys->precision, st))) return rc;
  std::vector<double> h((size_t)C.Qp * W);
  HIP_TRY(hipMemcpyAsync(h.data(), Y, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const int Pn = C.P, M = C.M;
  const std::vector<double> fnn = hf_nn(sys->h);
  std::vector<double> o((size_t)W * Pn), d((size_t)W * Pn), f((size_t)W * 3 * M), om((size_t)W * 3 * M), e(W);
  for (int64_t w = 0; w < W; ++w) {
    e[w] = h[(size_t)C.e() * W + w];
    for (int k = 0; k < Pn; ++k) {
      o[w * Pn + k] = h[(size_t)k * W + w];
      d[w * Pn + k] = h[(size_t)C.de(k) * W + w];
    }
    for (int q = 0; q < 3 * M; ++q) {
      om[w * 3 * M + q] = h[(size_t)(Pn + q) * W + w];
      f[w * 3 * M + q] = h[(size_t)C.f(q / 3, q % 3) * W + w] + fnn[q];
    }
  }
  if ((rc = put(eloc, e.data(), e.size())) || (rc = put(O, o.data(), o.size())) || (rc = put(dE, d.data(), d.size())) ||
      (rc = put(F, f.data(), f.size())) || (rc = put(Om, om.data(), om.size())))
    return rc;
  return QMC_OK;
}

int qmc_log_derivs(const qmc_system* s, int64_t W, const double* pos, double* O, void* st) {
  return qmc_derivatives(s, W, pos, nullptr, O, nullptr, nullptr, nullptr, st);
}
int qmc_dEL_dp(const qmc_system* s, int64_t W, const double* pos, double* dE, void* st) {
  return qmc_derivatives(s, W, pos, nullptr, nullptr, dE, nullptr, nullptr, st);
}
int qmc_hf_force(const qmc_system* s, int64_t W, const double* pos, double* F, void* st) {
  return qmc_derivatives(s, W, pos, nullptr, nullptr, nullptr, F, nullptr, st);
}
int qmc_nuc_grad_log(const qmc_system* s, int64_t W, const double* pos, double* Om, void* st) {
  return qmc_derivatives(s, W, pos, nullptr, nullptr, nullptr, nullptr, Om, st);
}

int qmc_derivatives_noalloc(const qmc_system* sys, int64_t W, const double* pos, double* eloc, double* O, double* dE,
                            double* F, double* Om, void* stream) {
  if (!sys || !pos || W < 0) return fail(QMC_ERR_INVALID, "bad arguments");
  if (W == 0) return QMC_OK;
  for (const void* p : {(const void*)pos, (const void*)eloc, (const void*)O, (const void*)dE, (const void*)F,
                        (const void*)Om})
    if (p && !is_device_ptr(p))
      return fail(QMC_ERR_INVALID, "qmc_derivatives_noalloc takes device pointers only (it stages nothing)");
  const Cols C(sys->h);
  const int64_t wmax = NOALLOC_SCRATCH_DOUBLES / C.Qp;
  if (W > wmax)
    return fail(QMC_ERR_INVALID, "qmc_derivatives_noalloc: " + std::to_string(W) + " walkers exceed the " +
                                     std::to_string(wmax) + " its " +
                                     std::to_string(NOALLOC_SCRATCH_DOUBLES * 8 >> 20) + " MiB scratch holds for " +
                                     std::to_string(C.Qp) + " feature rows");
  double* Y = nullptr;
  HIP_TRY(noalloc_scratch(&Y));
  hipStream_t st = (hipStream_t)stream;
  DerivArgs A{W, C.P, C.M, C.Qp, pos, nullptr, nullptr, 0, Y, W};
  if (int rc = launch_derivs(sys, A, sys->precision, st)) return rc;
  static_assert(3 * MAXN <= (int)(sizeof(DerivUnpack::fnn) / sizeof(double)), "fnn holds 3 MAXN forces");
  DerivUnpack U{W, C.P, C.M, Y, eloc, O, dE, F, Om, {}};
  hf_nn_fill(sys->h, U.fnn);
  HIP_TRY(launch_deriv_unpack(U, st));
  return QMC_OK;
}

int qmc_sr_solve(int32_t n, const double* S, const double* f, double eps, double* dp) {
  if (n < 1 || !S || !f || !dp) return fail(QMC_ERR_INVALID, "bad arguments");
  std::vector<double> a(S, S + (size_t)n * n), b(n), x;
  for (int i = 0; i < n; ++i) {
    a[(size_t)i * n + i] += eps;
    b[i] = -f[i];
  }
  if (!lu_solve(n, a, b, x)) {
    std::copy(b.begin(), b.end(), dp);  // sr_optimize.rs:239-242
    return fail(QMC_ERR_SINGULAR, "SR matrix singular; fell back to -f");
  }
  std::copy(x.begin(), x.end(), dp);
  return QMC_OK;
}

}  // extern "C"

// ===========================================================================
// Ensembles
// ===========================================================================
struct qmc_ensemble {
  const qmc_system* sys = nullptr;
  int64_t W = 0, ld = 0;
  int prec = QMC_FP64;
  void* pos = nullptr;
  void* ainv = nullptr;
  double* logdet = nullptr;
  double* sign = nullptr;
  double* ju = nullptr;
  double* eloc = nullptr;
  uint64_t seed = 0;
  uint32_t step = 0;
  bool inverse_stale = false;
  int since_refresh = 0;
  double* partial = nullptr;
  size_t partial_cap = 0;
  double* sums = nullptr;
  size_t sums_cap = 0;
  hipStream_t st = nullptr;
  bool timing = false;
  std::vector<hipEvent_t> ev;  // start/stop pairs, one pair per launch
  double kernel_ms = 0.0;
  int64_t launches = 0;
  uint64_t sys_version = 0;  // system parameters the cached state was built with
  // moment accumulator of qmc_ensemble_measure
  std::vector<double> gacc;
  int gq = 0;
  double* Y = nullptr;
  size_t Y_cap = 0;
  double* gpart = nullptr;
  size_t gpart_cap = 0;
  double* G = nullptr;
  size_t G_cap = 0;
  double* rwork = nullptr;  // Slater rows of the tile refresh
  size_t rwork_cap = 0;
  ~qmc_ensemble() {
    for (void* p : {pos, ainv, (void*)logdet, (void*)sign, (void*)ju, (void*)eloc, (void*)partial, (void*)sums,
                    (void*)Y, (void*)gpart, (void*)G, (void*)rwork})
      if (p) (void)hipFree(p);
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  size_t tsize() const { return prec == QMC_FP64 ? sizeof(double) : sizeof(float); }
  int nainv() const {
    const int nu = sys->h.n_up(), nd = sys->h.n_dn();
    return nu * nu + nd * nd;
  }
  template <class T>
  EnsView<T> view() const {
    return EnsView<T>{W, ld, (T*)pos, (T*)ainv, logdet, sign, ju, eloc};
  }
};

namespace {

// The periodic refresh of benzene-size ensembles (no local energy wanted)
// runs on tiles of 16 walkers with FP64 matrix cores; QMC_TILE_REFRESH=0 in
// the environment sends it to the wave engine's build instead.
bool tile_refresh_on() {
  const char* v = std::getenv("QMC_TILE_REFRESH");
  return !(v && v[0] == '0');
}

int ensure_cap(double** p, size_t* cap, size_t n);

int ens_refresh(qmc_ensemble* e, int want_eloc) {
  const qmc_system* s = e->sys;
  if (s->wave() && !want_eloc && s->tile.ok && tile_refresh_on()) {
    if (int rc = ensure_cap(&e->rwork, &e->rwork_cap, tile_refresh_work(e->W))) return rc;
    if (e->prec == QMC_FP64)
      HIP_TRY(tile_refresh<double>(s->d64, e->view<double>(), s->tile, e->rwork, nullptr, e->st));
    else
      HIP_TRY(tile_refresh<float>(s->d64, e->view<float>(), s->tile, e->rwork, nullptr, e->st));
  } else if (s->wave()) {
    if (e->prec == QMC_FP64)
      HIP_TRY(wave_build<double>(s->d64, s->d64, e->view<double>(), s->shape(), want_eloc, nullptr, e->st));
    else
      HIP_TRY(wave_build<float>(s->d64, s->d32, e->view<float>(), s->shape(), want_eloc, nullptr, e->st));
  } else if (e->prec == QMC_FP64) {
    HIP_TRY(s->ops->build64(s->d64, s->d64, e->view<double>(), want_eloc, nullptr, e->st));
  } else {
    HIP_TRY(s->ops->build32(s->d64, s->d32, e->view<float>(), want_eloc, nullptr, e->st));
  }
  e->inverse_stale = false;
  e->since_refresh = 0;
  e->sys_version = s->version;
  return QMC_OK;
}

int ensure_cap(double** p, size_t* cap, size_t n) {
  if (*cap >= n) return QMC_OK;
  if (*p) HIP_TRY(hipFree(*p));
  *p = nullptr;
  HIP_TRY(hipMalloc(p, n * sizeof(double)));
  *cap = n;
  return QMC_OK;
}

// blocks (partial-sum rows) of one single-electron sweep launch of e
int64_t sweep_blocks(const qmc_ensemble* e) {
  const qmc_system* s = e->sys;
  if (s->tile_sweeps(e->prec)) return tile_blocks(e->W);
  if (s->wave()) return e->W;
  return (e->W + 63) / 64;
}

// enqueue A.n_steps single-electron sweeps of e on st with the engine the
// system selects (partial sums into e->partial)
int enqueue_single_sweeps(qmc_ensemble* e, const SweepArgs& A, double* trace, hipStream_t st) {
  const qmc_system* s = e->sys;
  if (s->tile_sweeps(e->prec)) {
    if (e->prec == QMC_FP64)
      HIP_TRY(tile_sweep<double>(s->d64, e->view<double>(), s->tile, A, e->partial, trace, st));
    else
      HIP_TRY(tile_sweep<float>(s->d32, e->view<float>(), s->tile, A, e->partial, trace, st));
  } else if (s->wave()) {
    if (e->prec == QMC_FP64)
      HIP_TRY(wave_sweep<double>(s->d64, e->view<double>(), s->shape(), A, e->partial, trace, st));
    else
      HIP_TRY(wave_sweep<float>(s->d32, e->view<float>(), s->shape(), A, e->partial, trace, st));
  } else {
    const bool no_series = s->h.k_ee == 0 && s->h.k_en == 0;
    if (e->prec == QMC_FP64)
      HIP_TRY((no_series ? s->ops->sweep64_ns : s->ops->sweep64)(s->d64, e->view<double>(), A, e->partial, trace, st));
    else
      HIP_TRY((no_series ? s->ops->sweep32_ns : s->ops->sweep32)(s->d32, e->view<float>(), A, e->partial, trace, st));
  }
  return QMC_OK;
}

}  // namespace

extern "C" {

int qmc_ensemble_create(const qmc_system* sys, int64_t W, uint64_t seed, qmc_ensemble** out) {
  if (!sys || !out || W < 1 || W > (int64_t)1 << 31) return fail(QMC_ERR_INVALID, "bad arguments");
  int rc = need_engine(sys);
  if (rc) return rc;
  auto e = std::make_unique<qmc_ensemble>();
  e->sys = sys;
  e->W = W;
  e->ld = (W + 63) / 64 * 64;
  e->prec = sys->precision;
  e->seed = seed;
  const int N = sys->h.n_elec();
  HIP_TRY(hipMalloc(&e->pos, e->tsize() * 3 * N * e->ld));
  HIP_TRY(hipMalloc(&e->ainv, e->tsize() * std::max(1, e->nainv()) * e->ld));
  for (double** p : {&e->logdet, &e->sign, &e->ju, &e->eloc}) HIP_TRY(hipMalloc(p, sizeof(double) * e->ld));
  if (e->prec == QMC_FP64)
    HIP_TRY(launch_init64(sys->d64, e->view<double>(), seed, 0, e->st));
  else
    HIP_TRY(launch_init32(sys->d64, e->view<float>(), seed, 0, e->st));
  if ((rc = ens_refresh(e.get(), 1))) return rc;
  HIP_TRY(hipStreamSynchronize(e->st));
  *out = e.release();
  return QMC_OK;
}

int qmc_ensemble_destroy(qmc_ensemble* e) {
  delete e;
  return QMC_OK;
}

int qmc_ensemble_get_positions(const qmc_ensemble* e, double* pos) {
  if (!e || !pos) return fail(QMC_ERR_INVALID, "bad arguments");
  const int N = e->sys->h.n_elec();
  std::vector<char> buf(e->tsize() * 3 * N * e->ld);
  HIP_TRY(hipMemcpy(buf.data(), e->pos, buf.size(), hipMemcpyDeviceToHost));
  std::vector<double> tmp((size_t)e->W * N * 3);
  for (int q = 0; q < 3 * N; ++q)
    for (int64_t w = 0; w < e->W; ++w) {
      const size_t i = (size_t)q * e->ld + w;
      tmp[(size_t)w * 3 * N + q] = e->prec == QMC_FP64 ? ((double*)buf.data())[i] : ((float*)buf.data())[i];
    }
  if (is_device_ptr(pos))
    HIP_TRY(hipMemcpy(pos, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
  else
    std::copy(tmp.begin(), tmp.end(), pos);
  return QMC_OK;
}

int qmc_ensemble_set_positions(qmc_ensemble* e, const double* pos) {
  if (!e || !pos) return fail(QMC_ERR_INVALID, "bad arguments");
  const int N = e->sys->h.n_elec();
  std::vector<double> tmp((size_t)e->W * N * 3);
  if (is_device_ptr(pos))
    HIP_TRY(hipMemcpy(tmp.data(), pos, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
  else
    std::copy(pos, pos + tmp.size(), tmp.begin());
  std::vector<char> buf(e->tsize() * 3 * N * e->ld, 0);
  for (int q = 0; q < 3 * N; ++q)
    for (int64_t w = 0; w < e->W; ++w) {
      const size_t i = (size_t)q * e->ld + w;
      const double v = tmp[(size_t)w * 3 * N + q];
      if (e->prec == QMC_FP64)
        ((double*)buf.data())[i] = v;
      else
        ((float*)buf.data())[i] = (float)v;
    }
  HIP_TRY(hipMemcpy(e->pos, buf.data(), buf.size(), hipMemcpyHostToDevice));
  int rc = ens_refresh(e, 1);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(e->st));
  return QMC_OK;
}

int qmc_ensemble_get_state(const qmc_ensemble* e, double* logabs, double* sign, double* eloc) {
  if (!e) return fail(QMC_ERR_INVALID, "null ensemble");
  std::vector<double> a(e->W), b(e->W);
  if (logabs) {
    HIP_TRY(hipMemcpy(a.data(), e->logdet, sizeof(double) * e->W, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(b.data(), e->ju, sizeof(double) * e->W, hipMemcpyDeviceToHost));
    for (int64_t w = 0; w < e->W; ++w) logabs[w] = a[w] + b[w];
  }
  if (sign) HIP_TRY(hipMemcpy(sign, e->sign, sizeof(double) * e->W, hipMemcpyDeviceToHost));
  if (eloc) HIP_TRY(hipMemcpy(eloc, e->eloc, sizeof(double) * e->W, hipMemcpyDeviceToHost));
  return QMC_OK;
}

int qmc_ensemble_refresh(qmc_ensemble* e) {
  if (!e) return fail(QMC_ERR_INVALID, "null ensemble");
  int rc = ens_refresh(e, 1);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(e->st));
  return QMC_OK;
}

int qmc_ensemble_sweeps(qmc_ensemble* e, const qmc_sweep_params* p, double* e_sum, double* e2_sum,
                        double* accepted, double* eloc_trace) {
  if (!e || !p || p->n_steps < 0 || !(p->step_size >= 0.0)) return fail(QMC_ERR_INVALID, "bad arguments");
  if (p->move != QMC_MOVE_SINGLE && p->move != QMC_MOVE_ALL && p->move != QMC_MOVE_DRIFT)
    return fail(QMC_ERR_INVALID, "bad move kind");
  if (p->move == QMC_MOVE_DRIFT && !(p->step_size > 0.0 && p->max_drift > 0.0))
    return fail(QMC_ERR_INVALID, "drift-diffusion needs time_step > 0 and max_drift > 0");
  // single-electron moves keep the inverse by Sherman-Morrison updates
  const bool sm = p->move != QMC_MOVE_ALL;
  if (p->n_steps == 0) return QMC_OK;
  const qmc_system* s = e->sys;
  int rc;
  if (e->sys_version != s->version && (rc = ens_refresh(e, 1))) return rc;  // parameters or nuclei moved
  if (sm && e->inverse_stale && (rc = ens_refresh(e, 0))) return rc;
  const bool wave = s->wave();
  // drift-diffusion moves on the tile engine where it runs single-electron
  // sweeps; all-electron moves on the wave engine (also for tile-engine
  // systems): one block per walker
  const bool tdd = p->move == QMC_MOVE_DRIFT && s->tile_dd(e->prec);
  const int nblk = p->move == QMC_MOVE_SINGLE ? (int)sweep_blocks(e)
                   : tdd                       ? (int)tile_blocks(e->W)
                                               : (int)(wave ? e->W : (e->W + 63) / 64);
  const int refresh = p->refresh_interval > 0 ? p->refresh_interval : 0;
  int chunk = p->steps_per_launch > 0 ? p->steps_per_launch : 16;
  if ((rc = ensure_cap(&e->partial, &e->partial_cap, (size_t)chunk * nblk * 3))) return rc;
  if ((rc = ensure_cap(&e->sums, &e->sums_cap, (size_t)p->n_steps * 3))) return rc;
  Staged tr;
  if ((rc = tr.init(eloc_trace, sizeof(double) * (size_t)p->n_steps * e->W, false, e->st))) return rc;
  int done = 0;
  int64_t nl = 0;
  while (done < p->n_steps) {
    int ns = std::min(chunk, p->n_steps - done);
    if (sm && refresh > 0) {
      // an earlier call may have run past this call's (shorter) interval
      if (e->since_refresh >= refresh && (rc = ens_refresh(e, 0))) return rc;
      ns = std::min(ns, refresh - e->since_refresh);
    }
    SweepArgs A{ns, e->step, e->seed, 0, p->step_size, e->W, p->max_drift};
    double* trace = tr.dev ? (double*)tr.dev + (size_t)done * e->W : nullptr;
    if (e->timing) {
      while ((int64_t)e->ev.size() < 2 * (nl + 1)) {
        hipEvent_t ev;
        HIP_TRY(hipEventCreate(&ev));
        e->ev.push_back(ev);
      }
      HIP_TRY(hipEventRecord(e->ev[2 * nl], e->st));
    }
    if (p->move == QMC_MOVE_SINGLE) {
      if ((rc = enqueue_single_sweeps(e, A, trace, e->st))) return rc;
    } else if (p->move == QMC_MOVE_DRIFT) {
      if (tdd && e->prec == QMC_FP64)
        HIP_TRY(tile_dd<double>(s->d64, e->view<double>(), s->tile, A, e->partial, trace, e->st));
      else if (tdd)
        HIP_TRY(tile_dd<float>(s->d32, e->view<float>(), s->tile, A, e->partial, trace, e->st));
      else if (wave && e->prec == QMC_FP64)
        HIP_TRY(wave_dd<double>(s->d64, e->view<double>(), s->shape(), A, e->partial, trace, e->st));
      else if (wave)
        HIP_TRY(wave_dd<float>(s->d32, e->view<float>(), s->shape(), A, e->partial, trace, e->st));
      else if (e->prec == QMC_FP64)
        HIP_TRY((s->no_series() ? s->ops->dd64_ns : s->ops->dd64)(s->d64, e->view<double>(), A, e->partial, trace,
                                                                    e->st));
      else
        HIP_TRY((s->no_series() ? s->ops->dd32_ns : s->ops->dd32)(s->d32, e->view<float>(), A, e->partial, trace,
                                                                    e->st));
    } else {
      if (wave && e->prec == QMC_FP64)
        HIP_TRY(wave_ae<double>(s->d64, s->d64, e->view<double>(), s->shape(), A, e->partial, trace, e->st));
      else if (wave)
        HIP_TRY(wave_ae<float>(s->d64, s->d32, e->view<float>(), s->shape(), A, e->partial, trace, e->st));
      else if (e->prec == QMC_FP64)
        HIP_TRY(s->ops->ae64(s->d64, s->d64, e->view<double>(), A, e->partial, trace, e->st));
      else
        HIP_TRY(s->ops->ae32(s->d64, s->d32, e->view<float>(), A, e->partial, trace, e->st));
      e->inverse_stale = true;
    }
    if (e->timing) HIP_TRY(hipEventRecord(e->ev[2 * nl + 1], e->st));
    ++nl;
    HIP_TRY(launch_reduce(e->partial, ns, nblk, 3, e->sums + (size_t)done * 3, e->st));
    e->step += (uint32_t)ns;
    done += ns;
    if (sm) {
      e->since_refresh += ns;
      if (refresh > 0 && e->since_refresh >= refresh && (rc = ens_refresh(e, 0))) return rc;
    }
  }
  std::vector<double> h((size_t)p->n_steps * 3);
  HIP_TRY(hipMemcpyAsync(h.data(), e->sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, e->st));
  if ((rc = tr.copy_out())) return rc;
  HIP_TRY(hipStreamSynchronize(e->st));
  e->launches = nl;
  e->kernel_ms = 0.0;
  if (e->timing)
    for (int64_t i = 0; i < nl; ++i) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, e->ev[2 * i], e->ev[2 * i + 1]));
      e->kernel_ms += ms;
    }
  for (int q = 0; q < 3; ++q) {
    double* dst = q == 0 ? e_sum : (q == 1 ? e2_sum : accepted);
    if (!dst) continue;
    std::vector<double> col(p->n_steps);
    for (int i = 0; i < p->n_steps; ++i) col[i] = h[(size_t)i * 3 + q];
    if (is_device_ptr(dst))
      HIP_TRY(hipMemcpy(dst, col.data(), col.size() * sizeof(double), hipMemcpyHostToDevice));
    else
      std::copy(col.begin(), col.end(), dst);
  }
  return QMC_OK;
}

int qmc_ensemble_measure(qmc_ensemble* e, int32_t reset) {
  if (!e) return fail(QMC_ERR_INVALID, "null ensemble");
  const qmc_system* s = e->sys;
  int rc;
  if (e->sys_version != s->version && (rc = ens_refresh(e, 1))) return rc;
  const Cols C(s->h);
  const int nblk = gram_blocks(e->W);
  if ((rc = ensure_cap(&e->Y, &e->Y_cap, (size_t)C.Qp * e->ld))) return rc;
  if ((rc = ensure_cap(&e->gpart, &e->gpart_cap, (size_t)nblk * C.Qp * C.Qp))) return rc;
  if ((rc = ensure_cap(&e->G, &e->G_cap, (size_t)C.Qp * C.Qp))) return rc;
  DerivArgs A{e->W, C.P, C.M, C.Qp, nullptr, nullptr, nullptr, e->ld, e->Y, e->ld};
  if (e->prec == QMC_FP64)
    A.pos64 = (const double*)e->pos;
  else
    A.pos32 = (const float*)e->pos;
  if ((rc = launch_derivs(s, A, e->prec, e->st))) return rc;
  if (e->ld > e->W)  // the Gram reads the padding columns (up to W rounded to 16) as zeros
    HIP_TRY(hipMemset2DAsync(e->Y + e->W, sizeof(double) * e->ld, 0, sizeof(double) * (e->ld - e->W), C.Qp, e->st));
  HIP_TRY(launch_gram(e->Y, e->ld, e->W, C.Qp, e->gpart, e->G, e->st));
  std::vector<double> g((size_t)C.Qp * C.Qp);
  HIP_TRY(hipMemcpyAsync(g.data(), e->G, g.size() * sizeof(double), hipMemcpyDeviceToHost, e->st));
  HIP_TRY(hipStreamSynchronize(e->st));
  if (reset || e->gq != C.Qp) {
    e->gacc.assign(g.size(), 0.0);
    e->gq = C.Qp;
  }
  for (int i = 0; i < C.Qp; ++i)  // upper triangle is valid; mirror it
    for (int j = i; j < C.Qp; ++j) {
      e->gacc[(size_t)i * C.Qp + j] += g[(size_t)i * C.Qp + j];
      if (j != i) e->gacc[(size_t)j * C.Qp + i] += g[(size_t)i * C.Qp + j];
    }
  return QMC_OK;
}

int qmc_ensemble_moments(const qmc_ensemble* e, qmc_moments* out) {
  if (!e || !out) return fail(QMC_ERR_INVALID, "bad arguments");
  const Cols C(e->sys->h);
  if (e->gq != C.Qp || e->gacc.empty()) return fail(QMC_ERR_INVALID, "no measurements accumulated");
  auto g = [&](int a, int b) { return e->gacc[(size_t)a * C.Qp + b]; };
  const double n = g(C.one(), C.one());
  const int P = C.P, M = C.M, Q = P + 3 * M;
  out->n_samples = (int64_t)std::llround(n);
  out->P = P;
  out->M = M;
  const double em = g(C.e(), C.one()) / n;
  out->energy = em;
  out->variance = g(C.e(), C.e()) / n - em * em;
  std::vector<double> om(Q), S((size_t)Q * Q), f(Q), de(P), fhf(3 * M), fzv(3 * M), fv(M, 0.0);
  for (int q = 0; q < Q; ++q) om[q] = g(q, C.one()) / n;
  for (int q = 0; q < Q; ++q) {
    for (int r = 0; r < Q; ++r) S[(size_t)q * Q + r] = g(q, r) / n - om[q] * om[r];
    f[q] = g(q, C.e()) / n - em * om[q];
  }
  for (int k = 0; k < P; ++k) de[k] = g(C.de(k), C.one()) / n;
  const std::vector<double> fnn = hf_nn(e->sys->h);
  for (int I = 0; I < M; ++I)
    for (int c = 0; c < 3; ++c) {
      const int col = C.f(I, c);
      const double mean_el = g(col, C.one()) / n;
      fhf[3 * I + c] = mean_el + fnn[3 * I + c];
      fzv[3 * I + c] = fhf[3 * I + c] - 2.0 * f[P + 3 * I + c];
      fv[I] += g(col, col) / n - mean_el * mean_el;
    }
  for (double& v : fv) v = std::max(v, 0.0);
  int rc;
  if ((rc = put(out->o_mean, om.data(), om.size())) || (rc = put(out->S, S.data(), S.size())) ||
      (rc = put(out->f, f.data(), f.size())) || (rc = put(out->de_mean, de.data(), de.size())) ||
      (rc = put(out->f_hf, fhf.data(), fhf.size())) || (rc = put(out->f_zv, fzv.data(), fzv.size())) ||
      (rc = put(out->f_var, fv.data(), fv.size())))
    return rc;
  return QMC_OK;
}

// SROptimizer::optimize, sr_optimize.rs:252-348.  Walkers persist across
// iterations (the reference draws fresh ones and re-equilibrates each
// iteration); each iteration takes max(1, n_samples / n_walkers)
// measurement rounds separated by `decorrelation` sweeps (:141-168).
int qmc_sr_optimize(qmc_system* sys, const qmc_sr_params* p, uint64_t seed, qmc_sr_results* res,
                    double* energy_history, double* variance_history, double* param_history) {
  if (!sys || !p || !res || p->n_walkers < 1 || p->max_iterations < 1 || sys->h.jas_kind != QMC_JAS_PADE)
    return fail(QMC_ERR_INVALID, "bad arguments (SR needs the Pade Jastrow)");
  qmc_ensemble* e = nullptr;
  int rc = qmc_ensemble_create(sys, p->n_walkers, seed, &e);
  if (rc) return rc;
  std::unique_ptr<qmc_ensemble> guard(e);
  const int P = sys->h.n_params(), M = sys->h.n_nuc();
  const int Q = p->include_nuclear ? P + 3 * M : P;
  const int rounds = (int)std::max<int64_t>(1, p->n_samples / p->n_walkers);
  const int dec = p->decorrelation > 0 ? p->decorrelation : 5;
  auto sweeps = [&](int n) {
    qmc_sweep_params sp{p->move, n, p->step_size, p->refresh_interval, 0};
    return n > 0 ? qmc_ensemble_sweeps(e, &sp, nullptr, nullptr, nullptr, nullptr) : QMC_OK;
  };
  auto sample = [&](qmc_moments* m, std::vector<double>* S, std::vector<double>* f) {
    int r;
    for (int k = 0; k < rounds; ++k) {
      if ((r = sweeps(dec)) || (r = qmc_ensemble_measure(e, k == 0))) return r;
    }
    const int Qa = P + 3 * M;
    std::vector<double> Sa((size_t)Qa * Qa), fa(Qa);
    *m = qmc_moments{};
    m->S = Sa.data();
    m->f = fa.data();
    if ((r = qmc_ensemble_moments(e, m))) return r;
    m->S = nullptr;
    m->f = nullptr;
    if (S) {
      S->assign((size_t)Q * Q, 0.0);
      f->assign(Q, 0.0);
      for (int i = 0; i < Q; ++i) {
        (*f)[i] = fa[i];
        for (int j = 0; j < Q; ++j) (*S)[(size_t)i * Q + j] = Sa[(size_t)i * Qa + j];
      }
    }
    return QMC_OK;
  };
  if ((rc = sweeps((int)p->n_equilibrate))) return rc;
  std::vector<double> params(P);
  qmc_get_params(sys, params.data());
  if (param_history) std::copy(params.begin(), params.end(), param_history);
  double prev = std::numeric_limits<double>::max();
  int it = 0;
  for (; it < p->max_iterations; ++it) {
    qmc_moments m;
    std::vector<double> S, f, dp(Q);
    if ((rc = sample(&m, &S, &f))) return rc;
    if (energy_history) energy_history[it] = m.energy;
    if (variance_history) variance_history[it] = m.variance;
    qmc_sr_solve(Q, S.data(), f.data(), p->sr_epsilon, dp.data());
    // sr_optimize.rs:297-307: per-step clamp +-0.3, then the reference's
    // parameters to [0.05, 10]; series coefficients to [-10, 10]
    for (int k = 0; k < P; ++k) {
      params[k] += std::min(std::max(p->learning_rate * dp[k], -0.3), 0.3);
      params[k] = k < 4 ? std::min(std::max(params[k], 0.05), 10.0) : std::min(std::max(params[k], -10.0), 10.0);
    }
    if ((rc = qmc_set_params(sys, params.data(), P))) return rc;
    if (param_history) std::copy(params.begin(), params.end(), param_history + (size_t)(it + 1) * P);
    const double change = std::fabs(m.energy - prev);  // sr_optimize.rs:313-320
    if (it > 5 && change < p->tolerance) {
      ++it;
      break;
    }
    prev = m.energy;
  }
  qmc_moments m;
  if ((rc = sample(&m, nullptr, nullptr))) return rc;  // sr_optimize.rs:325-330
  res->final_energy = m.energy;
  res->final_variance = m.variance;
  res->iterations = it;
  return QMC_OK;
}

int qmc_sample_forces(const qmc_system* sys, const qmc_force_params* p, uint64_t seed, double* forces,
                      double* mean_energy, double* error, double* force_variance) {
  if (!sys || !p || p->n_walkers < 1) return fail(QMC_ERR_INVALID, "bad arguments");
  qmc_ensemble* e = nullptr;
  int rc = qmc_ensemble_create(sys, p->n_walkers, seed, &e);
  if (rc) return rc;
  std::unique_ptr<qmc_ensemble> guard(e);
  const int rounds = (int)std::max<int64_t>(1, p->n_samples / p->n_walkers);
  const int dec = p->decorrelation > 0 ? p->decorrelation : 5;
  qmc_sweep_params eq{p->move, (int32_t)p->n_equilibrate, p->mc_step_size, p->refresh_interval, 0};
  if (p->n_equilibrate > 0 && (rc = qmc_ensemble_sweeps(e, &eq, nullptr, nullptr, nullptr, nullptr))) return rc;
  qmc_sweep_params sp{p->move, dec, p->mc_step_size, p->refresh_interval, 0};
  for (int k = 0; k < rounds; ++k)
    if ((rc = qmc_ensemble_sweeps(e, &sp, nullptr, nullptr, nullptr, nullptr)) || (rc = qmc_ensemble_measure(e, k == 0)))
      return rc;
  const int M = sys->h.n_nuc();
  std::vector<double> fhf(3 * M), fzv(3 * M), fv(M);
  qmc_moments m{};
  m.f_hf = fhf.data();
  m.f_zv = fzv.data();
  m.f_var = fv.data();
  if ((rc = qmc_ensemble_moments(e, &m))) return rc;
  if (forces) std::copy_n(p->zero_variance ? fzv.data() : fhf.data(), 3 * M, forces);
  if (mean_energy) *mean_energy = m.energy;
  if (error) *error = std::sqrt(m.variance / (double)m.n_samples);  // geometry_opt.rs:198-201
  if (force_variance) std::copy(fv.begin(), fv.end(), force_variance);
  return QMC_OK;
}

int qmc_system_to_json(const qmc_system* s, char* buf, int64_t len, int64_t* needed) {
  if (!s || len < 0 || (len > 0 && !buf)) return fail(QMC_ERR_INVALID, "bad arguments");
  const HostSys& h = s->h;
  std::string o;
  char tmp[40];
  auto num = [&](double v) {
    std::snprintf(tmp, sizeof tmp, "%.17g", v);
    o += tmp;
  };
  auto nums = [&](const double* v, size_t n) {
    o += '[';
    for (size_t i = 0; i < n; ++i) {
      if (i) o += ", ";
      num(v[i]);
    }
    o += ']';
  };
  auto ints = [&](const int32_t* v, size_t n) {
    o += '[';
    for (size_t i = 0; i < n; ++i) o += (i ? ", " : "") + std::to_string(v[i]);
    o += ']';
  };
  auto rows = [&](const double* v, int nr, int nc) {
    o += '[';
    for (int r = 0; r < nr; ++r) {
      if (r) o += ", ";
      nums(v + (size_t)r * nc, nc);
    }
    o += ']';
  };
  const int A = h.n_ao();
  o += "{\"format\": \"rust_qmc_amd.system/1\", \"nuclei\": ";
  rows(h.nuc.data(), h.n_nuc(), 3);
  o += ", \"charges\": ";
  nums(h.Z.data(), h.Z.size());
  o += ", \"spins\": ";
  ints(h.spins.data(), h.spins.size());
  o += ", \"aos\": [";
  for (int a = 0; a < A; ++a) {
    const qmc_ao& q = h.aos[a];
    o += a ? ", " : "";
    o += std::string("{\"kind\": \"") +
         (q.kind == QMC_AO_GTO ? "gto" : q.kind == QMC_AO_STO ? "sto" : "slater1s") + "\", \"center\": " +
         std::to_string(q.center) + ", \"l\": " + std::to_string(q.l) + ", \"m\": " + std::to_string(q.m) +
         ", \"exp\": ";
    nums(h.pexp.data() + q.prim0, q.nprim);
    o += ", \"coef\": ";
    nums(h.pcoef.data() + q.prim0, q.nprim);
    o += ", \"pow\": ";
    ints(h.ppow.data() + q.prim0, q.nprim);
    o += '}';
  }
  o += "], \"mo_up\": ";
  rows(h.mo_up.data(), h.n_mo, A);
  o += ", \"mo_dn\": ";
  rows(h.mo_dn.data(), h.n_mo, A);
  o += std::string(", \"jastrow\": {\"kind\": \"") +
       (h.jas_kind == QMC_JAS_PAIR ? "pair" : h.jas_kind == QMC_JAS_PADE ? "pade" : "none") + "\", \"F\": ";
  num(h.F);
  o += ", \"b_ee\": ";
  num(h.b_ee);
  o += ", \"b_en\": ";
  num(h.b_en);
  o += ", \"a_anti\": ";
  num(h.a_anti);
  o += ", \"a_para\": ";
  num(h.a_para);
  if (h.k_ee > 0) {
    o += ", \"c_anti\": ";
    nums(h.c_anti.data(), h.k_ee);
    o += ", \"c_para\": ";
    nums(h.c_para.data(), h.k_ee);
  }
  if (h.k_en > 0) {
    o += ", \"species\": ";
    ints(h.species.data(), h.species.size());
    o += ", \"d_en\": ";
    rows(h.d_en.data(), h.n_species, h.k_en);
  }
  o += "}, \"init_nuc\": ";
  ints(h.init_nuc.data(), h.init_nuc.size());
  o += ", \"init_sigma\": ";
  num(h.init_sigma);
  o += ", \"flags\": " + std::to_string(h.flags);
  if (h.mo_rule == QMC_MO_BENZENE_CH) o += ", \"mo_rule\": \"benzene_ch\"";
  o += "}";
  if (needed) *needed = (int64_t)o.size() + 1;
  if (len == 0) return QMC_OK;  // size query
  if ((int64_t)o.size() + 1 > len) return fail(QMC_ERR_INVALID, "buffer too small");
  std::memcpy(buf, o.c_str(), o.size() + 1);
  return QMC_OK;
}

int qmc_geometry_optimize(qmc_system* sys, const qmc_force_params* fp, const qmc_geomopt_params* gp, uint64_t seed,
                          qmc_geomopt_results* res, double* energy_history, double* force_history,
                          double* geometry_history) {
  if (!sys || !fp || !gp || !res || gp->max_iterations < 0 || !(gp->max_displacement > 0.0) ||
      !(gp->force_threshold >= 0.0))
    return fail(QMC_ERR_INVALID, "bad arguments");
  const int M = sys->h.n_nuc();
  double qsum = 0.0;
  for (int I = 0; I < M; ++I) qsum += sys->h.Z[I];
  std::vector<double> R(sys->h.nuc), F(3 * M), fv(M);
  if (geometry_history) std::copy(R.begin(), R.end(), geometry_history);
  auto key = [&](int64_t k) { return seed + (uint64_t)k * 0x9E3779B97F4A7C15ull; };
  int rc;
  double e = 0.0, err = 0.0, fmax = 0.0;
  int64_t it = 0;
  res->converged = 0;
  for (; it < gp->max_iterations; ++it) {
    if ((rc = qmc_sample_forces(sys, fp, key(it), F.data(), &e, &err, fv.data()))) return rc;
    fmax = 0.0;
    for (int I = 0; I < M; ++I)
      fmax = std::max(fmax, std::sqrt(F[3 * I] * F[3 * I] + F[3 * I + 1] * F[3 * I + 1] + F[3 * I + 2] * F[3 * I + 2]));
    if (energy_history) energy_history[it] = e;
    if (force_history) force_history[it] = fmax;
    if (fmax < gp->force_threshold) {
      res->converged = 1;
      ++it;
      break;
    }
    // charge-weighted mean force (the charges serve as masses, geometry_opt.rs:283-289)
    double com[3] = {0.0, 0.0, 0.0};
    for (int I = 0; I < M; ++I)
      for (int c = 0; c < 3; ++c) com[c] += F[3 * I + c] * sys->h.Z[I];
    for (int c = 0; c < 3; ++c) com[c] /= qsum;
    for (int I = 0; I < M; ++I) {
      double d[3], n2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        d[c] = gp->step_size * (F[3 * I + c] - com[c]);
        n2 += d[c] * d[c];
      }
      const double n = std::sqrt(n2), sc = n > gp->max_displacement ? gp->max_displacement / n : 1.0;
      for (int c = 0; c < 3; ++c) R[3 * I + c] += d[c] * sc;
    }
    if ((rc = qmc_set_nuclei(sys, R.data(), M))) return rc;
    if (geometry_history) std::copy(R.begin(), R.end(), geometry_history + 3 * M * (it + 1));
  }
  res->iterations = it;
  // final energy at the final geometry (geometry_opt.rs:316-317)
  if ((rc = qmc_sample_forces(sys, fp, key(it), F.data(), &e, &err, fv.data()))) return rc;
  res->final_energy = e;
  res->final_max_force = it > 0 ? fmax : 0.0;
  return QMC_OK;
}

int qmc_ensemble_set_timing(qmc_ensemble* e, int32_t on) {
  if (!e) return fail(QMC_ERR_INVALID, "null ensemble");
  e->timing = on != 0;
  return QMC_OK;
}

int qmc_ensemble_timing(const qmc_ensemble* e, double* ms, int64_t* launches) {
  if (!e) return fail(QMC_ERR_INVALID, "null ensemble");
  if (ms) *ms = e->kernel_ms;
  if (launches) *launches = e->launches;
  return QMC_OK;
}

// vmc.rs:141-163
double qmc_autocorrelation_time(const double* e, int64_t n) {
  if (!e || n < 1) return 1.0;
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
    const double at = s / ((double)(n - t) * var);
    if (at < 0.0) break;
    ac += 2.0 * at;
  }
  return ac;
}

// vmc.rs:166-188
double qmc_blocking_error(const double* e, int64_t n, double tau) {
  const int64_t bs = (int64_t)std::ceil(2.0 * tau);
  if (!e || bs < 1) return 0.0;
  const int64_t nb = n / bs;
  if (nb < 2) return 0.0;
  std::vector<double> bm(nb);
  for (int64_t b = 0; b < nb; ++b) {
    double s = 0.0;
    for (int64_t i = b * bs; i < b * bs + bs; ++i) s += e[i];
    bm[b] = s / (double)bs;
  }
  double mean = 0.0;
  for (double v : bm) mean += v;
  mean /= (double)nb;
  double var = 0.0;
  for (double v : bm) var += (v - mean) * (v - mean);
  var /= (double)(nb - 1);
  return std::sqrt(var / (double)nb);
}

// MCMCSimulation::run (vmc.rs:69-92) + compute_results (vmc.rs:131-138)
int qmc_vmc_run(const qmc_system* sys, const qmc_mcmc_params* p, uint64_t seed, qmc_vmc_results* res,
                double* energy_trace) {
  if (!sys || !p || !res || p->n_walkers < 1 || p->n_steps < 1 || p->adaptation_interval < 1 ||
      p->n_burnin < 0 || p->n_burnin >= p->n_steps)
    return fail(QMC_ERR_INVALID, "bad arguments");
  qmc_ensemble* e = nullptr;
  int rc = qmc_ensemble_create(sys, p->n_walkers, seed, &e);
  if (rc) return rc;
  std::unique_ptr<qmc_ensemble> guard(e);
  const int N = sys->h.n_elec();
  const double per_step_moves = p->move == QMC_MOVE_ALL ? (double)p->n_walkers : (double)p->n_walkers * N;
  std::vector<double> energies(p->n_steps), es(p->adaptation_interval), acc(p->adaptation_interval);
  double step = p->initial_step_size, acc_total = 0.0;
  for (int64_t done = 0; done < p->n_steps;) {
    const int ns = (int)std::min<int64_t>(p->adaptation_interval, p->n_steps - done);
    qmc_sweep_params sp{p->move, ns, step, p->refresh_interval, 0};
    if ((rc = qmc_ensemble_sweeps(e, &sp, es.data(), nullptr, acc.data(), nullptr))) return rc;
    double a = 0.0;
    for (int i = 0; i < ns; ++i) {
      energies[done + i] = es[i] / (double)p->n_walkers;
      a += acc[i];
    }
    acc_total += a;
    done += ns;
    if (ns == p->adaptation_interval) {  // vmc.rs:85-88, 122-128
      const double rate = a / (per_step_moves * (double)p->adaptation_interval);
      step = std::min(std::max(step * std::sqrt(rate / p->target_acceptance), p->min_step_size), p->max_step_size);
    }
  }
  const double* e0 = energies.data() + p->n_burnin;
  const int64_t n = p->n_steps - p->n_burnin;
  double mean = 0.0;
  for (int64_t i = 0; i < n; ++i) mean += e0[i];
  res->energy = mean / (double)n;
  res->autocorrelation_time = qmc_autocorrelation_time(e0, n);
  res->error = qmc_blocking_error(e0, n, res->autocorrelation_time);
  res->acceptance = acc_total / (per_step_moves * (double)p->n_steps);
  res->final_step_size = step;
  if (energy_trace) std::copy(energies.begin(), energies.end(), energy_trace);
  return QMC_OK;
}

int qmc_ddvmc_run(const qmc_system* sys, const qmc_ddvmc_params* p, uint64_t seed, qmc_ddvmc_results* res,
                  double* energy_trace) {
  if (!sys || !p || !res || p->n_walkers < 1 || p->n_steps < 1 || p->n_burnin < 0 || p->adaptation_interval < 1 ||
      !(p->time_step > 0.0) || !(p->max_drift > 0.0) || !(p->target_acceptance > 0.0))
    return fail(QMC_ERR_INVALID, "bad arguments");
  qmc_ensemble* e = nullptr;
  int rc = qmc_ensemble_create(sys, p->n_walkers, seed, &e);
  if (rc) return rc;
  std::unique_ptr<qmc_ensemble> guard(e);
  const double moves_per_sweep = (double)p->n_walkers * sys->h.n_elec();
  double tau = p->time_step;
  auto adapt = [&](double rate) {  // DriftDiffusionVMC::adapt_time_step, vmc.rs:441-446
    const double adj = std::min(std::max(std::sqrt(rate / p->target_acceptance), 0.8), 1.2);
    tau = std::min(std::max(tau * adj, 0.001), 0.05);
  };
  const int64_t I = p->adaptation_interval;
  std::vector<double> es(I), acc(I), energies(p->n_steps);
  auto run = [&](int ns) {
    qmc_sweep_params sp{QMC_MOVE_DRIFT, ns, tau, p->refresh_interval, 0, p->max_drift};
    return qmc_ensemble_sweeps(e, &sp, es.data(), nullptr, acc.data(), nullptr);
  };
  // burn-in with adaptation every I sweeps (vmc.rs:462-476)
  double acc_b = 0.0, moves_b = 0.0;
  for (int64_t st = 0; st < p->n_burnin;) {
    const int ns = (int)std::min(I - st % I, p->n_burnin - st);
    if ((rc = run(ns))) return rc;
    for (int i = 0; i < ns; ++i) acc_b += acc[i];
    moves_b += ns * moves_per_sweep;
    st += ns;
    if (st % I == 0 && moves_b > 0.0) {
      adapt(acc_b / moves_b);
      acc_b = moves_b = 0.0;
    }
  }
  // production: mean E_L of every sweep, windowed adaptation (vmc.rs:478-507)
  double acc_p = 0.0, acc_w = 0.0, moves_w = 0.0;
  for (int64_t st = 0; st < p->n_steps;) {
    const int ns = (int)std::min(I - st % I, p->n_steps - st);
    if ((rc = run(ns))) return rc;
    for (int i = 0; i < ns; ++i) {
      energies[st + i] = es[i] / (double)p->n_walkers;
      acc_p += acc[i];
      acc_w += acc[i];
    }
    moves_w += ns * moves_per_sweep;
    st += ns;
    if (st % I == 0 && moves_w > 0.0) {
      adapt(acc_w / moves_w);
      acc_w = moves_w = 0.0;
    }
  }
  const int64_t n = p->n_steps;
  double mean = 0.0;
  for (double v : energies) mean += v;
  res->energy = mean / (double)n;
  res->autocorrelation_time = qmc_autocorrelation_time(energies.data(), n);
  res->error = qmc_blocking_error(energies.data(), n, res->autocorrelation_time);
  res->acceptance_rate = acc_p / (moves_per_sweep * (double)n);
  res->final_time_step = tau;
  if (energy_trace) std::copy(energies.begin(), energies.end(), energy_trace);
  return QMC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// heterogeneous batches
// ---------------------------------------------------------------------------

struct qmc_batch {
  std::vector<qmc_ensemble*> ens;
  std::vector<int> kind;  // fused-kernel segment kind per ensemble (-1: none)
  bool fusable = false;
  qmc_batch_config cfg{QMC_BATCH_FUSED, 0, 8, QMC_BATCH_ORDER_COST};
  hipStream_t st = nullptr;
  std::vector<hipStream_t> sst;  // one stream per ensemble (STREAMS mode)
  int* counter = nullptr;
  std::vector<hipEvent_t> ev;
  double kernel_ms = 0.0;
  int64_t launches = 0;
  ~qmc_batch() {
    if (st) (void)hipStreamDestroy(st);
    for (hipStream_t x : sst) (void)hipStreamDestroy(x);
    if (counter) (void)hipFree(counter);
    for (hipEvent_t x : ev) (void)hipEventDestroy(x);
  }
};

namespace {

int segment_kind(const qmc_ensemble* e) {
  const qmc_system* s = e->sys;
  if (e->prec != QMC_MIXED || s->h.k_ee != 0 || s->h.k_en != 0) return -1;
  if (s->tile_sweeps(e->prec)) return SEG_TILE_21_21;
  if (s->wave() || !s->ops || !s->ops->match) return -1;
  const std::string n = s->ops->name;
  if (n == "ch4-sto3g") return SEG_REG_CH4_STO3G;
  if (n == "h2-slater1s") return SEG_REG_H2;
  return -1;
}

// relative cost of one queue item of a segment kind (walkers x flop per walker-move)
double item_cost(int kind) {
  switch (kind) {
    case SEG_TILE_21_21: return 16.0 * 827643.0;
    case SEG_REG_CH4_STO3G: return 256.0 * 16338.0;
    default: return 256.0 * 524.0;
  }
}

hipEvent_t batch_event(qmc_batch* b, size_t i) {
  while (b->ev.size() <= i) {
    hipEvent_t x;
    if (hipEventCreate(&x) != hipSuccess) return nullptr;
    b->ev.push_back(x);
  }
  return b->ev[i];
}

// one chunk of ns sweeps of every ensemble with the current configuration
int batch_chunk(qmc_batch* b, int ns, const double* step_size, int done, size_t* nev) {
  const size_t n = b->ens.size();
  std::vector<SweepArgs> args(n);
  int rc;
  for (size_t i = 0; i < n; ++i) {
    qmc_ensemble* e = b->ens[i];
    if ((rc = ensure_cap(&e->partial, &e->partial_cap, (size_t)ns * sweep_blocks(e) * 3))) return rc;
    args[i] = SweepArgs{ns, e->step, e->seed, 0, step_size[i], e->W};
  }
  hipEvent_t e0 = batch_event(b, (*nev)++), e1 = batch_event(b, (*nev)++);
  if (!e0 || !e1) return fail(QMC_ERR_HIP, "hipEventCreate failed");
  HIP_TRY(hipEventRecord(e0, b->st));
  if (b->cfg.mode == QMC_BATCH_FUSED) {
    BatchLaunch L;
    L.nseg = (int)n;
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    if (b->cfg.order == QMC_BATCH_ORDER_COST)
      std::stable_sort(order.begin(), order.end(),
                       [&](size_t x, size_t y) { return item_cost(b->kind[x]) > item_cost(b->kind[y]); });
    L.start[0] = 0;
    for (size_t j = 0; j < n; ++j) {
      const size_t i = order[j];
      qmc_ensemble* e = b->ens[i];
      BatchSeg& G = L.seg[j];
      G.kind = b->kind[i];
      G.S = e->sys->d32;
      G.E = e->view<float>();
      G.A = args[i];
      G.partial = e->partial;
      G.nblk = sweep_blocks(e);
      G.items = G.kind == SEG_TILE_21_21 ? G.nblk : (G.nblk + 3) / 4;
      if (G.kind == SEG_TILE_21_21) L.tile = e->sys->tile;
      L.start[j + 1] = L.start[j] + G.items;
    }
    L.total = L.start[n];
    L.counter = b->counter;
    int per_cu = 0, dev = 0, cus = 0;
    HIP_TRY(batch_occupancy(batch_lds_bytes(L), &per_cu));
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (per_cu < 1) return fail(QMC_ERR_UNSUPPORTED, "the fused batch kernel does not fit a CU");
    const int bpc = b->cfg.blocks_per_cu > 0 ? std::min(b->cfg.blocks_per_cu, per_cu) : per_cu;
    const int64_t grid = std::min<int64_t>((int64_t)bpc * cus, L.total);
    HIP_TRY(hipMemsetAsync(b->counter, 0, sizeof(int), b->st));
    HIP_TRY(batch_launch(L, (int)std::max<int64_t>(grid, 1), b->st));
  } else {
    const bool par = b->cfg.mode == QMC_BATCH_STREAMS;
    for (size_t i = 0; i < n; ++i) {
      hipStream_t x = par ? b->sst[i] : b->st;
      if (par) HIP_TRY(hipStreamWaitEvent(x, e0, 0));
      if ((rc = enqueue_single_sweeps(b->ens[i], args[i], nullptr, x))) return rc;
      if (par) {
        hipEvent_t ei = batch_event(b, (*nev)++);
        if (!ei) return fail(QMC_ERR_HIP, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(ei, x));
        HIP_TRY(hipStreamWaitEvent(b->st, ei, 0));
      }
    }
  }
  HIP_TRY(hipEventRecord(e1, b->st));
  ++b->launches;
  for (size_t i = 0; i < n; ++i) {
    qmc_ensemble* e = b->ens[i];
    HIP_TRY(launch_reduce(e->partial, ns, (int)sweep_blocks(e), 3, e->sums + (size_t)done * 3, b->st));
    e->step += (uint32_t)ns;
    e->since_refresh += ns;
  }
  return QMC_OK;
}

int batch_run(qmc_batch* b, int32_t n_steps, const double* step_size, int32_t refresh, double* e_sum, double* e2_sum,
              double* accepted) {
  const size_t n = b->ens.size();
  int rc;
  // every ensemble's work goes to the batch stream while the batch runs
  std::vector<hipStream_t> saved(n);
  for (size_t i = 0; i < n; ++i) {
    saved[i] = b->ens[i]->st;
    HIP_TRY(hipStreamSynchronize(saved[i]));
    b->ens[i]->st = b->st;
  }
  struct Restore {
    qmc_batch* b;
    std::vector<hipStream_t>& s;
    ~Restore() {
      for (size_t i = 0; i < s.size(); ++i) b->ens[i]->st = s[i];
    }
  } restore{b, saved};
  for (qmc_ensemble* e : b->ens) {
    if (e->sys_version != e->sys->version && (rc = ens_refresh(e, 1))) return rc;
    if (e->inverse_stale && (rc = ens_refresh(e, 0))) return rc;
    if ((rc = ensure_cap(&e->sums, &e->sums_cap, (size_t)n_steps * 3))) return rc;
  }
  size_t nev = 0;
  b->launches = 0;
  const int chunk = b->cfg.steps_per_launch > 0 ? b->cfg.steps_per_launch : 8;
  for (int done = 0; done < n_steps;) {
    int ns = std::min(chunk, n_steps - done);
    if (refresh > 0)
      for (qmc_ensemble* e : b->ens) {
        if (e->since_refresh >= refresh && (rc = ens_refresh(e, 0))) return rc;
        ns = std::min(ns, refresh - e->since_refresh);
      }
    if ((rc = batch_chunk(b, ns, step_size, done, &nev))) return rc;
    done += ns;
    if (refresh > 0)
      for (qmc_ensemble* e : b->ens)
        if (e->since_refresh >= refresh && (rc = ens_refresh(e, 0))) return rc;
  }
  std::vector<double> h(n * (size_t)n_steps * 3);
  for (size_t i = 0; i < n; ++i)
    HIP_TRY(hipMemcpyAsync(h.data() + i * (size_t)n_steps * 3, b->ens[i]->sums, sizeof(double) * n_steps * 3,
                           hipMemcpyDeviceToHost, b->st));
  HIP_TRY(hipStreamSynchronize(b->st));
  b->kernel_ms = 0.0;
  for (size_t k = 0; k + 1 < nev;) {
    // chunk events come as (start, end, per-stream events ...); pair each start with its end
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[k], b->ev[k + 1]));
    b->kernel_ms += ms;
    k += 2 + (b->cfg.mode == QMC_BATCH_STREAMS ? n : 0);
  }
  for (size_t i = 0; i < n; ++i)
    for (int q = 0; q < 3; ++q) {
      double* dst = q == 0 ? e_sum : (q == 1 ? e2_sum : accepted);
      if (!dst) continue;
      for (int t = 0; t < n_steps; ++t) dst[i * (size_t)n_steps + t] = h[(i * (size_t)n_steps + t) * 3 + q];
    }
  return QMC_OK;
}

// Sherman-Morrison drift of the local energy after n sweeps from a fresh
// FP64 inverse: the E_L the sweeps computed minus the E_L of a fresh inverse
// at the same positions; mean and standard error over the walkers of each
// ensemble, the worst ensemble's |mean| + 2 sem returned in *bound (Ha).
int batch_drift(qmc_batch* b, const double* step_size, int n, double* bound) {
  int rc;
  for (qmc_ensemble* e : b->ens)
    if ((rc = ens_refresh(e, 0))) return rc;
  if ((rc = batch_run(b, n, step_size, 0, nullptr, nullptr, nullptr))) return rc;
  *bound = 0.0;
  for (qmc_ensemble* e : b->ens) {
    std::vector<double> a(e->W), f(e->W);
    HIP_TRY(hipStreamSynchronize(e->st));
    HIP_TRY(hipMemcpy(a.data(), e->eloc, sizeof(double) * e->W, hipMemcpyDeviceToHost));
    if ((rc = ens_refresh(e, 1))) return rc;
    HIP_TRY(hipStreamSynchronize(e->st));
    HIP_TRY(hipMemcpy(f.data(), e->eloc, sizeof(double) * e->W, hipMemcpyDeviceToHost));
    double m = 0.0, m2 = 0.0;
    for (int64_t w = 0; w < e->W; ++w) {
      const double d = a[w] - f[w];
      m += d;
      m2 += d * d;
    }
    m /= (double)e->W;
    const double var = std::max(0.0, m2 / (double)e->W - m * m);
    *bound = std::max(*bound, std::fabs(m) + 2.0 * std::sqrt(var / (double)e->W));
  }
  return QMC_OK;
}

// the longest refresh interval of 16 .. 256 (at most max_interval) whose
// drift bound stays under 1e-5 Ha in every ensemble
int batch_tune_refresh(qmc_batch* b, const double* step_size, int max_interval, int* picked) {
  const double guard = 1e-5;
  *picked = 16;
  for (int n : {16, 32, 64, 128, 256}) {
    if (n > max_interval) break;
    double bound = 0.0;
    if (int rc = batch_drift(b, step_size, n, &bound)) return rc;
    if (!(bound < guard)) break;
    *picked = n;
  }
  return QMC_OK;
}

}  // namespace

extern "C" {

int qmc_batch_create(qmc_ensemble* const* ens, int32_t n, qmc_batch** out) {
  if (!ens || !out || n < 1 || n > BATCH_MAXSEG) return fail(QMC_ERR_INVALID, "bad arguments (1 <= n <= 8)");
  auto b = std::make_unique<qmc_batch>();
  int tiles = 0;
  b->fusable = true;
  for (int i = 0; i < n; ++i) {
    if (!ens[i]) return fail(QMC_ERR_INVALID, "null ensemble");
    for (int j = 0; j < i; ++j)
      if (ens[j] == ens[i]) return fail(QMC_ERR_INVALID, "an ensemble appears twice");
    b->ens.push_back(ens[i]);
    const int k = segment_kind(ens[i]);
    b->kind.push_back(k);
    if (k < 0) b->fusable = false;
    if (k == SEG_TILE_21_21) ++tiles;
  }
  if (tiles > 1) b->fusable = false;  // one tile plan per fused launch
  if (!b->fusable) b->cfg.mode = QMC_BATCH_STREAMS;
  HIP_TRY(hipStreamCreate(&b->st));
  b->sst.resize(n);
  for (int i = 0; i < n; ++i) HIP_TRY(hipStreamCreate(&b->sst[i]));
  HIP_TRY(hipMalloc(&b->counter, sizeof(int)));
  *out = b.release();
  return QMC_OK;
}

int qmc_batch_destroy(qmc_batch* b) {
  if (b) (void)hipStreamSynchronize(b->st);
  delete b;
  return QMC_OK;
}

int qmc_batch_fusable(const qmc_batch* b, int32_t* fusable) {
  if (!b || !fusable) return fail(QMC_ERR_INVALID, "bad arguments");
  *fusable = b->fusable ? 1 : 0;
  return QMC_OK;
}

int qmc_batch_set_config(qmc_batch* b, const qmc_batch_config* c) {
  if (!b || !c) return fail(QMC_ERR_INVALID, "bad arguments");
  if (c->mode < QMC_BATCH_FUSED || c->mode > QMC_BATCH_SEQUENTIAL || c->blocks_per_cu < 0 ||
      c->steps_per_launch < 0 || c->refresh_interval < 0 ||
      (c->order != QMC_BATCH_ORDER_COST && c->order != QMC_BATCH_ORDER_GIVEN))
    return fail(QMC_ERR_INVALID, "bad batch configuration");
  if (c->mode == QMC_BATCH_FUSED && !b->fusable)
    return fail(QMC_ERR_UNSUPPORTED, "these ensembles cannot share the fused kernel");
  b->cfg = *c;
  return QMC_OK;
}

int qmc_batch_get_config(const qmc_batch* b, qmc_batch_config* c) {
  if (!b || !c) return fail(QMC_ERR_INVALID, "bad arguments");
  *c = b->cfg;
  return QMC_OK;
}

int qmc_batch_sweeps(qmc_batch* b, int32_t n_steps, const double* step_size, int32_t refresh_interval, double* e_sum,
                     double* e2_sum, double* accepted) {
  if (!b || !step_size || n_steps < 0) return fail(QMC_ERR_INVALID, "bad arguments");
  for (size_t i = 0; i < b->ens.size(); ++i)
    if (!(step_size[i] >= 0.0)) return fail(QMC_ERR_INVALID, "bad step size");
  if (n_steps == 0) return QMC_OK;
  if (refresh_interval < 0) refresh_interval = b->cfg.refresh_interval > 0 ? b->cfg.refresh_interval : 64;
  return batch_run(b, n_steps, step_size, refresh_interval, e_sum, e2_sum, accepted);
}

int qmc_batch_timing(const qmc_batch* b, double* kernel_ms, int64_t* launches) {
  if (!b) return fail(QMC_ERR_INVALID, "null batch");
  if (kernel_ms) *kernel_ms = b->kernel_ms;
  if (launches) *launches = b->launches;
  return QMC_OK;
}

int qmc_batch_tune(qmc_batch* b, int32_t trial_steps, const double* step_size, int32_t refresh_interval,
                   qmc_batch_config* best, double* best_ms_per_sweep) {
  if (!b || !step_size || trial_steps < 1 || refresh_interval == 0) return fail(QMC_ERR_INVALID, "bad arguments");
  int rc;
  int refresh = refresh_interval;
  if (refresh_interval < 0 && (rc = batch_tune_refresh(b, step_size, -refresh_interval, &refresh))) return rc;
  std::vector<qmc_batch_config> cand;
  for (int spl : {4, 8, 16}) {
    if (b->fusable)
      for (int bpc : {1, 2})
        for (int order : {QMC_BATCH_ORDER_COST, QMC_BATCH_ORDER_GIVEN})
          cand.push_back(qmc_batch_config{QMC_BATCH_FUSED, bpc, spl, order, 0});
    cand.push_back(qmc_batch_config{QMC_BATCH_STREAMS, 0, spl, QMC_BATCH_ORDER_GIVEN, 0});
    cand.push_back(qmc_batch_config{QMC_BATCH_SEQUENTIAL, 0, spl, QMC_BATCH_ORDER_GIVEN, 0});
  }
  for (auto& c : cand) c.refresh_interval = refresh_interval < 0 ? refresh : 0;
  refresh_interval = refresh;
  const qmc_batch_config keep = b->cfg;
  double best_ms = std::numeric_limits<double>::infinity();
  qmc_batch_config pick = keep;
  for (const auto& c : cand) {
    b->cfg = c;
    if ((rc = batch_run(b, c.steps_per_launch, step_size, refresh_interval, nullptr, nullptr, nullptr))) {
      b->cfg = keep;
      return rc;
    }
    const int steps = std::max(trial_steps, c.steps_per_launch);
    if ((rc = batch_run(b, steps, step_size, refresh_interval, nullptr, nullptr, nullptr))) {
      b->cfg = keep;
      return rc;
    }
    const double ms = b->kernel_ms / steps;
    if (ms < best_ms) {
      best_ms = ms;
      pick = c;
    }
  }
  b->cfg = pick;
  if (best) *best = pick;
  if (best_ms_per_sweep) *best_ms_per_sweep = best_ms;
  return QMC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// importance-sampled DMC (ImportanceSampledDMC, is_dmc.rs:176-607)
// ---------------------------------------------------------------------------

namespace {

constexpr uint64_t ISDMC_VMC_KEY = 0xD1B54A32D192ED03ull;

// device allocations of one run, freed together
struct DevArena {
  std::vector<void*> ptrs;
  ~DevArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class P>
  int get(P** p, size_t bytes) {
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, bytes > 0 ? bytes : 1));
    ptrs.push_back(q);
    *p = (P*)q;
    return QMC_OK;
  }
};

struct DmcHost {
  double step_ms = 0.0;
  int64_t collapses = 0;
  std::vector<double> e_trace, pop_trace, acc;  // [n_burnin + n_steps]
  std::vector<char> fresh;                      // population (re)initialised before the step
  DmcCtl ctl{};
};

// phases 2-4 of is_dmc.rs:376-546 on a population of capacity cap
template <class T>
int isdmc_steps(const qmc_system* s, const qmc_isdmc_params* p, uint64_t seed, int64_t cap, DmcHost* out) {
  const int N = s->h.n_elec(), M = s->h.n_nuc(), nu = s->h.n_up(), nd = s->h.n_dn();
  const int nainv = nu * nu + nd * nd;
  const int64_t T_all = p->n_burnin + p->n_steps;
  // sweep partial rows: one per 64 walkers (register engine) or per walker (wave engine)
  const bool wave = s->wave();
  const int nb = (int)(cap / DMC_BLOCK), sweep_bs = wave ? 1 : 64, sweep_rows = (int)(cap / sweep_bs);
  const bool f64 = sizeof(T) == 8;
  // a stream of its own (graphs cannot capture the null stream) and the
  // two-step graphs, one per parity of the buffer a pair of steps starts in
  hipStream_t st = nullptr;
  hipGraphExec_t gx[2] = {nullptr, nullptr};
  HIP_TRY(hipStreamCreate(&st));
  struct Release {
    hipStream_t& st;
    hipGraphExec_t* gx;
    ~Release() {
      for (int i = 0; i < 2; ++i)
        if (gx[i]) (void)hipGraphExecDestroy(gx[i]);
      (void)hipStreamDestroy(st);
    }
  } release{st, gx};
  DevArena A;
  int rc;
  T* pos[2];
  T* ainv[2];
  double *logdet[2], *sign[2], *ju[2], *eloc[2], *eprev[2];
  for (int b = 0; b < 2; ++b) {
    if ((rc = A.get(&pos[b], sizeof(T) * 3 * N * cap)) || (rc = A.get(&ainv[b], sizeof(T) * std::max(1, nainv) * cap)))
      return rc;
    for (double** q : {&logdet[b], &sign[b], &ju[b], &eloc[b], &eprev[b]})
      if ((rc = A.get(q, sizeof(double) * cap))) return rc;
  }
  DmcBufs B{};
  double *partial = nullptr, *acc_dev = nullptr;
  if ((rc = A.get(&B.ctl, sizeof(DmcCtl))) || (rc = A.get(&B.copies, sizeof(int) * cap)) ||
      (rc = A.get(&B.loff, sizeof(int) * cap)) || (rc = A.get(&B.bsum, sizeof(int64_t) * nb)) ||
      (rc = A.get(&B.part, sizeof(double) * nb * (1 + 3 * M))) || (rc = A.get(&B.e_trace, sizeof(double) * T_all)) ||
      (rc = A.get(&B.pop_trace, sizeof(double) * T_all)) || (rc = A.get(&partial, sizeof(double) * 3 * sweep_rows)) ||
      (rc = A.get(&acc_dev, sizeof(double) * T_all)))
    return rc;
  HIP_TRY(hipMemsetAsync(B.ctl, 0, sizeof(DmcCtl), st));
  EnsView<T> V[2];
  for (int b = 0; b < 2; ++b)
    V[b] = EnsView<T>{cap, cap, pos[b], ainv[b], logdet[b], sign[b], ju[b], eloc[b], &B.ctl->n_live[b]};
  const double tau_b = p->matched_branching ? 2.0 * p->time_step : p->time_step;
  const DmcArgs a{3 * N, nainv, M, cap, tau_b, p->e_ref_damping, (double)p->n_walkers, seed, p->n_burnin};
  // a fresh population of n_walkers in buffer b from walker ids w0 + w
  auto init_pop = [&](int b, int64_t w0) -> int {
    EnsView<T> E = V[b];
    E.W = p->n_walkers;
    E.n_live = nullptr;
    if constexpr (sizeof(T) == 8) {
      HIP_TRY(launch_init64(s->d64, E, seed, w0, st));
      if (wave)
        HIP_TRY(wave_build<double>(s->d64, s->d64, E, s->shape(), 1, nullptr, st));
      else
        HIP_TRY(s->ops->build64(s->d64, s->d64, E, 1, nullptr, st));
    } else {
      HIP_TRY(launch_init32(s->d64, E, seed, w0, st));
      if (wave)
        HIP_TRY(wave_build<float>(s->d64, s->d32, E, s->shape(), 1, nullptr, st));
      else
        HIP_TRY(s->ops->build32(s->d64, s->d32, E, 1, nullptr, st));
    }
    HIP_TRY(dmc_reset(B.ctl, b, p->n_walkers, eloc[b], eprev[b], st));
    return QMC_OK;
  };
  out->e_trace.assign(T_all, 0.0);
  out->pop_trace.assign(T_all, 0.0);
  out->fresh.assign(T_all, 0);
  std::vector<int> cur_after(T_all);
  int cur = 0;
  if ((rc = init_pop(cur, 0))) return rc;
  out->fresh[0] = 1;
  const int R = f64 ? 0 : std::max(0, p->refresh_interval);
  int since = 0;
  hipEvent_t ev0, ev1;
  HIP_TRY(hipEventCreate(&ev0));
  HIP_TRY(hipEventCreate(&ev1));
  HIP_TRY(hipEventRecord(ev0, st));
  const int64_t chunk = 64;
  // One step: the drift-diffusion sweep of buffer c, then branching into
  // 1 - c.  In a graph (from_ctl) the step, trace slot and force flag come
  // from ctl->step, which each step's finalize advances.
  auto enqueue_step = [&](int c, int64_t u, bool from_ctl) -> int {
    SweepArgs sa{1, (uint32_t)u, seed, 0, p->time_step, 0, p->max_drift};
    if (from_ctl) sa.step0_dev = &B.ctl->step;
    if constexpr (sizeof(T) == 8) {
      if (wave)
        HIP_TRY(wave_dd<double>(s->d64, V[c], s->shape(), sa, partial, nullptr, st));
      else
        HIP_TRY((s->no_series() ? s->ops->dd64_ns : s->ops->dd64)(s->d64, V[c], sa, partial, nullptr, st));
    } else {
      if (wave)
        HIP_TRY(wave_dd<float>(s->d32, V[c], s->shape(), sa, partial, nullptr, st));
      else
        HIP_TRY((s->no_series() ? s->ops->dd32_ns : s->ops->dd32)(s->d32, V[c], sa, partial, nullptr, st));
    }
    const int n = 1 - c;
    HIP_TRY(dmc_branch_step<T>(s->d64, a, B, c, V[c], eprev[c], V[n], eprev[n],
                               from_ctl ? DMC_STEP_CTL : (uint32_t)u, from_ctl ? DMC_TS_CTL : u,
                               from_ctl ? DMC_FORCE_CTL : (u >= p->n_burnin ? 1 : 0), partial, sweep_rows, sweep_bs,
                               acc_dev, st));
    return QMC_OK;
  };
  // pairs of steps replay a captured graph (QMC_DMC_GRAPH=0 keeps every
  // step an individual launch sequence; the results are the same)
  const char* genv = std::getenv("QMC_DMC_GRAPH");
  const bool use_graph = !(genv && std::strcmp(genv, "0") == 0);
  auto capture = [&](int c) -> int {
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    int rc1 = enqueue_step(c, 0, true);
    if (!rc1) rc1 = enqueue_step(1 - c, 0, true);
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (rc1) {
      if (ec == hipSuccess && g) (void)hipGraphDestroy(g);
      return rc1;
    }
    HIP_TRY(ec);
    const hipError_t ei = hipGraphInstantiate(&gx[c], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(ei);
    return QMC_OK;
  };
  int64_t t = 0;
  while (t < T_all) {
    const int64_t t_end = std::min(T_all, t + chunk);
    HIP_TRY(dmc_set_step(B.ctl, (uint32_t)t, st));  // the device step follows t
    for (int64_t u = t; u < t_end;) {
      const bool refresh_due = R > 0 && since >= R;
      if (use_graph && u + 1 < t_end && !refresh_due && !(R > 0 && since + 1 >= R)) {
        if (!gx[cur] && (rc = capture(cur))) return rc;
        HIP_TRY(hipGraphLaunch(gx[cur], st));
        cur_after[u] = 1 - cur;
        cur_after[u + 1] = cur;
        since += 2;
        u += 2;
        continue;
      }
      if constexpr (sizeof(T) == 4) {
        if (refresh_due) {
          if (wave)
            HIP_TRY(wave_build<float>(s->d64, s->d32, V[cur], s->shape(), 0, nullptr, st));
          else
            HIP_TRY(s->ops->build32(s->d64, s->d32, V[cur], 0, nullptr, st));
          since = 0;
        }
      }
      if ((rc = enqueue_step(cur, u, false))) return rc;
      cur = 1 - cur;
      cur_after[u] = cur;
      ++since;
      ++u;
    }
    HIP_TRY(hipMemcpyAsync(out->pop_trace.data() + t, B.pop_trace + t, sizeof(double) * (t_end - t),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int64_t dead = -1;
    for (int64_t u = t; u < t_end && dead < 0; ++u)
      if (out->pop_trace[u] == 0.0) dead = u;
    if (dead < 0) {
      t = t_end;
      continue;
    }
    // the population died out at step dead: start again from fresh walkers
    // (is_dmc.rs:424-433, 492-501) and redo the steps queued after it
    ++out->collapses;
    cur = cur_after[dead];
    if ((rc = init_pop(cur, out->collapses * p->n_walkers))) return rc;
    since = 0;
    t = dead + 1;
    if (t < T_all) out->fresh[t] = 1;
  }
  HIP_TRY(hipEventRecord(ev1, st));
  HIP_TRY(hipMemcpyAsync(out->e_trace.data(), B.e_trace, sizeof(double) * T_all, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(out->pop_trace.data(), B.pop_trace, sizeof(double) * T_all, hipMemcpyDeviceToHost, st));
  out->acc.resize(T_all);
  HIP_TRY(hipMemcpyAsync(out->acc.data(), acc_dev, sizeof(double) * T_all, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&out->ctl, B.ctl, sizeof(DmcCtl), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  out->step_ms = ms;
  return QMC_OK;
}

}  // namespace

extern "C" {

// ImportanceSampledDMC::block_error (is_dmc.rs:555-607)
double qmc_isdmc_block_error(const double* e, int64_t n) {
  if (!e || n < 20) return 0.0;
  double mean = 0.0;
  for (int64_t i = 0; i < n; ++i) mean += e[i];
  mean /= (double)n;
  double var = 0.0;
  for (int64_t i = 0; i < n; ++i) var += (e[i] - mean) * (e[i] - mean);
  var /= (double)n;
  if (var < 1e-30) return 0.0;
  const double ac = qmc_autocorrelation_time(e, n);  // the same positive-sequence sum (is_dmc.rs:569-580)
  const int64_t bs = (int64_t)std::ceil(2.0 * ac);
  const int64_t nb = n / bs;
  if (nb < 2) return std::sqrt(var * ac / (double)n);
  std::vector<double> bm(nb);
  for (int64_t b = 0; b < nb; ++b) {
    double sum = 0.0;
    for (int64_t i = b * bs; i < b * bs + bs; ++i) sum += e[i];
    bm[b] = sum / (double)bs;
  }
  double bmean = 0.0;
  for (double v : bm) bmean += v;
  bmean /= (double)nb;
  double bvar = 0.0;
  for (double v : bm) bvar += (v - bmean) * (v - bmean);
  bvar /= (double)(nb - 1);
  return std::sqrt(bvar / (double)nb);
}

int qmc_isdmc_run(const qmc_system* sys, const qmc_isdmc_params* p, uint64_t seed, qmc_isdmc_results* res,
                  double* forces_mixed, double* forces_vmc, double* forces_extrapolated, double* energy_trace,
                  double* population_trace) {
  if (!sys || !p || !res || p->n_walkers < 1 || p->n_steps < 1 || p->n_burnin < 0 || !(p->time_step > 0.0) ||
      !(p->max_drift > 0.0) || p->n_vmc_equilibrate < 0 || p->n_vmc_force_samples < 0 || p->capacity < 0 ||
      p->refresh_interval < 0 || p->n_burnin + p->n_steps > INT32_MAX || p->n_vmc_equilibrate > INT32_MAX ||
      (p->matched_branching != 0 && p->matched_branching != 1))
    return fail(QMC_ERR_INVALID, "bad arguments");
  int rc = need_engine(sys);
  if (rc) return rc;
  const int M = sys->h.n_nuc();
  if (M > MAXN) return fail(QMC_ERR_UNSUPPORTED, "too many nuclei");
  int64_t cap = p->capacity > 0 ? p->capacity : 4 * p->n_walkers;
  cap = (cap + DMC_BLOCK - 1) / DMC_BLOCK * DMC_BLOCK;
  if ((p->capacity > 0 && p->capacity < p->n_walkers) || cap > (int64_t)DMC_BLOCK * DMC_MAX_BLOCKS)
    return fail(QMC_ERR_INVALID, "capacity must hold n_walkers and be at most 1048576");
  const std::vector<double> fnn = hf_nn(sys->h);
  // phase 1: VMC forces (is_dmc.rs:305-370)
  std::vector<double> fv(3 * M, 0.0);
  {
    const int64_t nv = std::min<int64_t>(p->n_walkers, 50);
    qmc_ensemble* ev = nullptr;
    if ((rc = qmc_ensemble_create(sys, nv, seed ^ ISDMC_VMC_KEY, &ev))) return rc;
    std::unique_ptr<qmc_ensemble> guard(ev);
    qmc_sweep_params sp{QMC_MOVE_ALL, (int32_t)p->n_vmc_equilibrate, 0.5, 0, 0, 0.0};
    if (p->n_vmc_equilibrate > 0 && (rc = qmc_ensemble_sweeps(ev, &sp, nullptr, nullptr, nullptr, nullptr)))
      return rc;
    const int64_t nfb = std::max<int64_t>(p->n_vmc_force_samples / nv, 1);
    DevArena A;
    double *part = nullptr, *vout = nullptr;
    if ((rc = A.get(&part, sizeof(double) * 3 * M)) || (rc = A.get(&vout, sizeof(double) * 3 * M * nfb))) return rc;
    sp.n_steps = 5;  // decorrelation sweeps between force samples (is_dmc.rs:334-353)
    for (int64_t it = 0; it < nfb; ++it) {
      if ((rc = qmc_ensemble_sweeps(ev, &sp, nullptr, nullptr, nullptr, nullptr))) return rc;
      int nblk = 0;
      if (ev->prec == QMC_FP64)
        HIP_TRY(hf_partials<double>(sys->d64, ev->view<double>(), nv, 3 * sys->h.n_elec(), M, part, &nblk, ev->st));
      else
        HIP_TRY(hf_partials<float>(sys->d64, ev->view<float>(), nv, 3 * sys->h.n_elec(), M, part, &nblk, ev->st));
      HIP_TRY(launch_reduce(part, 1, nblk, 3 * M, vout + 3 * M * it, ev->st));
    }
    std::vector<double> h(3 * M * nfb);
    HIP_TRY(hipMemcpyAsync(h.data(), vout, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ev->st));
    HIP_TRY(hipStreamSynchronize(ev->st));
    for (int64_t it = 0; it < nfb; ++it)
      for (int q = 0; q < 3 * M; ++q) fv[q] += h[3 * M * it + q];
    for (int q = 0; q < 3 * M; ++q) fv[q] = fv[q] / (double)(nv * nfb) + fnn[q];
  }
  // phases 2-4: the branching population
  DmcHost H;
  rc = sys->precision == QMC_FP64 ? isdmc_steps<double>(sys, p, seed, cap, &H)
                                  : isdmc_steps<float>(sys, p, seed, cap, &H);
  if (rc) return rc;
  // phase 5: results (is_dmc.rs:510-545)
  const int64_t n = p->n_steps, b0 = p->n_burnin;
  const double* E = H.e_trace.data() + b0;
  double sum_e = 0.0, sum_pop = 0.0, acc = 0.0, moves = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    sum_e += E[i];
    sum_pop += H.pop_trace[b0 + i];
  }
  const double n_e = (double)sys->h.n_elec();
  for (int64_t u = b0; u < b0 + n; ++u) {
    const double before = H.fresh[u] ? (double)p->n_walkers : H.pop_trace[u - 1];
    acc += H.acc[u];
    moves += before * n_e;
  }
  res->energy = sum_e / (double)n;
  res->error = qmc_isdmc_block_error(E, n);
  res->acceptance_rate = moves > 0.0 ? acc / moves : 0.0;
  res->avg_population = sum_pop / (double)n;
  res->collapses = H.collapses;
  res->dropped = H.ctl.dropped;
  res->capacity = cap;
  res->step_ms = H.step_ms;
  std::vector<double> fm(3 * M);
  for (int q = 0; q < 3 * M; ++q) fm[q] = H.ctl.force_sum[q] / (double)H.ctl.n_force + fnn[q];
  for (int q = 0; q < 3 * M; ++q) {
    if (forces_mixed) forces_mixed[q] = fm[q];
    if (forces_vmc) forces_vmc[q] = fv[q];
    if (forces_extrapolated) forces_extrapolated[q] = 2.0 * fm[q] - fv[q];
  }
  if (energy_trace) std::copy(E, E + n, energy_trace);
  if (population_trace) std::copy(H.pop_trace.begin() + b0, H.pop_trace.end(), population_trace);
  return QMC_OK;
}

}  // extern "C"
