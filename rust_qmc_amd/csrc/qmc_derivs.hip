// qmc_derivs.hip -- per-walker derivatives for stochastic reconfiguration
// and forces, and their ensemble moments.
//
// derivs_kernel writes, for every walker, one column of the feature matrix
// Y [Qp][ld] (feature-major, walker fastest):
//   O_k = d ln psi / d p_k        (Jastrow parameters, methane.rs:1212-1219
//                                  plus the Pade series)          cols [0, P)
//   Omega_I = d ln psi / d R_I    (methane.rs:990-1101)             [P, P+3M)
//   E_L                                                              P+3M
//   1                                                                P+3M+1
//   d E_L / d p_k                                                    [P+3M+2, 2P+3M+2)
//   F_I electron part of hellmann_feynman_force (methane.rs:1247-1257)
//                                                                    [2P+3M+2, 2P+6M+2)
// gram_kernel then forms G = Y Y^T with FP64 MFMA, which holds every mean,
// covariance and cross moment SR (sr_optimize.rs:176-246) and the ZV force
// estimator (geometry_opt.rs:174-226) need.
#include "qmc_derivs.hpp"

#include <type_traits>

#include "qmc_engine.hpp"
#include "qmc_wave.hpp"
#include "qmc_wave_dev.hpp"

namespace qmc {

struct DerivLayout {
  int P, M, N, Qp;
  __device__ __forceinline__ int o(int k) const { return k; }
  __device__ __forceinline__ int om(int I, int c) const { return P + 3 * I + c; }
  __device__ __forceinline__ int e() const { return P + 3 * M; }
  __device__ __forceinline__ int one() const { return P + 3 * M + 1; }
  __device__ __forceinline__ int de(int k) const { return P + 3 * M + 2 + k; }
  __device__ __forceinline__ int f(int I, int c) const { return 2 * P + 3 * M + 2 + 3 * I + c; }
  __device__ __forceinline__ int g(int q, int c) const { return Qp + 3 * q + c; }
};

// per-thread accumulator rows in LDS: element q of lane t at lds[q*64 + t],
// in the arithmetic type (FP32 rows take half the LDS, so twice the
// workgroups fit a CU); E_L is kept in FP64 outside the rows
template <class T>
struct Row {
  T* base;
  __device__ __forceinline__ T& operator[](int q) const { return base[q * 64]; }
};

// Omega and HF-force sums per nucleus.  With a compiled shell structure
// (NM > 0 nuclei, indices known at compile time) they stay in registers
// over all electrons and reach the LDS rows once; otherwise every term goes
// to the rows.
template <class T, int NM>
struct NucAcc {
  T om[NM > 0 ? NM : 1][3], f[NM > 0 ? NM : 1][3];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int I = 0; I < (NM > 0 ? NM : 1); ++I)
#pragma unroll
      for (int c = 0; c < 3; ++c) om[I][c] = f[I][c] = T(0);
  }
  template <class RowT>
  __device__ __forceinline__ void add_om(RowT row, const DerivLayout& L, int I, int c, T v) {
    if constexpr (NM > 0)
      om[I][c] += v;
    else
      row[L.om(I, c)] += v;
  }
  template <class RowT>
  __device__ __forceinline__ void add_f(RowT row, const DerivLayout& L, int I, int c, T v) {
    if constexpr (NM > 0)
      f[I][c] += v;
    else
      row[L.f(I, c)] += v;
  }
  // the register sums into the rows: all reads first, then all writes
  template <class RowT>
  __device__ __forceinline__ void flush(RowT row, const DerivLayout& L) {
    if constexpr (NM > 0) {
      T a[NM][6];
#pragma unroll
      for (int I = 0; I < NM; ++I)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          a[I][c] = row[L.om(I, c)];
          a[I][3 + c] = row[L.f(I, c)];
        }
#pragma unroll
      for (int I = 0; I < NM; ++I)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          row[L.om(I, c)] = a[I][c] + om[I][c];
          row[L.f(I, c)] = a[I][3 + c] + f[I][c];
        }
    }
  }
};

// Pass 1 for the electron in slot 0 of sector A (combined index q):
// determinant and Jastrow gradients, E_L terms, Omega, HF force.
template <int NA, int NB, class T, class Bs, int NM>
__device__ __forceinline__ double electron_pass(const DevSys<T>& S, int spin, int q, const T ra[Cap<NA>::v][3],
                                                const T ai[Cap<NA>::v][Cap<NA>::v], const T rb[Cap<NB>::v][3],
                                                const DerivLayout& L, Row<T> row, NucAcc<T, NM>& acc, double& pot) {
  const T x = ra[0][0], y = ra[0][1], z = ra[0][2];
  T gdx, gdy, gdz, ld;
  auto on_shell = [&](int c, T sx, T sy, T sz) {
    acc.add_om(row, L, c, 0, -sx);
    acc.add_om(row, L, c, 1, -sy);
    acc.add_om(row, L, c, 2, -sz);
  };
  det_terms<NA, T, decltype(on_shell), Bs>(S, spin, x, y, z, ai, gdx, gdy, gdz, ld, on_shell);
  T gux = 0, guy = 0, guz = 0, lu = 0, pairsq = 0;
  double pe = 0.0;
  auto ee = [&](T px, T py, T pz, bool same) {
    T dx = x - px, dy = y - py, dz = z - pz;
    T r, ir;
    M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
    T u1, lp;
    uee_d(S, r, ir, same, u1, lp);
    if (r >= T(1e-10)) {
      T f = u1 * ir;
      gux += f * dx;
      guy += f * dy;
      guz += f * dz;
      lu += lp;
      pairsq += u1 * u1;
    }
    pe += 0.5 * (double)ir;
  };
#pragma unroll
  for (int j = 1; j < NA; ++j) ee(ra[j][0], ra[j][1], ra[j][2], true);
#pragma unroll
  for (int j = 0; j < NB; ++j) ee(rb[j][0], rb[j][1], rb[j][2], false);
  Bs::nuclei(S, [&](int I) {
    T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
    T r, ir;
    M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
    pe -= (double)(S.Z[I] * ir);
    if (r > T(1e-10)) {  // methane.rs:1253
      const T f3 = S.Z[I] * ir * ir * ir;
      acc.add_f(row, L, I, 0, f3 * dx);
      acc.add_f(row, L, I, 1, f3 * dy);
      acc.add_f(row, L, I, 2, f3 * dz);
    }
    if (S.en_on) {
      T u1, lp;
      uen_d(S, r, ir, I, u1, lp);
      if (r >= T(1e-10)) {
        T f = u1 * ir;
        gux += f * dx;
        guy += f * dy;
        guz += f * dz;
        lu += lp;
        // d u_en / d R_I = -u1 (r_i - R_I)/r   (methane.rs:1071-1090)
        acc.add_om(row, L, I, 0, -f * dx);
        acc.add_om(row, L, I, 1, -f * dy);
        acc.add_om(row, L, I, 2, -f * dz);
      }
    }
  });
  const T gs = S.grad_scale;
  T gjx = gs * gux, gjy = gs * guy, gjz = gs * guz;
  T lj = (S.flags & F_J2_LAP_QUIRK) ? (lu + pairsq) : (lu + gux * gux + guy * guy + guz * guz);
  const T lap = ld + lj + T(2) * (gdx * gjx + gdy * gjy + gdz * gjz);
  row[L.g(q, 0)] = gdx + gux;
  row[L.g(q, 1)] = gdy + guy;
  row[L.g(q, 2)] = gdz + guz;
  pot += pe;
  return -0.5 * (double)lap;
}

// O_k and d E_L / d p_k terms of one electron pair (Pade Jastrow only),
// handed to addO(slot, v) and addE(slot, v) by parameter slot: 0 = b_ee,
// 1 = a (a_para for same spins, a_anti otherwise), 2 + k = series term k.
// (dx, dy, dz) = r_i - r_j and gd(c) = G_i[c] - G_j[c], the difference of
// the two electrons' d ln psi / d r (read only when the pair is apart).
template <class T, class GD, class AO, class AE>
__device__ __forceinline__ void ee_pair_terms(const DevSys<T>& S, T dx, T dy, T dz, bool same, GD gd, AO addO,
                                              AE addE) {
  const int K = S.k_ee;
  T r, ir;
  M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
  const T a = same ? S.a_para : S.a_anti;
  const T* c = same ? S.c_para : S.c_anti;
  PairDerivs<T> pd;
  pd.init(r, S.b_ee, a, K, c);
  // O: d u / d b_ee = -(a + s') x^2, d u / d a = x, d u / d c_k = x^(k+2)
  addO(0, -pd.A1 * pd.x * pd.x);
  addO(1, pd.x);
  if (r < T(1e-10)) return;
  // Delta = (G_i - G_j) . rhat_ij
  const T dlt = gd(0) * (dx * ir) + gd(1) * (dy * ir) + gd(2) * (dz * ir);
  T du1, dl;
  pd.d_b(r, ir, S.b_ee, du1, dl);
  addE(0, -(dl + dlt * du1));
  pd.d_a(ir, S.b_ee, du1, dl);
  addE(1, -(dl + dlt * du1));
  T xk = T(1);
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {  // unrolled, so the slot is a constant
    if (k >= K) break;
    addO(2 + k, xk * pd.x * pd.x);
    pd.d_c(k, ir, S.b_ee, xk, du1, dl);
    addE(2 + k, -(dl + dlt * du1));
    xk *= pd.x;
  }
}

// The same for electron i and nucleus I (slot 0 = b_en, 2 + k = the term k
// of the nucleus' species); (dx, dy, dz) = r_i - R_I and g(c) = G_i[c].
template <class T, class GI, class AO, class AE>
__device__ __forceinline__ void en_pair_terms(const DevSys<T>& S, int I, T dx, T dy, T dz, GI g, AO addO, AE addE) {
  const int Ke = S.k_en;
  T r, ir;
  M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
  const T* c = S.d_en[S.species[I]];
  PairDerivs<T> pd;
  pd.init(r, S.b_en, -S.Z[I], Ke, c);
  // d u_en / d b_en = (Z - s') y^2 = -A1 y^2
  addO(0, -pd.A1 * pd.x * pd.x);
  if (r < T(1e-10)) {
    T xk = T(1);
#pragma unroll
    for (int k = 0; k < MAXK; ++k) {
      if (k >= Ke) break;
      addO(2 + k, xk * pd.x * pd.x);
      xk *= pd.x;
    }
    return;
  }
  const T dlt = g(0) * (dx * ir) + g(1) * (dy * ir) + g(2) * (dz * ir);
  T du1, dl;
  pd.d_b(r, ir, S.b_en, du1, dl);
  addE(0, -(T(0.5) * dl + dlt * du1));
  T xk = T(1);
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {
    if (k >= Ke) break;
    addO(2 + k, xk * pd.x * pd.x);
    pd.d_c(k, ir, S.b_en, xk, du1, dl);
    addE(2 + k, -(T(0.5) * dl + dlt * du1));
    xk *= pd.x;
  }
}

// feature columns of a pair slot: e-e (class 0 anti, 1 para) and e-n
// (species sp) Jastrow parameters in the order [b_ee, b_en, a_anti,
// a_para, c_anti[K], c_para[K], d_en[species][Ke]]
template <class T>
__device__ __forceinline__ int ee_slot_col(const DevSys<T>& S, int cls, int slot) {
  return slot == 0 ? 0 : (slot == 1 ? 2 + cls : 4 + cls * S.k_ee + slot - 2);
}
template <class T>
__device__ __forceinline__ int en_slot_col(const DevSys<T>& S, int sp, int slot) {
  return slot == 0 ? 1 : 4 + 2 * S.k_ee + sp * S.k_en + slot - 2;
}

// Pass 2: unique pairs -> O_k and d E_L / d p_k (Pade Jastrow only).  The
// e-e terms are summed in registers per spin class over all pairs and the
// e-n terms per nucleus over all electrons; each sum reaches the LDS rows
// once (all reads of a flush first, then the writes).
template <int NU, int ND, class T, class Bs>
__device__ __forceinline__ void pair_pass(const DevSys<T>& S, const T ru[Cap<NU>::v][3],
                                          const T rd[Cap<ND>::v][3], const DerivLayout& L, Row<T> row) {
  constexpr int N = NU + ND;
  constexpr int NS = 2 + MAXK;
  auto pos = [&](int q, int c) -> T { return q < NU ? ru[q][c] : rd[q - NU][c]; };
  // flush of the slot sums [first, nslot) (slot 1 skipped unless has_a)
  // into the columns col(slot)
  auto flush = [&](const T (&o)[NS], const T (&e)[NS], int first, int nslot, bool has_a, auto col) {
    T ro[NS], re[NS];
#pragma unroll
    for (int sl = 0; sl < NS; ++sl)
      if (sl >= first && sl < nslot && (sl != 1 || has_a)) {
        ro[sl] = row[L.o(col(sl))];
        re[sl] = row[L.de(col(sl))];
      }
#pragma unroll
    for (int sl = 0; sl < NS; ++sl)
      if (sl >= first && sl < nslot && (sl != 1 || has_a)) {
        row[L.o(col(sl))] = ro[sl] + o[sl];
        row[L.de(col(sl))] = re[sl] + e[sl];
      }
  };
  T eo[2][NS], ee_[2][NS];
#pragma unroll
  for (int cl = 0; cl < 2; ++cl)
#pragma unroll
    for (int sl = 0; sl < NS; ++sl) eo[cl][sl] = ee_[cl][sl] = T(0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i + 1; j < N; ++j) {
      const int cl = ((i < NU) == (j < NU)) ? 1 : 0;
      ee_pair_terms(S, pos(i, 0) - pos(j, 0), pos(i, 1) - pos(j, 1), pos(i, 2) - pos(j, 2), cl == 1,
                    [=](int c) { return row[L.g(i, c)] - row[L.g(j, c)]; },
                    [&](int sl, T v) { eo[cl][sl] += v; }, [&](int sl, T v) { ee_[cl][sl] += v; });
    }
  // b_ee (slot 0) is one column for both classes
  eo[1][0] += eo[0][0];
  ee_[1][0] += ee_[0][0];
  const int nse = 2 + S.k_ee;
#pragma unroll
  for (int cl = 0; cl < 2; ++cl)
    flush(eo[cl], ee_[cl], cl == 0 ? 1 : 0, nse, true, [&](int sl) { return ee_slot_col(S, cl, sl); });
  const int nsn = 2 + S.k_en;
  Bs::nuclei(S, [&](int I) {
    T no[NS], ne[NS];
#pragma unroll
    for (int sl = 0; sl < NS; ++sl) no[sl] = ne[sl] = T(0);
#pragma unroll
    for (int i = 0; i < N; ++i)
      en_pair_terms(S, I, pos(i, 0) - S.nuc[I][0], pos(i, 1) - S.nuc[I][1], pos(i, 2) - S.nuc[I][2],
                    [=](int c) { return row[L.g(i, c)]; }, [&](int sl, T v) { no[sl] += v; },
                    [&](int sl, T v) { ne[sl] += v; });
    const int sp = S.species[I];
    flush(no, ne, 0, nsn, false, [&](int sl) { return en_slot_col(S, sp, sl); });
  });
}

template <int NU, int ND, class Bs, class T>
__global__ void __launch_bounds__(64) derivs_kernel(const DevSys<double>* __restrict__ Sd,
                                                    const DevSys<T>* __restrict__ St, DerivArgs A) {
  extern __shared__ double lds_raw[];
  T* lds = (T*)lds_raw;
  const DevSys<T>& S = *St;
  const int64_t w = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const bool live = w < A.W;
  const int64_t wl = live ? w : A.W - 1;
  DerivLayout L{A.P, A.M, NU + ND, A.Qp};
  Row<T> row{lds + threadIdx.x};
  const int nq = A.Qp + 3 * (NU + ND);
  for (int q = 0; q < nq; ++q) row[q] = T(0);
  // positions in FP64, from AoS input or from the ensemble's SoA state
  WalkerRegs<NU, ND, double> D;
  constexpr int N = NU + ND;
#pragma unroll
  for (int k = 0; k < NU; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int e = Sd->sector_elec[0][k];
      D.ru[k][c] = A.pos_aos ? A.pos_aos[(wl * N + e) * 3 + c]
                             : (A.pos64 ? A.pos64[(e * 3 + c) * A.ld + wl] : (double)A.pos32[(e * 3 + c) * A.ld + wl]);
    }
#pragma unroll
  for (int k = 0; k < ND; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int e = Sd->sector_elec[1][k];
      D.rd[k][c] = A.pos_aos ? A.pos_aos[(wl * N + e) * 3 + c]
                             : (A.pos64 ? A.pos64[(e * 3 + c) * A.ld + wl] : (double)A.pos32[(e * 3 + c) * A.ld + wl]);
    }
  double lu, su, ldn, sd;
  sector_inverse<NU, double, Bs>(*Sd, 0, D.ru, D.au, lu, su);
  sector_inverse<ND, double, Bs>(*Sd, 1, D.rd, D.ad, ldn, sd);
  WalkerRegs<NU, ND, T> R;
#pragma unroll
  for (int k = 0; k < Cap<NU>::v; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) R.ru[k][c] = (T)D.ru[k][c];
#pragma unroll
    for (int m = 0; m < Cap<NU>::v; ++m) R.au[k][m] = (T)D.au[k][m];
  }
#pragma unroll
  for (int k = 0; k < Cap<ND>::v; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) R.rd[k][c] = (T)D.rd[k][c];
#pragma unroll
    for (int m = 0; m < Cap<ND>::v; ++m) R.ad[k][m] = (T)D.ad[k][m];
  }
  double kin = 0.0, pot = 0.0;
  NucAcc<T, Bs::n_nuc_ct> acc;
  acc.zero();
#pragma unroll 1
  for (int k = 0; k < NU; ++k) {
    kin += electron_pass<NU, ND, T, Bs>(S, 0, k, R.ru, R.au, R.rd, L, row, acc, pot);
    rotate<NU>(R.ru, R.au);
  }
#pragma unroll 1
  for (int k = 0; k < ND; ++k) {
    kin += electron_pass<ND, NU, T, Bs>(S, 1, NU + k, R.rd, R.ad, R.ru, L, row, acc, pot);
    rotate<ND>(R.rd, R.ad);
  }
  acc.flush(row, L);
  const double el = kin + pot + S.vnn;
  row[L.one()] = T(1);
  if (S.jas_kind == JAS_PADE) pair_pass<NU, ND, T, Bs>(S, R.ru, R.rd, L, row);
  if (live)
    for (int q = 0; q < A.Qp; ++q) A.Y[(int64_t)q * A.ldy + w] = q == L.e() ? el : (double)row[q];
}

// ---------------------------------------------------------------------------
// Two wavefronts per 64 walkers for equal spin sectors (NS up, NS down) with
// a compiled shell structure: wave s owns sector s -- its inverse, the
// electron pass of its NS electrons, its same-spin pairs, half of the
// opposite-spin pairs and the electron-nucleus pairs of its electrons -- so
// each SIMD holds two waves instead of one.  The sector is wave-uniform, so
// every parameter load stays a scalar load.  The two waves share one LDS
// row per walker (element q of lane t at lds[q*64 + t]): each writes the
// gradients G of its own electrons, and the register sums reach the row in
// wave order (wave 0's read-modify-write, a barrier, wave 1's), so the
// result does not depend on timing.
//
// Opposite-spin pair (u_a, d_b) has offset (b - a) mod NS seen from wave 0
// and (a - b) mod NS from wave 1.  With m = NS / 2, wave 0 takes offsets
// 0..m and wave 1 offsets 1..m, except that for even NS the offset-m class
// (seen from both waves) is split by k < m: every pair is taken once.
// ---------------------------------------------------------------------------
template <int NS>
__host__ __device__ constexpr bool anti_pair_taken(int wave, int k, int d) {
  constexpr int m = NS / 2;
  if (d > m) return false;
  if (wave == 1 && d == 0) return false;
  if (NS % 2 == 0 && d == m) return k < m;
  return true;
}

// every opposite-spin pair (u_a, d_b) is taken by exactly one wave
template <int NS>
constexpr bool anti_pairs_covered_once() {
  for (int a = 0; a < NS; ++a)
    for (int b = 0; b < NS; ++b) {
      const int n = (anti_pair_taken<NS>(0, a, (b - a + NS) % NS) ? 1 : 0) +
                    (anti_pair_taken<NS>(1, b, (a - b + NS) % NS) ? 1 : 0);
      if (n != 1) return false;
    }
  return true;
}
static_assert(anti_pairs_covered_once<1>() && anti_pairs_covered_once<2>() && anti_pairs_covered_once<3>() &&
                  anti_pairs_covered_once<4>() && anti_pairs_covered_once<5>() && anti_pairs_covered_once<6>() &&
                  anti_pairs_covered_once<7>() && anti_pairs_covered_once<8>(),
              "the two waves must share the opposite-spin pairs exactly");

template <int NS, class Bs, class T>
__global__ void __launch_bounds__(128) derivs2_kernel(const DevSys<double>* __restrict__ Sd,
                                                     const DevSys<T>* __restrict__ St, DerivArgs A) {
  static_assert(Bs::n_nuc_ct > 0, "the nucleus sums must be compile-time register arrays");
  extern __shared__ double lds_raw[];
  const DevSys<T>& S = *St;
  constexpr int N = 2 * NS, NSL = 2 + MAXK;
  const int lane = threadIdx.x & 63;
  const int sp = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // this wave's sector
  const int64_t w = (int64_t)blockIdx.x * 64 + lane;
  const bool live = w < A.W;
  const int64_t wl = live ? w : A.W - 1;
  DerivLayout L{A.P, A.M, N, A.Qp};
  const int nq = A.Qp + 3 * N;
  T* lds = (T*)lds_raw;
  double* xch = lds_raw + ((size_t)nq * 64 * sizeof(T) + 7) / 8;  // [2][64] E_L parts
  Row<T> row{lds + lane};
  for (int q = sp; q < nq; q += 2) row[q] = T(0);
  // positions: this wave's sector (a) and the other one (b), in FP64
  double pa[Cap<NS>::v][3], pb[Cap<NS>::v][3];
#pragma unroll
  for (int k = 0; k < NS; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int ea = Sd->sector_elec[sp][k], eb = Sd->sector_elec[1 - sp][k];
      auto ld = [&](int e) {
        return A.pos_aos ? A.pos_aos[(wl * N + e) * 3 + c]
                         : (A.pos64 ? A.pos64[(e * 3 + c) * A.ld + wl] : (double)A.pos32[(e * 3 + c) * A.ld + wl]);
      };
      pa[k][c] = ld(ea);
      pb[k][c] = ld(eb);
    }
  double a64[Cap<NS>::v][Cap<NS>::v], la, sg;
  sector_inverse<NS, double, Bs>(*Sd, sp, pa, a64, la, sg);
  T ra[Cap<NS>::v][3], rb[Cap<NS>::v][3], ai[Cap<NS>::v][Cap<NS>::v];
#pragma unroll
  for (int k = 0; k < Cap<NS>::v; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ra[k][c] = (T)pa[k][c];
      rb[k][c] = (T)pb[k][c];
    }
#pragma unroll
    for (int m = 0; m < Cap<NS>::v; ++m) ai[k][m] = (T)a64[k][m];
  }
  __syncthreads();  // the row is zeroed before any read-modify-write
  NucAcc<T, Bs::n_nuc_ct> acc;
  acc.zero();
  double kin = 0.0, pot = 0.0;
#pragma unroll 1
  for (int k = 0; k < NS; ++k) {
    kin += electron_pass<NS, NS, T, Bs>(S, sp, sp * NS + k, ra, ai, rb, L, row, acc, pot);
    rotate<NS>(ra, ai);
  }
  // nucleus sums in wave order; the barriers also publish every G
  for (int v = 0; v < 2; ++v) {
    if (sp == v) acc.flush(row, L);
    __syncthreads();
  }
  xch[sp * 64 + lane] = kin + pot;
  if (S.jas_kind == JAS_PADE) {
    // flush of slot sums [first, nslot) (slot 1 skipped unless has_a)
    auto flush = [&](const T (&o)[NSL], const T (&e)[NSL], int first, int nslot, bool has_a, auto col) {
      T ro[NSL], re[NSL];
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl)
        if (sl >= first && sl < nslot && (sl != 1 || has_a)) {
          ro[sl] = row[L.o(col(sl))];
          re[sl] = row[L.de(col(sl))];
        }
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl)
        if (sl >= first && sl < nslot && (sl != 1 || has_a)) {
          row[L.o(col(sl))] = ro[sl] + o[sl];
          row[L.de(col(sl))] = re[sl] + e[sl];
        }
    };
    // electron q of the pair-pass order: q < NS up, NS + k down
    auto gq = [&](int q, int c) { return row[L.g(q, c)]; };
    const int qa = sp * NS, qb = (1 - sp) * NS;
    T eo[2][NSL], ee_[2][NSL];
#pragma unroll
    for (int cl = 0; cl < 2; ++cl)
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl) eo[cl][sl] = ee_[cl][sl] = T(0);
    // same-spin pairs of this sector
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = i + 1; j < NS; ++j)
        ee_pair_terms(S, ra[i][0] - ra[j][0], ra[i][1] - ra[j][1], ra[i][2] - ra[j][2], true,
                      [&](int c) { return gq(qa + i, c) - gq(qa + j, c); },
                      [&](int sl, T v) { eo[1][sl] += v; }, [&](int sl, T v) { ee_[1][sl] += v; });
    // this wave's share of the opposite-spin pairs
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
      for (int d = 0; d < NS; ++d) {
        const int l = (k + d) % NS;
        if (anti_pair_taken<NS>(sp, k, d))
          ee_pair_terms(S, ra[k][0] - rb[l][0], ra[k][1] - rb[l][1], ra[k][2] - rb[l][2], false,
                        [&](int c) { return gq(qa + k, c) - gq(qb + l, c); },
                        [&](int sl, T v) { eo[0][sl] += v; }, [&](int sl, T v) { ee_[0][sl] += v; });
      }
    eo[1][0] += eo[0][0];  // b_ee (slot 0) is one column for both classes
    ee_[1][0] += ee_[0][0];
    const int nse = 2 + S.k_ee;
    for (int v = 0; v < 2; ++v) {
      if (sp == v) {
#pragma unroll
        for (int cl = 0; cl < 2; ++cl)
          flush(eo[cl], ee_[cl], cl == 0 ? 1 : 0, nse, true, [&](int sl) { return ee_slot_col(S, cl, sl); });
      }
      __syncthreads();
    }
    const int nsn = 2 + S.k_en;
    Bs::nuclei(S, [&](int I) {
      T no[NSL], ne[NSL];
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl) no[sl] = ne[sl] = T(0);
#pragma unroll
      for (int k = 0; k < NS; ++k)
        en_pair_terms(S, I, ra[k][0] - S.nuc[I][0], ra[k][1] - S.nuc[I][1], ra[k][2] - S.nuc[I][2],
                      [&](int c) { return gq(qa + k, c); }, [&](int sl, T v) { no[sl] += v; },
                      [&](int sl, T v) { ne[sl] += v; });
      const int spc = S.species[I];
      for (int v = 0; v < 2; ++v) {
        if (sp == v) flush(no, ne, 0, nsn, false, [&](int sl) { return en_slot_col(S, spc, sl); });
        __syncthreads();
      }
    });
  } else {
    __syncthreads();
  }
  if (sp == 0) row[L.one()] = T(1);
  __syncthreads();
  const double el = xch[lane] + xch[64 + lane] + S.vnn;
  if (live)
    for (int q = sp; q < A.Qp; q += 2) A.Y[(int64_t)q * A.ldy + w] = q == L.e() ? el : (double)row[q];
}

// ---------------------------------------------------------------------------
// The wave version: one wavefront per walker, for shapes without a register
// kernel (benzene: 21 + 21 electrons, 66 AOs).  The walker's positions and
// inverses sit in LDS (the inverses from the FP64 Gauss-Jordan build, cast
// to the arithmetic type).  Pass 1 puts electron i on lane i (determinant
// and Jastrow gradients, E_L terms, HF force, Omega, and the Omega terms of
// MO coefficients that follow the geometry); pass 2 spreads the electron
// pairs and electron-nucleus pairs over the lanes (O_k, dE_L/dp_k).  Every
// lane accumulates into its own LDS row (stride Qp + 1, so the rows' banks
// differ); the rows are summed in lane order at the end.
// ---------------------------------------------------------------------------
template <class T>
struct WRow {
  T* base;
  __device__ __forceinline__ T& operator[](int q) const { return base[q]; }
};

template <class T>
__global__ void __launch_bounds__(64) wave_derivs_kernel(const DevSys<double>* __restrict__ Sd,
                                                         const DevSys<T>* __restrict__ St, DerivArgs A, WaveLayout Y) {
  extern __shared__ double lds_d[];
  const DevSys<T>& S = *St;
  const int64_t w = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = Y.N, Mn = A.M, rs = A.Qp + 1;
  DerivLayout Ly{A.P, A.M, N, A.Qp};
  T* L = (T*)lds_d;
  T* G = L + Y.chi();  // d ln psi / d r_i, [N][3]
  T* rows = G + 3 * N; // [64][rs]
  double* pd = lds_d + ((size_t)(Y.chi() + 3 * N + 64 * rs) * sizeof(T) + 7) / 8;  // FP64 positions [N][3]
  double* a = pd + 3 * N;
  double* inv = a + Y.nmax * Y.nmax;
  for (int q = lane; q < 3 * N; q += 64) {
    const double v = A.pos_aos ? A.pos_aos[w * 3 * N + q]
                               : (A.pos64 ? A.pos64[(int64_t)q * A.ld + w] : (double)A.pos32[(int64_t)q * A.ld + w]);
    pd[q] = v;
    L[Y.pos() + q] = (T)v;
  }
  for (int q = lane; q < 64 * rs; q += 64) rows[q] = T(0);
  __syncthreads();
  for (int spin = 0; spin < 2; ++spin) {
    const int n = spin ? Y.nd : Y.nu;
    double la, sg;
    wave_sector_inverse(*Sd, spin, pd, n, a, inv, la, sg);
    for (int q = lane; q < n * n; q += 64) L[Y.ainv(spin) + q] = (T)inv[q];
    __syncthreads();
  }
  WRow<T> row{rows + lane * rs};
  double kin = 0.0, pot = 0.0;
  for (int i = lane; i < N; i += 64) {
    const int spin = S.elec_spin[i], k = S.elec_slot[i];
    const int n = spin ? Y.nd : Y.nu;
    const T* ai = L + Y.ainv(spin);
    const T x = L[3 * i], y = L[3 * i + 1], z = L[3 * i + 2];
    auto wgt = [&](int ao) {
      T v = 0;
      for (int m = 0; m < n; ++m) v += S.moT[spin][ao][m] * ai[m * n + k];
      return v;
    };
    auto on_shell = [&](int c, T sx, T sy, T sz) {
      row[Ly.om(c, 0)] -= sx;
      row[Ly.om(c, 1)] -= sy;
      row[Ly.om(c, 2)] -= sz;
    };
    T gdx, gdy, gdz, ld;
    det_terms_w<T, decltype(wgt), decltype(on_shell), RtBasis>(S, x, y, z, wgt, gdx, gdy, gdz, ld, on_shell);
    T gux = 0, guy = 0, guz = 0, lu = 0, pairsq = 0;
    double pe = 0.0;
    for (int j = 0; j < N; ++j) {
      if (j == i) continue;
      T dx = x - L[3 * j], dy = y - L[3 * j + 1], dz = z - L[3 * j + 2];
      T r, ir;
      M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
      T u1, lp;
      uee_d(S, r, ir, S.elec_spin[j] == spin, u1, lp);
      if (r >= T(1e-10)) {
        T f = u1 * ir;
        gux += f * dx;
        guy += f * dy;
        guz += f * dz;
        lu += lp;
        pairsq += u1 * u1;
      }
      pe += 0.5 * (double)ir;
    }
    for (int I = 0; I < Mn; ++I) {
      T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
      T r, ir;
      M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
      pe -= (double)(S.Z[I] * ir);
      if (r > T(1e-10)) {  // methane.rs:1253
        const T f3 = S.Z[I] * ir * ir * ir;
        row[Ly.f(I, 0)] += f3 * dx;
        row[Ly.f(I, 1)] += f3 * dy;
        row[Ly.f(I, 2)] += f3 * dz;
      }
      if (S.en_on) {
        T u1, lp;
        uen_d(S, r, ir, I, u1, lp);
        if (r >= T(1e-10)) {
          T f = u1 * ir;
          gux += f * dx;
          guy += f * dy;
          guz += f * dz;
          lu += lp;
          row[Ly.om(I, 0)] -= f * dx;
          row[Ly.om(I, 1)] -= f * dy;
          row[Ly.om(I, 2)] -= f * dz;
        }
      }
    }
    // d ln D / d C[m][ao] = sum_k ainv[m][k] chi_ao(r_k), times dC/dR
    int last = -1;
    T chi[4] = {T(0), T(0), T(0), T(0)};
    for (int p = A.mdp0[spin]; p < A.mdp0[spin + 1]; ++p) {
      const MoDerivPair P = A.mdp[p];
      if (P.sh != last) {
        shell_values(S, P.sh, x, y, z, chi, true);
        last = P.sh;
      }
      const T cq = P.q == 0 ? chi[0] : (P.q == 1 ? chi[1] : (P.q == 2 ? chi[2] : chi[3]));
      const T v = ai[P.m * n + k] * cq;
      for (int e = P.e0; e < P.e1; ++e) row[Ly.om(A.mde[e].I, A.mde[e].c)] += (T)A.mde[e].d * v;
    }
    const T gs = S.grad_scale;
    T gjx = gs * gux, gjy = gs * guy, gjz = gs * guz;
    T lj = (S.flags & F_J2_LAP_QUIRK) ? (lu + pairsq) : (lu + gux * gux + guy * guy + guz * guz);
    const T lap = ld + lj + T(2) * (gdx * gjx + gdy * gjy + gdz * gjz);
    G[3 * i] = gdx + gux;
    G[3 * i + 1] = gdy + guy;
    G[3 * i + 2] = gdz + guz;
    kin += -0.5 * (double)lap;
    pot += pe;
  }
  __syncthreads();
  if (S.jas_kind == JAS_PADE) {
    const int npair = N * (N - 1) / 2;
    for (int p = lane; p < npair; p += 64) {
      int i = 0, r = p;  // p -> (i, j > i), rows of the upper triangle in order
      while (r >= N - 1 - i) {
        r -= N - 1 - i;
        ++i;
      }
      const int j = i + 1 + r;
      const int cl = S.elec_spin[i] == S.elec_spin[j] ? 1 : 0;
      ee_pair_terms(S, L[3 * i] - L[3 * j], L[3 * i + 1] - L[3 * j + 1], L[3 * i + 2] - L[3 * j + 2], cl == 1,
                    [&](int c) { return G[3 * i + c] - G[3 * j + c]; },
                    [&](int sl, T v) { row[Ly.o(ee_slot_col(S, cl, sl))] += v; },
                    [&](int sl, T v) { row[Ly.de(ee_slot_col(S, cl, sl))] += v; });
    }
    for (int p = lane; p < N * Mn; p += 64) {
      const int i = p / Mn, I = p % Mn;
      const int sp = S.species[I];
      en_pair_terms(S, I, L[3 * i] - S.nuc[I][0], L[3 * i + 1] - S.nuc[I][1], L[3 * i + 2] - S.nuc[I][2],
                    [&](int c) { return G[3 * i + c]; },
                    [&](int sl, T v) { row[Ly.o(en_slot_col(S, sp, sl))] += v; },
                    [&](int sl, T v) { row[Ly.de(en_slot_col(S, sp, sl))] += v; });
    }
  }
  const double el = wsum(kin) + wsum(pot) + S.vnn;
  __syncthreads();
  for (int q = lane; q < A.Qp; q += 64) {
    double v;
    if (q == Ly.e()) {
      v = el;
    } else if (q == Ly.one()) {
      v = 1.0;
    } else {
      v = 0.0;
      for (int l = 0; l < 64; ++l) v += (double)rows[l * rs + q];
    }
    A.Y[(int64_t)q * A.ldy + w] = v;
  }
}

template <class T>
hipError_t wave_derivs(const void* Sd, const void* St, DerivArgs A, const WaveShape& sh, hipStream_t st) {
  WaveLayout Y;
  Y.N = sh.N;
  Y.nu = sh.nu;
  Y.nd = sh.nd;
  Y.nmax = sh.nu > sh.nd ? sh.nu : sh.nd;
  Y.A = sh.A;
  if (Y.N > 64 || Y.nmax > 32) return hipErrorInvalidValue;
  const size_t b = ((size_t)(Y.chi() + 3 * Y.N + 64 * (A.Qp + 1)) * sizeof(T) + 7) / 8 * 8 +
                   sizeof(double) * (3 * (size_t)Y.N + 2 * (size_t)Y.nmax * Y.nmax);
  if (b + 512 > 160 * 1024) return hipErrorInvalidValue;
  // the kernel also holds a static 512 B pivot column, so the dynamic
  // limit is set to what this shape needs rather than the CU's 160 KiB
  hipError_t e = hipFuncSetAttribute((const void*)wave_derivs_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)b);
  if (e != hipSuccess) return e;
  wave_derivs_kernel<T><<<(unsigned)A.W, 64, b, st>>>((const DevSys<double>*)Sd, (const DevSys<T>*)St, A, Y);
  return hipGetLastError();
}

template hipError_t wave_derivs<double>(const void*, const void*, DerivArgs, const WaveShape&, hipStream_t);
template hipError_t wave_derivs<float>(const void*, const void*, DerivArgs, const WaveShape&, hipStream_t);

// ---------------------------------------------------------------------------
// G = Y Y^T over a walker chunk per wave with v_mfma_f64_16x16x4_f64.
// Fragment convention (16x16x4 f64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator register r of lane l is D[4r + l/16][l%16].
// A tile (I, J) pairs feature block I (rows) with block J (columns); both
// fragments are "feature (l%16) of walker k0 + l/16", so one load per block
// serves every tile.  Only I <= J is formed; the host mirrors.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int GRAM_WAVES = 4;  // waves per workgroup, one per SIMD

template <int NB>
__global__ void __launch_bounds__(64 * GRAM_WAVES) gram_kernel(const double* __restrict__ Y, int64_t ldy, int64_t Wp,
                                                               int64_t chunk, double* __restrict__ partial) {
  constexpr int NT = NB * (NB + 1) / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the workgroup's walkers [w0, w1), both multiples of 16; at step s wave
  // v takes walkers w0 + 16 s + 4 v .. + 3, so the four waves read 16
  // consecutive walkers.  Columns up to Wp (W rounded up to 16) are
  // readable and zero past W, so no load needs a guard.
  const int64_t w0 = (int64_t)blockIdx.x * chunk;
  const int64_t w1 = w0 + chunk < Wp ? w0 + chunk : Wp;
  const int nsteps = w1 > w0 ? (int)((w1 - w0) / 16) : 0;
  dbl4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
  const int fr = lane & 15, kk = lane >> 4;
  const double* yl = Y + (int64_t)fr * ldy + 4 * wave + kk;
  auto load = [&](double* f, int s) {
#pragma unroll
    for (int b = 0; b < NB; ++b) f[b] = yl[(int64_t)16 * b * ldy + w0 + 16 * s];
  };
  auto mfmas = [&](const double* f) {
    int t = 0;
#pragma unroll
    for (int I = 0; I < NB; ++I)
#pragma unroll
      for (int J = I; J < NB; ++J) {
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(f[I], f[J], acc[t], 0, 0, 0);
        ++t;
      }
  };
  // double buffer: each step's fragments are loaded one step ahead, and
  // the scheduling barriers keep the loads in front of the MFMAs that do
  // not need them (a last, unused load re-reads the final step)
  double fa[NB], fb[NB];
  if (nsteps > 0) load(fa, 0);
  int s = 0;
  for (; s + 2 <= nsteps; s += 2) {
    load(fb, s + 1);
    __builtin_amdgcn_sched_barrier(0);
    mfmas(fa);
    __builtin_amdgcn_sched_barrier(0);
    load(fa, s + 2 < nsteps ? s + 2 : nsteps - 1);
    __builtin_amdgcn_sched_barrier(0);
    mfmas(fb);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (s < nsteps) mfmas(fa);
  // the four waves' accumulators summed in LDS in wave order (deterministic)
  extern __shared__ double red[];  // [NT][4][64]
  for (int v = 0; v < GRAM_WAVES; ++v) {
    if (wave == v) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double* q = red + (t * 4 + r) * 64 + lane;
          *q = v == 0 ? acc[t][r] : *q + acc[t][r];
        }
    }
    __syncthreads();
  }
  // partial [block][Qp][Qp] (upper tiles only); slot (t, r, lane) is
  // row 16 I + 4 r + lane / 16, column 16 J + lane % 16
  const int Qp = 16 * NB;
  double* out = partial + (int64_t)blockIdx.x * Qp * Qp;
  for (int i = threadIdx.x; i < NT * 4 * 64; i += 64 * GRAM_WAVES) {
    const int t = i >> 8, r = (i >> 6) & 3, l = i & 63;
    int I = 0, tt = t;  // tile t -> (I, J >= I) in the order the MFMA loop forms them
    while (tt >= NB - I) {
      tt -= NB - I;
      ++I;
    }
    const int J = I + tt;
    out[(16 * I + 4 * r + (l >> 4)) * Qp + 16 * J + (l & 15)] = red[i];
  }
}

// deterministic sum of the per-workgroup partial Grams: four threads per
// element each sum every fourth partial in order, then the four sums are
// added in a fixed order
__global__ void __launch_bounds__(256) gram_reduce_kernel(const double* __restrict__ partial, int nblk, int Qp,
                                                          double* __restrict__ G) {
  __shared__ double sh[4][64];
  const int part = threadIdx.x >> 6, el = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 64 + el;
  const int64_t n = (int64_t)Qp * Qp;
  double s = 0.0;
  if (i < n) {
#pragma unroll 16
    for (int b = part; b < nblk; b += 4) s += partial[(int64_t)b * n + i];  // loads in flight together
  }
  sh[part][el] = s;
  __syncthreads();
  if (part == 0 && i < n) G[i] = ((sh[0][el] + sh[1][el]) + sh[2][el]) + sh[3][el];
}

template <int NU, int ND, class Bs, class T>
static hipError_t launch_derivs(const void* Sd, const void* St, DerivArgs A, hipStream_t st) {
  // equal sectors and a compiled shell structure: two waves per 64 walkers
  // (in FP64 a few registers spill, and the 64 KiB of rows let two
  // workgroups share a CU; measured 1.6x faster than one wave all the same)
  if constexpr (NU == ND && Bs::n_nuc_ct > 0) {
    const size_t shm2 = (sizeof(T) * 64 * (A.Qp + 3 * (NU + ND)) + 7) / 8 * 8 + sizeof(double) * 128;
    if (shm2 > 160 * 1024) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute((const void*)derivs2_kernel<NU, Bs, T>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm2);
    if (e != hipSuccess) return e;
    derivs2_kernel<NU, Bs, T><<<(unsigned)((A.W + 63) / 64), 128, shm2, st>>>((const DevSys<double>*)Sd,
                                                                               (const DevSys<T>*)St, A);
    return hipGetLastError();
  }
  const size_t shm = sizeof(T) * 64 * (A.Qp + 3 * (NU + ND));
  if (shm > 160 * 1024) return hipErrorInvalidValue;
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)derivs_kernel<NU, ND, Bs, T>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  derivs_kernel<NU, ND, Bs, T><<<(unsigned)((A.W + 63) / 64), 64, shm, st>>>((const DevSys<double>*)Sd,
                                                                              (const DevSys<T>*)St, A);
  return hipGetLastError();
}

template <int NU, int ND, class Bs>
static DerivOps make_dops(bool (*match)(const DevSys<double>&)) {
  DerivOps o;
  o.nu = NU;
  o.nd = ND;
  o.match = match;
  o.derivs64 = &launch_derivs<NU, ND, Bs, double>;
  o.derivs32 = &launch_derivs<NU, ND, Bs, float>;
  return o;
}

template <class Bd>
static DerivOps compiled_dops() {
  return make_dops<Bd::nu, Bd::nd, CtBasis<Bd>>(&basis_matches<Bd, double>);
}

// shell structures first (as the register engine), then generic shapes
static const DerivOps kDOps[] = {
    compiled_dops<BdH2>(),           compiled_dops<BdLiSeeded>(),    compiled_dops<BdLiRef>(),
    compiled_dops<BdCH4Sto3g>(),     compiled_dops<BdCH4631g>(),     make_dops<1, 0, RtBasis>(nullptr),
    make_dops<1, 1, RtBasis>(nullptr), make_dops<2, 1, RtBasis>(nullptr), make_dops<2, 2, RtBasis>(nullptr),
    make_dops<3, 2, RtBasis>(nullptr), make_dops<3, 3, RtBasis>(nullptr), make_dops<4, 4, RtBasis>(nullptr),
    make_dops<5, 5, RtBasis>(nullptr),
};

const DerivOps* find_derivs(const DevSys<double>& D, bool generic_only) {
  for (const auto& o : kDOps) {
    if (o.nu != D.n_up || o.nd != D.n_dn) continue;
    if (o.match && (generic_only || !o.match(D))) continue;
    return &o;
  }
  return nullptr;
}

int gram_blocks(int64_t W) {
  const int64_t n = (W + 255) / 256;
  return (int)(n < 1 ? 1 : (n > 256 ? 256 : n));
}

hipError_t launch_gram(const double* Y, int64_t ldy, int64_t W, int Qp, double* partial, double* G, hipStream_t st) {
  const int64_t Wp = (W + 15) / 16 * 16;
  if (W < 1 || ldy < Wp) return hipErrorInvalidValue;
  const int nblk = gram_blocks(W);
  // walkers per workgroup: a multiple of the 16 one step of its four waves takes
  const int64_t chunk = ((Wp + nblk - 1) / nblk + 15) / 16 * 16;
  const size_t lds = sizeof(double) * (size_t)((Qp / 16) * (Qp / 16 + 1) / 2) * 4 * 64;
  switch (Qp / 16) {
#define QMC_GRAM_CASE(nb)                                                                                    \
  case nb:                                                                                                   \
    if (lds > 64 * 1024) {                                                                                   \
      hipError_t ea = hipFuncSetAttribute((const void*)gram_kernel<nb>,                                     \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);             \
      if (ea != hipSuccess) return ea;                                                                       \
    }                                                                                                        \
    gram_kernel<nb><<<nblk, 64 * GRAM_WAVES, lds, st>>>(Y, ldy, Wp, chunk, partial);                         \
    break;
    QMC_GRAM_CASE(1)
    QMC_GRAM_CASE(2)
    QMC_GRAM_CASE(3)
    QMC_GRAM_CASE(4)
    QMC_GRAM_CASE(5)
    QMC_GRAM_CASE(6)
    QMC_GRAM_CASE(7)
    QMC_GRAM_CASE(8)
#undef QMC_GRAM_CASE
    default:
      return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t n = (int64_t)Qp * Qp;
  gram_reduce_kernel<<<(unsigned)((n + 63) / 64), 256, 0, st>>>(partial, nblk, Qp, G);
  return hipGetLastError();
}

}  // namespace qmc
