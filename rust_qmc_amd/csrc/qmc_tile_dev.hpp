// qmc_tile_dev.hpp -- device code of the tile engine (see qmc_tile.hip):
// a 256-thread workgroup owns a tile of 16 walkers, 16 lanes per walker,
// in FP32 (mixed precision) or FP64; T is the walker state's type.
#pragma once

#include "qmc_engine.hpp"
#include "qmc_tile.hpp"
#include "qmc_wave_dev.hpp"

namespace qmc {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// One 16x16x4 matrix-core step in T, D += A B: lane l supplies A[l%16][l/16]
// and B[l/16][l%16].  The two types lay the result out differently: register
// r of lane l is D[4 (l/16) + r][l%16] for FP32 (v_mfma_f32_16x16x4f32) and
// D[4 r + l/16][l%16] for FP64 (v_mfma_f64_16x16x4f64).
template <class T>
struct TileMfma;
template <>
struct TileMfma<float> {
  typedef f32x4 V;
  static __device__ __forceinline__ V step(float a, float b, V d) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, d, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};
template <>
struct TileMfma<double> {
  typedef f64x4 V;
  static __device__ __forceinline__ V step(double a, double b, V d) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * r + (l >> 4); }
};

constexpr int TILE_T = 256;   // threads per tile
constexpr int TILE_MO = 32;   // MO columns of Phi (two 16-wide MFMA tiles)
// Row stride of Phi in LDS: a walker's row is read by its 16 lanes at one
// address (a broadcast), the four walkers of a wavefront at 33-element
// offsets that fall in different banks; a stride of 32 put all four in the
// same bank.
constexpr int TILE_PS = 33;
constexpr int TILE_MAXP = 6;  // primitives per shell
// elements per shell record in LDS: 33 fields, an odd stride so that the
// 16 lanes of a group, reading the same field of 16 different records, hit
// 16 different banks (a stride of 32 elements put them all in one bank)
constexpr int TILE_REC = 33;
// row stride of an inverse column or MO-coefficient row in LDS: odd, so
// the 16 lanes of a group reading rows a = g + 16 j hit 16 different banks
constexpr int TILE_CK = 25;
constexpr int TILE_AJ = 5;    // AOs per lane in the AO-weight pass: a = g + 16 j, j < 5
constexpr int TILE_WS = 16 * TILE_AJ;  // row stride of a walker's AO weights

struct TileArgs {
  int N, A, XS, n_nuc, n_shell;
  int lane_sh[16][TILE_MAXSH];  // shells evaluated by group lane g (-1: none)
  int gauss, n_slot, slot_np[TILE_MAXSH];  // the Gaussian path of a move's AO values (TilePlan)
  // LDS carve-out in elements of the state type; o_cmo, o_ck and o_w only
  // in a drift-diffusion launch
  int o_pos, o_x, o_phi, o_spin, o_sh, o_sum, o_park, o_rng, o_cmo, o_ck, o_w;
  int lds_floats;  // total carve-out, in elements of the state type
};

// sum over the 16 lanes of a group (butterfly inside the 16-lane row)
template <class V>
__device__ __forceinline__ V gsum(V v) {
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int as_int(float f) { return __float_as_int(f); }
__device__ __forceinline__ float as_float(int i) { return __int_as_float(i); }

// Integers in a shell record of T: bit patterns in a float record, exact
// values in a double record.
__device__ __forceinline__ float rec_of(int i, float) { return as_float(i); }
__device__ __forceinline__ double rec_of(int i, double) { return (double)i; }
__device__ __forceinline__ int rec_int(float v) { return as_int(v); }
__device__ __forceinline__ int rec_int(double v) { return (int)v; }

// One shell record: kind, first AO, primitive count, P1 component (ints),
// centre, the Gaussian path's s destination, then exponents, s and p
// coefficients and STO powers of up to TILE_MAXP primitives (the DevSys
// shell and primitive tables, per shell), and the Gaussian path's p
// destination.
template <class T>
__device__ __forceinline__ void tile_write_record(const DevSys<T>& S, int sh, T* rec, int pad) {
  rec[0] = rec_of(S.sh_kind[sh], T());
  rec[1] = rec_of(S.sh_ao0[sh], T());
  rec[2] = rec_of(S.sh_nprim[sh], T());
  rec[3] = rec_of(S.sh_m[sh], T());
  const int c = S.sh_center[sh], p0 = S.sh_prim0[sh], np = S.sh_nprim[sh];
  rec[4] = S.nuc[c][0];
  rec[5] = S.nuc[c][1];
  rec[6] = S.nuc[c][2];
  for (int k = 0; k < TILE_MAXP; ++k) {
    const bool v = k < np;
    rec[8 + k] = v ? S.prim_exp[p0 + k] : T(0);
    rec[14 + k] = v ? S.coef_s[p0 + k] : T(0);
    rec[20 + k] = v ? S.coef_p[p0 + k] : T(0);
    rec[26 + k] = rec_of(v ? S.prim_pow[p0 + k] : 0, T());
  }
  // the Gaussian path's destinations in a row of X (pad: its last column,
  // whose MO coefficients are zero): the s value's, and the first of the
  // p values' (s shells write zeros to the pad)
  const int kind = S.sh_kind[sh], a0 = S.sh_ao0[sh];
  rec[7] = rec_of(kind == SH_P ? pad : a0, T());
  rec[32] = rec_of(kind == SH_S ? pad : (kind == SH_SP ? a0 + 1 : a0), T());
}

// the Gaussian path's record for lanes without a shell in a slot: zero
// coefficients, both destinations the pad
template <class T>
__device__ __forceinline__ void tile_null_record(T* rec, int pad) {
  for (int i = 0; i < TILE_REC; ++i) rec[i] = T(0);
  rec[0] = rec_of(SH_S, T());
  rec[7] = rec_of(pad, T());
  rec[32] = rec_of(pad, T());
}

// AO values of one s, p or sp shell from its record, branch-free: npu
// primitives (the slot's count, zero-padded records), the s value to
// rec[7] and the p values to rec[32] + 0, 1, 2 (+ 0, 0, 0 for s shells,
// whose p part is zero); the same arithmetic as tile_shell_values
template <class T>
__device__ __forceinline__ void tile_gauss_values(const T* rec, T x, T y, T z, T* chi, int npu) {
  const T dx = x - rec[4], dy = y - rec[5], dz = z - rec[6];
  const T r2 = dx * dx + dy * dy + dz * dz;
  T rs = T(0), rp = T(0);
  for (int k = 0; k < npu; ++k) {
    const T e = M<T>::exn(-rec[8 + k] * r2);
    rs += rec[14 + k] * e;
    rp += rec[20 + k] * e;
  }
  const int sd = rec_int(rec[7]), pd = rec_int(rec[32]), ps = rec_int(rec[0]) == SH_S ? 0 : 1;
  chi[sd] = rs;
  chi[pd] = dx * rp;
  chi[pd + ps] = dy * rp;
  chi[pd + 2 * ps] = dz * rp;
}

// AO values of one shell from its LDS record, written to chi[ao]; the same
// arithmetic as shell_values (methane.rs:498-590, slater.rs:17-41, 92-144)
template <class T>
__device__ __forceinline__ void tile_shell_values(const T* rec, T x, T y, T z, T* chi) {
  const int kind = rec_int(rec[0]), a0 = rec_int(rec[1]), np = rec_int(rec[2]);
  const T dx = x - rec[4], dy = y - rec[5], dz = z - rec[6];
  const T r2 = dx * dx + dy * dy + dz * dz;
  if (kind <= SH_P1) {
    T rs = T(0), rp = T(0);
    for (int k = 0; k < np; ++k) {
      const T e = M<T>::exn(-rec[8 + k] * r2);
      rs += rec[14 + k] * e;
      rp += rec[20 + k] * e;
    }
    if (kind == SH_S) {
      chi[a0] = rs;
    } else if (kind == SH_P1) {
      const int m = rec_int(rec[3]);
      chi[a0] = (m == 0 ? dx : (m == 1 ? dy : dz)) * rp;
    } else {
      int b = a0;
      if (kind == SH_SP) {
        chi[a0] = rs;
        b = a0 + 1;
      }
      chi[b] = dx * rp;
      chi[b + 1] = dy * rp;
      chi[b + 2] = dz * rp;
    }
  } else {
    const T sr = M<T>::sq(r2);
    T v = T(0);
    for (int k = 0; k < np; ++k) {
      T sp = T(1);
      for (int q = 0; q < rec_int(rec[26 + k]); ++q) sp *= sr;
      v += rec[14 + k] * sp * M<T>::exn(-rec[8 + k] * sr);
    }
    chi[a0] = v;
  }
}

// Values and gradients of the AOs of one shell from its LDS record, handed
// to f(ao, value, d/dx, d/dy, d/dz); the same arithmetic as
// shell_values_grad (contracted Gaussians with s / p / sp angular parts,
// radial STO sums whose gradient at the centre is taken as zero).
template <class T, class F>
__device__ __forceinline__ void tile_shell_grad(const T* rec, T x, T y, T z, F f) {
  const int kind = rec_int(rec[0]), a0 = rec_int(rec[1]), np = rec_int(rec[2]);
  const T dx = x - rec[4], dy = y - rec[5], dz = z - rec[6];
  const T r2 = dx * dx + dy * dy + dz * dz;
  if (kind <= SH_P1) {
    T rs = T(0), rsa = T(0), rp = T(0), rpa = T(0);
    for (int k = 0; k < np; ++k) {
      const T a = rec[8 + k];
      const T e = M<T>::exn(-a * r2);
      const T cs = rec[14 + k] * e, cp = rec[20 + k] * e;
      rs += cs;
      rsa += cs * a;
      rp += cp;
      rpa += cp * a;
    }
    if (kind == SH_S || kind == SH_SP) {  // grad = -2 Ra d
      const T t = T(-2) * rsa;
      f(a0, rs, t * dx, t * dy, t * dz);
    }
    if (kind == SH_S) return;
    // p_m = d_m R: grad_j = delta_jm R - 2 d_m Ra d_j
    auto pm = [&](int ao, int m) {
      const T dm = m == 0 ? dx : (m == 1 ? dy : dz);
      const T t = T(2) * dm * rpa;
      f(ao, dm * rp, (m == 0 ? rp : T(0)) - t * dx, (m == 1 ? rp : T(0)) - t * dy, (m == 2 ? rp : T(0)) - t * dz);
    };
    if (kind == SH_P1) {
      pm(a0, rec_int(rec[3]));
    } else {
      const int b = kind == SH_SP ? a0 + 1 : a0;
      pm(b, 0);
      pm(b + 1, 1);
      pm(b + 2, 2);
    }
    return;
  }
  // radial STO: f(s) = sum c s^p e^{-zeta s}; grad = f'(s) d / s
  const T s = M<T>::sq(r2);
  T v = T(0), d1 = T(0);
  for (int k = 0; k < np; ++k) {
    const T zt = rec[8 + k];
    const int p = rec_int(rec[26 + k]);
    T sp = T(1);
    for (int q = 0; q < p; ++q) sp *= s;
    const T spm1 = p >= 1 ? sp / s : T(0);
    const T e = rec[14 + k] * M<T>::exn(-zt * s);
    v += e * sp;
    d1 += e * (-zt * sp + T(p) * spm1);
  }
  const T gs = s == T(0) ? T(0) : d1 / s;
  f(a0, v, gs * dx, gs * dy, gs * dz);
}

// values and gradients of one s, p or sp shell from its record, branch-free
// as tile_gauss_values: f(ao, value, d/dx, d/dy, d/dz) for the s part at
// rec[7] and the three p components at rec[32] + 0, 1, 2 (s shells and the
// null record hand zeros to the pad); the same arithmetic as
// tile_shell_grad, in the same order
template <class T, class F>
__device__ __forceinline__ void tile_gauss_grad(const T* rec, T x, T y, T z, int npu, F f) {
  const T dx = x - rec[4], dy = y - rec[5], dz = z - rec[6];
  const T r2 = dx * dx + dy * dy + dz * dz;
  T rs = T(0), rsa = T(0), rp = T(0), rpa = T(0);
  for (int k = 0; k < npu; ++k) {
    const T a = rec[8 + k];
    const T e = M<T>::exn(-a * r2);
    const T cs = rec[14 + k] * e, cp = rec[20 + k] * e;
    rs += cs;
    rsa += cs * a;
    rp += cp;
    rpa += cp * a;
  }
  const int sd = rec_int(rec[7]), pd = rec_int(rec[32]), ps = rec_int(rec[0]) == SH_S ? 0 : 1;
  const T t = T(-2) * rsa;  // grad = -2 Ra d
  f(sd, rs, t * dx, t * dy, t * dz);
  // p_m = d_m R: grad_j = delta_jm R - 2 d_m Ra d_j
  auto pm = [&](int ao, int m) {
    const T dm = m == 0 ? dx : (m == 1 ? dy : dz);
    const T u = T(2) * dm * rpa;
    f(ao, dm * rp, (m == 0 ? rp : T(0)) - u * dx, (m == 1 ? rp : T(0)) - u * dy, (m == 2 ? rp : T(0)) - u * dz);
  };
  pm(pd, 0);
  pm(pd + ps, 1);
  pm(pd + 2 * ps, 2);
}

// e-n Jastrow value and u'(r) with the charge already in a register
// (uen_val, uen_d)
template <class Js, class T>
__device__ __forceinline__ void tile_uen_vd(const DevSys<T>& S, T r, T ir, T Z, int I, T& v, T& u1) {
  if constexpr (Js::no_series) {
    const T iD = M<T>::rcp(T(1) + S.b_en * r);
    v = -Z * (r * iD);
    u1 = -Z * (iD * iD);
  } else {
    T lp;
    uen_d<T, Js>(S, r, ir, I, u1, lp);
    v = uen_val<T, Js>(S, r, I);
  }
}

// e-n Jastrow value with the charge already in a register (uen_val)
template <class Js, class T>
__device__ __forceinline__ T tile_uen(const DevSys<T>& S, T r, T Z, int I) {
  const T D = T(1) + S.b_en * r;
  const T y = r * M<T>::rcp(D);
  T v = -Z * y;
  const int K = Js::ken(S);
  if (K > 0) {
    T s, s1, s2;
    series(K, S.d_en[S.species[I]], y, s, s1, s2);
    v += s;
  }
  return v;
}

// n_steps single-electron sweeps of tile blk (16 walkers) by a 256-thread
// workgroup; lds is the workgroup's LDS carve-out described by TA, in T.
// DD selects drift-diffusion moves (DriftDiffusionVMC::sweep,
// vmc.rs:382-438) over Metropolis moves (MCMCSimulation::
// single_electron_step): the proposal from the quantum force of the moved
// electron, the Green's function ratio in the acceptance, with the random
// streams of the register and wave engines.
template <int NU, int ND, int KS, class Js, class T, bool DD = false>
__device__ __forceinline__ void tile_sweep_block(const DevSys<T>& S, EnsView<T> E, const TileArgs& TA, SweepArgs A,
                                                 double* __restrict__ partial, double* __restrict__ el_trace,
                                                 int64_t blk, int64_t nblk, int tid, T* lds) {
  constexpr int KH = (KS + 1) / 2;  // MFMA steps per K half
  typedef TileMfma<T> MF;
  const int q = tid >> 6, l = tid & 63;   // wavefront, lane
  const int wt = tid >> 4, g = tid & 15;  // walker in the tile, lane in its group
  const int64_t w = blk * TILE_W + wt;
  const bool live = w < E.W;
  const int64_t wl = live ? w : E.W - 1;  // dead walkers shadow the last one and write nothing
  const int N = TA.N, M_ = TA.n_nuc;
  T* pos = lds + TA.o_pos;  // [TILE_W][3N]
  T* X = lds + TA.o_x;      // [TILE_W][XS], AO values of the proposals
  T* Phi = lds + TA.o_phi;  // [2 K halves][TILE_W][TILE_PS]
  int* spin_of = (int*)(lds + TA.o_spin);
  T* shrec = lds + TA.o_sh;  // [n_shell][TILE_REC]
  double* sums = (double*)(lds + TA.o_sum);  // [TILE_W][3], per-step walker sums
  T* pw = pos + wt * 3 * N;

  // ---- state in: positions and tables to LDS, owned inverse columns to VGPRs ----
  // HBM holds the ensemble element-major ([element][walker]): the tile's 16
  // walkers sit in consecutive lanes for every load and store of the state
  const int64_t w0 = blk * TILE_W;
  auto wclamp = [&](int ww) { return w0 + ww < E.W ? w0 + ww : E.W - 1; };
  for (int i = tid; i < 3 * N * TILE_W; i += TILE_T) {
    const int el = i / TILE_W, ww = i - el * TILE_W;
    pos[ww * 3 * N + el] = E.pos[(int64_t)el * E.ld + wclamp(ww)];
  }
  for (int i = g; i < TA.XS; i += 16) X[wt * TA.XS + i] = T(0);
  for (int i = tid; i < N; i += TILE_T) spin_of[i] = S.elec_spin[i];
  for (int i = tid; i < TA.n_shell; i += TILE_T) tile_write_record(S, i, shrec + i * TILE_REC, TA.XS - 1);
  if (tid == 0) tile_null_record(shrec + TA.n_shell * TILE_REC, TA.XS - 1);
  if constexpr (DD) {
    // AO weights past the AOs stay zero (the Gaussian path reads the pad's)
    for (int i = TA.A + g; i < TILE_WS; i += 16) lds[TA.o_w + wt * TILE_WS + i] = T(0);
  }
  if constexpr (DD) {  // MO coefficients, [spin][ao][mo] with rows of TILE_CK
    static_assert(NU <= TILE_CK && ND <= TILE_CK, "sector too large for the AO-weight rows");
    T* cmo = lds + TA.o_cmo;
    for (int i = tid; i < 2 * TA.A * TILE_CK; i += TILE_T) {
      const int sp = i / (TA.A * TILE_CK), rem = i - sp * TA.A * TILE_CK;
      const int a = rem / TILE_CK, m = rem - a * TILE_CK;
      cmo[i] = m < (sp ? ND : NU) ? S.moT[sp][a][m] : T(0);
    }
  }
  int my_sh[TILE_MAXSH];
#pragma unroll
  for (int j = 0; j < TILE_MAXSH; ++j) my_sh[j] = TA.lane_sh[g][j];
  // this lane's Jastrow partners: electrons p = g + 16 t (t < TILE_EE), then
  // nucleus g (preloaded), so that every lane of the group takes the e-e
  // rounds and then the one e-n round without divergence between the two
  const bool nuc_on = g < M_;
  const T pnx = nuc_on ? S.nuc[g][0] : T(0), pny = nuc_on ? S.nuc[g][1] : T(0);
  const T pnz = nuc_on ? S.nuc[g][2] : T(0), pnZ = nuc_on ? S.Z[g] : T(0);
  // Inverse columns g and g + 16 of the sector that moves live in VGPRs
  // (cur); the other sector's wait in this lane's slots of the LDS park
  // [walker][column][row], swapped at each sector boundary.
  constexpr int NX = NU > ND ? NU : ND;
  T* park0 = lds + TA.o_park;
  T* park = park0 + wt * NX * NX;
  // a sector's inverse between HBM (element m n + k of walker w0 + ww) and
  // the park (park0[(ww NX + k) NX + m]), all threads of the tile
  auto park_in = [&](int sp) {
    const int n = sp ? ND : NU;
    const int64_t base = sp ? (int64_t)NU * NU : 0;
    if constexpr (NU != ND)
      for (int i = tid; i < TILE_W * NX * NX; i += TILE_T) park0[i] = T(0);
    __syncthreads();
    for (int i = tid; i < n * n * TILE_W; i += TILE_T) {
      const int el = i / TILE_W, ww = i - el * TILE_W, m = el / n, k = el - m * n;
      park0[(ww * NX + k) * NX + m] = E.ainv[(base + el) * E.ld + wclamp(ww)];
    }
    __syncthreads();
  };
  auto park_out = [&](int sp) {
    const int n = sp ? ND : NU;
    const int64_t base = sp ? (int64_t)NU * NU : 0;
    __syncthreads();
    for (int i = tid; i < n * n * TILE_W; i += TILE_T) {
      const int el = i / TILE_W, ww = i - el * TILE_W, m = el / n, k = el - m * n;
      if (w0 + ww < E.W) E.ainv[(base + el) * E.ld + w0 + ww] = park0[(ww * NX + k) * NX + m];
    }
  };
  T cur[2][NX];
  park_in(0);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int k = g + 16 * t;
#pragma unroll
    for (int m = 0; m < NX; ++m) cur[t][m] = k < NU && m < NU ? park[k * NX + m] : T(0);
  }
  park_in(1);
  // MFMA B operands: wavefront q computes MO tile q & 1 over K half q >> 1;
  // lane l holds C[mo = 16 tile + l%16][ao = (l/16) KS + half KH + s]
  const int tb = q & 1, kh = q >> 1;
  T Creg[2][KH];
#pragma unroll
  for (int sp = 0; sp < 2; ++sp)
#pragma unroll
    for (int s = 0; s < KH; ++s) {
      const int kk = kh * KH + s;
      const int ao = (l >> 4) * KS + kk, mo = tb * 16 + (l & 15);
      Creg[sp][s] = (kk < KS && ao < TA.A && mo < (sp ? ND : NU)) ? S.moT[sp][ao][mo] : T(0);
    }
  double logdet = E.logdet[wl], sgn = E.sign[wl], ju = E.ju[wl], el = 0.0;
  LogAcc<T> lacc;  // ln|R| of the accepted moves, folded into ln|D| at the end
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t wid = (uint32_t)(A.walker0 + wl);
  const T sigma = (T)A.step_size;
  const bool en_on = S.en_on;
  // drift-diffusion: time step, drift cap, sqrt(2 tau), 1/(4 tau)
  const T tau = (T)A.step_size, md = (T)A.max_drift;
  const T s2t = M<T>::sq(T(2) * tau), i4t = T(0.25) / tau;
  __syncthreads();

  for (int st = 0; st < A.n_steps; ++st) {
    const uint32_t step = A.step0 + (uint32_t)st;
    int acc = 0;
#pragma unroll
    for (int spin = 0; spin < 2; ++spin) {
      const int n = spin ? ND : NU;
      // the sector's random numbers (three normals and the acceptance
      // uniform of each move) drawn ahead, lane g for slots g and g + 16,
      // into the walker's LDS rows: two Philox blocks and Box-Muller
      // transforms per lane instead of one per move on every lane
      T* rg = lds + TA.o_rng + wt * 4 * NX;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int k = g + 16 * t;
        if (k < n) {
          T z[3], u;
          Rng<T>::move(philox(wid, step, (uint32_t)S.sector_elec[spin][k], DD ? RNG_DRIFT : RNG_SINGLE, k0, k1), z,
                       u);
          rg[4 * k] = z[0];
          rg[4 * k + 1] = z[1];
          rg[4 * k + 2] = z[2];
          rg[4 * k + 3] = u;
        }
      }
      wave_sync();
#pragma unroll 1
      for (int k = 0; k < n; ++k) {
        const int e = S.sector_elec[spin][k];
        const int owner = (l & 48) | (k & 15);
        const bool hi = k >= 16;
        const T ox = pw[3 * e], oy = pw[3 * e + 1], oz = pw[3 * e + 2];
        T nx, ny, nz, U;
        [[maybe_unused]] T Rd, dud, pd;  // drift-diffusion: ratio, Jastrow difference, acceptance
        if constexpr (DD) {
          // W: AO weights w[a] = sum_m C[a][m] Ainv[m][k] of the moved
          // electron's inverse column, staged through the group's LDS rows
          T* ck = lds + TA.o_ck + wt * TILE_CK;
          T* wa = lds + TA.o_w + wt * TILE_WS;
          const T* cmo = lds + TA.o_cmo + spin * TA.A * TILE_CK;
          if (g == (k & 15)) {
#pragma unroll
            for (int m = 0; m < n; ++m) ck[m] = hi ? cur[1][m] : cur[0][m];
          }
          wave_sync();
          {
            T c[NU > ND ? NU : ND];
#pragma unroll
            for (int m = 0; m < n; ++m) c[m] = ck[m];
#pragma unroll
            for (int j = 0; j < TILE_AJ; ++j) {
              const int a = g + 16 * j;
              if (a < TA.A) {
                const T* cr = cmo + a * TILE_CK;
                T wq = T(0);
#pragma unroll
                for (int m = 0; m < n; ++m) wq += cr[m] * c[m];
                wa[a] = wq;
              }
            }
          }
          wave_sync();
          // the Jastrow value and gradient of electron e at (x, y, z) over
          // this lane's partners: electrons g + 16 t, then nucleus g
          // (branch-free: every lane evaluates its partner slots, a pair that
          // does not count -- the electron itself, past the electrons, no
          // nucleus -- adds a selected zero)
          auto jpart = [&](T x, T y, T z, T& u, T& jx, T& jy, T& jz) {
            u = jx = jy = jz = T(0);
            auto add = [&](bool on, T dx, T dy, T dz, T r, T ir, T v, T u1) {
              u += on ? v : T(0);
              const T f = on && r >= T(1e-10) ? u1 * ir : T(0);
              jx += f * dx;
              jy += f * dy;
              jz += f * dz;
            };
#pragma unroll
            for (int t = 0; t < TILE_EE; ++t) {
              const int p = g + 16 * t, pc = p < N ? p : N - 1;
              const T dx = x - pw[3 * pc], dy = y - pw[3 * pc + 1], dz = z - pw[3 * pc + 2];
              T r, ir, u1, lp;
              M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
              const bool same = spin_of[pc] == spin;
              uee_d<T, Js>(S, r, ir, same, u1, lp);
              add(p < N && p != e, dx, dy, dz, r, ir, uee_val<T, Js>(S, r, same), u1);
            }
            if (Js::no_series ? en_on : en_on && nuc_on) {  // no series: pnZ = 0 past the nuclei
              const T dx = x - pnx, dy = y - pny, dz = z - pnz;
              T r, ir, v, u1;
              M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
              tile_uen_vd<Js>(S, r, ir, pnZ, g, v, u1);
              add(nuc_on, dx, dy, dz, r, ir, v, u1);
            }
          };
          // G: quantum force at r (grad ln|D| from the weighted AO
          // gradients, the Jastrow gradient), the Langevin proposal, then at
          // r' the ratio R = sum w chi, grad D'/D = sum w grad chi / R and the
          // Jastrow terms; chi(r') goes to X for the update
          T gx = T(0), gy = T(0), gz = T(0);
          auto grad_at_r = [&](int a, T, T ddx, T ddy, T ddz) {
            const T wq = wa[a];
            gx += wq * ddx;
            gy += wq * ddy;
            gz += wq * ddz;
          };
          if (TA.gauss) {
#pragma unroll
            for (int j = 0; j < TILE_MAXSH; ++j)
              if (j < TA.n_slot)
                tile_gauss_grad(shrec + (my_sh[j] >= 0 ? my_sh[j] : TA.n_shell) * TILE_REC, ox, oy, oz,
                                TA.slot_np[j], grad_at_r);
          } else {
#pragma unroll
            for (int j = 0; j < TILE_MAXSH; ++j)
              if (my_sh[j] >= 0) tile_shell_grad(shrec + my_sh[j] * TILE_REC, ox, oy, oz, grad_at_r);
          }
          T uo, jx, jy, jz;
          jpart(ox, oy, oz, uo, jx, jy, jz);
          gx = rsum(gx);
          gy = rsum(gy);
          gz = rsum(gz);
          uo = rsum(uo);
          jx = rsum(jx);
          jy = rsum(jy);
          jz = rsum(jz);
          const T gs = S.grad_scale;
          T dold[3];
          cap_drift(T(2) * (gx + gs * jx), T(2) * (gy + gs * jy), T(2) * (gz + gs * jz), tau, md, dold);
          const T z[3] = {rg[4 * k], rg[4 * k + 1], rg[4 * k + 2]};
          U = rg[4 * k + 3];
          nx = ox + dold[0] + s2t * z[0];
          ny = oy + dold[1] + s2t * z[1];
          nz = oz + dold[2] + s2t * z[2];
          T R = T(0), hx = T(0), hy = T(0), hz = T(0);
          T* xw = X + wt * TA.XS;
          auto at_rp = [&](int a, T v, T ddx, T ddy, T ddz) {
            const T wq = wa[a];
            xw[a] = v;
            R += wq * v;
            hx += wq * ddx;
            hy += wq * ddy;
            hz += wq * ddz;
          };
          if (TA.gauss) {
#pragma unroll
            for (int j = 0; j < TILE_MAXSH; ++j)
              if (j < TA.n_slot)
                tile_gauss_grad(shrec + (my_sh[j] >= 0 ? my_sh[j] : TA.n_shell) * TILE_REC, nx, ny, nz,
                                TA.slot_np[j], at_rp);
          } else {
#pragma unroll
            for (int j = 0; j < TILE_MAXSH; ++j)
              if (my_sh[j] >= 0) tile_shell_grad(shrec + my_sh[j] * TILE_REC, nx, ny, nz, at_rp);
          }
          T un, vx, vy, vz;
          jpart(nx, ny, nz, un, vx, vy, vz);
          R = rsum(R);
          hx = rsum(hx);
          hy = rsum(hy);
          hz = rsum(hz);
          un = rsum(un);
          vx = rsum(vx);
          vy = rsum(vy);
          vz = rsum(vz);
          const T iR = M<T>::rcp(R);
          T dnew[3];
          cap_drift(T(2) * (hx * iR + gs * vx), T(2) * (hy * iR + gs * vy), T(2) * (hz * iR + gs * vz), tau, md,
                    dnew);
          const T f0 = nx - ox - dold[0], f1 = ny - oy - dold[1], f2 = nz - oz - dold[2];
          const T b0 = ox - nx - dnew[0], b1 = oy - ny - dnew[1], b2 = oz - nz - dnew[2];
          const T gf = -(f0 * f0 + f1 * f1 + f2 * f2) * i4t;
          const T gb = -(b0 * b0 + b1 * b1 + b2 * b2) * i4t;
          Rd = R;
          dud = un - uo;
          pd = R * R * M<T>::ex(T(2) * dud + gb - gf);
        } else {
          // A: proposal
          const T z[3] = {rg[4 * k], rg[4 * k + 1], rg[4 * k + 2]};
          U = rg[4 * k + 3];
          nx = ox + sigma * z[0];
          ny = oy + sigma * z[1];
          nz = oz + sigma * z[2];
          // B: AO values of the proposal
          if (TA.gauss) {  // uniform branches and loop counts only
#pragma unroll
            for (int j = 0; j < TILE_MAXSH; ++j)
              if (j < TA.n_slot)
                tile_gauss_values(shrec + (my_sh[j] >= 0 ? my_sh[j] : TA.n_shell) * TILE_REC, nx, ny, nz,
                                  X + wt * TA.XS, TA.slot_np[j]);
          } else {
#pragma unroll
            for (int j = 0; j < TILE_MAXSH; ++j)
              if (my_sh[j] >= 0) tile_shell_values(shrec + my_sh[j] * TILE_REC, nx, ny, nz, X + wt * TA.XS);
          }
        }
        __syncthreads();
        // C: Phi[half][walker][mo] = sum over the half's AOs of X[walker][ao] C[mo][ao]
        if (tb * 16 < n) {
          typename MF::V d0 = {T(0), T(0), T(0), T(0)}, d1 = {T(0), T(0), T(0), T(0)};
          const T* xr = X + (l & 15) * TA.XS + (l >> 4) * KS + kh * KH;
          // past K (kh KH + s >= KS) the B operand is zero and X is in its
          // zero padding: no early exit, so the loop unrolls and Creg stays
          // statically indexed
#pragma unroll
          for (int s = 0; s < KH; ++s) {
            const T b = spin ? Creg[1][s] : Creg[0][s];
            if (s & 1)
              d1 = MF::step(xr[s], b, d1);
            else
              d0 = MF::step(xr[s], b, d0);
          }
          T* ph = Phi + kh * TILE_W * TILE_PS;
#pragma unroll
          for (int r = 0; r < 4; ++r) ph[MF::row(l, r) * TILE_PS + tb * 16 + (l & 15)] = d0[r] + d1[r];
        }
        __syncthreads();
        // D: ratio, Jastrow difference, Metropolis test, Sherman-Morrison
        const T* ph0 = Phi + wt * TILE_PS;
        const T* ph1 = ph0 + TILE_W * TILE_PS;
        T v0 = T(0), v1 = T(0);
#pragma unroll
        for (int m = 0; m < n; ++m) {
          const T pm = ph0[m] + ph1[m];
          v0 += pm * cur[0][m];
          v1 += pm * cur[1][m];
        }
        T R, du, pacc;
        if constexpr (DD) {
          R = Rd;
          du = dud;
          pacc = pd;
        } else {
          R = __shfl(hi ? v1 : v0, owner, 64);
          du = T(0);
#pragma unroll
          // branch-free: a partner slot that does not count (the electron
          // itself, past the electrons) adds a selected zero
          for (int t = 0; t < TILE_EE; ++t) {
            const int p = g + 16 * t, pc = p < N ? p : N - 1;
            const T px = pw[3 * pc], py = pw[3 * pc + 1], pz = pw[3 * pc + 2];
            const bool same = spin_of[pc] == spin;
            const T ax = nx - px, ay = ny - py, az = nz - pz;
            const T bx = ox - px, by = oy - py, bz = oz - pz;
            const T d = uee_diff<T, Js>(S, M<T>::dist(ax * ax + ay * ay + az * az),
                                        M<T>::dist(bx * bx + by * by + bz * bz), same);
            du += p < N && p != e ? d : T(0);
          }
          if constexpr (Js::no_series) {
            if (en_on) {  // -Z (r' - r) / ((1+b r')(1+b r)), one reciprocal; pnZ = 0 past the nuclei
              const T ax = nx - pnx, ay = ny - pny, az = nz - pnz;
              const T bx = ox - pnx, by = oy - pny, bz = oz - pnz;
              const T rn = M<T>::dist(ax * ax + ay * ay + az * az), ro = M<T>::dist(bx * bx + by * by + bz * bz);
              const T d = -pnZ * (rn - ro) * M<T>::rcp((T(1) + S.b_en * rn) * (T(1) + S.b_en * ro));
              du += nuc_on ? d : T(0);
            }
          } else if (en_on && nuc_on) {
            const T ax = nx - pnx, ay = ny - pny, az = nz - pnz;
            const T bx = ox - pnx, by = oy - pny, bz = oz - pnz;
            const T rn = M<T>::dist(ax * ax + ay * ay + az * az), ro = M<T>::dist(bx * bx + by * by + bz * bz);
            du += tile_uen<Js>(S, rn, pnZ, g) - tile_uen<Js>(S, ro, pnZ, g);
          }
          du = gsum(du);
          pacc = R * R * M<T>::ex(T(2) * du);
        }
        if (U < pacc) {  // uniform within the walker's group
          const T iR = M<T>::rcp(R);
          const T f0 = v0 * iR, f1 = v1 * iR;
          // column j takes Ainv_j - c f_j, f_j = v_j / R, and the moved
          // electron's column c becomes c / R: on its owner lane a scale 0
          // and a factor -1/R make the same fused multiply-add that product
          const bool own = g == (k & 15);
          const T sk = own ? T(0) : T(1), fk = own ? -iR : (hi ? f1 : f0);
          // the moved electron's old inverse column: staged in LDS by a
          // drift-diffusion move, else read from its owner lane; which half
          // holds it is uniform over the wavefront (hi), so a branch picks
          // the half instead of a select per element
          const T* ck = lds + TA.o_ck + wt * TILE_CK;
          auto sm_update = [&](auto H) {
            constexpr int h = decltype(H)::value;
            const T fo = h ? f0 : f1;
#pragma unroll
            for (int m = 0; m < n; ++m) {
              const T c = DD ? ck[m] : __shfl(cur[h][m], owner, 64);
              cur[h][m] = fma(-c, fk, cur[h][m] * sk);
              cur[1 - h][m] = fma(-c, fo, cur[1 - h][m]);
            }
          };
          if (hi)
            sm_update(std::integral_constant<int, 1>{});
          else
            sm_update(std::integral_constant<int, 0>{});
          if (g == 0) {
            pw[3 * e] = nx;
            pw[3 * e + 1] = ny;
            pw[3 * e + 2] = nz;
          }
          lacc.add(logdet, R);
          if (R < T(0)) sgn = -sgn;
          ju += (double)du;
          ++acc;
        }
      }
      // sector boundary: the moved sector's columns to the park, the other's back
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int k = g + 16 * t;
#pragma unroll
        for (int m = 0; m < NX; ++m) {
          const T o = k < NX ? park[k * NX + m] : T(0);
          if (k < NX) park[k * NX + m] = cur[t][m];
          cur[t][m] = o;
        }
      }
    }
    __syncthreads();

    // local energy in three passes over the lanes: slot g of sector 0, slot
    // g of sector 1, then the slots from 16 up of both sectors (lane g below
    // NU - 16 takes slot 16 + g of sector 0 with its own column, the next
    // ND - 16 lanes take sector 1's from the park) -- three passes where
    // slot-by-sector order would take four, two of them on five lanes
    double kin = 0.0, pot = 0.0;
    auto el_add = [&](bool on, int spin, int k, auto wgt) {
      if (!on) return;
      const int i = S.sector_elec[spin][k];
      const T x = pw[3 * i], y = pw[3 * i + 1], z = pw[3 * i + 2];
      T gdx, gdy, gdz, ld;
      det_terms_w(S, x, y, z, wgt, gdx, gdy, gdz, ld, NoShell{});
      T gux = T(0), guy = T(0), guz = T(0), lu = T(0), pairsq = T(0), pee = T(0), pen = T(0);
      // every lane walks all N electrons: its own (r = 0, rejected by the
      // r >= 1e-10 guard and the selected potential term) adds zeros, so the
      // lanes stay in step
      for (int j = 0; j < N; ++j) {
        const T dx = x - pw[3 * j], dy = y - pw[3 * j + 1], dz = z - pw[3 * j + 2];
        T r, ir;
        M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
        T u1, lp;
        uee_d<T, Js>(S, r, ir, spin_of[j] == spin, u1, lp);
        const bool ok = r >= T(1e-10);
        const T f = ok ? u1 * ir : T(0);
        gux += f * dx;
        guy += f * dy;
        guz += f * dz;
        lu += ok ? lp : T(0);
        pairsq += ok ? u1 * u1 : T(0);
        pee += j != i ? ir : T(0);
      }
      for (int I = 0; I < M_; ++I) {
        const T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
        T r, ir;
        M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
        pen += S.Z[I] * ir;
        if (en_on) {
          T u1, lp;
          uen_d<T, Js>(S, r, ir, I, u1, lp);
          const bool ok = r >= T(1e-10);
          const T f = ok ? u1 * ir : T(0);
          gux += f * dx;
          guy += f * dy;
          guz += f * dz;
          lu += ok ? lp : T(0);
        }
      }
      const T gs = S.grad_scale;
      const T gjx = gs * gux, gjy = gs * guy, gjz = gs * guz;
      const T lj = (S.flags & F_J2_LAP_QUIRK) ? (lu + pairsq) : (lu + gux * gux + guy * guy + guz * guz);
      const T lap = ld + lj + T(2) * (gdx * gjx + gdy * gjy + gdz * gjz);
      kin += -0.5 * (double)lap;
      pot += 0.5 * (double)pee - (double)pen;
    };
    // after both sectors moved, spin 0's columns are in cur, spin 1's parked
    el_add(g < NU, 0, g, [&](int ao) {
      T s = T(0);
#pragma unroll
      for (int m = 0; m < NU; ++m) s += S.moT[0][ao][m] * cur[0][m];
      return s;
    });
    {
      T col[NX];
#pragma unroll
      for (int m = 0; m < NX; ++m) col[m] = park[g * NX + m];
      el_add(g < ND, 1, g, [&](int ao) {
        T s = T(0);
#pragma unroll
        for (int m = 0; m < ND; ++m) s += S.moT[1][ao][m] * col[m];
        return s;
      });
    }
    constexpr int XU = NU > 16 ? NU - 16 : 0, XD = ND > 16 ? ND - 16 : 0;
    static_assert(XU + XD <= 16, "more extra slots than lanes");
    if constexpr (XU + XD > 0) {
      // mixed sectors across the lanes: both coefficient sets stay scalar
      // operands, each lane's other column zero
      const bool s1 = g >= XU, on = g < XU + XD;
      const int k = on ? 16 + (s1 ? g - XU : g) : 0;
      T c0[NX], c1[NX];
#pragma unroll
      for (int m = 0; m < NX; ++m) {
        const T pv = park[(s1 ? k : 0) * NX + m];
        c0[m] = s1 ? T(0) : cur[1][m];
        c1[m] = s1 ? pv : T(0);
      }
      el_add(on, s1 ? 1 : 0, k, [&](int ao) {
        T s = T(0);
#pragma unroll
        for (int m = 0; m < NU; ++m) s += S.moT[0][ao][m] * c0[m];
#pragma unroll
        for (int m = 0; m < ND; ++m) s += S.moT[1][ao][m] * c1[m];
        return s;
      });
    }
    el = gsum(kin) + gsum(pot) + S.vnn;
    // per-step sums over the tile's walkers, in walker order
    if (g == 0) {
      if (el_trace && live) el_trace[(int64_t)st * A.trace_ld + w] = el;
      sums[wt * 3 + 0] = live ? el : 0.0;
      sums[wt * 3 + 1] = live ? el * el : 0.0;
      sums[wt * 3 + 2] = live ? (double)acc : 0.0;
    }
    __syncthreads();
    if (tid < 3) {
      double v = 0.0;
      for (int j = 0; j < TILE_W; ++j) v += sums[j * 3 + tid];
      partial[((int64_t)st * nblk + blk) * 3 + tid] = v;
    }
  }
  __syncthreads();
  // ---- state out ----
  for (int i = tid; i < 3 * N * TILE_W; i += TILE_T) {
    const int el = i / TILE_W, ww = i - el * TILE_W;
    if (w0 + ww < E.W) E.pos[(int64_t)el * E.ld + w0 + ww] = pos[ww * 3 * N + el];
  }
  park_out(1);  // spin 1 rests in the park, spin 0 is in cur
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int k = g + 16 * t;
    if (k < NU)
#pragma unroll
      for (int m = 0; m < NU; ++m) park[k * NX + m] = cur[t][m];
  }
  park_out(0);
  if (live) {
    if (g == 0) {
      lacc.fold(logdet);
      E.logdet[w] = logdet;
      E.sign[w] = sgn;
      E.ju[w] = ju;
      E.eloc[w] = el;
    }
  }
}

// kernel arguments of a tile launch: LDS carve-out (in elements of the
// state type, of esz bytes) and lane shell lists
template <int KS>
inline TileArgs make_tile_args(const TilePlan& P, int esz = (int)sizeof(float), bool dd = false) {
  TileArgs T;
  T.N = P.N;
  T.A = P.A;
  T.XS = 4 * KS + 1;
  T.n_nuc = P.n_nuc;
  T.n_shell = P.n_shell;
  for (int g = 0; g < 16; ++g)
    for (int j = 0; j < TILE_MAXSH; ++j) T.lane_sh[g][j] = P.lane_sh[g][j];
  T.gauss = P.gauss ? 1 : 0;
  T.n_slot = P.n_slot;
  for (int j = 0; j < TILE_MAXSH; ++j) T.slot_np[j] = P.slot_np[j];
  const int per8 = 8 / esz;  // elements per double
  T.o_pos = 0;
  T.o_x = T.o_pos + TILE_W * 3 * P.N;
  T.o_phi = T.o_x + TILE_W * T.XS;
  T.o_spin = T.o_phi + 2 * TILE_W * TILE_PS;
  T.o_sh = T.o_spin + ((P.N + 3) / 4 * 4 * 4 + esz - 1) / esz;  // N ints
  T.o_sum = T.o_phi;  // the per-step walker sums reuse Phi, idle during the local energy
  const int nx = P.nu > P.nd ? P.nu : P.nd;
  const int per16 = 16 / esz;  // the park and the drift-diffusion rows start 16-byte aligned
  auto al16 = [&](int o) { return (o + per16 - 1) / per16 * per16; };
  T.o_park = al16(T.o_sh + (P.n_shell + 1) * TILE_REC);  // the shells' records and the null record
  T.o_rng = al16(T.o_park + TILE_W * nx * nx);
  T.o_cmo = al16(T.o_rng + TILE_W * 4 * nx);
  T.o_ck = T.o_cmo + (dd ? 2 * P.A * TILE_CK : 0);
  T.o_w = T.o_ck + (dd ? TILE_W * TILE_CK : 0);
  T.lds_floats = T.o_w + (dd ? TILE_W * TILE_WS : 0);
  return T;
}

}  // namespace qmc
