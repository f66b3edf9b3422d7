// qmc_device.hpp -- device-side building blocks of the walker engine for
// gfx950: system parameters, Philox4x32-10, orbital / Jastrow evaluation,
// register-resident Gauss-Jordan inversion and the per-electron local-energy
// terms.  One thread owns one walker; every array below is indexed with
// compile-time indices only, so it lives in VGPRs.
#pragma once

#include "qmc_basis.hpp"
#include "qmc_sys.hpp"
#define QMC_EXPN_QUAL __host__ __device__ __forceinline__
#include "qmc_expn.hpp"

namespace qmc {

// ---------------------------------------------------------------------------
// precision-dependent math: FP64 uses the library routines, the
// exponentials of qmc_expn.hpp and the reciprocals Newton-refined hardware
// estimates; FP32 uses the hardware transcendental instructions.
// ---------------------------------------------------------------------------
template <class T>
struct M;
// dist(d2) = sqrt(d2) and dist_inv(d2, r, ir) = (sqrt(d2), 1/sqrt(d2)) are
// the distance forms of a squared distance d2 >= 0.
template <>
struct M<double> {
  // e^x of an acceptance exponent (either sign): exp_clamped, the library
  // exponential's reduction without its range checks (NaN gives 0, a
  // rejection, as the library's NaN does)
  static __device__ __forceinline__ double ex(double x) { return exp_clamped(x); }
  // e^x of a primitive's argument (finite, not positive): exp_nonpos, the
  // library exponential without its range checks
  static __device__ __forceinline__ double exn(double x) { return exp_nonpos(x); }
  static __device__ __forceinline__ double sq(double x) { return sqrt(x); }
  static __device__ __forceinline__ double lg(double x) { return log(x); }
  // 1/x and 1/sqrt(x): the hardware estimates (v_rcp_f64, v_rsq_f64)
  // refined by two Newton-Raphson steps each.  Each step squares the
  // relative error e of the estimate (e -> ~e^2), so two steps reach FP64
  // rounding from any e < 2^-14, at about a third of the instructions of
  // the IEEE division and square-root sequences.  Measured on gfx950 over
  // 4.2 M log-uniform x in [1e-12, 1e12]: rcp(x) equal to 1.0 / x in every
  // case, rsq(x) within 2 ulp of 1.0 / sqrt(x).
  static __device__ __forceinline__ double rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(e, y, y);
    e = fma(-x, y, 1.0);
    return fma(e, y, y);
  }
  // x >= 0, clamped to 1e-300 so that x = 0 gives a finite 1e150 and
  // x * rsq(x) = 0
  static __device__ __forceinline__ double rsq(double x) {
    x = fmax(x, 1e-300);
    const double hx = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-hx * y, y, 0.5);  // 1/2 - x y^2 / 2
    y = fma(y, e, y);
    e = fma(-hx * y, y, 0.5);
    return fma(y, e, y);
  }
  static __device__ __forceinline__ double dist(double d2) { return d2 * rsq(d2); }
  static __device__ __forceinline__ void dist_inv(double d2, double& r, double& ir) {
    ir = rsq(d2);
    r = d2 * ir;
  }
};
template <>
struct M<float> {
  static __device__ __forceinline__ float ex(float x) { return __expf(x); }
  static __device__ __forceinline__ float exn(float x) { return __expf(x); }
  static __device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float lg(float x) { return __logf(x); }
  static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
  static __device__ __forceinline__ float dist(float d2) { return sq(d2); }
  static __device__ __forceinline__ void dist_inv(float d2, float& r, float& ir) {
    r = sq(d2);
    ir = rcp(r);
  }
};

// ln|D| gains ln|R| per accepted move.  FP64 keeps a running product of
// the ratios as a mantissa in [1/2, 1) and a binary exponent (frexp) and
// folds it into ln|D| with one logarithm; FP32 adds the hardware logarithm
// per move.
template <class T>
struct LogAcc;
template <>
struct LogAcc<double> {
  double m = 1.0;
  int e = 0;
  __device__ __forceinline__ void add(double&, double R) {
    int k;
    m = frexp(m * fabs(R), &k);
    e += k;
  }
  __device__ __forceinline__ void fold(double& ld) {  // + ln m + e ln 2, ln 2 split in two doubles
    ld += fma((double)e, 0x1.62e42fefa39efp-1, fma((double)e, 0x1.abc9e3b39803fp-56, log(m)));
    m = 1.0;
    e = 0;
  }
};
template <>
struct LogAcc<float> {
  __device__ __forceinline__ void add(double& ld, float R) { ld += (double)__logf(R < 0.0f ? -R : R); }
  __device__ __forceinline__ void fold(double&) {}
};

// e^(-a x) of primitive k with exponent a (x = r^2 of a Gaussian, r of a
// Slater term): FP64 exp_nonpos(-a x); FP32 the hardware 2^y (v_exp_f32) at
// y = x (-a log2 e), the factor rounded once on the host (prim_nl2e)
// instead of every argument scaled by log2 e on the device
__device__ __forceinline__ double prim_e(const DevSys<double>& S, int k, double x) {
  return M<double>::exn(-S.prim_exp[k] * x);
}
__device__ __forceinline__ float prim_e(const DevSys<float>& S, int k, float x) {
  return __builtin_amdgcn_exp2f(S.prim_nl2e[k] * x);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller normals
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                     uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32 -> 64-bit product gives both halves
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

__device__ __forceinline__ double u01d(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }
__device__ __forceinline__ double u24d(uint32_t w) { return ((double)(w >> 8) + 0.5) * 5.9604644775390625e-08; }
// the low bytes of a Philox block, as one 32-bit word
__device__ __forceinline__ uint32_t low_bytes(U4 o) {
  return (o.x & 0xFFu) | ((o.y & 0xFFu) << 8) | ((o.z & 0xFFu) << 16) | ((o.w & 0xFFu) << 24);
}
__device__ __forceinline__ float u01f(uint32_t w) { return ((float)(w >> 8) + 0.5f) * 5.9604644775390625e-08f; }

template <class T>
struct Rng;
template <>
struct Rng<double> {
  // the Box-Muller angles 2 pi u from sincos_turns (exact quarter-turn
  // reduction of u, no rounding of 2 pi u) instead of the library's sincos;
  // the radii sqrt(-2 ln u) from log_unit and x rsq(x) (x > 0 for u in
  // (0, 1))
  static __device__ __forceinline__ void box_muller(double u1, double u2, double u3, double u4, double z[3]) {
    const double x1 = -2.0 * log_unit(u1), x3 = -2.0 * log_unit(u3);
    const double r1 = x1 * M<double>::rsq(x1), r2 = x3 * M<double>::rsq(x3);
    double s, c, s4, c4;
    sincos_turns(u2, s, c);
    sincos_turns(u4, s4, c4);
    z[0] = r1 * c;
    z[1] = r1 * s;
    z[2] = r2 * c4;
  }
  static __device__ __forceinline__ void normals3(U4 o, double z[3]) {
    box_muller(u01d(o.x), u01d(o.y), u01d(o.z), u01d(o.w), z);
  }
  static __device__ __forceinline__ double uniform(uint32_t w) { return u01d(w); }
  // One single-electron move from one Philox block: three Box-Muller normals
  // from the top 24 bits of the four words and the acceptance uniform from
  // their low bytes (independent bits), as the FP32 path does.
  static __device__ __forceinline__ void move(U4 o, double z[3], double& U) {
    box_muller(u24d(o.x), u24d(o.y), u24d(o.z), u24d(o.w), z);
    U = u01d(low_bytes(o));
  }
};
template <>
struct Rng<float> {
  static __device__ __forceinline__ void normals3(U4 o, float z[3]) {
    const float twopi = 6.2831853071795865f;
    float u1 = u01f(o.x), u2 = u01f(o.y), u3 = u01f(o.z), u4 = u01f(o.w);
    float r1 = __builtin_amdgcn_sqrtf(-2.0f * __logf(u1)), r2 = __builtin_amdgcn_sqrtf(-2.0f * __logf(u3));
    z[0] = r1 * __cosf(twopi * u2);
    z[1] = r1 * __sinf(twopi * u2);
    z[2] = r2 * __cosf(twopi * u4);
  }
  static __device__ __forceinline__ float uniform(uint32_t w) { return u01f(w); }
  static __device__ __forceinline__ void move(U4 o, float z[3], float& U) {
    normals3(o, z);  // u01f already takes the top 24 bits
    U = u01f(low_bytes(o));
  }
};

// ---------------------------------------------------------------------------
// Atomic orbitals contracted straight into MO accumulators.
// GTO shells follow methane.rs:498-590 (unnormalised primitives, s and p);
// STO shells follow slater.rs:92-144 (and slater.rs:17-41 for Slater1s).
// ---------------------------------------------------------------------------
// LDS table of the MO coefficients for bases with lds_mo: entry
// [(spin * MO_TAB_AO + ao) * MO_TAB_MO + m], zero past the system's AO and
// sector sizes; filled by fill_mo_tab at the start of a kernel.
constexpr int MO_TAB_AO = 32, MO_TAB_MO = 8;
__shared__ float g_mo_tab32[2 * MO_TAB_AO * MO_TAB_MO];
__shared__ double g_mo_tab64[2 * MO_TAB_AO * MO_TAB_MO];
__device__ __forceinline__ float* lds_mo_tab(float) { return g_mo_tab32; }
__device__ __forceinline__ double* lds_mo_tab(double) { return g_mo_tab64; }
template <class Bs, class T>
__device__ __forceinline__ T mo_coef(const DevSys<T>& S, int spin, int ao, int m) {
  if constexpr (Bs::lds_mo)
    return lds_mo_tab(T())[(spin * MO_TAB_AO + ao) * MO_TAB_MO + m];
  else
    return S.moT[spin][ao][m];
}
// every thread of the block; ends with a block barrier
template <int NU, int ND, class T>
__device__ __forceinline__ void fill_mo_tab(const DevSys<T>& S) {
  static_assert(NU <= MO_TAB_MO && ND <= MO_TAB_MO, "sector too large for the MO table");
  T* tab = lds_mo_tab(T());
  for (int q = threadIdx.x; q < 2 * MO_TAB_AO * MO_TAB_MO; q += blockDim.x) {
    const int sp = q / (MO_TAB_AO * MO_TAB_MO), ao = (q / MO_TAB_MO) % MO_TAB_AO, m = q % MO_TAB_MO;
    tab[q] = (ao < S.n_ao && m < (sp ? ND : NU)) ? S.moT[sp][ao][m] : T(0);
  }
  __syncthreads();
}

template <int NS, class Bs = RtBasis, class T>
__device__ __forceinline__ void mo_add(const DevSys<T>& S, int spin, int ao, T v, T phi[Cap<NS>::v]) {
#pragma unroll
  for (int m = 0; m < NS; ++m) phi[m] += mo_coef<Bs>(S, spin, ao, m) * v;
}

// phi[m] = sum_ao C[spin][m][ao] chi_ao(r), m < NS
template <int NS, class T, class Bs = RtBasis>
__device__ __forceinline__ void mo_values(const DevSys<T>& S, int spin, T x, T y, T z, T phi[Cap<NS>::v]) {
#pragma unroll
  for (int m = 0; m < NS; ++m) phi[m] = T(0);
  Bs::shells(S, [&](auto sh) {
    const int c = sh.center(), kind = sh.kind(), p0 = sh.p0(), np = sh.nprim(), a0 = sh.a0();
    T dx = x - S.nuc[c][0], dy = y - S.nuc[c][1], dz = z - S.nuc[c][2];
    T r2 = dx * dx + dy * dy + dz * dz;
    if (kind <= SH_P1) {
      T rs = T(0), rp = T(0);
      for (int k = 0; k < np; ++k) {
        T e = prim_e(S, p0 + k, r2);
        if (kind == SH_S || kind == SH_SP) rs += S.coef_s[p0 + k] * e;
        if (kind != SH_S) rp += S.coef_p[p0 + k] * e;
      }
      if (kind == SH_S) {
        mo_add<NS, Bs>(S, spin, a0, rs, phi);
      } else if (kind == SH_P1) {
        const int m = sh.m();
        mo_add<NS, Bs>(S, spin, a0, (m == 0 ? dx : (m == 1 ? dy : dz)) * rp, phi);
      } else {
        int b = a0;
        if (kind == SH_SP) {
          mo_add<NS, Bs>(S, spin, a0, rs, phi);
          b = a0 + 1;
        }
        mo_add<NS, Bs>(S, spin, b, dx * rp, phi);
        mo_add<NS, Bs>(S, spin, b + 1, dy * rp, phi);
        mo_add<NS, Bs>(S, spin, b + 2, dz * rp, phi);
      }
    } else {
      T s = M<T>::sq(r2);
      T v = T(0);
      for (int k = 0; k < np; ++k) {
        T sp = T(1);
        for (int q = 0; q < sh.pw(k); ++q) sp *= s;
        v += S.coef_s[p0 + k] * sp * prim_e(S, p0 + k, s);
      }
      mo_add<NS, Bs>(S, spin, a0, v, phi);
    }
  });
}

// Determinant terms of one electron from AO derivatives contracted with the
// inverse column: w_ao = sum_m C[spin][m][ao] ainv[m][0], then
// grad D / D = sum_ao w_ao grad chi_ao and lap D / D = sum_ao w_ao lap chi_ao.
// on_shell(center, sx, sy, sz) receives each shell's sum_ao w_ao grad chi_ao
// (minus d ln|D| / dR_center for this electron).
// The same with the AO weights w_ao supplied by wgt(ao):
template <class T, class Wgt, class OnShell, class Bs = RtBasis>
__device__ __forceinline__ void det_terms_w(const DevSys<T>& S, T x, T y, T z, Wgt wgt, T& gdx, T& gdy, T& gdz,
                                            T& ld, OnShell on_shell) {
  gdx = gdy = gdz = ld = T(0);
  Bs::shells(S, [&](auto sh) {
    const int c = sh.center(), kind = sh.kind(), p0 = sh.p0(), np = sh.nprim(), a0 = sh.a0();
    T dx = x - S.nuc[c][0], dy = y - S.nuc[c][1], dz = z - S.nuc[c][2];
    T r2 = dx * dx + dy * dy + dz * dz;
    T sx = 0, sy = 0, sz = 0, sl = 0;
    if (kind <= SH_P1) {
      // R = sum c e, Ra = sum c a e, Raa = sum c a^2 e (s and p coefficient sets)
      T rs = 0, rsa = 0, rsaa = 0, rp = 0, rpa = 0, rpaa = 0;
      for (int k = 0; k < np; ++k) {
        const T a = S.prim_exp[p0 + k];
        T e = prim_e(S, p0 + k, r2);
        if (kind == SH_S || kind == SH_SP) {
          T cs = S.coef_s[p0 + k] * e;
          rs += cs;
          rsa += cs * a;
          rsaa += cs * a * a;
        }
        if (kind != SH_S) {
          T cp = S.coef_p[p0 + k] * e;
          rp += cp;
          rpa += cp * a;
          rpaa += cp * a * a;
        }
      }
      if (kind == SH_S || kind == SH_SP) {
        // grad = -2 Ra d; lap = -6 Ra + 4 r^2 Raa
        const T w = wgt(a0);
        const T g = T(-2) * rsa * w;
        sx += g * dx;
        sy += g * dy;
        sz += g * dz;
        sl += w * (T(-6) * rsa + T(4) * r2 * rsaa);
      }
      if (kind != SH_S) {
        // p_m = d_m R: grad_j = delta_jm R - 2 d_m Ra d_j; lap = d_m (-10 Ra + 4 r^2 Raa)
        const T lr = T(-10) * rpa + T(4) * r2 * rpaa;
        T wx, wy, wz;
        if (kind == SH_P1) {
          const int m = sh.m();
          const T w = wgt(a0);
          wx = m == 0 ? w : T(0);
          wy = m == 1 ? w : T(0);
          wz = m == 2 ? w : T(0);
        } else {
          const int b = (kind == SH_SP) ? a0 + 1 : a0;
          wx = wgt(b);
          wy = wgt(b + 1);
          wz = wgt(b + 2);
        }
        const T wd = wx * dx + wy * dy + wz * dz;
        const T g = T(-2) * rpa * wd;
        sx += g * dx + wx * rp;
        sy += g * dy + wy * rp;
        sz += g * dz + wz * rp;
        sl += wd * lr;
      }
    } else {
      // radial STO: R' and R''; grad = R' d/s, lap = R'' + 2 R'/s
      T s = M<T>::sq(r2);
      T d1 = 0, d2 = 0;
      for (int k = 0; k < np; ++k) {
        const T zt = S.prim_exp[p0 + k];
        const int p = sh.pw(k);
        const T pf = T(p);
        // s^p, s^(p-1), s^(p-2) with the p = 0, 1 terms multiplied by zero
        T sp = T(1);
        for (int q = 0; q < p; ++q) sp *= s;
        T spm1 = p >= 1 ? sp / s : T(0);
        T spm2 = p >= 2 ? spm1 / s : T(0);
        T e = S.coef_s[p0 + k] * prim_e(S, p0 + k, s);
        d1 += e * (-zt * sp + pf * spm1);
        d2 += e * (zt * zt * sp - T(2) * zt * pf * spm1 + pf * (pf - T(1)) * spm2);
      }
      const T w = wgt(a0);
      if (s == T(0)) {
        sl = (kind == SH_SLATER1S) ? (w == T(0) ? T(0) : -INFINITY * w) : T(0);
      } else {
        T is = T(1) / s;
        T g = w * d1 * is;
        sx = g * dx;
        sy = g * dy;
        sz = g * dz;
        sl = w * (d2 + T(2) * d1 * is);
      }
    }
    gdx += sx;
    gdy += sy;
    gdz += sz;
    ld += sl;
    on_shell(c, sx, sy, sz);
  });
}

template <int NS, class T, class OnShell, class Bs = RtBasis>
__device__ __forceinline__ void det_terms(const DevSys<T>& S, int spin, T x, T y, T z,
                                          const T ai[Cap<NS>::v][Cap<NS>::v], T& gdx, T& gdy, T& gdz, T& ld,
                                          OnShell on_shell) {
  auto wgt = [&](int ao) {
    T w = 0;
#pragma unroll
    for (int m = 0; m < NS; ++m) w += mo_coef<Bs>(S, spin, ao, m) * ai[m][0];
    return w;
  };
  det_terms_w<T, decltype(wgt), OnShell, Bs>(S, x, y, z, wgt, gdx, gdy, gdz, ld, on_shell);
}

// AO values of one shell at (x, y, z), written to chi[ao - base] for the
// shell's AOs (base = 0: absolute AO index; base = first AO: shell-local)
template <class T>
__device__ __forceinline__ void shell_values(const DevSys<T>& S, int sh, T x, T y, T z, T* chi, bool local = false) {
  const int c = S.sh_center[sh], kind = S.sh_kind[sh], p0 = S.sh_prim0[sh], np = S.sh_nprim[sh];
  const int a0 = local ? 0 : S.sh_ao0[sh];
  T dx = x - S.nuc[c][0], dy = y - S.nuc[c][1], dz = z - S.nuc[c][2];
  T r2 = dx * dx + dy * dy + dz * dz;
  if (kind <= SH_P1) {
    T rs = T(0), rp = T(0);
    for (int k = 0; k < np; ++k) {
      T e = prim_e(S, p0 + k, r2);
      rs += S.coef_s[p0 + k] * e;
      rp += S.coef_p[p0 + k] * e;
    }
    if (kind == SH_S) {
      chi[a0] = rs;
    } else if (kind == SH_P1) {
      const int m = S.sh_m[sh];
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
    T s = M<T>::sq(r2);
    T v = T(0);
    for (int k = 0; k < np; ++k) {
      T sp = T(1);
      for (int q = 0; q < S.prim_pow[p0 + k]; ++q) sp *= s;
      v += S.coef_s[p0 + k] * sp * prim_e(S, p0 + k, s);
    }
    chi[a0] = v;
  }
}

struct NoShell {
  template <class T>
  __device__ __forceinline__ void operator()(int, T, T, T) const {}
};

// ---------------------------------------------------------------------------
// Jastrow pair functions (jastrow.rs:218-321 and the Pade series):
//   u_ee = a x + cst + sum_k c_k x^(k+2),  x = r/(1+b r)
//   u_en = -Z y + sum_k d_k y^(k+2),        y = r/(1+b_en r)
// Jastrow1/Jastrow2 (jastrow.rs:30-125) are u_ee with a = 1/2, b = 1/F,
// cst = -F/2, since -F/(2(1+r/F)) = r/(2(1+r/F)) - F/2.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void series(int K, const T* c, T x, T& s, T& s1, T& s2) {
  T v = 0, v1 = 0, v2 = 0, xp = T(1);
  // unrolled to the maximum length: the coefficient addresses are constants,
  // so their scalar loads are issued together and can be hoisted
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {
    if (k >= K) break;
    const T n = T(k + 2);
    v += c[k] * xp * x * x;
    v1 += c[k] * n * xp * x;
    v2 += c[k] * n * (n - T(1)) * xp;
    xp *= x;
  }
  s = v;
  s1 = v1;
  s2 = v2;
}

// s'''(x) of the series, needed by d/db of the pair Laplacian
template <class T>
__device__ __forceinline__ T series3(int K, const T* c, T x) {
  T v = 0, xp = T(1);  // x^(k-1) for k >= 1
#pragma unroll
  for (int k = 1; k < MAXK; ++k) {
    if (k >= K) break;
    const T n = T(k + 2);
    v += c[k] * n * (n - T(1)) * (n - T(2)) * xp;
    xp *= x;
  }
  return v;
}

// Parameter derivatives of one Pade pair term u = A x + s(x), x = r/(1+b r),
// where A is a (e-e) or -Z (e-n) and s the power series with coefficients c.
// Gives du'/dq and d(u''+2u'/r)/dq for q = b, A and each c_k; callers
// combine them into d E_L / dq.
template <class T>
struct PairDerivs {
  T x, iD, A1, s2, s3;
  __device__ __forceinline__ void init(T r, T b, T A, int K, const T* c) {
    T D = T(1) + b * r;
    iD = M<T>::rcp(D);
    x = r * iD;
    T s, s1;
    s1 = 0;
    s2 = 0;
    s3 = 0;
    if (K > 0) {
      series(K, c, x, s, s1, s2);
      s3 = series3(K, c, x);
    }
    A1 = A + s1;
  }
  // d/db of u' and of lap = u'' + 2u'/r (ir = 1/r)
  __device__ __forceinline__ void d_b(T r, T ir, T b, T& du1, T& dlap) const {
    const T iD2 = iD * iD, iD3 = iD2 * iD, iD4 = iD2 * iD2, x2 = x * x;
    du1 = -x2 * s2 * iD2 - T(2) * r * A1 * iD3;
    const T t1 = -x2 * s3 * iD4 - T(4) * r * s2 * iD4 * iD;
    const T t2 = -T(2) * A1 * iD3 + T(2) * b * x2 * s2 * iD3 + T(6) * b * r * A1 * iD4;
    dlap = t1 + t2 + T(2) * ir * du1;
  }
  // d/dA (A enters linearly)
  __device__ __forceinline__ void d_a(T ir, T b, T& du1, T& dlap) const {
    du1 = iD * iD;
    dlap = -T(2) * b * iD * iD * iD + T(2) * du1 * ir;
  }
  // d/dc_k for the term c_k x^(k+2)
  __device__ __forceinline__ void d_c(int k, T ir, T b, T xk, T& du1, T& dlap) const {
    // xk = x^k
    const T n = T(k + 2);
    const T p1 = n * xk * x, p2 = n * (n - T(1)) * xk;
    const T iD2 = iD * iD;
    du1 = p1 * iD2;
    dlap = p2 * iD2 * iD2 - T(2) * b * p1 * iD2 * iD + T(2) * du1 * ir;
  }
};

// Pade series lengths: JsRt reads them from DevSys at run time; JsNone
// compiles the series out (the reference Jastrow3 has none).
struct JsRt {
  static constexpr bool no_series = false;
  template <class T>
  static __device__ __forceinline__ int kee(const DevSys<T>& S) { return S.k_ee; }
  template <class T>
  static __device__ __forceinline__ int ken(const DevSys<T>& S) { return S.k_en; }
};
struct JsNone {
  static constexpr bool no_series = true;
  template <class T>
  static __host__ __device__ constexpr int kee(const DevSys<T>&) { return 0; }
  template <class T>
  static __host__ __device__ constexpr int ken(const DevSys<T>&) { return 0; }
};

// value only (cst excluded: it cancels in every ratio and is added to ln J separately)
template <class T, class Js = JsRt>
__device__ __forceinline__ T uee_val(const DevSys<T>& S, T r, bool same) {
  T D = T(1) + S.b_ee * r;
  T x = r * M<T>::rcp(D);
  T v = (same ? S.a_para : S.a_anti) * x;
  const int K = Js::kee(S);
  if (K > 0) {
    T s, s1, s2;
    series(K, same ? S.c_para : S.c_anti, x, s, s1, s2);
    v += s;
  }
  return v;
}

template <class T, class Js = JsRt>
__device__ __forceinline__ T uen_val(const DevSys<T>& S, T r, int I) {
  T D = T(1) + S.b_en * r;
  T y = r * M<T>::rcp(D);
  T v = -S.Z[I] * y;
  const int K = Js::ken(S);
  if (K > 0) {
    T s, s1, s2;
    series(K, S.d_en[S.species[I]], y, s, s1, s2);
    v += s;
  }
  return v;
}

// u(r_new) - u(r_old) of an e-e pair.  Without a series the Pade terms
// combine over one reciprocal: a r'/(1+b r') - a r/(1+b r) =
// a (r' - r) / ((1+b r')(1+b r)).
template <class T, class Js = JsRt>
__device__ __forceinline__ T uee_diff(const DevSys<T>& S, T rn, T ro, bool same) {
  if constexpr (Js::no_series) {
    const T a = same ? S.a_para : S.a_anti;
    return a * (rn - ro) * M<T>::rcp((T(1) + S.b_ee * rn) * (T(1) + S.b_ee * ro));
  } else {
    return uee_val<T, Js>(S, rn, same) - uee_val<T, Js>(S, ro, same);
  }
}

// the same for the e-n pair with nucleus I: -Z (r' - r) / ((1+b r')(1+b r))
template <class T, class Js = JsRt>
__device__ __forceinline__ T uen_diff(const DevSys<T>& S, T rn, T ro, int I) {
  if constexpr (Js::no_series) {
    return -S.Z[I] * (rn - ro) * M<T>::rcp((T(1) + S.b_en * rn) * (T(1) + S.b_en * ro));
  } else {
    return uen_val<T, Js>(S, rn, I) - uen_val<T, Js>(S, ro, I);
  }
}

// u'(r) and u''(r) + 2 u'(r)/r for an e-e pair
template <class T, class Js = JsRt>
__device__ __forceinline__ void uee_d(const DevSys<T>& S, T r, T ir, bool same, T& u1, T& lap) {
  T D = T(1) + S.b_ee * r;
  T iD = M<T>::rcp(D);
  T iD2 = iD * iD;
  T a = same ? S.a_para : S.a_anti;
  T s1 = 0, s2 = 0;
  const int K = Js::kee(S);
  if (K > 0) {
    T s, x = r * iD;
    series(K, same ? S.c_para : S.c_anti, x, s, s1, s2);
  }
  T ap = a + s1;
  u1 = ap * iD2;
  lap = s2 * iD2 * iD2 - T(2) * S.b_ee * ap * iD2 * iD + T(2) * u1 * ir;
}

template <class T, class Js = JsRt>
__device__ __forceinline__ void uen_d(const DevSys<T>& S, T r, T ir, int I, T& u1, T& lap) {
  T D = T(1) + S.b_en * r;
  T iD = M<T>::rcp(D);
  T iD2 = iD * iD;
  T s1 = 0, s2 = 0;
  const int K = Js::ken(S);
  if (K > 0) {
    T s, y = r * iD;
    series(K, S.d_en[S.species[I]], y, s, s1, s2);
  }
  T ap = -S.Z[I] + s1;
  u1 = ap * iD2;
  lap = s2 * iD2 * iD2 - T(2) * S.b_en * ap * iD2 * iD + T(2) * u1 * ir;
}

// ---------------------------------------------------------------------------
// Gauss-Jordan inversion with partial pivoting, fully register resident
// (row swaps by selects).  Returns ln|det| and sign; a singular matrix
// (an exactly zero pivot) yields the identity, as methane.rs:829 does.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ bool gj_invert(double a[N][N], double inv[N][N], double& logabs, double& sgn) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) inv[i][j] = (i == j) ? 1.0 : 0.0;
  bool singular = false;
  double la = 0.0, sg = 1.0;
#pragma unroll
  for (int c = 0; c < N; ++c) {
    int p = c;
    double best = fabs(a[c][c]);
#pragma unroll
    for (int i = c + 1; i < N; ++i) {
      double v = fabs(a[i][c]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    if (best == 0.0) singular = true;
    if (p != c) sg = -sg;
#pragma unroll
    for (int i = c + 1; i < N; ++i) {
      const bool sw = (p == i);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double t = a[c][j], u = a[i][j];
        a[c][j] = sw ? u : t;
        a[i][j] = sw ? t : u;
        double t2 = inv[c][j], u2 = inv[i][j];
        inv[c][j] = sw ? u2 : t2;
        inv[i][j] = sw ? t2 : u2;
      }
    }
    const double pv = a[c][c];
    la += log(fabs(pv));
    if (pv < 0.0) sg = -sg;
    const double ip = singular ? 0.0 : 1.0 / pv;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      a[c][j] *= ip;
      inv[c][j] *= ip;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i == c) continue;
      const double f = a[i][c];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        a[i][j] -= f * a[c][j];
        inv[i][j] -= f * inv[c][j];
      }
    }
  }
  if (singular) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) inv[i][j] = (i == j) ? 1.0 : 0.0;
    la = -INFINITY;
    sg = 0.0;
  }
  logabs = la;
  sgn = sg;
  return singular;
}

// V_nn (methane.rs:908-924), computed on the host and carried in DevSys.

}  // namespace qmc
