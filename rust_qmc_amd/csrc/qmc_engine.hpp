// qmc_engine.hpp -- the per-walker engine: one thread owns one walker whose
// electron positions and per-spin inverse Slater matrices live in VGPRs
// for the whole launch.  Spin sectors are walked with a cyclic rotation
// of the sector's position rows and inverse columns, so the electron being
// moved (or measured) always sits in slot 0 and every register array is
// indexed with compile-time indices.
#pragma once

#include "qmc_device.hpp"
#include "qmc_kernels.hpp"

namespace qmc {

// one walker's register-resident state
template <int NU, int ND, class T>
struct WalkerRegs {
  T ru[Cap<NU>::v][3], rd[Cap<ND>::v][3];
  T au[Cap<NU>::v][Cap<NU>::v], ad[Cap<ND>::v][Cap<ND>::v];
};

// rotate sector rows: slot k <- slot k+1, slot N-1 <- old slot 0
template <int NA, class T>
__device__ __forceinline__ void rotate(T ra[Cap<NA>::v][3], T ai[Cap<NA>::v][Cap<NA>::v]) {
  if (NA <= 1) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T t = ra[0][c];
#pragma unroll
    for (int k = 0; k + 1 < NA; ++k) ra[k][c] = ra[k + 1][c];
    ra[NA - 1][c] = t;
  }
#pragma unroll
  for (int m = 0; m < NA; ++m) {
    T t = ai[m][0];
#pragma unroll
    for (int k = 0; k + 1 < NA; ++k) ai[m][k] = ai[m][k + 1];
    ai[m][NA - 1] = t;
  }
}

// Per-electron local-energy terms of the electron in slot 0 of sector A.
// Returns the contribution to -1/2 sum lap(psi)/psi and adds the electron's
// share of the potential (1/2 of each e-e pair) to pot.  Coincident
// particles (r < 1e-10) contribute no Jastrow terms; the selects keep the
// pair loops free of branches.
template <int NA, int NB, class T, class Bs = RtBasis, class Js = JsRt>
__device__ __forceinline__ void electron_terms(const DevSys<T>& S, int spin, const T ra[Cap<NA>::v][3],
                                               const T ai[Cap<NA>::v][Cap<NA>::v],
                                               const T rb[Cap<NB>::v][3], T& gx, T& gy, T& gz, T& lap,
                                               double& pot) {
  const T x = ra[0][0], y = ra[0][1], z = ra[0][2];
  // determinant part: grad D / D and lap D / D from the inverse column
  T gdx, gdy, gdz, ld;
  det_terms<NA, T, NoShell, Bs>(S, spin, x, y, z, ai, gdx, gdy, gdz, ld, NoShell{});
  // Jastrow part
  T gux = 0, guy = 0, guz = 0, lu = 0, pairsq = 0;
  T pee = 0, pen = 0;
  auto ee = [&](T px, T py, T pz, bool same) {
    T dx = x - px, dy = y - py, dz = z - pz;
    T r, ir;
    M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
    T u1, lp;
    uee_d<T, Js>(S, r, ir, same, u1, lp);
    const bool ok = r >= T(1e-10);
    const T f = ok ? u1 * ir : T(0);
    gux += f * dx;
    guy += f * dy;
    guz += f * dz;
    lu += ok ? lp : T(0);
    pairsq += ok ? u1 * u1 : T(0);
    pee += ir;
  };
#pragma unroll
  for (int j = 1; j < NA; ++j) ee(ra[j][0], ra[j][1], ra[j][2], true);
#pragma unroll
  for (int j = 0; j < NB; ++j) ee(rb[j][0], rb[j][1], rb[j][2], false);
  const bool en_on = S.en_on;
  Bs::nuclei(S, [&](int I) {
    T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
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
  });
  // Jastrow2's gradient carries an extra factor F (jastrow.rs:102) and its
  // Laplacian omits the cross terms (jastrow.rs:110-125) under the flags.
  const T gs = S.grad_scale;
  T gjx = gs * gux, gjy = gs * guy, gjz = gs * guz;
  T lj = (S.flags & F_J2_LAP_QUIRK) ? (lu + pairsq) : (lu + gux * gux + guy * guy + guz * guz);
  gx = gdx + gjx;
  gy = gdy + gjy;
  gz = gdz + gjz;
  lap = ld + lj + T(2) * (gdx * gjx + gdy * gjy + gdz * gjz);
  pot += 0.5 * (double)pee - (double)pen;
}

// shift a sector's per-electron vector one slot, as rotate() does
template <int NA, class T>
__device__ __forceinline__ void rotate_vec(T* v) {
  if (NA <= 1) return;
  T t = v[0];
#pragma unroll
  for (int k = 0; k + 1 < NA; ++k) v[k] = v[k + 1];
  v[NA - 1] = t;
}

// Local energy E_L = -1/2 sum_i lap_i psi / psi + V, V = -sum Z/r + sum 1/r_ij + V_nn
// (methane.rs:1154-1182), summing the Jastrow and potential terms over unique
// pairs first (each e-e pair once, with compile-time electron indices), which
// gives every electron its Jastrow gradient and Laplacian terms; the
// determinant terms then walk each sector with the rotation, the Jastrow
// gradients rotating along.  Sectors are left in their canonical order.  The
// per-electron arrays fit the register file in FP32 only.
template <int NU, int ND, class T, class Bs, class Js>
__device__ __forceinline__ double local_energy_pairs(const DevSys<T>& S, T ru[Cap<NU>::v][3],
                                                     T au[Cap<NU>::v][Cap<NU>::v], T rd[Cap<ND>::v][3],
                                                     T ad[Cap<ND>::v][Cap<ND>::v], double* gout, double* lout,
                                                     int64_t ld) {
  constexpr int N = NU + ND;
  constexpr int NC = N > 0 ? N : 1;
  T gx[NC], gy[NC], gz[NC], lu[NC];
#pragma unroll
  for (int i = 0; i < N; ++i) gx[i] = gy[i] = gz[i] = lu[i] = T(0);
  // Jastrow2's Laplacian omits the cross terms (jastrow.rs:110-125): the
  // squared pair derivatives replace |grad u|^2 under the flag
  const bool quirk = S.flags & F_J2_LAP_QUIRK;
  const T qw = quirk ? T(1) : T(0);
  auto px = [&](int i, int c) -> T { return i < NU ? ru[i][c] : rd[i - NU][c]; };
  // potential sums: FP32 within an electron's row, FP64 across rows
  double pot = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T pee = 0;
#pragma unroll
    for (int j = i + 1; j < N; ++j) {
      const T dx = px(i, 0) - px(j, 0), dy = px(i, 1) - px(j, 1), dz = px(i, 2) - px(j, 2);
      // 1/r from one reciprocal square root; r = 0 gives r = NaN, which the
      // r >= 1e-10 guards reject like the coincident pair they are
      const T d2 = dx * dx + dy * dy + dz * dz;
      const T ir = M<T>::rsq(d2);
      const T r = d2 * ir;
      T u1, lp;
      uee_d<T, Js>(S, r, ir, (i < NU) == (j < NU), u1, lp);
      const bool ok = r >= T(1e-10);
      const T f = ok ? u1 * ir : T(0);
      const T fx = f * dx, fy = f * dy, fz = f * dz;
      gx[i] += fx;
      gy[i] += fy;
      gz[i] += fz;
      gx[j] -= fx;
      gy[j] -= fy;
      gz[j] -= fz;
      const T l = ok ? lp + qw * u1 * u1 : T(0);
      lu[i] += l;
      lu[j] += l;
      pee += ir;
    }
    pot += (double)pee;
  }
  // e-n pairs: one uniform branch on the e-n Jastrow for the whole pass
  auto en_pass = [&](auto en_tag) {
    constexpr bool EN = decltype(en_tag)::value;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const T x = px(i, 0), y = px(i, 1), z = px(i, 2);
      T pen = 0;
      Bs::nuclei(S, [&](int I) {
        const T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
        const T d2 = dx * dx + dy * dy + dz * dz;
        const T ir = M<T>::rsq(d2);
        const T r = d2 * ir;
        pen += S.Z[I] * ir;
        if constexpr (EN) {
          T u1, lp;
          uen_d<T, Js>(S, r, ir, I, u1, lp);
          const bool ok = r >= T(1e-10);
          const T f = ok ? u1 * ir : T(0);
          gx[i] += f * dx;
          gy[i] += f * dy;
          gz[i] += f * dz;
          lu[i] += ok ? lp : T(0);
        }
      });
      pot -= (double)pen;
    }
  };
  if (S.en_on)
    en_pass(std::true_type{});
  else
    en_pass(std::false_type{});
  // |grad u|^2 completes the Jastrow Laplacian; Jastrow2's gradient carries
  // an extra factor F (jastrow.rs:102) under the flags
  const T gs = S.grad_scale;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (!quirk) lu[i] += gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i];
    gx[i] *= gs;
    gy[i] *= gs;
    gz[i] *= gs;
  }
  double kin = 0.0;
  auto sector = [&](auto na_tag, int spin, auto& ra, auto& ai, T* gxs, T* gys, T* gzs, T* ls) {
    constexpr int NA = decltype(na_tag)::value;
    // unrolled: the rotations become register renames
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      T gdx, gdy, gdz, ldd;
      det_terms<NA, T, NoShell, Bs>(S, spin, ra[0][0], ra[0][1], ra[0][2], ai, gdx, gdy, gdz, ldd, NoShell{});
      const T lap = ldd + ls[0] + T(2) * (gdx * gxs[0] + gdy * gys[0] + gdz * gzs[0]);
      kin += -0.5 * (double)lap;
      if (gout) {
        const int e = S.sector_elec[spin][k];
        gout[(e * 3 + 0) * ld] = gdx + gxs[0];
        gout[(e * 3 + 1) * ld] = gdy + gys[0];
        gout[(e * 3 + 2) * ld] = gdz + gzs[0];
        lout[e * ld] = lap;
      }
      rotate<NA>(ra, ai);
      rotate_vec<NA>(gxs);
      rotate_vec<NA>(gys);
      rotate_vec<NA>(gzs);
      rotate_vec<NA>(ls);
    }
  };
  sector(std::integral_constant<int, NU>{}, 0, ru, au, gx, gy, gz, lu);
  sector(std::integral_constant<int, ND>{}, 1, rd, ad, gx + NU, gy + NU, gz + NU, lu + NU);
  return kin + pot + S.vnn;
}

// The same E_L electron by electron (each e-e pair visited from both ends),
// for FP64 state whose per-electron arrays would not fit the register file.
template <int NU, int ND, class T, class Bs, class Js>
__device__ __forceinline__ double local_energy_rows(const DevSys<T>& S, T ru[Cap<NU>::v][3],
                                                    T au[Cap<NU>::v][Cap<NU>::v], T rd[Cap<ND>::v][3],
                                                    T ad[Cap<ND>::v][Cap<ND>::v], double* gout, double* lout,
                                                    int64_t ld) {
  double kin = 0.0, pot = 0.0;
  auto sector = [&](auto na_tag, auto nb_tag, int spin, auto& ra, auto& ai, auto& rb) {
    constexpr int NA = decltype(na_tag)::value, NB = decltype(nb_tag)::value;
#pragma unroll 1
    for (int k = 0; k < NA; ++k) {
      T gx, gy, gz, lap;
      electron_terms<NA, NB, T, Bs, Js>(S, spin, ra, ai, rb, gx, gy, gz, lap, pot);
      kin += -0.5 * (double)lap;
      if (gout) {
        const int e = S.sector_elec[spin][k];
        gout[(e * 3 + 0) * ld] = gx;
        gout[(e * 3 + 1) * ld] = gy;
        gout[(e * 3 + 2) * ld] = gz;
        lout[e * ld] = lap;
      }
      rotate<NA>(ra, ai);
    }
  };
  sector(std::integral_constant<int, NU>{}, std::integral_constant<int, ND>{}, 0, ru, au, rd);
  sector(std::integral_constant<int, ND>{}, std::integral_constant<int, NU>{}, 1, rd, ad, ru);
  return kin + pot + S.vnn;
}

// Local energy of the register state; when gout/lout are given, per-electron
// grad ln psi and lap psi / psi are written to gout[(e*3+c)*ld], lout[e*ld]
// for global electron index e.
template <int NU, int ND, class T, class Bs = RtBasis, class Js = JsRt>
__device__ __forceinline__ double local_energy(const DevSys<T>& S, T ru[Cap<NU>::v][3],
                                               T au[Cap<NU>::v][Cap<NU>::v], T rd[Cap<ND>::v][3],
                                               T ad[Cap<ND>::v][Cap<ND>::v], double* gout, double* lout,
                                               int64_t ld) {
  if constexpr (sizeof(T) == 4)
    return local_energy_pairs<NU, ND, T, Bs, Js>(S, ru, au, rd, ad, gout, lout, ld);
  else
    return local_energy_rows<NU, ND, T, Bs, Js>(S, ru, au, rd, ad, gout, lout, ld);
}

// Full Jastrow exponent ln J (FP64), including the pair constant.
template <int NU, int ND, class T, class Bs = RtBasis, class Js = JsRt>
__device__ __forceinline__ double jastrow_u(const DevSys<double>& S, const T ru[Cap<NU>::v][3],
                                            const T rd[Cap<ND>::v][3]) {
  constexpr int N = NU + ND;
  double r[N > 0 ? N : 1][3];
#pragma unroll
  for (int k = 0; k < NU; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) r[k][c] = (double)ru[k][c];
#pragma unroll
  for (int k = 0; k < ND; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) r[NU + k][c] = (double)rd[k][c];
  double u = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i + 1; j < N; ++j) {
      double dx = r[i][0] - r[j][0], dy = r[i][1] - r[j][1], dz = r[i][2] - r[j][2];
      double rr = sqrt(dx * dx + dy * dy + dz * dz);
      u += uee_val<double, Js>(S, rr, (i < NU) == (j < NU)) + S.cst_ee;
    }
  if (S.en_on)
#pragma unroll
    for (int i = 0; i < N; ++i)
      Bs::nuclei(S, [&](int I) {
        double dx = r[i][0] - S.nuc[I][0], dy = r[i][1] - S.nuc[I][1], dz = r[i][2] - S.nuc[I][2];
        u += uen_val<double, Js>(S, sqrt(dx * dx + dy * dy + dz * dz), I);
      });
  return u;
}

// Build one sector's Slater matrix in FP64 and invert it.
template <int NA, class T, class Bs = RtBasis>
__device__ __forceinline__ bool sector_inverse(const DevSys<double>& S, int spin, const T ra[Cap<NA>::v][3],
                                               T ai[Cap<NA>::v][Cap<NA>::v], double& logabs, double& sgn) {
  if (NA == 0) {
    logabs = 0.0;
    sgn = 1.0;
    return false;
  }
  constexpr int C = Cap<NA>::v;
  double a[C][C], inv[C][C];
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    double phi[C];
    mo_values<C, double, Bs>(S, spin, (double)ra[k][0], (double)ra[k][1], (double)ra[k][2], phi);
#pragma unroll
    for (int m = 0; m < NA; ++m) a[k][m] = phi[m];
  }
  bool sing = gj_invert<C>(a, inv, logabs, sgn);
#pragma unroll
  for (int m = 0; m < NA; ++m)
#pragma unroll
    for (int k = 0; k < NA; ++k) ai[m][k] = (T)inv[m][k];
  return sing;
}

// e-n Jastrow sum sum_I u_en(|r - R_I|) of a position
template <class T, class Bs, class Js>
__device__ __forceinline__ T en_sum(const DevSys<T>& S, T x, T y, T z) {
  T u = 0;
  Bs::nuclei(S, [&](int I) {
    const T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
    u += uen_val<T, Js>(S, M<T>::dist(dx * dx + dy * dy + dz * dz), I);
  });
  return u;
}

// One sector's single-electron Metropolis moves (Gaussian proposal, ratio
// from the inverse column, Jastrow difference from the moved electron's
// pairs, Sherman-Morrison row update on acceptance).  uen holds each
// sector electron's e-n Jastrow sum at its current position (slot order,
// rotating with the sector), so a move evaluates the e-n terms at the
// proposal only.
template <int NA, int NB, class T, class Bs = RtBasis, class Js = JsRt>
__device__ __forceinline__ void sector_moves(const DevSys<T>& S, int spin, T ra[Cap<NA>::v][3],
                                             T ai[Cap<NA>::v][Cap<NA>::v], const T rb[Cap<NB>::v][3],
                                             T uen[Cap<NA>::v], uint32_t w, uint32_t step, uint32_t k0,
                                             uint32_t k1, T sigma, double& logdet, double& sgn, double& ju,
                                             int& acc) {
  LogAcc<T> lacc;
#pragma unroll 1
  for (int k = 0; k < NA; ++k) {
    const uint32_t e = (uint32_t)S.sector_elec[spin][k];
    T z[3], U;
    Rng<T>::move(philox(w, step, e, RNG_SINGLE, k0, k1), z, U);
    const T ox = ra[0][0], oy = ra[0][1], oz = ra[0][2];
    const T nx = ox + sigma * z[0], ny = oy + sigma * z[1], nz = oz + sigma * z[2];
    T phi[Cap<NA>::v];
    mo_values<NA, T, Bs>(S, spin, nx, ny, nz, phi);
    T R = 0;
#pragma unroll
    for (int m = 0; m < NA; ++m) R += phi[m] * ai[m][0];
    // Jastrow difference
    T du = 0;
    auto pair = [&](T px, T py, T pz, bool same) {
      T ax = nx - px, ay = ny - py, az = nz - pz;
      T bx = ox - px, by = oy - py, bz = oz - pz;
      du += uee_diff<T, Js>(S, M<T>::dist(ax * ax + ay * ay + az * az), M<T>::dist(bx * bx + by * by + bz * bz), same);
    };
#pragma unroll
    for (int j = 1; j < NA; ++j) pair(ra[j][0], ra[j][1], ra[j][2], true);
#pragma unroll
    for (int j = 0; j < NB; ++j) pair(rb[j][0], rb[j][1], rb[j][2], false);
    T un = 0;
    if (S.en_on) {
      un = en_sum<T, Bs, Js>(S, nx, ny, nz);
      du += un - uen[0];
    }
    const T p = R * R * M<T>::ex(T(2) * du);
    if (U < p) {
      // Sherman-Morrison: A'^-1 = A^-1 - (A^-1 e_0)(phi^T A^-1 - e_0^T)/R
      const T iR = M<T>::rcp(R);
      T c[Cap<NA>::v];
#pragma unroll
      for (int m = 0; m < NA; ++m) c[m] = ai[m][0];
#pragma unroll
      for (int j = 1; j < NA; ++j) {
        T wj = 0;
#pragma unroll
        for (int m = 0; m < NA; ++m) wj += phi[m] * ai[m][j];
        const T f = wj * iR;
#pragma unroll
        for (int m = 0; m < NA; ++m) ai[m][j] -= c[m] * f;
      }
#pragma unroll
      for (int m = 0; m < NA; ++m) ai[m][0] = c[m] * iR;
      ra[0][0] = nx;
      ra[0][1] = ny;
      ra[0][2] = nz;
      lacc.add(logdet, R);
      if (R < T(0)) sgn = -sgn;
      ju += (double)du;
      ++acc;
      uen[0] = un;
    }
    rotate<NA>(ra, ai);
    rotate_vec<NA>(uen);
  }
  lacc.fold(logdet);
}

// wave-wide sum (deterministic butterfly order)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Ensemble state in HBM (structure-of-arrays, walker index fastest) <->
// one walker's registers.
// ---------------------------------------------------------------------------
template <int NU, int ND, class T>
__device__ __forceinline__ void load_positions(const DevSys<T>& S, const T* pos, int64_t ld, int64_t w,
                                               WalkerRegs<NU, ND, T>& W) {
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    const int e = S.sector_elec[0][k];
#pragma unroll
    for (int c = 0; c < 3; ++c) W.ru[k][c] = pos[(e * 3 + c) * ld + w];
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const int e = S.sector_elec[1][k];
#pragma unroll
    for (int c = 0; c < 3; ++c) W.rd[k][c] = pos[(e * 3 + c) * ld + w];
  }
}

template <int NU, int ND, class T>
__device__ __forceinline__ void store_positions(const DevSys<T>& S, T* pos, int64_t ld, int64_t w,
                                                const WalkerRegs<NU, ND, T>& W) {
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    const int e = S.sector_elec[0][k];
#pragma unroll
    for (int c = 0; c < 3; ++c) pos[(e * 3 + c) * ld + w] = W.ru[k][c];
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const int e = S.sector_elec[1][k];
#pragma unroll
    for (int c = 0; c < 3; ++c) pos[(e * 3 + c) * ld + w] = W.rd[k][c];
  }
}

template <int NU, int ND, class T>
__device__ __forceinline__ void load_inverse(const T* ainv, int64_t ld, int64_t w, WalkerRegs<NU, ND, T>& W) {
#pragma unroll
  for (int m = 0; m < NU; ++m)
#pragma unroll
    for (int k = 0; k < NU; ++k) W.au[m][k] = ainv[(m * NU + k) * ld + w];
#pragma unroll
  for (int m = 0; m < ND; ++m)
#pragma unroll
    for (int k = 0; k < ND; ++k) W.ad[m][k] = ainv[(NU * NU + m * ND + k) * ld + w];
}

template <int NU, int ND, class T>
__device__ __forceinline__ void store_inverse(T* ainv, int64_t ld, int64_t w, const WalkerRegs<NU, ND, T>& W) {
#pragma unroll
  for (int m = 0; m < NU; ++m)
#pragma unroll
    for (int k = 0; k < NU; ++k) ainv[(m * NU + k) * ld + w] = W.au[m][k];
#pragma unroll
  for (int m = 0; m < ND; ++m)
#pragma unroll
    for (int k = 0; k < ND; ++k) ainv[(NU * NU + m * ND + k) * ld + w] = W.ad[m][k];
}

// ---------------------------------------------------------------------------
// n_steps single-electron sweeps of the 64 walkers of block blk (one
// walker per lane), E_L after each, and per-(step, block) partial sums
// {sum E_L, sum E_L^2, accepted moves} into partial[step][nblk][3].
// ---------------------------------------------------------------------------
template <int NU, int ND, class Bs, class Js, class T>
__device__ __forceinline__ void sweep_block(const DevSys<T>& S, EnsView<T> E, SweepArgs A,
                                            double* __restrict__ partial, double* __restrict__ el_trace,
                                            int64_t blk, int64_t nblk, int lane) {
  const int64_t w = blk * 64 + lane;
  const bool live = w < E.W;
  const int64_t wl = live ? w : E.W - 1;  // dead lanes shadow the last walker and write nothing
  WalkerRegs<NU, ND, T> R;
  load_positions<NU, ND>(S, E.pos, E.ld, wl, R);
  load_inverse<NU, ND>(E.ainv, E.ld, wl, R);
  double logdet = E.logdet[wl], sgn = E.sign[wl], ju = E.ju[wl];
  double el = 0.0;
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t wid = (uint32_t)(A.walker0 + wl);
  const T sigma = (T)A.step_size;
  T uen_u[Cap<NU>::v], uen_d[Cap<ND>::v];
  const bool en_on = S.en_on;
#pragma unroll
  for (int k = 0; k < NU; ++k) uen_u[k] = en_on ? en_sum<T, Bs, Js>(S, R.ru[k][0], R.ru[k][1], R.ru[k][2]) : T(0);
#pragma unroll
  for (int k = 0; k < ND; ++k) uen_d[k] = en_on ? en_sum<T, Bs, Js>(S, R.rd[k][0], R.rd[k][1], R.rd[k][2]) : T(0);
  for (int s = 0; s < A.n_steps; ++s) {
    const uint32_t step = A.step0 + (uint32_t)s;
    int acc = 0;
    sector_moves<NU, ND, T, Bs, Js>(S, 0, R.ru, R.au, R.rd, uen_u, wid, step, k0, k1, sigma, logdet, sgn, ju, acc);
    sector_moves<ND, NU, T, Bs, Js>(S, 1, R.rd, R.ad, R.ru, uen_d, wid, step, k0, k1, sigma, logdet, sgn, ju, acc);
    el = local_energy<NU, ND, T, Bs, Js>(S, R.ru, R.au, R.rd, R.ad, nullptr, nullptr, 0);
    if (el_trace && live) el_trace[(int64_t)s * A.trace_ld + w] = el;
    double e1 = live ? el : 0.0, e2 = live ? el * el : 0.0, a = live ? (double)acc : 0.0;
    e1 = wave_sum(e1);
    e2 = wave_sum(e2);
    a = wave_sum(a);
    if (lane == 0) {
      double* p = partial + ((int64_t)s * nblk + blk) * 3;
      p[0] = e1;
      p[1] = e2;
      p[2] = a;
    }
  }
  if (live) {
    store_positions<NU, ND>(S, E.pos, E.ld, w, R);
    store_inverse<NU, ND>(E.ainv, E.ld, w, R);
    E.logdet[w] = logdet;
    E.sign[w] = sgn;
    E.ju[w] = ju;
    E.eloc[w] = el;
  }
}


// Jastrow value of the slot-0 electron of sector A placed at (x, y, z)
// (its pairs with the other electrons and the nuclei) and the gradient of
// that sum with respect to its position.
template <int NA, int NB, class T, class Bs, class Js>
__device__ __forceinline__ void slot0_jastrow(const DevSys<T>& S, T x, T y, T z, const T ra[Cap<NA>::v][3],
                                              const T rb[Cap<NB>::v][3], T& u, T& gx, T& gy, T& gz) {
  u = gx = gy = gz = T(0);
  auto ee = [&](T px, T py, T pz, bool same) {
    const T dx = x - px, dy = y - py, dz = z - pz;
    T r, ir;
    M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
    T u1, lp;
    uee_d<T, Js>(S, r, ir, same, u1, lp);
    u += uee_val<T, Js>(S, r, same);
    const T f = r >= T(1e-10) ? u1 * ir : T(0);
    gx += f * dx;
    gy += f * dy;
    gz += f * dz;
  };
#pragma unroll
  for (int j = 1; j < NA; ++j) ee(ra[j][0], ra[j][1], ra[j][2], true);
#pragma unroll
  for (int j = 0; j < NB; ++j) ee(rb[j][0], rb[j][1], rb[j][2], false);
  if (S.en_on)
    Bs::nuclei(S, [&](int I) {
      const T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
      T r, ir;
      M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
      T u1, lp;
      uen_d<T, Js>(S, r, ir, I, u1, lp);
      u += uen_val<T, Js>(S, r, I);
      const T f = r >= T(1e-10) ? u1 * ir : T(0);
      gx += f * dx;
      gy += f * dy;
      gz += f * dz;
    });
}

// tau F capped at max_drift in norm (DriftDiffusionVMC::limit_drift, vmc.rs:313-321)
template <class T>
__device__ __forceinline__ void cap_drift(T fx, T fy, T fz, T tau, T max_drift, T d[3]) {
  d[0] = tau * fx;
  d[1] = tau * fy;
  d[2] = tau * fz;
  const T m = M<T>::sq(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const T s = m > max_drift ? max_drift / m : T(1);
  d[0] *= s;
  d[1] *= s;
  d[2] *= s;
}

// One sector's drift-diffusion single-electron moves (DriftDiffusionVMC::
// sweep, vmc.rs:382-438): Langevin proposal r' = r + cap(tau F(r)) +
// sqrt(2 tau) z with the quantum force F = 2 grad ln|psi| of the moved
// electron, acceptance R^2 e^{2 du} e^{G_b - G_f} with the Green's function
// exponents of vmc.rs:323-346.  grad ln|D| at both positions comes from the
// AO gradients weighted by w_ao = sum_m C[m][ao] Ainv[m][0] (at r' divided
// by R, the new inverse column being the old one over R); the Jastrow part
// carries the Jastrow2 gradient factor of the flags.  Sherman-Morrison on
// acceptance as in sector_moves.
template <int NA, int NB, class T, class Bs = RtBasis, class Js = JsRt>
__device__ __forceinline__ void sector_moves_dd(const DevSys<T>& S, int spin, T ra[Cap<NA>::v][3],
                                                T ai[Cap<NA>::v][Cap<NA>::v], const T rb[Cap<NB>::v][3],
                                                uint32_t w, uint32_t step, uint32_t k0, uint32_t k1, T tau,
                                                T max_drift, double& logdet, double& sgn, double& ju, int& acc) {
  const T s2t = M<T>::sq(T(2) * tau), i4t = T(0.25) / tau;
  const T gs = S.grad_scale;
  LogAcc<T> lacc;
#pragma unroll 1
  for (int k = 0; k < NA; ++k) {
    const uint32_t e = (uint32_t)S.sector_elec[spin][k];
    auto wgt = [&](int ao) {
      T v = 0;
#pragma unroll
      for (int m = 0; m < NA; ++m) v += S.moT[spin][ao][m] * ai[m][0];
      return v;
    };
    const T ox = ra[0][0], oy = ra[0][1], oz = ra[0][2];
    // quantum force at the current position
    T gdx, gdy, gdz, ld, uo, gux, guy, guz;
    det_terms_w<T, decltype(wgt), NoShell, Bs>(S, ox, oy, oz, wgt, gdx, gdy, gdz, ld, NoShell{});
    slot0_jastrow<NA, NB, T, Bs, Js>(S, ox, oy, oz, ra, rb, uo, gux, guy, guz);
    T dold[3];
    cap_drift(T(2) * (gdx + gs * gux), T(2) * (gdy + gs * guy), T(2) * (gdz + gs * guz), tau, max_drift, dold);
    // Langevin proposal
    T z[3];
    T U;
    Rng<T>::move(philox(w, step, e, RNG_DRIFT, k0, k1), z, U);
    const T nx = ox + dold[0] + s2t * z[0], ny = oy + dold[1] + s2t * z[1], nz = oz + dold[2] + s2t * z[2];
    T phi[Cap<NA>::v];
    mo_values<NA, T, Bs>(S, spin, nx, ny, nz, phi);
    T R = 0;
#pragma unroll
    for (int m = 0; m < NA; ++m) R += phi[m] * ai[m][0];
    T hx, hy, hz, hl, un, vx, vy, vz;
    det_terms_w<T, decltype(wgt), NoShell, Bs>(S, nx, ny, nz, wgt, hx, hy, hz, hl, NoShell{});
    slot0_jastrow<NA, NB, T, Bs, Js>(S, nx, ny, nz, ra, rb, un, vx, vy, vz);
    const T iR = M<T>::rcp(R);
    T dnew[3];
    cap_drift(T(2) * (hx * iR + gs * vx), T(2) * (hy * iR + gs * vy), T(2) * (hz * iR + gs * vz), tau, max_drift,
              dnew);
    T gf = 0, gb = 0;
    {
      const T a0 = nx - ox - dold[0], a1 = ny - oy - dold[1], a2 = nz - oz - dold[2];
      const T b0 = ox - nx - dnew[0], b1 = oy - ny - dnew[1], b2 = oz - nz - dnew[2];
      gf = -(a0 * a0 + a1 * a1 + a2 * a2) * i4t;
      gb = -(b0 * b0 + b1 * b1 + b2 * b2) * i4t;
    }
    const T du = un - uo;
    const T p = R * R * M<T>::ex(T(2) * du + gb - gf);
    if (U < p) {
      T c[Cap<NA>::v];
#pragma unroll
      for (int m = 0; m < NA; ++m) c[m] = ai[m][0];
#pragma unroll
      for (int j = 1; j < NA; ++j) {
        T wj = 0;
#pragma unroll
        for (int m = 0; m < NA; ++m) wj += phi[m] * ai[m][j];
        const T f = wj * iR;
#pragma unroll
        for (int m = 0; m < NA; ++m) ai[m][j] -= c[m] * f;
      }
#pragma unroll
      for (int m = 0; m < NA; ++m) ai[m][0] = c[m] * iR;
      ra[0][0] = nx;
      ra[0][1] = ny;
      ra[0][2] = nz;
      lacc.add(logdet, R);
      if (R < T(0)) sgn = -sgn;
      ju += (double)du;
      ++acc;
    }
    rotate<NA>(ra, ai);
  }
  lacc.fold(logdet);
}

// n_steps drift-diffusion sweeps of block blk, as sweep_block
template <int NU, int ND, class Bs, class Js, class T>
__device__ __forceinline__ void dd_sweep_block(const DevSys<T>& S, EnsView<T> E, SweepArgs A,
                                               double* __restrict__ partial, double* __restrict__ el_trace,
                                               int64_t blk, int64_t nblk, int lane) {
  const int64_t W = E.n_live ? *E.n_live : E.W;
  if (blk * 64 >= W) {  // block past the live population: zero partial sums only
    if (lane == 0)
      for (int s = 0; s < A.n_steps; ++s)
        for (int q = 0; q < 3; ++q) partial[((int64_t)s * nblk + blk) * 3 + q] = 0.0;
    return;
  }
  const int64_t w = blk * 64 + lane;
  const bool live = w < W;
  const int64_t wl = live ? w : W - 1;
  WalkerRegs<NU, ND, T> R;
  load_positions<NU, ND>(S, E.pos, E.ld, wl, R);
  load_inverse<NU, ND>(E.ainv, E.ld, wl, R);
  double logdet = E.logdet[wl], sgn = E.sign[wl], ju = E.ju[wl];
  double el = 0.0;
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t wid = (uint32_t)(A.walker0 + wl);
  const T tau = (T)A.step_size, md = (T)A.max_drift;
  const uint32_t step0 = A.step0_dev ? *A.step0_dev : A.step0;
  for (int s = 0; s < A.n_steps; ++s) {
    const uint32_t step = step0 + (uint32_t)s;
    int acc = 0;
    sector_moves_dd<NU, ND, T, Bs, Js>(S, 0, R.ru, R.au, R.rd, wid, step, k0, k1, tau, md, logdet, sgn, ju, acc);
    sector_moves_dd<ND, NU, T, Bs, Js>(S, 1, R.rd, R.ad, R.ru, wid, step, k0, k1, tau, md, logdet, sgn, ju, acc);
    el = local_energy<NU, ND, T, Bs, Js>(S, R.ru, R.au, R.rd, R.ad, nullptr, nullptr, 0);
    if (el_trace && live) el_trace[(int64_t)s * A.trace_ld + w] = el;
    double e1 = live ? el : 0.0, e2 = live ? el * el : 0.0, a = live ? (double)acc : 0.0;
    e1 = wave_sum(e1);
    e2 = wave_sum(e2);
    a = wave_sum(a);
    if (lane == 0) {
      double* p = partial + ((int64_t)s * nblk + blk) * 3;
      p[0] = e1;
      p[1] = e2;
      p[2] = a;
    }
  }
  if (live) {
    store_positions<NU, ND>(S, E.pos, E.ld, w, R);
    store_inverse<NU, ND>(E.ainv, E.ld, w, R);
    E.logdet[w] = logdet;
    E.sign[w] = sgn;
    E.ju[w] = ju;
    E.eloc[w] = el;
  }
}

}  // namespace qmc
