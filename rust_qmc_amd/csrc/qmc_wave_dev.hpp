// qmc_wave_dev.hpp -- device pieces of the wave engine shared by its
// sweep / build / evaluation kernels and the wave derivative kernel: the
// wave-wide sum, the LDS carve-out of one walker, and the FP64 Gauss-Jordan
// build of a spin sector's inverse Slater matrix.
#pragma once

#include "qmc_engine.hpp"

namespace qmc {

// Wave-wide sum (all 64 lanes active), the same value in every lane and
// the same order on every call: DPP moves within each row of 16 lanes
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror:
// after them every lane of a row holds the row's sum), then the four row
// sums read into scalar registers and added as (r0 + r16) + (r32 + r48).
// The DPP moves are register-to-register; the butterfly of LDS permutes it
// replaces cost a round trip through LDS per step.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float read_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
// Orders one wavefront's LDS accesses across its lanes: a release fence,
// the wavefront barrier and an acquire fence at wavefront scope, so a value
// one lane stores is what another lane of the same wavefront loads after
// it (the compiler may not move either access across).  No s_barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over each row of 16 lanes (all lanes active): the first four DPP
// steps of wsum; every lane of a row gets the bit-identical row sum.
template <class T>
__device__ __forceinline__ T rsum(T v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return v;
}

template <class T>
__device__ __forceinline__ T wsum(T v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Visits the elements idx = lane, lane + 64, ... of an n x n row-major
// matrix as f(idx, row, col) with one division per call: each stride of 64
// elements advances the column by 64 mod n and the row by 64 / n, with the
// carry (n <= 64).
template <class F>
__device__ __forceinline__ void for_each_elem(int n, F f) {
  if (n <= 0) return;
  const int lane = threadIdx.x & 63;
  const int dr = 64 / n, dc = 64 % n;
  int r = lane / n, c = lane % n;
  for (int idx = lane; idx < n * n; idx += 64) {
    f(idx, r, c);
    r += dr;
    c += dc;
    if (c >= n) {
      c -= n;
      ++r;
    }
  }
}

// LDS carve-out of one walker (element counts in T unless noted)
struct WaveLayout {
  int N, nu, nd, nmax, A;
  __device__ __host__ int pos() const { return 0; }
  __device__ __host__ int ainv(int spin) const { return 3 * N + (spin ? nu * nu : 0); }
  __device__ __host__ int chi() const { return 3 * N + nu * nu + nd * nd; }
  __device__ __host__ int phi() const { return chi() + A; }
  __device__ __host__ int cvec() const { return phi() + nmax; }
  __device__ __host__ int wvec() const { return cvec() + nmax; }
  __device__ __host__ int wao() const { return wvec() + nmax; }  // AO weights of a drift-diffusion move
  __device__ __host__ int mpart() const { return wao() + A; }    // 64 partial MO sums of a move
  __device__ __host__ int total() const { return mpart() + 64; }
};

// Values and gradients of the AOs of shell sh at (x, y, z), shell-local
// (v[q], g[q][0..2] for q below the shell's AO count); returns that count.
// The same functions as shell_values and det_terms_w: contracted Gaussians
// with s / p / sp angular parts, and radial STO sums whose gradient at the
// centre is taken as zero.
template <class T>
__device__ __forceinline__ int shell_values_grad(const DevSys<T>& S, int sh, T x, T y, T z, T v[4], T g[4][3]) {
  const int c = S.sh_center[sh], kind = S.sh_kind[sh], p0 = S.sh_prim0[sh], np = S.sh_nprim[sh];
  const T dx = x - S.nuc[c][0], dy = y - S.nuc[c][1], dz = z - S.nuc[c][2];
  const T r2 = dx * dx + dy * dy + dz * dz;
  const T d[3] = {dx, dy, dz};
  if (kind <= SH_P1) {
    // R = sum c e, Ra = sum c a e for the s and p coefficient sets
    T rs = 0, rsa = 0, rp = 0, rpa = 0;
    for (int k = 0; k < np; ++k) {
      const T a = S.prim_exp[p0 + k];
      const T e = M<T>::exn(-a * r2);
      const T cs = S.coef_s[p0 + k] * e, cp = S.coef_p[p0 + k] * e;
      rs += cs;
      rsa += cs * a;
      rp += cp;
      rpa += cp * a;
    }
    int q = 0;
    if (kind == SH_S || kind == SH_SP) {  // grad = -2 Ra d
      v[0] = rs;
      for (int j = 0; j < 3; ++j) g[0][j] = T(-2) * rsa * d[j];
      q = 1;
    }
    if (kind == SH_S) return 1;
    // p_m = d_m R: grad_j = delta_jm R - 2 d_m Ra d_j
    const int m0 = kind == SH_P1 ? S.sh_m[sh] : 0, m1 = kind == SH_P1 ? m0 + 1 : 3;
    for (int m = m0; m < m1; ++m, ++q) {
      v[q] = d[m] * rp;
      for (int j = 0; j < 3; ++j) g[q][j] = (j == m ? rp : T(0)) - T(2) * d[m] * rpa * d[j];
    }
    return q;
  }
  // radial STO: f(s) = sum c s^p e^{-zeta s}; grad = f'(s) d / s
  const T s = M<T>::sq(r2);
  T f = 0, d1 = 0;
  for (int k = 0; k < np; ++k) {
    const T zt = S.prim_exp[p0 + k];
    const int p = S.prim_pow[p0 + k];
    T sp = T(1);
    for (int q = 0; q < p; ++q) sp *= s;
    const T spm1 = p >= 1 ? sp / s : T(0);
    const T e = S.coef_s[p0 + k] * M<T>::exn(-zt * s);
    f += e * sp;
    d1 += e * (-zt * sp + T(p) * spm1);
  }
  v[0] = f;
  const T gs = s == T(0) ? T(0) : d1 / s;
  for (int j = 0; j < 3; ++j) g[0][j] = gs * d[j];
  return 1;
}

template <class T>
__device__ __forceinline__ T* ainv_at(T* L, const WaveLayout& Y, int spin, int n, int m, int k) {
  return L + Y.ainv(spin) + m * n + k;
}

// ---------------------------------------------------------------------------
// FP64 build of one spin sector: Slater matrix from the positions in LDS,
// Gauss-Jordan inversion with partial pivoting by the wave.  a and inv are
// FP64 LDS scratch of n*n each.  Returns ln|det| and sign (wave-uniform).
// ---------------------------------------------------------------------------
template <class T>
__device__ bool wave_sector_inverse(const DevSys<double>& S, int spin, const T* pos, int n, double* a, double* inv,
                                    double& logabs, double& sgn) {
  const int lane = threadIdx.x;
  logabs = 0.0;
  sgn = 1.0;
  if (n == 0) return false;
  // rows: lane k evaluates all MOs at electron k of the sector
  for (int k = lane; k < n; k += 64) {
    const int e = S.sector_elec[spin][k];
    double xr = (double)pos[3 * e], yr = (double)pos[3 * e + 1], zr = (double)pos[3 * e + 2];
    for (int m = 0; m < n; ++m) a[k * n + m] = 0.0;
    for (int sh = 0; sh < S.n_shell; ++sh) {
      double chi[4];
      const int a0 = S.sh_ao0[sh];
      const int kind = S.sh_kind[sh];
      const int nao = kind == SH_SP ? 4 : (kind == SH_P ? 3 : 1);
      shell_values(S, sh, xr, yr, zr, chi, true);
      for (int q = 0; q < nao; ++q)
        for (int m = 0; m < n; ++m) a[k * n + m] += S.moT[spin][a0 + q][m] * chi[q];
    }
  }
  for (int i = lane; i < n * n; i += 64) inv[i] = (i / n == i % n) ? 1.0 : 0.0;
  __syncthreads();
  bool singular = false;
  __shared__ double fcol[64];
  for (int c = 0; c < n; ++c) {
    // pivot: largest |a[i][c]|, i >= c (ties to the lowest row)
    double best = -1.0;
    int p = c;
    for (int i = c + lane; i < n; i += 64) {
      double v = fabs(a[i * n + c]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      double ob = __shfl_xor(best, off, 64);
      int op = __shfl_xor(p, off, 64);
      if (ob > best || (ob == best && op < p)) {
        best = ob;
        p = op;
      }
    }
    if (best == 0.0) singular = true;
    if (p != c) {
      sgn = -sgn;
      for (int j = lane; j < n; j += 64) {
        double t = a[c * n + j];
        a[c * n + j] = a[p * n + j];
        a[p * n + j] = t;
        t = inv[c * n + j];
        inv[c * n + j] = inv[p * n + j];
        inv[p * n + j] = t;
      }
    }
    __syncthreads();
    const double pv = a[c * n + c];
    logabs += log(fabs(pv));
    if (pv < 0.0) sgn = -sgn;
    const double ip = singular ? 0.0 : 1.0 / pv;
    __syncthreads();
    for (int j = lane; j < n; j += 64) {
      a[c * n + j] *= ip;
      inv[c * n + j] *= ip;
    }
    for (int i = lane; i < n; i += 64) fcol[i] = a[i * n + c];
    __syncthreads();
    for_each_elem(n, [&](int idx, int i, int j) {
      if (i == c) return;
      const double f = fcol[i];
      a[idx] -= f * a[c * n + j];
      inv[idx] -= f * inv[c * n + j];
    });
    __syncthreads();
  }
  if (singular) {
    for (int i = lane; i < n * n; i += 64) inv[i] = (i / n == i % n) ? 1.0 : 0.0;
    logabs = -INFINITY;
    sgn = 0.0;
    __syncthreads();
  }
  return singular;
}

}  // namespace qmc
