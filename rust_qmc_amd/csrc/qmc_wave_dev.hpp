// qmc_wave_dev.hpp -- device pieces of the wave engine shared by its
// sweep / build / evaluation kernels and the wave derivative kernel: the
// wave-wide sum, the LDS carve-out of one walker, and the FP64 Gauss-Jordan
// build of a spin sector's inverse Slater matrix.
#pragma once

#include "qmc_engine.hpp"

namespace qmc {

// wave-wide sum, deterministic butterfly
template <class T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
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
  __device__ __host__ int total() const { return wvec() + nmax; }
};

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
    for (int idx = lane; idx < n * n; idx += 64) {
      const int i = idx / n, j = idx % n;
      if (i == c) continue;
      const double f = fcol[i];
      a[i * n + j] -= f * a[c * n + j];
      inv[i * n + j] -= f * inv[c * n + j];
    }
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
