// qmc_dmc.hip -- branching, compaction and population statistics of the
// importance-sampled DMC population (is_dmc.rs:272-300, 460-507).
//
// One step after the drift-diffusion sweep of buffer cur:
//   dmc_branch_kernel   copies m_w = min(floor(W_w + xi_w), 3) per walker with
//                       W_w = exp(tau (E_ref - (E_L' + E_L) / 2)), and the
//                       in-workgroup exclusive scan of m
//   dmc_scan_kernel     one workgroup scans the workgroup totals, giving the
//                       new population size (capped at the capacity)
//   dmc_copy_kernel     writes every survivor and clone into buffer nxt in
//                       source order (clones adjacent) with per-workgroup sums
//                       of E_L and of the HF force over the new population
//   dmc_finalize_kernel mean E_L, population, force sums, the sweep's accepted
//                       moves and the E_ref feedback
//                       E_ref = <E_L> - damping ln(N / N_target) / tau
// The live counts, E_ref and the sums stay in device memory.
#include <hip/hip_runtime.h>

#include "qmc_device.hpp"
#include "qmc_dmc.hpp"

namespace qmc {

namespace {

__device__ __forceinline__ double dmc_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// electronic HF force of walker w, Z_I sum_i (r_i - R_I) / |r_i - R_I|^3
// over electrons at distance > 1e-10 (hydrogen.rs:227-243,
// methane.rs:1241-1257), times the weight c
template <class T>
__device__ __forceinline__ void hf_electronic(const DevSys<double>& S, const T* pos, int64_t ld, int64_t w, int n3,
                                              int n_nuc, double c, double f[MAXN][3]) {
#pragma unroll
  for (int I = 0; I < MAXN; ++I) f[I][0] = f[I][1] = f[I][2] = 0.0;
  for (int q = 0; q < n3; q += 3) {
    const double x = (double)pos[(int64_t)q * ld + w], y = (double)pos[(int64_t)(q + 1) * ld + w],
                 z = (double)pos[(int64_t)(q + 2) * ld + w];
#pragma unroll
    for (int I = 0; I < MAXN; ++I) {
      if (I < n_nuc) {
        const double dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (d > 1e-10) {
          const double s = S.Z[I] / (d * d * d);
          f[I][0] += s * dx;
          f[I][1] += s * dy;
          f[I][2] += s * dz;
        }
      }
    }
  }
#pragma unroll
  for (int I = 0; I < MAXN; ++I) {
    f[I][0] *= c;
    f[I][1] *= c;
    f[I][2] *= c;
  }
}

// workgroup sums of the 3 n_nuc force components into sh[wave][q0 + ...]
__device__ __forceinline__ void wave_sums_force(double f[MAXN][3], int n_nuc, double (*sh)[1 + 3 * MAXN], int q0) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int I = 0; I < MAXN; ++I) {
    if (I < n_nuc) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double v = dmc_wave_sum(f[I][c]);
        if (lane == 0) sh[wv][q0 + 3 * I + c] = v;
      }
    }
  }
}

__global__ void __launch_bounds__(DMC_BLOCK) dmc_branch_kernel(DmcArgs a, DmcBufs b, int cur,
                                                               const double* __restrict__ eloc,
                                                               const double* __restrict__ eprev, uint32_t t) {
  const int64_t W = b.ctl->n_live[cur];
  const int64_t w = (int64_t)blockIdx.x * DMC_BLOCK + threadIdx.x;
  int m = 0;
  if (w < W) {
    // is_dmc.rs:284-289; the saturating float->usize cast of the reference
    // maps NaN to 0 and anything >= 3 to MAX_CLONES
    const double wgt = exp(a.tau_b * (b.ctl->e_ref - 0.5 * (eloc[w] + eprev[w])));
    const double xi = u01d(philox((uint32_t)w, t, 0u, RNG_BRANCH, (uint32_t)a.seed, (uint32_t)(a.seed >> 32)).x);
    const double v = wgt + xi;
    m = v >= (double)DMC_MAX_CLONES ? DMC_MAX_CLONES : (v >= 0.0 ? (int)v : 0);
  }
  b.copies[w] = m;
  // exclusive scan of m over the workgroup
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = m;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  __shared__ int wt[DMC_BLOCK / 64];
  if (lane == 63) wt[wv] = x;
  __syncthreads();
  int base = 0;
  for (int i = 0; i < wv; ++i) base += wt[i];
  b.loff[w] = base + x - m;
  if (threadIdx.x == DMC_BLOCK - 1) b.bsum[blockIdx.x] = (int64_t)(base + x);
}

__global__ void __launch_bounds__(1024) dmc_scan_kernel(DmcArgs a, DmcBufs b, int nxt, int nb) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int64_t v[4], s = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = t * 4 + i;
    v[i] = k < nb ? b.bsum[k] : 0;
    s += v[i];
  }
  int64_t x = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  __shared__ int64_t wt[16];
  if (lane == 63) wt[wv] = x;
  __syncthreads();
  int64_t base = 0;
  for (int i = 0; i < wv; ++i) base += wt[i];
  int64_t run = base + x - s;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = t * 4 + i;
    if (k < nb) b.bsum[k] = run;
    run += v[i];
  }
  if (t == 1023) {
    const int64_t total = base + x;
    const int64_t n = total < a.cap ? total : a.cap;
    b.ctl->n_live[nxt] = n;
    b.ctl->dropped += total - n;
  }
}

template <class T>
__global__ void __launch_bounds__(DMC_BLOCK) dmc_copy_kernel(const DevSys<double>* __restrict__ Sd, DmcArgs a,
                                                             DmcBufs b, int cur, EnsView<T> src, EnsView<T> dst,
                                                             double* __restrict__ eprev_dst, int with_force) {
  const DevSys<double>& S = *Sd;
  const int64_t W = b.ctl->n_live[cur];
  const int64_t w = (int64_t)blockIdx.x * DMC_BLOCK + threadIdx.x;
  int c = 0;
  int64_t d0 = 0;
  if (w < W) {
    d0 = b.bsum[blockIdx.x] + b.loff[w];
    c = b.copies[w];
    if (d0 + c > a.cap) c = d0 < a.cap ? (int)(a.cap - d0) : 0;
  }
  if (c > 0) {
    for (int q = 0; q < a.n3; ++q) {
      const T v = src.pos[(int64_t)q * src.ld + w];
      for (int k = 0; k < c; ++k) dst.pos[(int64_t)q * dst.ld + d0 + k] = v;
    }
    for (int q = 0; q < a.nainv; ++q) {
      const T v = src.ainv[(int64_t)q * src.ld + w];
      for (int k = 0; k < c; ++k) dst.ainv[(int64_t)q * dst.ld + d0 + k] = v;
    }
    const double ld = src.logdet[w], sg = src.sign[w], ju = src.ju[w], el = src.eloc[w];
    for (int k = 0; k < c; ++k) {
      dst.logdet[d0 + k] = ld;
      dst.sign[d0 + k] = sg;
      dst.ju[d0 + k] = ju;
      dst.eloc[d0 + k] = el;
      eprev_dst[d0 + k] = el;  // the next step's pre-sweep E_L
    }
  }
  __shared__ double sh[DMC_BLOCK / 64][1 + 3 * MAXN];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double e = dmc_wave_sum(c > 0 ? (double)c * src.eloc[w] : 0.0);
  if (lane == 0) sh[wv][0] = e;
  if (with_force) {
    double f[MAXN][3];
    hf_electronic<T>(S, src.pos, src.ld, w, c > 0 ? a.n3 : 0, a.n_nuc, (double)c, f);
    wave_sums_force(f, a.n_nuc, sh, 1);
  }
  __syncthreads();
  const int Q = 1 + 3 * a.n_nuc;
  if ((int)threadIdx.x < Q) {
    double s = 0.0;
    if (threadIdx.x == 0 || with_force)
      for (int k = 0; k < DMC_BLOCK / 64; ++k) s += sh[k][threadIdx.x];
    b.part[(int64_t)blockIdx.x * Q + threadIdx.x] = s;
  }
}

// One workgroup: sums the copy kernel's partials (E_L, HF force) and the
// sweep's accepted-move partials in a single pass, then updates the control
// block.  acc_out receives the step's accepted moves.
__global__ void __launch_bounds__(256) dmc_finalize_kernel(DmcArgs a, DmcBufs b, int nxt, int nb, int64_t ts,
                                                           int with_force, const double* __restrict__ sweep_part,
                                                           int sweep_nblk, double* __restrict__ acc_out) {
  constexpr int QM = 2 + 3 * MAXN;  // E_L, accepted moves, forces
  const int Q = 1 + 3 * a.n_nuc;
  const int nq = with_force ? Q : 1;
  double v[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) v[q] = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) {
    const double* p = b.part + (int64_t)k * Q;
    v[0] += p[0];
#pragma unroll
    for (int q = 1; q < 1 + 3 * MAXN; ++q)
      if (q < nq) v[q + 1] += p[q];
  }
  for (int k = threadIdx.x; k < sweep_nblk; k += 256) v[1] += sweep_part[(int64_t)k * 3 + 2];
  __shared__ double sh[256 / 64][QM];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    if (q < nq + 1) {
      const double x = dmc_wave_sum(v[q]);
      if (lane == 0) sh[wv][q] = x;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[QM];
    for (int q = 0; q < nq + 1; ++q) tot[q] = sh[0][q] + sh[1][q] + sh[2][q] + sh[3][q];
    DmcCtl* c = b.ctl;
    const int64_t N = c->n_live[nxt];
    const double e_mean = tot[0] / (double)N;  // 0/0 = NaN when the population died, as is_dmc.rs:474-475
    if (ts >= 0) {
      b.e_trace[ts] = e_mean;
      b.pop_trace[ts] = (double)N;
      acc_out[ts] = tot[1];
    }
    if (with_force) {
      for (int q = 0; q < 3 * a.n_nuc; ++q) c->force_sum[q] += tot[2 + q];
      c->n_force += N;
    }
    c->e_ref = e_mean - a.damping * log((double)N / a.n_target) / a.tau_b;  // is_dmc.rs:487-489
  }
}

__global__ void __launch_bounds__(1024) dmc_reset_kernel(DmcCtl* ctl, int b0, int64_t n0,
                                                         const double* __restrict__ eloc,
                                                         double* __restrict__ eprev) {
  __shared__ double sh[1024];
  double v = 0.0;
  for (int64_t w = threadIdx.x; w < n0; w += 1024) {
    v += eloc[w];
    eprev[w] = eloc[w];
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ctl->n_live[b0] = n0;
    ctl->e_ref = sh[0] / (double)n0;  // is_dmc.rs:406-407
  }
}

template <class T>
__global__ void __launch_bounds__(DMC_BLOCK) hf_partials_kernel(const DevSys<double>* __restrict__ Sd, EnsView<T> E,
                                                                int64_t W, int n3, int n_nuc,
                                                                double* __restrict__ part) {
  const int64_t w = (int64_t)blockIdx.x * DMC_BLOCK + threadIdx.x;
  __shared__ double sh[DMC_BLOCK / 64][1 + 3 * MAXN];
  double f[MAXN][3];
  hf_electronic<T>(*Sd, E.pos, E.ld, w, w < W ? n3 : 0, n_nuc, 1.0, f);
  wave_sums_force(f, n_nuc, sh, 0);
  __syncthreads();
  const int Q = 3 * n_nuc;
  if ((int)threadIdx.x < Q) {
    double s = 0.0;
    for (int k = 0; k < DMC_BLOCK / 64; ++k) s += sh[k][threadIdx.x];
    part[(int64_t)blockIdx.x * Q + threadIdx.x] = s;
  }
}

}  // namespace

template <class T>
hipError_t dmc_branch_step(const void* Sd, const DmcArgs& a, const DmcBufs& b, int cur, EnsView<T> src,
                           const double* eprev_src, EnsView<T> dst, double* eprev_dst, uint32_t t, int64_t ts,
                           int with_force, const double* sweep_part, int sweep_nblk, double* acc_out,
                           hipStream_t st) {
  if (a.cap <= 0 || a.cap % DMC_BLOCK || a.cap / DMC_BLOCK > DMC_MAX_BLOCKS || a.n_nuc > MAXN || src.ld < a.cap ||
      dst.ld < a.cap)
    return hipErrorInvalidValue;
  const int nb = (int)(a.cap / DMC_BLOCK), nxt = 1 - cur;
  dmc_branch_kernel<<<nb, DMC_BLOCK, 0, st>>>(a, b, cur, src.eloc, eprev_src, t);
  dmc_scan_kernel<<<1, 1024, 0, st>>>(a, b, nxt, nb);
  dmc_copy_kernel<T><<<nb, DMC_BLOCK, 0, st>>>((const DevSys<double>*)Sd, a, b, cur, src, dst, eprev_dst, with_force);
  dmc_finalize_kernel<<<1, 256, 0, st>>>(a, b, nxt, nb, ts, with_force, sweep_part, sweep_nblk, acc_out);
  return hipGetLastError();
}

hipError_t dmc_reset(DmcCtl* ctl, int b0, int64_t n0, const double* eloc, double* eprev, hipStream_t st) {
  if (n0 < 1) return hipErrorInvalidValue;
  dmc_reset_kernel<<<1, 1024, 0, st>>>(ctl, b0, n0, eloc, eprev);
  return hipGetLastError();
}

template <class T>
hipError_t hf_partials(const void* Sd, EnsView<T> E, int64_t W, int n3, int n_nuc, double* part, int* nblk,
                       hipStream_t st) {
  if (W < 1 || W > E.ld || n_nuc > MAXN) return hipErrorInvalidValue;
  const int nb = (int)((W + DMC_BLOCK - 1) / DMC_BLOCK);  // lanes past W read nothing
  *nblk = nb;
  hf_partials_kernel<T><<<nb, DMC_BLOCK, 0, st>>>((const DevSys<double>*)Sd, E, W, n3, n_nuc, part);
  return hipGetLastError();
}

template hipError_t dmc_branch_step<double>(const void*, const DmcArgs&, const DmcBufs&, int, EnsView<double>,
                                            const double*, EnsView<double>, double*, uint32_t, int64_t, int,
                                            const double*, int, double*, hipStream_t);
template hipError_t dmc_branch_step<float>(const void*, const DmcArgs&, const DmcBufs&, int, EnsView<float>,
                                           const double*, EnsView<float>, double*, uint32_t, int64_t, int,
                                           const double*, int, double*, hipStream_t);
template hipError_t hf_partials<double>(const void*, EnsView<double>, int64_t, int, int, double*, int*, hipStream_t);
template hipError_t hf_partials<float>(const void*, EnsView<float>, int64_t, int, int, double*, int*, hipStream_t);

}  // namespace qmc
