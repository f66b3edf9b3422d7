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
//                       source order (clones adjacent) with per-chunk sums of
//                       E_L and of the HF force over the new population
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

// Wave sums of c times the electronic HF force of each lane's walker w,
// F_I = Z_I sum_i (r_i - R_I) / |r_i - R_I|^3 over electrons at distance
// > 1e-10 (hydrogen.rs:227-243, methane.rs:1241-1257), written by lane 0
// to sh[wave][q0 + 3 I + c].  w must index a valid column for every lane;
// lanes with c = 0 add exactly nothing.  One nucleus at a time, electrons
// four at a time with their twelve loads issued together (the distance guard
// is a select, so no branch separates them).
template <class T>
__device__ __forceinline__ void hf_wave_sums(const DevSys<double>& S, const T* pos, int64_t ld, int64_t w, int n3,
                                             int n_nuc, double c, double (*sh)[1 + 3 * MAXN], int q0) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int I = 0; I < n_nuc; ++I) {
    const double X = S.nuc[I][0], Y = S.nuc[I][1], Zc = S.nuc[I][2], Z = S.Z[I];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int e0 = 0; e0 < n3; e0 += 12) {
      double x[4], y[4], z[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int q = e0 + 3 * k < n3 ? e0 + 3 * k : n3 - 3;  // past the last electron: reread it, weight 0
        x[k] = (double)pos[(int64_t)q * ld + w];
        y[k] = (double)pos[(int64_t)(q + 1) * ld + w];
        z[k] = (double)pos[(int64_t)(q + 2) * ld + w];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double dx = x[k] - X, dy = y[k] - Y, dz = z[k] - Zc;
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const double sc = (e0 + 3 * k < n3 && d > 1e-10) ? Z / (d * d * d) : 0.0;
        fx += sc * dx;
        fy += sc * dy;
        fz += sc * dz;
      }
    }
    fx = dmc_wave_sum(c > 0.0 ? fx * c : 0.0);
    fy = dmc_wave_sum(c > 0.0 ? fy * c : 0.0);
    fz = dmc_wave_sum(c > 0.0 ? fz * c : 0.0);
    if (lane == 0) {
      sh[wv][q0 + 3 * I] = fx;
      sh[wv][q0 + 3 * I + 1] = fy;
      sh[wv][q0 + 3 * I + 2] = fz;
    }
  }
}

// c times the HF force of each lane's walker on nucleus I, wave-summed into
// sh[wave][0..2] by lane 0 (the same electron order and select as above)
template <class T>
__device__ __forceinline__ void hf_wave_sums_one(const DevSys<double>& S, const T* pos, int64_t ld, int64_t w, int n3,
                                                 int I, double c, double (*sh)[1 + 3 * MAXN]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double X = S.nuc[I][0], Y = S.nuc[I][1], Zc = S.nuc[I][2], Z = S.Z[I];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int e0 = 0; e0 < n3; e0 += 12) {
    double x[4], y[4], z[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = e0 + 3 * k < n3 ? e0 + 3 * k : n3 - 3;
      x[k] = (double)pos[(int64_t)q * ld + w];
      y[k] = (double)pos[(int64_t)(q + 1) * ld + w];
      z[k] = (double)pos[(int64_t)(q + 2) * ld + w];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double dx = x[k] - X, dy = y[k] - Y, dz = z[k] - Zc;
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const double sc = (e0 + 3 * k < n3 && d > 1e-10) ? Z / (d * d * d) : 0.0;
      fx += sc * dx;
      fy += sc * dy;
      fz += sc * dz;
    }
  }
  fx = dmc_wave_sum(c > 0.0 ? fx * c : 0.0);
  fy = dmc_wave_sum(c > 0.0 ? fy * c : 0.0);
  fz = dmc_wave_sum(c > 0.0 ? fz * c : 0.0);
  if (lane == 0) {
    sh[wv][0] = fx;
    sh[wv][1] = fy;
    sh[wv][2] = fz;
  }
}

__global__ void __launch_bounds__(DMC_BLOCK) dmc_branch_kernel(DmcArgs a, DmcBufs b, int cur,
                                                               const double* __restrict__ eloc,
                                                               const double* __restrict__ eprev, uint32_t t) {
  const int64_t W = b.ctl->n_live[cur];
  const int64_t w = (int64_t)blockIdx.x * DMC_BLOCK + threadIdx.x;
  if (t == DMC_STEP_CTL) t = b.ctl->step;
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

// Compaction as a grid-stride loop over items (kind, chunk) of 256 source
// walkers: kind 0 copies the scalar state and sums E_L over the copies,
// kinds 1..n3+nainv copy one row of positions or inverse elements (one load,
// c coalesced stores per walker), and with forces the n_nuc last kinds sum
// c times the HF force on one nucleus.  Items exist only for the live
// source population, so the work follows it rather than the capacity.
template <class T>
__global__ void __launch_bounds__(DMC_BLOCK) dmc_copy_kernel(const DevSys<double>* __restrict__ Sd, DmcArgs a,
                                                             DmcBufs b, int cur, EnsView<T> src, EnsView<T> dst,
                                                             double* __restrict__ eprev_dst, int with_force) {
  const DevSys<double>& S = *Sd;
  const int64_t W = b.ctl->n_live[cur];
  if (with_force == DMC_FORCE_CTL) with_force = (int64_t)b.ctl->step >= a.n_burnin ? 1 : 0;
  const int64_t nch = (W + DMC_BLOCK - 1) / DMC_BLOCK;
  const int rows = a.n3 + a.nainv, Q = 1 + 3 * a.n_nuc;
  const int64_t n_items = nch * (1 + rows + (with_force ? a.n_nuc : 0));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ double sh[DMC_BLOCK / 64][1 + 3 * MAXN];
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int kind = (int)(item / nch);
    const int64_t ch = item - (int64_t)kind * nch;
    const int64_t w = ch * DMC_BLOCK + threadIdx.x;
    int c = 0;
    int64_t d0 = 0;
    if (w < W) {
      d0 = b.bsum[ch] + b.loff[w];
      c = b.copies[w];
      if (d0 + c > a.cap) c = d0 < a.cap ? (int)(a.cap - d0) : 0;
    }
    if (kind == 0) {
      double e = 0.0;
      if (c > 0) {
        const double ld = src.logdet[w], sg = src.sign[w], ju = src.ju[w], el = src.eloc[w];
        for (int k = 0; k < c; ++k) {
          dst.logdet[d0 + k] = ld;
          dst.sign[d0 + k] = sg;
          dst.ju[d0 + k] = ju;
          dst.eloc[d0 + k] = el;
          eprev_dst[d0 + k] = el;  // the next step's pre-sweep E_L
        }
        e = (double)c * el;
      }
      e = dmc_wave_sum(e);
      __syncthreads();  // the previous item's readers of sh are done
      if (lane == 0) sh[wv][0] = e;
      __syncthreads();
      if (threadIdx.x == 0) b.part[ch * Q] = sh[0][0] + sh[1][0] + sh[2][0] + sh[3][0];
    } else if (kind <= rows) {
      const int q = kind - 1;
      if (c > 0) {
        if (q < a.n3) {
          const T v = src.pos[(int64_t)q * src.ld + w];
          for (int k = 0; k < c; ++k) dst.pos[(int64_t)q * dst.ld + d0 + k] = v;
        } else {
          const int64_t r = q - a.n3;
          const T v = src.ainv[r * src.ld + w];
          for (int k = 0; k < c; ++k) dst.ainv[r * dst.ld + d0 + k] = v;
        }
      }
    } else {
      const int I = kind - 1 - rows;
      // every lane of a live chunk reads a valid column (w < cap <= ld); c = 0 adds nothing
      __syncthreads();
      hf_wave_sums_one<T>(S, src.pos, src.ld, w, a.n3, I, (double)c, sh);
      __syncthreads();
      if (threadIdx.x < 3)
        b.part[ch * Q + 1 + 3 * I + threadIdx.x] =
            sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
    }
  }
}

// One workgroup of 16 waves: wave j sums column j of the copy kernel's
// partials (E_L, then the HF force components) or, for j = nq, the sweep's
// accepted moves, then lane 0 of thread 0 updates the control block.
// acc_out receives the step's accepted moves.
__global__ void __launch_bounds__(1024) dmc_finalize_kernel(DmcArgs a, DmcBufs b, int cur, int64_t ts,
                                                            int with_force, const double* __restrict__ sweep_part,
                                                            int sweep_bs, double* __restrict__ acc_out) {
  const uint32_t step = b.ctl->step;  // read by every thread before thread 0 advances it
  if (ts == DMC_TS_CTL) ts = (int64_t)step;
  if (with_force == DMC_FORCE_CTL) with_force = (int64_t)step >= a.n_burnin ? 1 : 0;
  const int nxt = 1 - cur;
  const int Q = 1 + 3 * a.n_nuc;
  const int nq = with_force ? Q : 1;
  // rows written for the source population: its 256-walker chunks and sweep blocks
  const int64_t W = b.ctl->n_live[cur];
  const int nb = (int)((W + DMC_BLOCK - 1) / DMC_BLOCK), sweep_nblk = (int)((W + sweep_bs - 1) / sweep_bs);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ double tot[2 + 3 * MAXN];
  for (int j = wv; j <= nq; j += 16) {
    double v = 0.0;
    // unrolled so that the loads of one lane are in flight together
    if (j < nq) {
#pragma unroll 16
      for (int k = lane; k < nb; k += 64) v += b.part[(int64_t)k * Q + j];
    } else {
#pragma unroll 16
      for (int k = lane; k < sweep_nblk; k += 64) v += sweep_part[(int64_t)k * 3 + 2];
    }
    v = dmc_wave_sum(v);
    if (lane == 0) tot[j] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    DmcCtl* c = b.ctl;
    const int64_t N = c->n_live[nxt];
    const double e_mean = tot[0] / (double)N;  // 0/0 = NaN when the population died, as is_dmc.rs:474-475
    if (ts >= 0) {
      b.e_trace[ts] = e_mean;
      b.pop_trace[ts] = (double)N;
      acc_out[ts] = tot[nq];
    }
    if (with_force) {
      for (int q = 0; q < 3 * a.n_nuc; ++q) c->force_sum[q] += tot[1 + q];
      c->n_force += N;
    }
    c->e_ref = e_mean - a.damping * log((double)N / a.n_target) / a.tau_b;  // is_dmc.rs:487-489
    c->step = step + 1;
  }
}

__global__ void dmc_set_step_kernel(DmcCtl* ctl, uint32_t t) {
  if (threadIdx.x == 0) ctl->step = t;
}

__global__ void __launch_bounds__(1024) dmc_reset_kernel(DmcCtl* ctl, int b0, int64_t n0,
                                                         const double* __restrict__ eloc,
                                                         double* __restrict__ eprev) {
  __shared__ double sh[1024];
  double v = 0.0;
#pragma unroll 8
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
  // lanes past W reread walker W - 1 with weight 0
  hf_wave_sums<T>(*Sd, E.pos, E.ld, w < W ? w : W - 1, n3, n_nuc, w < W ? 1.0 : 0.0, sh, 0);
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
                           int with_force, const double* sweep_part, int sweep_nblk, int sweep_bs, double* acc_out,
                           hipStream_t st) {
  if (a.cap <= 0 || a.cap % DMC_BLOCK || a.cap / DMC_BLOCK > DMC_MAX_BLOCKS || a.n_nuc > MAXN || src.ld < a.cap ||
      dst.ld < a.cap)
    return hipErrorInvalidValue;
  const int nb = (int)(a.cap / DMC_BLOCK), nxt = 1 - cur;
  if (sweep_bs < 1 || (int64_t)sweep_nblk * sweep_bs < a.cap)  // the sweep's partial rows cover the capacity
    return hipErrorInvalidValue;
  dmc_branch_kernel<<<nb, DMC_BLOCK, 0, st>>>(a, b, cur, src.eloc, eprev_src, t);
  dmc_scan_kernel<<<1, 1024, 0, st>>>(a, b, nxt, nb);
  dmc_copy_kernel<T><<<nb < 1024 ? nb : 1024, DMC_BLOCK, 0, st>>>((const DevSys<double>*)Sd, a, b, cur, src, dst,
                                                                  eprev_dst, with_force);
  dmc_finalize_kernel<<<1, 1024, 0, st>>>(a, b, cur, ts, with_force, sweep_part, sweep_bs, acc_out);
  return hipGetLastError();
}

hipError_t dmc_set_step(DmcCtl* ctl, uint32_t t, hipStream_t st) {
  dmc_set_step_kernel<<<1, 64, 0, st>>>(ctl, t);
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
                                            const double*, int, int, double*, hipStream_t);
template hipError_t dmc_branch_step<float>(const void*, const DmcArgs&, const DmcBufs&, int, EnsView<float>,
                                           const double*, EnsView<float>, double*, uint32_t, int64_t, int,
                                           const double*, int, int, double*, hipStream_t);
template hipError_t hf_partials<double>(const void*, EnsView<double>, int64_t, int, int, double*, int*, hipStream_t);
template hipError_t hf_partials<float>(const void*, EnsView<float>, int64_t, int, int, double*, int*, hipStream_t);

}  // namespace qmc
