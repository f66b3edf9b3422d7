// qmc_refresh.hip -- the FP64 state refresh of benzene-size ensembles
// (21 + 21 electrons, up to 68 AOs) by tiles of 16 walkers, one 256-thread
// workgroup (four wavefronts) per tile, 16 lanes per walker, in two
// kernels.
//
// tile_slater_kernel, resident workgroups looping over units (tile, spin
// sector, batch of four electron slots); each wavefront works alone on its
// four walkers, synchronised within itself only:
//   A  AO values of the batch's four electrons of each walker, in FP64, the
//      shells spread over the walker's 16 lanes (the tile plan's lists),
//      into the wavefront's LDS matrix X[4 walkers x 4 slots][ao]
//   B  the 16 Slater-matrix rows, Phi = X C^T, on the FP64 matrix cores
//      (v_mfma_f64_16x16x4f64), both MO tiles over the full AO range, the MO
//      coefficients held as the B operand
//   C  the rows from the accumulators to a work buffer in HBM, walker-major.
// tile_invert_kernel, one workgroup per tile: lane g of a walker's group
// loads rows g and g + 16 into VGPRs, then Gauss-Jordan with partial
// pivoting in place on the register rows, the inverse written with the
// pivots' relabelling undone.
// tile_lnj_kernel, one wavefront per walker: ln J over the electron pairs
// and electron-nucleus pairs, balanced over the lanes.
// Three kernels instead of one, so each holds fewer registers and more
// waves per SIMD hide the FP64 latencies.  The arithmetic is the wave
// engine's build (wave_sector_inverse, wave_jastrow_u) in another summation
// and pivoting order.
#include "qmc_refresh.hpp"

#include <algorithm>

#include "qmc_tile_dev.hpp"
#include "qmc_wave_dev.hpp"

namespace qmc {

typedef double rdbl4 __attribute__((ext_vector_type(4)));

constexpr int RF_REC = 33;  // doubles per shell record: 32 used, odd stride against bank conflicts

struct RefreshArgs {
  int N, A, n_nuc, n_shell;
  int lane_sh[16][TILE_MAXSH];  // shells evaluated by group lane g (-1: none)
  // LDS carve-outs in doubles: the Slater kernel's four X matrices and the
  // shell records; the inversion's pivot rows, pivot lists and the stage of
  // its output
  int o_x, o_sh, o_prow, o_piv, o_stg;
  int lds_slater, lds_invert;
  double* work;  // [W][2][NS][NS] Slater rows, walker-major
};

// One FP64 shell record: kind, first AO, primitive count, P1 component,
// centre, then exponents, s and p coefficients and STO powers of up to
// TILE_MAXP primitives (small integers are exact in a double).
__device__ __forceinline__ void rf_write_record(const DevSys<double>& S, int sh, double* rec) {
  rec[0] = (double)S.sh_kind[sh];
  rec[1] = (double)S.sh_ao0[sh];
  rec[2] = (double)S.sh_nprim[sh];
  rec[3] = (double)S.sh_m[sh];
  const int c = S.sh_center[sh], p0 = S.sh_prim0[sh], np = S.sh_nprim[sh];
  rec[4] = S.nuc[c][0];
  rec[5] = S.nuc[c][1];
  rec[6] = S.nuc[c][2];
  rec[7] = 0.0;
  for (int k = 0; k < TILE_MAXP; ++k) {
    const bool v = k < np;
    rec[8 + k] = v ? S.prim_exp[p0 + k] : 0.0;
    rec[14 + k] = v ? S.coef_s[p0 + k] : 0.0;
    rec[20 + k] = v ? S.coef_p[p0 + k] : 0.0;
    rec[26 + k] = v ? (double)S.prim_pow[p0 + k] : 0.0;
  }
}

// AO values of one shell from its record, written to chi[ao]: the FP64
// arithmetic of shell_values (methane.rs:498-590, slater.rs:17-41, 92-144)
__device__ __forceinline__ void rf_shell_values(const double* rec, double x, double y, double z, double* chi) {
  const int kind = (int)rec[0], a0 = (int)rec[1], np = (int)rec[2];
  const double dx = x - rec[4], dy = y - rec[5], dz = z - rec[6];
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (kind <= SH_P1) {
    double rs = 0.0, rp = 0.0;
    // unrolled to the maximum count, so the record's loads issue together
#pragma unroll
    for (int k = 0; k < TILE_MAXP; ++k) {
      if (k >= np) break;
      const double e = M<double>::exn(-rec[8 + k] * r2);
      rs += rec[14 + k] * e;
      rp += rec[20 + k] * e;
    }
    if (kind == SH_S) {
      chi[a0] = rs;
    } else if (kind == SH_P1) {
      const int m = (int)rec[3];
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
    const double s = M<double>::sq(r2);
    double v = 0.0;
    for (int k = 0; k < np; ++k) {
      double sp = 1.0;
      for (int q = 0; q < (int)rec[26 + k]; ++q) sp *= s;
      v += rec[14 + k] * sp * M<double>::exn(-rec[8 + k] * s);
    }
    chi[a0] = v;
  }
}

template <int NS, int KS, class T>
__global__ void __launch_bounds__(TILE_T, 3) tile_slater_kernel(const DevSys<double>* __restrict__ Sp, EnsView<T> E,
                                                             RefreshArgs RA, int n_units) {
  extern __shared__ double lds_d[];
  constexpr int XS = 4 * KS + 1;           // X row stride
  constexpr int NB = (NS + 3) / 4;         // batches of four electron slots
  const DevSys<double>& S = *Sp;
  const int tid = threadIdx.x;
  const int q = tid >> 6, l = tid & 63;    // wavefront, lane
  const int ww = l >> 4, g = l & 15;       // walker of the wavefront, lane in its group
  double* X = lds_d + RA.o_x + q * 16 * XS;  // this wavefront's [16 rows][XS], row 4 ww + slot
  double* shrec = lds_d + RA.o_sh;           // [n_shell][RF_REC]
  for (int i = tid; i < RA.n_shell; i += TILE_T) rf_write_record(S, i, shrec + i * RF_REC);
  for (int i = l; i < 16 * XS; i += 64) X[i] = 0.0;
  int my_sh[TILE_MAXSH];
#pragma unroll
  for (int j = 0; j < TILE_MAXSH; ++j) my_sh[j] = RA.lane_sh[g][j];
  __syncthreads();
  // units (tile, spin, batch), spread over the resident workgroups
  for (int u = blockIdx.x; u < n_units; u += gridDim.x) {
    const int b = u % NB, spin = (u / NB) & 1, tile = u / (2 * NB);
    const int64_t w = (int64_t)tile * TILE_W + 4 * q + ww;
    const int64_t wl = w < E.W ? w : E.W - 1;  // dead walkers shadow the last one
    // B operand: lane l holds C[ao = (l / 16) KS + s][mo = 16 t + l % 16]
    double Creg[2][KS];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s2 = 0; s2 < KS; ++s2) {
        const int ao = (l >> 4) * KS + s2, mo = t * 16 + g;
        Creg[t][s2] = (ao < RA.A && mo < NS) ? S.moT[spin][ao][mo] : 0.0;
      }
    // A: AO values of electron slots 4 b .. 4 b + 3 of the group's walker,
    // the shells spread over its 16 lanes, into rows 4 ww + i
    wave_sync();  // the previous unit's reads of X are done
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = 4 * b + i;
      if (k < NS) {
        const int e = S.sector_elec[spin][k];
        const double x = (double)E.pos[(int64_t)(3 * e) * E.ld + wl];
        const double y = (double)E.pos[(int64_t)(3 * e + 1) * E.ld + wl];
        const double z = (double)E.pos[(int64_t)(3 * e + 2) * E.ld + wl];
#pragma unroll
        for (int j = 0; j < TILE_MAXSH; ++j)
          if (my_sh[j] >= 0) rf_shell_values(shrec + my_sh[j] * RF_REC, x, y, z, X + (4 * ww + i) * XS);
      }
    }
    wave_sync();
    // B: Phi[row][mo] = sum_ao X[row][ao] C[mo][ao] on the FP64 matrix cores,
    // two MO tiles, each over two interleaved accumulators
    rdbl4 acc[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t) acc[t][0] = acc[t][1] = rdbl4{0.0, 0.0, 0.0, 0.0};
    const double* xr = X + g * XS + (l >> 4) * KS;
#pragma unroll
    for (int s2 = 0; s2 < KS; ++s2) {
      const double a = xr[s2];
#pragma unroll
      for (int t = 0; t < 2; ++t) acc[t][s2 & 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Creg[t][s2], acc[t][s2 & 1], 0, 0, 0);
    }
    // C: register r of lane l holds Phi[4 r + l / 16][16 t + l % 16]: walker
    // 4 q + r of the tile, electron slot 4 b + l / 16
    const int k = 4 * b + (l >> 4);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t wr = (int64_t)tile * TILE_W + 4 * q + r;
      if (wr < E.W && k < NS)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int m = 16 * t + g;
          if (m < NS) RA.work[(wr * 2 + spin) * NS * NS + k * NS + m] = acc[t][0][r] + acc[t][1][r];
        }
    }
  }
}

template <int NS, class T>
__global__ void __launch_bounds__(TILE_T, 2) tile_invert_kernel(const DevSys<double>* __restrict__ Sp, EnsView<T> E,
                                                             RefreshArgs RA, int* singular_count) {
  extern __shared__ double lds_d[];
  const DevSys<double>& S = *Sp;
  const int tid = threadIdx.x;
  const int wt = tid >> 4, g = tid & 15;  // walker in the tile, lane in its group
  const int64_t w = (int64_t)blockIdx.x * TILE_W + wt;
  const bool live = w < E.W;
  const int64_t wl = live ? w : E.W - 1;  // dead walkers shadow the last one and write nothing
  double* prow = lds_d + RA.o_prow;  // [TILE_W][NS] pivot rows
  int* piv = (int*)(lds_d + RA.o_piv);  // [TILE_W][NS + 1] pivots and sign, then [TILE_W][2 NS] kinv and marks
  int* pv_of = piv + wt * (NS + 1);                  // pivots p_c, then the permutation's sign
  int* dst = piv + TILE_W * (NS + 1) + wt * 2 * NS;  // kinv[r], then cycle marks
  double ltot = 0.0, stot = 1.0;
  bool sing_any = false;
  __syncthreads();
#pragma unroll 1
  for (int spin = 0; spin < 2; ++spin) {
    // rows g and g + 16 of this walker's Slater matrix a[k][m] = phi_m(r_k)
    double r0[NS], r1[NS];
    const double* row = RA.work + (wl * 2 + spin) * NS * NS;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      r0[m] = row[g * NS + m];
      r1[m] = g + 16 < NS ? row[(g + 16) * NS + m] : 0.0;
    }
    // Gauss-Jordan in place on the register rows, without row
    // interchanges: the pivot of column c is the largest |a[i][c]| over the
    // rows not yet used as pivots (ties to the lowest row), found by DPP
    // moves inside the 16-lane group (one DPP row); row p_c is scaled and
    // column c leaves every other row.  The result B is the inverse with
    // its rows and columns relabelled by the pivots:
    // A^-1[k][p_l] = B[p_k][l].  |det| is kept as a mantissa in [1/2, 1)
    // and a binary exponent (one log per sector), the sign as the pivots'
    // signs times the sign of the permutation c -> p_c.
    double mant = 1.0, ssec = 1.0, lsec;
    int bexp = 0;
    bool sing = false;
    bool used0 = false, used1 = g + 16 >= NS;  // slot 1 exists for g < NS - 16
    double* pr_l = prow + wt * NS;
#pragma unroll
    for (int c = 0; c < NS; ++c) {
      double best = used0 ? -1.0 : fabs(r0[c]);
      int p = g;
      if (!used1) {
        const double v = fabs(r1[c]);
        if (v > best) {
          best = v;
          p = g + 16;
        }
      }
      auto take = [&](double ob, int op) {
        if (ob > best || (ob == best && op < p)) {
          best = ob;
          p = op;
        }
      };
      take(dpp_mov<0xB1>(best), __builtin_amdgcn_update_dpp(0, p, 0xB1, 0xF, 0xF, false));    // quad_perm [1,0,3,2]
      take(dpp_mov<0x4E>(best), __builtin_amdgcn_update_dpp(0, p, 0x4E, 0xF, 0xF, false));    // quad_perm [2,3,0,1]
      take(dpp_mov<0x141>(best), __builtin_amdgcn_update_dpp(0, p, 0x141, 0xF, 0xF, false));  // row_half_mirror
      take(dpp_mov<0x140>(best), __builtin_amdgcn_update_dpp(0, p, 0x140, 0xF, 0xF, false));  // row_mirror
      // the pivot row to the whole group (a walker's group sits in one
      // wavefront: a wavefront-scope sync orders its LDS accesses)
      const bool own0 = p == g, own1 = p == g + 16;
      if (own0) {
#pragma unroll
        for (int j = 0; j < NS; ++j) pr_l[j] = r0[j];
      }
      if (own1) {
#pragma unroll
        for (int j = 0; j < NS; ++j) pr_l[j] = r1[j];
      }
      if (g == 0) pv_of[c] = p;
      wave_sync();
      double pr[NS];
#pragma unroll
      for (int j = 0; j < NS; ++j) pr[j] = pr_l[j];
      const double pv = pr[c];
      if (pv == 0.0) sing = true;
      {
        int e;
        mant = frexp(mant * fabs(pv), &e);
        bexp += e;
      }
      if (pv < 0.0) ssec = -ssec;
      const double ip = sing ? 0.0 : 1.0 / pv;
#pragma unroll
      for (int j = 0; j < NS; ++j) pr[j] = j == c ? ip : pr[j] * ip;
      if (own0) {
#pragma unroll
        for (int j = 0; j < NS; ++j) r0[j] = pr[j];
        used0 = true;
      } else {
        const double f = r0[c];
#pragma unroll
        for (int j = 0; j < NS; ++j) r0[j] = j == c ? -f * pr[j] : r0[j] - f * pr[j];
      }
      if (own1) {
#pragma unroll
        for (int j = 0; j < NS; ++j) r1[j] = pr[j];
        used1 = true;
      } else if (g + 16 < NS) {
        const double f = r1[c];
#pragma unroll
        for (int j = 0; j < NS; ++j) r1[j] = j == c ? -f * pr[j] : r1[j] - f * pr[j];
      }
    }
    // kinv[r]: the step whose pivot was row r; the permutation's sign from
    // its cycles
    if (g == 0) {
      int cycles = 0;
      for (int c = 0; c < NS; ++c) {
        dst[pv_of[c]] = c;
        dst[NS + c] = 0;
      }
      for (int c = 0; c < NS; ++c) {
        if (dst[NS + c]) continue;
        ++cycles;
        for (int k = c; !dst[NS + k]; k = pv_of[k]) dst[NS + k] = 1;
      }
      pv_of[NS] = ((NS - cycles) & 1) ? -1 : 1;
    }
    wave_sync();
    ssec *= (double)pv_of[NS];
    lsec = log(mant) + (double)bexp * 0.69314718055994530942;
    if (sing) {
      lsec = -INFINITY;
      ssec = 0.0;
    }
    // the inverse staged element-major in LDS (a stride of TILE_W + 1 keeps
    // the group's 16 rows in different banks), then written out with the
    // tile's walkers in consecutive lanes: walker-contiguous HBM stores
    T* stg = (T*)(lds_d + RA.o_stg);  // [NS * NS][TILE_W + 1]
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int r = g + 16 * t;  // storage row r holds inverse row kinv[r]
      if (r >= NS) continue;
      if (!sing) {
        const int m = dst[r];
#pragma unroll
        for (int l = 0; l < NS; ++l) stg[(m * NS + pv_of[l]) * (TILE_W + 1) + wt] = (T)(t ? r1[l] : r0[l]);
      } else {  // the identity, as the wave engine's build leaves it
#pragma unroll
        for (int k = 0; k < NS; ++k) stg[(r * NS + k) * (TILE_W + 1) + wt] = (T)(r == k ? 1.0 : 0.0);
      }
    }
    __syncthreads();
    {
      const int64_t base = spin ? (int64_t)NS * NS : 0, w0 = (int64_t)blockIdx.x * TILE_W;
      for (int i = tid; i < NS * NS * TILE_W; i += TILE_T) {
        const int el = i / TILE_W, ww = i - el * TILE_W;
        if (w0 + ww < E.W) E.ainv[(base + el) * E.ld + w0 + ww] = stg[el * (TILE_W + 1) + ww];
      }
    }
    ltot += lsec;
    stot *= ssec;
    sing_any = sing_any || sing;
    __syncthreads();  // dst and the LDS buffers are reused by the next sector
  }
  if (live && g == 0) {
    E.logdet[w] = ltot;
    E.sign[w] = stot;
    if (sing_any && singular_count) atomicAdd(singular_count, 1);
  }
}

// ln J of every walker from its positions (wave_jastrow_u's terms), one
// wavefront per walker: the N (N - 1) / 2 electron pairs (i, j > i) and the
// N M electron-nucleus pairs spread evenly over the lanes.  Electron rows i
// and N - 1 - i are folded together (N - 1 pair terms per folded row), and
// each folded row is split over P = 64 / ceil(N / 2) lanes.
template <class T>
__global__ void __launch_bounds__(TILE_T) tile_lnj_kernel(const DevSys<double>* __restrict__ Sp, EnsView<T> E,
                                                          int N) {
  extern __shared__ double lds_d[];
  const DevSys<double>& S = *Sp;
  const int q = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * 4 + q;
  const int64_t wl = w < E.W ? w : E.W - 1;  // dead walkers shadow the last one and write nothing
  double* pw = lds_d + q * 3 * N;
  for (int i = l; i < 3 * N; i += 64) pw[i] = (double)E.pos[(int64_t)i * E.ld + wl];
  wave_sync();
  const int F = (N + 1) / 2, P = 64 / F;
  const int f = l / P, part = l - f * P;
  double u = 0.0;
  if (f < F) {
    const int i1 = f, i2 = N - 1 - f;
    const int n1 = N - 1 - i1, n2 = i2 != i1 ? N - 1 - i2 : 0;
    for (int t = part; t < n1 + n2; t += P) {
      const int i = t < n1 ? i1 : i2, j = t < n1 ? i1 + 1 + t : i2 + 1 + (t - n1);
      const double dx = pw[3 * i] - pw[3 * j], dy = pw[3 * i + 1] - pw[3 * j + 1], dz = pw[3 * i + 2] - pw[3 * j + 2];
      const bool same = S.elec_spin[i] == S.elec_spin[j];
      u += uee_val(S, sqrt(dx * dx + dy * dy + dz * dz), same) + S.cst_ee;
    }
    if (S.en_on) {
      const int M = S.n_nuc, ne = (i2 != i1 ? 2 : 1) * M;
      for (int t = part; t < ne; t += P) {
        const int i = t < M ? i1 : i2, I = t < M ? t : t - M;
        const double dx = pw[3 * i] - S.nuc[I][0], dy = pw[3 * i + 1] - S.nuc[I][1], dz = pw[3 * i + 2] - S.nuc[I][2];
        u += uen_val(S, sqrt(dx * dx + dy * dy + dz * dz), I);
      }
    }
  }
  u = wsum(u);
  if (l == 0 && w < E.W) E.ju[w] = u;
}

static RefreshArgs make_refresh_args(const TilePlan& P, int NS, int KS, double* work) {
  RefreshArgs R;
  R.N = P.N;
  R.A = P.A;
  R.n_nuc = P.n_nuc;
  R.n_shell = P.n_shell;
  for (int g = 0; g < 16; ++g)
    for (int j = 0; j < TILE_MAXSH; ++j) R.lane_sh[g][j] = P.lane_sh[g][j];
  R.o_x = 0;  // Slater kernel: four wavefronts' X, then the shell records
  R.o_sh = R.o_x + 4 * 16 * (4 * KS + 1);
  R.lds_slater = R.o_sh + P.n_shell * RF_REC;
  R.o_prow = 0;
  R.o_piv = R.o_prow + TILE_W * NS;
  R.o_stg = R.o_piv + (TILE_W * (NS + 1) + TILE_W * 2 * NS + 1) / 2;
  R.lds_invert = R.o_stg + NS * NS * (TILE_W + 1);  // in doubles: room for a double stage
  R.work = work;
  return R;
}

size_t tile_refresh_work(int64_t W) { return (size_t)W * 2 * 21 * 21; }

template <class T>
hipError_t tile_refresh(const void* Sd, EnsView<T> E, const TilePlan& P, double* work, int* singular_count,
                        hipStream_t st) {
  constexpr int NS = 21, KS = 17;
  if (!P.ok || P.nu != NS || P.nd != NS || P.A > 4 * KS || E.n_live || !work) return hipErrorInvalidValue;
  if (E.W <= 0) return hipSuccess;
  const RefreshArgs R = make_refresh_args(P, NS, KS, work);
  const unsigned tiles = (unsigned)tile_blocks(E.W);
  const size_t ba = sizeof(double) * (size_t)R.lds_slater, bb = sizeof(double) * (size_t)R.lds_invert;
  auto ka = tile_slater_kernel<NS, KS, T>;
  auto kb = tile_invert_kernel<NS, T>;
  hipError_t e = hipFuncSetAttribute((const void*)ka, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ba);
  if (e == hipSuccess) e = hipFuncSetAttribute((const void*)kb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bb);
  if (e != hipSuccess) return e;
  // the Slater kernel's units over as many workgroups as are resident at once
  const int n_units = (int)tiles * 2 * ((NS + 3) / 4);
  int dev = 0, n_cu = 0, occ = 0;
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)ka, TILE_T, ba);
  if (e != hipSuccess) return e;
  const int grid = std::max(1, std::min(n_units, n_cu * std::max(occ, 1)));
  ka<<<grid, TILE_T, ba, st>>>((const DevSys<double>*)Sd, E, R, n_units);
  kb<<<tiles, TILE_T, bb, st>>>((const DevSys<double>*)Sd, E, R, singular_count);
  tile_lnj_kernel<T><<<(unsigned)((E.W + 3) / 4), TILE_T, sizeof(double) * 4 * 3 * P.N, st>>>(
      (const DevSys<double>*)Sd, E, P.N);
  return hipGetLastError();
}

template hipError_t tile_refresh<float>(const void*, EnsView<float>, const TilePlan&, double*, int*, hipStream_t);
template hipError_t tile_refresh<double>(const void*, EnsView<double>, const TilePlan&, double*, int*, hipStream_t);

}  // namespace qmc
