// qmc_refresh.hip -- the FP64 state refresh of benzene-size ensembles
// (21 + 21 electrons, up to 68 AOs) by tiles of 16 walkers, one 256-thread
// workgroup (four wavefronts) per tile, 16 lanes per walker, in two
// kernels.
//
// tile_slater_kernel, one workgroup per (tile, spin sector), electron slot
// by electron slot:
//   A  AO values of the slot's electron for the 16 walkers, in FP64, the
//      shells spread over each walker's 16 lanes (the tile plan's lists),
//      into the LDS matrix X[walker][ao]
//   B  the slot's Slater-matrix row of every walker, Phi = X C^T, on the FP64
//      matrix cores (v_mfma_f64_16x16x4f64): wavefront q takes MO tile q & 1
//      over K half q >> 1, the MO coefficients held as the B operand
//   C  the rows to a work buffer in HBM, walker-major.
// tile_invert_kernel, one workgroup per tile: lane g of a walker's group
// loads rows g and g + 16 into VGPRs, then Gauss-Jordan with partial
// pivoting in place on the register rows, the inverse written with the
// pivots' relabelling undone; ln J summed over the electron pairs and
// electron-nucleus pairs spread on the 16 lanes.  Two kernels instead of
// one, so each holds fewer registers and more waves per SIMD hide the FP64
// latencies.  The arithmetic is the wave engine's build (wave_sector_inverse,
// wave_jastrow_u) in another summation and pivoting order.
#include "qmc_refresh.hpp"

#include "qmc_tile_dev.hpp"
#include "qmc_wave_dev.hpp"

namespace qmc {

typedef double rdbl4 __attribute__((ext_vector_type(4)));

constexpr int RF_REC = 32;  // doubles per shell record
constexpr int RF_MO = 32;   // MO columns of Phi (two 16-wide MFMA tiles)

struct RefreshArgs {
  int N, A, n_nuc, n_shell;
  int lane_sh[16][TILE_MAXSH];  // shells evaluated by group lane g (-1: none)
  // LDS carve-outs in doubles: positions first in both kernels, then the
  // Slater kernel's X, Phi and shell records, or the inversion's pivot rows
  // and pivot lists
  int o_pos, o_x, o_phi, o_sh, o_prow, o_piv;
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
      const double e = M<double>::ex(-rec[8 + k] * r2);
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
      v += rec[14 + k] * sp * M<double>::ex(-rec[8 + k] * s);
    }
    chi[a0] = v;
  }
}

template <int NS, int KS, class T>
__global__ void __launch_bounds__(TILE_T) tile_slater_kernel(const DevSys<double>* __restrict__ Sp, EnsView<T> E,
                                                             RefreshArgs RA) {
  extern __shared__ double lds_d[];
  constexpr int KH = (KS + 1) / 2;  // MFMA steps per K half
  constexpr int XS = 4 * KS + 1;    // X row stride
  const DevSys<double>& S = *Sp;
  const int tid = threadIdx.x, spin = blockIdx.y;
  const int q = tid >> 6, l = tid & 63;   // wavefront, lane
  const int wt = tid >> 4, g = tid & 15;  // walker in the tile, lane in its group
  const int64_t w0 = (int64_t)blockIdx.x * TILE_W;
  const int64_t w = w0 + wt;
  const int64_t wl = w < E.W ? w : E.W - 1;  // dead walkers shadow the last one
  const int N = RA.N;
  double* pos = lds_d + RA.o_pos;  // [TILE_W][3N]
  double* X = lds_d + RA.o_x;      // [TILE_W][XS]
  double* Phi = lds_d + RA.o_phi;  // [2 K halves][TILE_W][RF_MO]
  double* shrec = lds_d + RA.o_sh; // [n_shell][RF_REC]
  double* pw = pos + wt * 3 * N;
  for (int i = g; i < 3 * N; i += 16) pw[i] = (double)E.pos[(int64_t)i * E.ld + wl];
  for (int i = g; i < XS; i += 16) X[wt * XS + i] = 0.0;
  for (int i = tid; i < RA.n_shell; i += TILE_T) rf_write_record(S, i, shrec + i * RF_REC);
  int my_sh[TILE_MAXSH];
#pragma unroll
  for (int j = 0; j < TILE_MAXSH; ++j) my_sh[j] = RA.lane_sh[g][j];
  const int tb = q & 1, kh = q >> 1;
  // B operand: lane l holds C[mo = 16 tb + l % 16][ao = (l / 16) KS + kh KH + s]
  double Creg[KH];
#pragma unroll
  for (int s = 0; s < KH; ++s) {
    const int kk = kh * KH + s;
    const int ao = (l >> 4) * KS + kk, mo = tb * 16 + (l & 15);
    Creg[s] = (kk < KS && ao < RA.A && mo < NS) ? S.moT[spin][ao][mo] : 0.0;
  }
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < NS; ++k) {
    const int e = S.sector_elec[spin][k];
    const double x = pw[3 * e], y = pw[3 * e + 1], z = pw[3 * e + 2];
#pragma unroll
    for (int j = 0; j < TILE_MAXSH; ++j)
      if (my_sh[j] >= 0) rf_shell_values(shrec + my_sh[j] * RF_REC, x, y, z, X + wt * XS);
    __syncthreads();
    {
      rdbl4 acc = {0.0, 0.0, 0.0, 0.0};
      const double* xr = X + (l & 15) * XS + (l >> 4) * KS + kh * KH;
      // past K (the second half's last step when KS is odd) Creg is zero and
      // X stays inside its padded row
#pragma unroll
      for (int s = 0; s < KH; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[s], Creg[s], acc, 0, 0, 0);
      // register r of lane l holds D[4 r + l / 16][l % 16]
      double* ph = Phi + kh * TILE_W * RF_MO;
#pragma unroll
      for (int r = 0; r < 4; ++r) ph[(4 * r + (l >> 4)) * RF_MO + tb * 16 + (l & 15)] = acc[r];
    }
    __syncthreads();
    // row k of the 16 walkers: a[k][m] = phi_m(r_k), the two K halves summed
    for (int i = tid; i < TILE_W * NS; i += TILE_T) {
      const int ww = i / NS, m = i - ww * NS;
      if (w0 + ww < E.W)
        RA.work[((w0 + ww) * 2 + spin) * NS * NS + k * NS + m] =
            Phi[ww * RF_MO + m] + Phi[TILE_W * RF_MO + ww * RF_MO + m];
    }
  }
}

template <int NS, class T>
__global__ void __launch_bounds__(TILE_T) tile_invert_kernel(const DevSys<double>* __restrict__ Sp, EnsView<T> E,
                                                             RefreshArgs RA, int* singular_count) {
  extern __shared__ double lds_d[];
  const DevSys<double>& S = *Sp;
  const int tid = threadIdx.x;
  const int wt = tid >> 4, g = tid & 15;  // walker in the tile, lane in its group
  const int64_t w = (int64_t)blockIdx.x * TILE_W + wt;
  const bool live = w < E.W;
  const int64_t wl = live ? w : E.W - 1;  // dead walkers shadow the last one and write nothing
  const int N = RA.N;
  double* pos = lds_d + RA.o_pos;    // [TILE_W][3N]
  double* prow = lds_d + RA.o_prow;  // [TILE_W][NS] pivot rows
  int* piv = (int*)(lds_d + RA.o_piv);  // [TILE_W][NS + 1] pivots and sign, then [TILE_W][2 NS] kinv and marks
  double* pw = pos + wt * 3 * N;
  for (int i = g; i < 3 * N; i += 16) pw[i] = (double)E.pos[(int64_t)i * E.ld + wl];
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
    if (live) {
      const int64_t base = spin ? (int64_t)NS * NS : 0;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int r = g + 16 * t;  // storage row r holds inverse row kinv[r]
        if (r >= NS) continue;
        if (!sing) {
          const int m = dst[r];
#pragma unroll
          for (int l = 0; l < NS; ++l) E.ainv[(base + m * NS + pv_of[l]) * E.ld + w] = (T)(t ? r1[l] : r0[l]);
        } else {  // the identity, as the wave engine's build leaves it
#pragma unroll
          for (int k = 0; k < NS; ++k) E.ainv[(base + r * NS + k) * E.ld + w] = (T)(r == k ? 1.0 : 0.0);
        }
      }
    }
    ltot += lsec;
    stot *= ssec;
    sing_any = sing_any || sing;
    __syncthreads();  // dst and the LDS buffers are reused by the next sector
  }
  // ln J: electron pairs (i, j > i) and electron-nucleus pairs of the
  // electrons g, g + 16, g + 32
  double u = 0.0;
  for (int i = g; i < N; i += 16) {
    const double xi = pw[3 * i], yi = pw[3 * i + 1], zi = pw[3 * i + 2];
    for (int j = i + 1; j < N; ++j) {
      const double dx = xi - pw[3 * j], dy = yi - pw[3 * j + 1], dz = zi - pw[3 * j + 2];
      const bool same = S.elec_spin[i] == S.elec_spin[j];
      u += uee_val(S, sqrt(dx * dx + dy * dy + dz * dz), same) + S.cst_ee;
    }
    if (S.en_on)
      for (int I = 0; I < RA.n_nuc; ++I) {
        const double dx = xi - S.nuc[I][0], dy = yi - S.nuc[I][1], dz = zi - S.nuc[I][2];
        u += uen_val(S, sqrt(dx * dx + dy * dy + dz * dz), I);
      }
  }
  u = gsum(u);
  if (live && g == 0) {
    E.logdet[w] = ltot;
    E.sign[w] = stot;
    E.ju[w] = u;
    if (sing_any && singular_count) atomicAdd(singular_count, 1);
  }
}

static RefreshArgs make_refresh_args(const TilePlan& P, int NS, int KS, double* work) {
  RefreshArgs R;
  R.N = P.N;
  R.A = P.A;
  R.n_nuc = P.n_nuc;
  R.n_shell = P.n_shell;
  for (int g = 0; g < 16; ++g)
    for (int j = 0; j < TILE_MAXSH; ++j) R.lane_sh[g][j] = P.lane_sh[g][j];
  R.o_pos = 0;
  R.o_x = R.o_pos + TILE_W * 3 * P.N;
  R.o_phi = R.o_x + TILE_W * (4 * KS + 1);
  R.o_sh = R.o_phi + 2 * TILE_W * RF_MO;
  R.lds_slater = R.o_sh + P.n_shell * RF_REC;
  R.o_prow = R.o_pos + TILE_W * 3 * P.N;
  R.o_piv = R.o_prow + TILE_W * NS;
  R.lds_invert = R.o_piv + (TILE_W * (NS + 1) + TILE_W * 2 * NS + 1) / 2;
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
  ka<<<dim3(tiles, 2), TILE_T, ba, st>>>((const DevSys<double>*)Sd, E, R);
  kb<<<tiles, TILE_T, bb, st>>>((const DevSys<double>*)Sd, E, R, singular_count);
  return hipGetLastError();
}

template hipError_t tile_refresh<float>(const void*, EnsView<float>, const TilePlan&, double*, int*, hipStream_t);
template hipError_t tile_refresh<double>(const void*, EnsView<double>, const TilePlan&, double*, int*, hipStream_t);

}  // namespace qmc
