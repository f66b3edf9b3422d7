// qmc_tile_dev.hpp -- device code of the tile engine (see qmc_tile.hip):
// a 256-thread workgroup owns a tile of 16 walkers, 16 lanes per walker.
#pragma once

#include "qmc_engine.hpp"
#include "qmc_tile.hpp"

namespace qmc {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TILE_T = 256;   // threads per tile
constexpr int TILE_MO = 32;   // MO columns of Phi (two 16-wide MFMA tiles)
constexpr int TILE_MAXP = 6;  // primitives per shell
constexpr int TILE_REC = 32;  // floats per shell record in LDS

struct TileArgs {
  int N, A, XS, n_nuc, n_shell;
  int lane_sh[16][TILE_MAXSH];  // shells evaluated by group lane g (-1: none)
  // LDS carve-out in floats
  int o_pos, o_x, o_phi, o_spin, o_sh, o_sum;
  int lds_floats;  // total carve-out
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

// One shell record: kind, first AO, primitive count, P1 component (ints),
// centre, then exponents, s and p coefficients and STO powers of up to
// TILE_MAXP primitives (the DevSys shell and primitive tables, per shell).
__device__ __forceinline__ void tile_write_record(const DevSys<float>& S, int sh, float* rec) {
  rec[0] = as_float(S.sh_kind[sh]);
  rec[1] = as_float(S.sh_ao0[sh]);
  rec[2] = as_float(S.sh_nprim[sh]);
  rec[3] = as_float(S.sh_m[sh]);
  const int c = S.sh_center[sh], p0 = S.sh_prim0[sh], np = S.sh_nprim[sh];
  rec[4] = S.nuc[c][0];
  rec[5] = S.nuc[c][1];
  rec[6] = S.nuc[c][2];
  rec[7] = 0.f;
  for (int k = 0; k < TILE_MAXP; ++k) {
    const bool v = k < np;
    rec[8 + k] = v ? S.prim_exp[p0 + k] : 0.f;
    rec[14 + k] = v ? S.coef_s[p0 + k] : 0.f;
    rec[20 + k] = v ? S.coef_p[p0 + k] : 0.f;
    rec[26 + k] = as_float(v ? S.prim_pow[p0 + k] : 0);
  }
}

// AO values of one shell from its LDS record, written to chi[ao]; the same
// arithmetic as shell_values (methane.rs:498-590, slater.rs:17-41, 92-144)
__device__ __forceinline__ void tile_shell_values(const float* rec, float x, float y, float z, float* chi) {
  const int kind = as_int(rec[0]), a0 = as_int(rec[1]), np = as_int(rec[2]);
  const float dx = x - rec[4], dy = y - rec[5], dz = z - rec[6];
  const float r2 = dx * dx + dy * dy + dz * dz;
  if (kind <= SH_P1) {
    float rs = 0.f, rp = 0.f;
    for (int k = 0; k < np; ++k) {
      const float e = M<float>::ex(-rec[8 + k] * r2);
      rs += rec[14 + k] * e;
      rp += rec[20 + k] * e;
    }
    if (kind == SH_S) {
      chi[a0] = rs;
    } else if (kind == SH_P1) {
      const int m = as_int(rec[3]);
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
    const float sr = M<float>::sq(r2);
    float v = 0.f;
    for (int k = 0; k < np; ++k) {
      float sp = 1.f;
      for (int q = 0; q < as_int(rec[26 + k]); ++q) sp *= sr;
      v += rec[14 + k] * sp * M<float>::ex(-rec[8 + k] * sr);
    }
    chi[a0] = v;
  }
}

// e-n Jastrow value with the charge already in a register (uen_val)
template <class Js>
__device__ __forceinline__ float tile_uen(const DevSys<float>& S, float r, float Z, int I) {
  const float D = 1.f + S.b_en * r;
  const float y = r * M<float>::rcp(D);
  float v = -Z * y;
  const int K = Js::ken(S);
  if (K > 0) {
    float s, s1, s2;
    series(K, S.d_en[S.species[I]], y, s, s1, s2);
    v += s;
  }
  return v;
}

// n_steps single-electron sweeps of tile blk (16 walkers) by a 256-thread
// workgroup; lds_f is the workgroup's LDS carve-out described by TA.
template <int NU, int ND, int KS, class Js>
__device__ __forceinline__ void tile_sweep_block(const DevSys<float>& S, EnsView<float> E, const TileArgs& TA,
                                                 SweepArgs A, double* __restrict__ partial,
                                                 double* __restrict__ el_trace, int64_t blk, int64_t nblk, int tid,
                                                 float* lds_f) {
  constexpr int KH = (KS + 1) / 2;  // MFMA steps per K half
  const int q = tid >> 6, l = tid & 63;   // wavefront, lane
  const int wt = tid >> 4, g = tid & 15;  // walker in the tile, lane in its group
  const int64_t w = blk * TILE_W + wt;
  const bool live = w < E.W;
  const int64_t wl = live ? w : E.W - 1;  // dead walkers shadow the last one and write nothing
  const int N = TA.N, M_ = TA.n_nuc;
  float* pos = lds_f + TA.o_pos;  // [TILE_W][3N]
  float* X = lds_f + TA.o_x;      // [TILE_W][XS], AO values of the proposals
  float* Phi = lds_f + TA.o_phi;  // [2 K halves][TILE_W][TILE_MO]
  int* spin_of = (int*)(lds_f + TA.o_spin);
  float* shrec = lds_f + TA.o_sh;  // [n_shell][TILE_REC]
  double* sums = (double*)(lds_f + TA.o_sum);  // [TILE_W][3], per-step walker sums
  float* pw = pos + wt * 3 * N;

  // ---- state in: positions and tables to LDS, owned inverse columns to VGPRs ----
  for (int i = g; i < 3 * N; i += 16) pw[i] = E.pos[(int64_t)i * E.ld + wl];
  for (int i = g; i < TA.XS; i += 16) X[wt * TA.XS + i] = 0.f;
  for (int i = tid; i < N; i += TILE_T) spin_of[i] = S.elec_spin[i];
  for (int i = tid; i < TA.n_shell; i += TILE_T) tile_write_record(S, i, shrec + i * TILE_REC);
  int my_sh[TILE_MAXSH];
#pragma unroll
  for (int j = 0; j < TILE_MAXSH; ++j) my_sh[j] = TA.lane_sh[g][j];
  // this lane's Jastrow partners p = g + 16 t: electrons, nuclei (preloaded) or none
  float pnx[4], pny[4], pnz[4], pnZ[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int I = g + 16 * t - N;
    const bool nuc = I >= 0 && I < M_;
    pnx[t] = nuc ? S.nuc[I][0] : 0.f;
    pny[t] = nuc ? S.nuc[I][1] : 0.f;
    pnz[t] = nuc ? S.nuc[I][2] : 0.f;
    pnZ[t] = nuc ? S.Z[I] : 0.f;
  }
  float cu[2][NU], cd[2][ND];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int k = g + 16 * t;
#pragma unroll
    for (int m = 0; m < NU; ++m) cu[t][m] = k < NU ? E.ainv[(int64_t)(m * NU + k) * E.ld + wl] : 0.f;
#pragma unroll
    for (int m = 0; m < ND; ++m) cd[t][m] = k < ND ? E.ainv[(int64_t)(NU * NU + m * ND + k) * E.ld + wl] : 0.f;
  }
  // MFMA B operands: wavefront q computes MO tile q & 1 over K half q >> 1;
  // lane l holds C[mo = 16 tile + l%16][ao = (l/16) KS + half KH + s]
  const int tb = q & 1, kh = q >> 1;
  float Creg[2][KH];
#pragma unroll
  for (int sp = 0; sp < 2; ++sp)
#pragma unroll
    for (int s = 0; s < KH; ++s) {
      const int kk = kh * KH + s;
      const int ao = (l >> 4) * KS + kk, mo = tb * 16 + (l & 15);
      Creg[sp][s] = (kk < KS && ao < TA.A && mo < (sp ? ND : NU)) ? S.moT[sp][ao][mo] : 0.f;
    }
  double logdet = E.logdet[wl], sgn = E.sign[wl], ju = E.ju[wl], el = 0.0;
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t wid = (uint32_t)(A.walker0 + wl);
  const float sigma = (float)A.step_size;
  const bool en_on = S.en_on;
  __syncthreads();

  for (int st = 0; st < A.n_steps; ++st) {
    const uint32_t step = A.step0 + (uint32_t)st;
    int acc = 0;
#pragma unroll
    for (int spin = 0; spin < 2; ++spin) {
      const int n = spin ? ND : NU;
#pragma unroll 1
      for (int k = 0; k < n; ++k) {
        const int e = S.sector_elec[spin][k];
        // A: proposal
        float z[3];
        float U;
        Rng<float>::move(philox(wid, step, (uint32_t)e, RNG_SINGLE, k0, k1), z, U);
        const float ox = pw[3 * e], oy = pw[3 * e + 1], oz = pw[3 * e + 2];
        const float nx = ox + sigma * z[0], ny = oy + sigma * z[1], nz = oz + sigma * z[2];
        // B: AO values of the proposal
#pragma unroll
        for (int j = 0; j < TILE_MAXSH; ++j)
          if (my_sh[j] >= 0) tile_shell_values(shrec + my_sh[j] * TILE_REC, nx, ny, nz, X + wt * TA.XS);
        __syncthreads();
        // C: Phi[half][walker][mo] = sum over the half's AOs of X[walker][ao] C[mo][ao]
        if (tb * 16 < n) {
          f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
          const float* xr = X + (l & 15) * TA.XS + (l >> 4) * KS + kh * KH;
#pragma unroll
          for (int s = 0; s < KH; ++s) {
            if (kh * KH + s >= KS) break;
            const float b = spin ? Creg[1][s] : Creg[0][s];
            if (s & 1)
              d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[s], b, d1, 0, 0, 0);
            else
              d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[s], b, d0, 0, 0, 0);
          }
          // lane l, register r holds D[4 (l/16) + r][l % 16]
          float* ph = Phi + kh * TILE_W * TILE_MO;
#pragma unroll
          for (int r = 0; r < 4; ++r) ph[(4 * (l >> 4) + r) * TILE_MO + tb * 16 + (l & 15)] = d0[r] + d1[r];
        }
        __syncthreads();
        // D: ratio, Jastrow difference, Metropolis test, Sherman-Morrison
        const float* ph0 = Phi + wt * TILE_MO;
        const float* ph1 = ph0 + TILE_W * TILE_MO;
        float v0 = 0.f, v1 = 0.f;
        if (spin == 0) {
#pragma unroll
          for (int m = 0; m < NU; ++m) {
            const float pm = ph0[m] + ph1[m];
            v0 += pm * cu[0][m];
            v1 += pm * cu[1][m];
          }
        } else {
#pragma unroll
          for (int m = 0; m < ND; ++m) {
            const float pm = ph0[m] + ph1[m];
            v0 += pm * cd[0][m];
            v1 += pm * cd[1][m];
          }
        }
        const int owner = (l & 48) | (k & 15);
        const bool hi = k >= 16;
        const float R = __shfl(hi ? v1 : v0, owner, 64);
        float du = 0.f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int p = g + 16 * t;
          if (p < N) {
            if (p != e) {
              const float px = pw[3 * p], py = pw[3 * p + 1], pz = pw[3 * p + 2];
              const bool same = spin_of[p] == spin;
              const float ax = nx - px, ay = ny - py, az = nz - pz;
              const float bx = ox - px, by = oy - py, bz = oz - pz;
              du += uee_diff<float, Js>(S, M<float>::sq(ax * ax + ay * ay + az * az),
                                        M<float>::sq(bx * bx + by * by + bz * bz), same);
            }
          } else if (en_on && p < N + M_) {
            const float ax = nx - pnx[t], ay = ny - pny[t], az = nz - pnz[t];
            const float bx = ox - pnx[t], by = oy - pny[t], bz = oz - pnz[t];
            const float rn = M<float>::sq(ax * ax + ay * ay + az * az), ro = M<float>::sq(bx * bx + by * by + bz * bz);
            if constexpr (Js::no_series)  // -Z (r' - r) / ((1+b r')(1+b r)), one reciprocal
              du += -pnZ[t] * (rn - ro) * M<float>::rcp((1.f + S.b_en * rn) * (1.f + S.b_en * ro));
            else
              du += tile_uen<Js>(S, rn, pnZ[t], p - N) - tile_uen<Js>(S, ro, pnZ[t], p - N);
          }
        }
        du = gsum(du);
        const float pacc = R * R * M<float>::ex(2.f * du);
        if (U < pacc) {  // uniform within the walker's group
          const float iR = M<float>::rcp(R);
          const bool own0 = g == (k & 15) && !hi, own1 = g == (k & 15) && hi;
          const float f0 = v0 * iR, f1 = v1 * iR;
          if (spin == 0) {
#pragma unroll
            for (int m = 0; m < NU; ++m) {
              const float c = __shfl(hi ? cu[1][m] : cu[0][m], owner, 64);
              cu[0][m] = own0 ? c * iR : cu[0][m] - c * f0;
              cu[1][m] = own1 ? c * iR : cu[1][m] - c * f1;
            }
          } else {
#pragma unroll
            for (int m = 0; m < ND; ++m) {
              const float c = __shfl(hi ? cd[1][m] : cd[0][m], owner, 64);
              cd[0][m] = own0 ? c * iR : cd[0][m] - c * f0;
              cd[1][m] = own1 ? c * iR : cd[1][m] - c * f1;
            }
          }
          if (g == 0) {
            pw[3 * e] = nx;
            pw[3 * e + 1] = ny;
            pw[3 * e + 2] = nz;
          }
          logdet += (double)M<float>::lg(R < 0.f ? -R : R);
          if (R < 0.f) sgn = -sgn;
          ju += (double)du;
          ++acc;
        }
      }
    }
    __syncthreads();
    // local energy: lane g takes the electrons of sector slots g and g + 16
    double kin = 0.0, pot = 0.0;
#pragma unroll
    for (int spin = 0; spin < 2; ++spin) {
      const int n = spin ? ND : NU;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int k = g + 16 * t;
        if (k >= n) continue;
        const int i = S.sector_elec[spin][k];
        const float x = pw[3 * i], y = pw[3 * i + 1], z = pw[3 * i + 2];
        auto wgt = [&](int ao) {
          float s = 0.f;
          if (spin == 0) {
#pragma unroll
            for (int m = 0; m < NU; ++m) s += S.moT[0][ao][m] * cu[t][m];
          } else {
#pragma unroll
            for (int m = 0; m < ND; ++m) s += S.moT[1][ao][m] * cd[t][m];
          }
          return s;
        };
        float gdx, gdy, gdz, ld;
        det_terms_w(S, x, y, z, wgt, gdx, gdy, gdz, ld, NoShell{});
        float gux = 0.f, guy = 0.f, guz = 0.f, lu = 0.f, pairsq = 0.f, pee = 0.f, pen = 0.f;
        for (int j = 0; j < N; ++j) {
          if (j == i) continue;
          const float dx = x - pw[3 * j], dy = y - pw[3 * j + 1], dz = z - pw[3 * j + 2];
          const float r = M<float>::sq(dx * dx + dy * dy + dz * dz);
          const float ir = M<float>::rcp(r);
          float u1, lp;
          uee_d<float, Js>(S, r, ir, spin_of[j] == spin, u1, lp);
          const bool ok = r >= 1e-10f;
          const float f = ok ? u1 * ir : 0.f;
          gux += f * dx;
          guy += f * dy;
          guz += f * dz;
          lu += ok ? lp : 0.f;
          pairsq += ok ? u1 * u1 : 0.f;
          pee += ir;
        }
        for (int I = 0; I < M_; ++I) {
          const float dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
          const float r = M<float>::sq(dx * dx + dy * dy + dz * dz);
          const float ir = M<float>::rcp(r);
          pen += S.Z[I] * ir;
          if (en_on) {
            float u1, lp;
            uen_d<float, Js>(S, r, ir, I, u1, lp);
            const bool ok = r >= 1e-10f;
            const float f = ok ? u1 * ir : 0.f;
            gux += f * dx;
            guy += f * dy;
            guz += f * dz;
            lu += ok ? lp : 0.f;
          }
        }
        const float gs = S.grad_scale;
        const float gjx = gs * gux, gjy = gs * guy, gjz = gs * guz;
        const float lj = (S.flags & F_J2_LAP_QUIRK) ? (lu + pairsq) : (lu + gux * gux + guy * guy + guz * guz);
        const float lap = ld + lj + 2.f * (gdx * gjx + gdy * gjy + gdz * gjz);
        kin += -0.5 * (double)lap;
        pot += 0.5 * (double)pee - (double)pen;
      }
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
  if (live) {
    for (int i = g; i < 3 * N; i += 16) E.pos[(int64_t)i * E.ld + w] = pw[i];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int k = g + 16 * t;
      if (k < NU)
#pragma unroll
        for (int m = 0; m < NU; ++m) E.ainv[(int64_t)(m * NU + k) * E.ld + w] = cu[t][m];
      if (k < ND)
#pragma unroll
        for (int m = 0; m < ND; ++m) E.ainv[(int64_t)(NU * NU + m * ND + k) * E.ld + w] = cd[t][m];
    }
    if (g == 0) {
      E.logdet[w] = logdet;
      E.sign[w] = sgn;
      E.ju[w] = ju;
      E.eloc[w] = el;
    }
  }
}

// kernel arguments of a tile launch: LDS carve-out and lane shell lists
template <int KS>
inline TileArgs make_tile_args(const TilePlan& P) {
  TileArgs T;
  T.N = P.N;
  T.A = P.A;
  T.XS = 4 * KS + 1;
  T.n_nuc = P.n_nuc;
  T.n_shell = P.n_shell;
  for (int g = 0; g < 16; ++g)
    for (int j = 0; j < TILE_MAXSH; ++j) T.lane_sh[g][j] = P.lane_sh[g][j];
  T.o_pos = 0;
  T.o_x = T.o_pos + TILE_W * 3 * P.N;
  T.o_phi = T.o_x + TILE_W * T.XS;
  T.o_spin = T.o_phi + 2 * TILE_W * TILE_MO;
  T.o_sh = T.o_spin + (P.N + 3) / 4 * 4;
  T.o_sum = (T.o_sh + P.n_shell * TILE_REC + 1) / 2 * 2;  // 8-byte aligned
  T.lds_floats = T.o_sum + 2 * 3 * TILE_W;
  return T;
}

}  // namespace qmc
