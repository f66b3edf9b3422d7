// qmc_wave.hip -- the wave engine: one 64-lane wavefront owns one walker
// whose positions and inverse Slater matrices live in LDS.  Used for
// systems too large for the register engine (benzene: 21+21 electrons, 66
// AOs) and for any electron count the register engine was not compiled
// for.  Within a single-electron move the lanes split the work:
//   shells      -> AO values of the proposed position
//   MOs         -> MO values and the determinant ratio (wave reduction)
//   electrons   -> Jastrow difference (wave reduction)
//   (mo, row)   -> Sherman-Morrison update of the inverse
// and the local energy puts one electron on each lane.
#include "qmc_wave.hpp"

#include "qmc_engine.hpp"
#include "qmc_wave_dev.hpp"

namespace qmc {

// full ln J (FP64): lane i sums pairs (i, j > i) and its e-n terms
template <class T>
__device__ double wave_jastrow_u(const DevSys<double>& S, const T* pos, int N) {
  const int lane = threadIdx.x;
  double u = 0.0;
  for (int i = lane; i < N; i += 64) {
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    for (int j = i + 1; j < N; ++j) {
      double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
      const bool same = S.elec_spin[i] == S.elec_spin[j];
      u += uee_val(S, sqrt(dx * dx + dy * dy + dz * dz), same) + S.cst_ee;
    }
    if (S.en_on)
      for (int I = 0; I < S.n_nuc; ++I) {
        double dx = xi - S.nuc[I][0], dy = yi - S.nuc[I][1], dz = zi - S.nuc[I][2];
        u += uen_val(S, sqrt(dx * dx + dy * dy + dz * dz), I);
      }
  }
  return wsum(u);
}

// local energy: lane i handles electron i; returns the wave-uniform E_L
template <class T>
__device__ double wave_local_energy(const DevSys<T>& S, const WaveLayout& Y, T* L, double* gout, double* lout) {
  const int lane = threadIdx.x;
  const T* pos = L + Y.pos();
  double kin = 0.0, pot = 0.0;
  for (int i = lane; i < Y.N; i += 64) {
    const int spin = S.elec_spin[i], k = S.elec_slot[i];
    const int n = spin ? Y.nd : Y.nu;
    const T* ai = L + Y.ainv(spin);
    const T x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    auto wgt = [&](int ao) {
      T w = 0;
      for (int m = 0; m < n; ++m) w += S.moT[spin][ao][m] * ai[m * n + k];
      return w;
    };
    T gdx, gdy, gdz, ld;
    det_terms_w(S, x, y, z, wgt, gdx, gdy, gdz, ld, NoShell{});
    T gux = 0, guy = 0, guz = 0, lu = 0, pairsq = 0;
    double pe = 0.0;
    for (int j = 0; j < Y.N; ++j) {
      if (j == i) continue;
      T dx = x - pos[3 * j], dy = y - pos[3 * j + 1], dz = z - pos[3 * j + 2];
      T r, ir;
      M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
      T u1, lp;
      uee_d(S, r, ir, S.elec_spin[j] == spin, u1, lp);
      if (r >= T(1e-10)) {
        T f = u1 * ir;
        gux += f * dx;
        guy += f * dy;
        guz += f * dz;
        lu += lp;
        pairsq += u1 * u1;
      }
      pe += 0.5 * (double)ir;
    }
    for (int I = 0; I < S.n_nuc; ++I) {
      T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
      T r, ir;
      M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
      pe -= (double)(S.Z[I] * ir);
      if (S.en_on) {
        T u1, lp;
        uen_d(S, r, ir, I, u1, lp);
        if (r >= T(1e-10)) {
          T f = u1 * ir;
          gux += f * dx;
          guy += f * dy;
          guz += f * dz;
          lu += lp;
        }
      }
    }
    const T gs = S.grad_scale;
    T gjx = gs * gux, gjy = gs * guy, gjz = gs * guz;
    T lj = (S.flags & F_J2_LAP_QUIRK) ? (lu + pairsq) : (lu + gux * gux + guy * guy + guz * guz);
    const T lap = ld + lj + T(2) * (gdx * gjx + gdy * gjy + gdz * gjz);
    kin += -0.5 * (double)lap;
    pot += pe;
    if (gout) {
      gout[3 * i] = (double)(gdx + gjx);
      gout[3 * i + 1] = (double)(gdy + gjy);
      gout[3 * i + 2] = (double)(gdz + gjz);
    }
    if (lout) lout[i] = (double)lap;
  }
  return wsum(kin) + wsum(pot) + S.vnn;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void load_walker(const EnsView<T>& E, const WaveLayout& Y, int64_t w, T* L) {
  for (int q = threadIdx.x; q < 3 * Y.N; q += 64) L[Y.pos() + q] = E.pos[(int64_t)q * E.ld + w];
  for (int q = threadIdx.x; q < Y.nu * Y.nu + Y.nd * Y.nd; q += 64) L[Y.ainv(0) + q] = E.ainv[(int64_t)q * E.ld + w];
  __syncthreads();
}

template <class T>
__device__ __forceinline__ void store_walker(const EnsView<T>& E, const WaveLayout& Y, int64_t w, const T* L) {
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * Y.N; q += 64) E.pos[(int64_t)q * E.ld + w] = L[Y.pos() + q];
  for (int q = threadIdx.x; q < Y.nu * Y.nu + Y.nd * Y.nd; q += 64) E.ainv[(int64_t)q * E.ld + w] = L[Y.ainv(0) + q];
}

// build / refresh: inverse (FP64) from the positions, ln|D|, ln J, E_L
template <class T>
__global__ void __launch_bounds__(64) wave_build_kernel(const DevSys<double>* __restrict__ Sd,
                                                        const DevSys<T>* __restrict__ St, EnsView<T> E,
                                                        WaveLayout Y, int want_eloc, int* singular_count) {
  extern __shared__ double lds_d[];
  const int64_t w = blockIdx.x;
  if (E.n_live && w >= *E.n_live) return;  // past a branching population's live walkers
  T* L = (T*)lds_d;
  double* a = lds_d + (Y.total() * sizeof(T) + 7) / 8;
  double* inv = a + Y.nmax * Y.nmax;
  for (int q = threadIdx.x; q < 3 * Y.N; q += 64) L[Y.pos() + q] = E.pos[(int64_t)q * E.ld + w];
  __syncthreads();
  double ltot = 0.0, stot = 1.0;
  bool sing = false;
  for (int spin = 0; spin < 2; ++spin) {
    const int n = spin ? Y.nd : Y.nu;
    double la, sg;
    sing |= wave_sector_inverse(*Sd, spin, L + Y.pos(), n, a, inv, la, sg);
    ltot += la;
    stot *= sg;
    for (int q = threadIdx.x; q < n * n; q += 64) L[Y.ainv(spin) + q] = (T)inv[q];
    __syncthreads();
  }
  const double ju = wave_jastrow_u(*Sd, L + Y.pos(), Y.N);
  double el = 0.0;
  if (want_eloc) el = wave_local_energy(*St, Y, L, nullptr, nullptr);
  store_walker(E, Y, w, L);
  if (threadIdx.x == 0) {
    E.logdet[w] = ltot;
    E.sign[w] = stot;
    E.ju[w] = ju;
    if (want_eloc) E.eloc[w] = el;
    if (sing && singular_count) atomicAdd(singular_count, 1);
  }
}

// persistent single-electron sweeps
template <class T>
__global__ void __launch_bounds__(64) wave_sweep_kernel(const DevSys<T>* __restrict__ Sp, EnsView<T> E, WaveLayout Y,
                                                        SweepArgs A, double* __restrict__ partial,
                                                        double* __restrict__ el_trace) {
  extern __shared__ double lds_d[];
  const DevSys<T>& S = *Sp;
  const int64_t w = blockIdx.x;
  const int lane = threadIdx.x;
  T* L = (T*)lds_d;
  load_walker(E, Y, w, L);
  T* pos = L + Y.pos();
  T* chi = L + Y.chi();
  T* phi = L + Y.phi();
  T* cv = L + Y.cvec();
  T* wv = L + Y.wvec();
  double logdet = E.logdet[w], sgn = E.sign[w], ju = E.ju[w], el = 0.0;
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t wid = (uint32_t)(A.walker0 + w);
  const T sigma = (T)A.step_size;
  // MO values in chunks: lane m + n c (c < 64 / n) sums MO m over the AOs
  // [lo, hi) of chunk c, so no lane runs the whole AO sum alone
  T* mp = L + Y.mpart();
  int mc_C[2], mc_c[2], mc_m[2], mc_lo[2], mc_hi[2];
  for (int sp = 0; sp < 2; ++sp) {
    const int n = sp ? Y.nd : Y.nu;
    const int C = n > 0 ? 64 / n : 0, c = n > 0 ? lane / n : 64;
    mc_C[sp] = C;
    mc_c[sp] = c;
    mc_m[sp] = n > 0 ? lane - c * n : 0;
    mc_lo[sp] = c < C ? c * Y.A / C : 0;
    mc_hi[sp] = c < C ? (c + 1) * Y.A / C : 0;
  }
  for (int s = 0; s < A.n_steps; ++s) {
    const uint32_t step = A.step0 + (uint32_t)s;
    int acc = 0;
    for (int spin = 0; spin < 2; ++spin) {
      const int n = spin ? Y.nd : Y.nu;
      T* ai = L + Y.ainv(spin);
      for (int k = 0; k < n; ++k) {
        const int e = S.sector_elec[spin][k];
        T z[3];
        T U;
        Rng<T>::move(philox(wid, step, (uint32_t)e, RNG_SINGLE, k0, k1), z, U);
        const T ox = pos[3 * e], oy = pos[3 * e + 1], oz = pos[3 * e + 2];
        const T nx = ox + sigma * z[0], ny = oy + sigma * z[1], nz = oz + sigma * z[2];
        // AO values of the proposal: one shell per lane
        for (int sh = lane; sh < S.n_shell; sh += 64) shell_values(S, sh, nx, ny, nz, chi);
        __syncthreads();
        {
          const int mi = mc_m[spin];
          T v = 0;
          for (int ao = mc_lo[spin]; ao < mc_hi[spin]; ++ao) v += S.moT[spin][ao][mi] * chi[ao];
          if (mc_c[spin] < mc_C[spin]) mp[mc_c[spin] * n + mi] = v;
        }
        __syncthreads();
        // MO values (chunk sums in chunk order) and R = sum_m phi_m ainv[m][k]
        T part = 0;
        for (int m = lane; m < n; m += 64) {
          T v = 0;
          for (int c = 0; c < mc_C[spin]; ++c) v += mp[c * n + m];
          phi[m] = v;
          part += v * ai[m * n + k];
        }
        const T R = wsum(part);
        // Jastrow difference: one partner electron or nucleus per lane
        T du = 0;
        for (int j = lane; j < Y.N; j += 64) {
          if (j == e) continue;
          const bool same = S.elec_spin[j] == spin;
          T ax = nx - pos[3 * j], ay = ny - pos[3 * j + 1], az = nz - pos[3 * j + 2];
          T bx = ox - pos[3 * j], by = oy - pos[3 * j + 1], bz = oz - pos[3 * j + 2];
          du += uee_val(S, M<T>::dist(ax * ax + ay * ay + az * az), same) -
                uee_val(S, M<T>::dist(bx * bx + by * by + bz * bz), same);
        }
        if (S.en_on)
          for (int I = lane; I < S.n_nuc; I += 64) {
            T ax = nx - S.nuc[I][0], ay = ny - S.nuc[I][1], az = nz - S.nuc[I][2];
            T bx = ox - S.nuc[I][0], by = oy - S.nuc[I][1], bz = oz - S.nuc[I][2];
            du += uen_val(S, M<T>::dist(ax * ax + ay * ay + az * az), I) -
                  uen_val(S, M<T>::dist(bx * bx + by * by + bz * bz), I);
          }
        du = wsum(du);
        const T p = R * R * M<T>::ex(T(2) * du);
        __syncthreads();  // phi complete, chi reads done
        if (U < p) {       // wave-uniform
          const T iR = M<T>::rcp(R);
          for (int j = lane; j < n; j += 64) {
            T v = 0;
            for (int m = 0; m < n; ++m) v += phi[m] * ai[m * n + j];
            wv[j] = v * iR;
            cv[j] = ai[j * n + k];
          }
          __syncthreads();
          for_each_elem(n, [&](int idx, int m, int j) { ai[idx] = (j == k) ? cv[m] * iR : ai[idx] - cv[m] * wv[j]; });
          if (lane == 0) {
            pos[3 * e] = nx;
            pos[3 * e + 1] = ny;
            pos[3 * e + 2] = nz;
          }
          logdet += (double)M<T>::lg(R < T(0) ? -R : R);
          if (R < T(0)) sgn = -sgn;
          ju += (double)du;
          ++acc;
          __syncthreads();
        }
      }
    }
    el = wave_local_energy(S, Y, L, nullptr, nullptr);
    if (lane == 0) {
      if (el_trace) el_trace[(int64_t)s * A.trace_ld + w] = el;
      double* pp = partial + ((int64_t)s * gridDim.x + blockIdx.x) * 3;
      pp[0] = el;
      pp[1] = el * el;
      pp[2] = (double)acc;
    }
    __syncthreads();
  }
  store_walker(E, Y, w, L);
  if (lane == 0) {
    E.logdet[w] = logdet;
    E.sign[w] = sgn;
    E.ju[w] = ju;
    E.eloc[w] = el;
  }
}

// Persistent drift-diffusion sweeps (DriftDiffusionVMC::sweep,
// vmc.rs:382-438), one wavefront per walker: Langevin proposal
// r' = r + cap(tau F(r)) + sqrt(2 tau) z for each electron in turn, with
// F = 2 grad ln|psi| of the moved electron, acceptance R^2 e^{2 du}
// e^{G_b - G_f} (vmc.rs:323-346), Sherman-Morrison on acceptance.  The
// moves and random streams are those of the register engine's
// sector_moves_dd.  Lanes own shells for the determinant part: the AO
// weights w_ao = sum_m C[m][ao] Ainv[m][k] give grad ln|D| at r, and at r'
// R = sum w chi(r') and grad D'/D = sum w grad chi(r'); lanes own partner
// electrons and nuclei for the Jastrow part.  Blocks past a branching
// population's live count write zero partial sums.
template <class T>
__global__ void __launch_bounds__(64) wave_dd_kernel(const DevSys<T>* __restrict__ Sp, EnsView<T> E, WaveLayout Y,
                                                     SweepArgs A, double* __restrict__ partial,
                                                     double* __restrict__ el_trace) {
  extern __shared__ double lds_d[];
  const DevSys<T>& S = *Sp;
  const int64_t w = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t W = E.n_live ? *E.n_live : E.W;
  if (w >= W) {
    if (lane == 0)
      for (int s = 0; s < A.n_steps; ++s)
        for (int q = 0; q < 3; ++q) partial[((int64_t)s * gridDim.x + blockIdx.x) * 3 + q] = 0.0;
    return;
  }
  T* L = (T*)lds_d;
  load_walker(E, Y, w, L);
  T* pos = L + Y.pos();
  T* chi = L + Y.chi();
  T* phi = L + Y.phi();
  T* cv = L + Y.cvec();
  T* wv = L + Y.wvec();
  T* wao = L + Y.wao();
  double logdet = E.logdet[w], sgn = E.sign[w], ju = E.ju[w], el = 0.0;
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t wid = (uint32_t)(A.walker0 + w);
  const T tau = (T)A.step_size, md = (T)A.max_drift;
  const T s2t = M<T>::sq(T(2) * tau), i4t = T(0.25) / tau;
  const T gs = S.grad_scale;
  // the Jastrow value and gradient terms of electron e at (x, y, z) over
  // the partners this lane owns
  auto jastrow_part = [&](int e, int spin, T x, T y, T z, T& u, T& gx, T& gy, T& gz) {
    u = gx = gy = gz = T(0);
    for (int j = lane; j < Y.N; j += 64) {
      if (j == e) continue;
      const T dx = x - pos[3 * j], dy = y - pos[3 * j + 1], dz = z - pos[3 * j + 2];
      T r, ir;
      M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
      const bool same = S.elec_spin[j] == spin;
      T u1, lp;
      uee_d(S, r, ir, same, u1, lp);
      u += uee_val(S, r, same);
      const T f = r >= T(1e-10) ? u1 * ir : T(0);
      gx += f * dx;
      gy += f * dy;
      gz += f * dz;
    }
    if (S.en_on)
      for (int I = lane; I < S.n_nuc; I += 64) {
        const T dx = x - S.nuc[I][0], dy = y - S.nuc[I][1], dz = z - S.nuc[I][2];
        T r, ir;
        M<T>::dist_inv(dx * dx + dy * dy + dz * dz, r, ir);
        T u1, lp;
        uen_d(S, r, ir, I, u1, lp);
        u += uen_val(S, r, I);
        const T f = r >= T(1e-10) ? u1 * ir : T(0);
        gx += f * dx;
        gy += f * dy;
        gz += f * dz;
      }
  };
  const uint32_t step0 = A.step0_dev ? *A.step0_dev : A.step0;
  for (int s = 0; s < A.n_steps; ++s) {
    const uint32_t step = step0 + (uint32_t)s;
    int acc = 0;
    for (int spin = 0; spin < 2; ++spin) {
      const int n = spin ? Y.nd : Y.nu;
      T* ai = L + Y.ainv(spin);
      for (int k = 0; k < n; ++k) {
        const int e = S.sector_elec[spin][k];
        const T ox = pos[3 * e], oy = pos[3 * e + 1], oz = pos[3 * e + 2];
        // AO weights of the moved electron, one AO per lane
        for (int a = lane; a < Y.A; a += 64) {
          T wq = 0;
          for (int m = 0; m < n; ++m) wq += S.moT[spin][a][m] * ai[m * n + k];
          wao[a] = wq;
        }
        __syncthreads();
        // quantum force at r: grad ln|D| from the shells, Jastrow gradient
        T gx = 0, gy = 0, gz = 0;
        for (int sh = lane; sh < S.n_shell; sh += 64) {
          T v[4], g[4][3];
          const int nq = shell_values_grad(S, sh, ox, oy, oz, v, g);
          const int a0 = S.sh_ao0[sh];
          for (int q = 0; q < nq; ++q) {
            const T wq = wao[a0 + q];
            gx += wq * g[q][0];
            gy += wq * g[q][1];
            gz += wq * g[q][2];
          }
        }
        T uo, jx, jy, jz;
        jastrow_part(e, spin, ox, oy, oz, uo, jx, jy, jz);
        gx = wsum(gx);
        gy = wsum(gy);
        gz = wsum(gz);
        uo = wsum(uo);
        jx = wsum(jx);
        jy = wsum(jy);
        jz = wsum(jz);
        T dold[3];
        cap_drift(T(2) * (gx + gs * jx), T(2) * (gy + gs * jy), T(2) * (gz + gs * jz), tau, md, dold);
        T z[3];
        T U;
        Rng<T>::move(philox(wid, step, (uint32_t)e, RNG_DRIFT, k0, k1), z, U);
        const T nx = ox + dold[0] + s2t * z[0], ny = oy + dold[1] + s2t * z[1], nz = oz + dold[2] + s2t * z[2];
        __syncthreads();  // AO weights complete
        // at r': R = sum w chi, grad D'/D = sum w grad chi; chi(r') kept for the update
        T R = 0, hx = 0, hy = 0, hz = 0;
        for (int sh = lane; sh < S.n_shell; sh += 64) {
          T v[4], g[4][3];
          const int nq = shell_values_grad(S, sh, nx, ny, nz, v, g);
          const int a0 = S.sh_ao0[sh];
          for (int q = 0; q < nq; ++q) {
            const T wq = wao[a0 + q];
            chi[a0 + q] = v[q];
            R += wq * v[q];
            hx += wq * g[q][0];
            hy += wq * g[q][1];
            hz += wq * g[q][2];
          }
        }
        T un, vx, vy, vz;
        jastrow_part(e, spin, nx, ny, nz, un, vx, vy, vz);
        R = wsum(R);
        hx = wsum(hx);
        hy = wsum(hy);
        hz = wsum(hz);
        un = wsum(un);
        vx = wsum(vx);
        vy = wsum(vy);
        vz = wsum(vz);
        const T iR = M<T>::rcp(R);
        T dnew[3];
        cap_drift(T(2) * (hx * iR + gs * vx), T(2) * (hy * iR + gs * vy), T(2) * (hz * iR + gs * vz), tau, md,
                  dnew);
        const T f0 = nx - ox - dold[0], f1 = ny - oy - dold[1], f2 = nz - oz - dold[2];
        const T b0 = ox - nx - dnew[0], b1 = oy - ny - dnew[1], b2 = oz - nz - dnew[2];
        const T gf = -(f0 * f0 + f1 * f1 + f2 * f2) * i4t;
        const T gb = -(b0 * b0 + b1 * b1 + b2 * b2) * i4t;
        const T du = un - uo;
        const T p = R * R * M<T>::ex(T(2) * du + gb - gf);
        __syncthreads();  // chi(r') complete, AO weight reads done
        if (U < p) {      // wave-uniform
          for (int m = lane; m < n; m += 64) {
            T v = 0;
            for (int ao = 0; ao < Y.A; ++ao) v += S.moT[spin][ao][m] * chi[ao];
            phi[m] = v;
          }
          __syncthreads();
          for (int j = lane; j < n; j += 64) {
            T v = 0;
            for (int m = 0; m < n; ++m) v += phi[m] * ai[m * n + j];
            wv[j] = v * iR;
            cv[j] = ai[j * n + k];
          }
          __syncthreads();
          for_each_elem(n, [&](int idx, int m, int j) { ai[idx] = (j == k) ? cv[m] * iR : ai[idx] - cv[m] * wv[j]; });
          if (lane == 0) {
            pos[3 * e] = nx;
            pos[3 * e + 1] = ny;
            pos[3 * e + 2] = nz;
          }
          logdet += (double)M<T>::lg(R < T(0) ? -R : R);
          if (R < T(0)) sgn = -sgn;
          ju += (double)du;
          ++acc;
          __syncthreads();
        }
      }
    }
    el = wave_local_energy(S, Y, L, nullptr, nullptr);
    if (lane == 0) {
      if (el_trace) el_trace[(int64_t)s * A.trace_ld + w] = el;
      double* pp = partial + ((int64_t)s * gridDim.x + blockIdx.x) * 3;
      pp[0] = el;
      pp[1] = el * el;
      pp[2] = (double)acc;
    }
    __syncthreads();
  }
  store_walker(E, Y, w, L);
  if (lane == 0) {
    E.logdet[w] = logdet;
    E.sign[w] = sgn;
    E.ju[w] = ju;
    E.eloc[w] = el;
  }
}

// All-electron Metropolis (MCMCSimulation::metropolis_step, vmc.rs:95-119),
// one wavefront per walker: every electron displaced by sigma z with the
// register engine's random streams, ln|psi'| from the FP64 Gauss-Jordan
// build of both sectors and the full Jastrow sum, acceptance
// U < exp(2 (ln|psi'| - ln|psi|)), E_L recomputed on acceptance and cached
// otherwise.  Only the positions persist: the proposal's inverses serve
// its E_L, and the host refreshes the inverses before single-electron
// sweeps.
template <class T>
__global__ void __launch_bounds__(64) wave_ae_kernel(const DevSys<double>* __restrict__ Sd,
                                                     const DevSys<T>* __restrict__ St, EnsView<T> E, WaveLayout Y,
                                                     SweepArgs A, double* __restrict__ partial,
                                                     double* __restrict__ el_trace) {
  extern __shared__ double lds_d[];
  const DevSys<T>& S = *St;
  const int64_t w = blockIdx.x;
  const int lane = threadIdx.x;
  T* L = (T*)lds_d;
  double* a = lds_d + (Y.total() * sizeof(T) + 7) / 8;
  double* inv = a + Y.nmax * Y.nmax;
  T* qn = (T*)(inv + Y.nmax * Y.nmax);  // the proposal, [N][3]
  T* pos = L + Y.pos();
  for (int q = lane; q < 3 * Y.N; q += 64) pos[q] = E.pos[(int64_t)q * E.ld + w];
  __syncthreads();
  double ldet = E.logdet[w], ju = E.ju[w], sgn = E.sign[w], el = E.eloc[w];
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t wid = (uint32_t)(A.walker0 + w);
  const T sigma = (T)A.step_size;
  for (int s = 0; s < A.n_steps; ++s) {
    const uint32_t step = A.step0 + (uint32_t)s;
    for (int e = lane; e < Y.N; e += 64) {
      T z[3];
      Rng<T>::normals3(philox(wid, step, (uint32_t)e, RNG_ALL, k0, k1), z);
      for (int c = 0; c < 3; ++c) qn[3 * e + c] = pos[3 * e + c] + sigma * z[c];
    }
    __syncthreads();
    double lnew = 0.0, snew = 1.0;
    for (int spin = 0; spin < 2; ++spin) {
      const int n = spin ? Y.nd : Y.nu;
      double la, sg;
      wave_sector_inverse(*Sd, spin, qn, n, a, inv, la, sg);
      lnew += la;
      snew *= sg;
      for (int q = lane; q < n * n; q += 64) L[Y.ainv(spin) + q] = (T)inv[q];
      __syncthreads();
    }
    const double jun = wave_jastrow_u(*Sd, qn, Y.N);
    const double U = u01d(philox(wid, step, (uint32_t)Y.N, RNG_ALL, k0, k1).x);
    int acc = 0;
    if (U < exp(2.0 * ((lnew + jun) - (ldet + ju)))) {  // wave-uniform
      for (int q = lane; q < 3 * Y.N; q += 64) pos[q] = qn[q];
      __syncthreads();
      ldet = lnew;
      ju = jun;
      sgn = snew;
      el = wave_local_energy(S, Y, L, nullptr, nullptr);
      acc = 1;
    }
    if (lane == 0) {
      if (el_trace) el_trace[(int64_t)s * A.trace_ld + w] = el;
      double* pp = partial + ((int64_t)s * gridDim.x + blockIdx.x) * 3;
      pp[0] = el;
      pp[1] = el * el;
      pp[2] = (double)acc;
    }
    __syncthreads();
  }
  for (int q = lane; q < 3 * Y.N; q += 64) E.pos[(int64_t)q * E.ld + w] = pos[q];
  if (lane == 0) {
    E.logdet[w] = ldet;
    E.sign[w] = sgn;
    E.ju[w] = ju;
    E.eloc[w] = el;
  }
}

// per-configuration evaluation from AoS input
template <class T>
__global__ void __launch_bounds__(64) wave_eval_kernel(const DevSys<double>* __restrict__ Sd,
                                                       const DevSys<T>* __restrict__ St, int64_t W,
                                                       const double* __restrict__ in, WaveLayout Y, EvalOut O,
                                                       int* singular_count) {
  extern __shared__ double lds_d[];
  const int64_t w = blockIdx.x;
  T* L = (T*)lds_d;
  double* a = lds_d + (Y.total() * sizeof(T) + 7) / 8;
  double* inv = a + Y.nmax * Y.nmax;
  double* pd = inv + Y.nmax * Y.nmax;  // FP64 copy of the positions
  for (int q = threadIdx.x; q < 3 * Y.N; q += 64) {
    pd[q] = in[w * 3 * Y.N + q];
    L[Y.pos() + q] = (T)pd[q];
  }
  __syncthreads();
  double ltot = 0.0, stot = 1.0;
  bool sing = false;
  for (int spin = 0; spin < 2; ++spin) {
    const int n = spin ? Y.nd : Y.nu;
    double la, sg;
    sing |= wave_sector_inverse(*Sd, spin, pd, n, a, inv, la, sg);
    ltot += la;
    stot *= sg;
    for (int q = threadIdx.x; q < n * n; q += 64) L[Y.ainv(spin) + q] = (T)inv[q];
    __syncthreads();
  }
  const double ju = wave_jastrow_u(*Sd, pd, Y.N);
  double* g = O.grad ? O.grad + w * 3 * Y.N : nullptr;
  double* l = O.lap ? O.lap + w * Y.N : nullptr;
  double el = 0.0;
  if (O.eloc || O.grad || O.lap) el = wave_local_energy(*St, Y, L, g, l);
  if (threadIdx.x == 0) {
    if (O.logabs) O.logabs[w] = ltot + ju;
    if (O.sign) O.sign[w] = stot;
    if (O.eloc) O.eloc[w] = el;
    if (sing && singular_count) atomicAdd(singular_count, 1);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static WaveLayout layout(int N, int nu, int nd, int A) {
  WaveLayout Y;
  Y.N = N;
  Y.nu = nu;
  Y.nd = nd;
  Y.nmax = nu > nd ? nu : nd;
  Y.A = A;
  return Y;
}

template <class T>
static size_t lds_bytes(const WaveLayout& Y, bool fp64_scratch) {
  size_t b = ((size_t)Y.total() * sizeof(T) + 7) / 8 * 8;
  if (fp64_scratch) b += sizeof(double) * (2 * (size_t)Y.nmax * Y.nmax + 3 * (size_t)Y.N);
  return b;
}

template <class K>
static hipError_t set_lds(K kernel, size_t bytes) {
  if (bytes > 160 * 1024) return hipErrorInvalidValue;
  return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <class T>
hipError_t wave_build(const void* Sd, const void* St, EnsView<T> E, WaveShape sh, int want_eloc, int* sing,
                      hipStream_t st) {
  WaveLayout Y = layout(sh.N, sh.nu, sh.nd, sh.A);
  const size_t b = lds_bytes<T>(Y, true);
  hipError_t e = set_lds(wave_build_kernel<T>, b);
  if (e != hipSuccess) return e;
  wave_build_kernel<T><<<(unsigned)E.W, 64, b, st>>>((const DevSys<double>*)Sd, (const DevSys<T>*)St, E, Y,
                                                     want_eloc, sing);
  return hipGetLastError();
}

template <class T>
hipError_t wave_sweep(const void* St, EnsView<T> E, WaveShape sh, SweepArgs A, double* partial, double* trace,
                      hipStream_t st) {
  WaveLayout Y = layout(sh.N, sh.nu, sh.nd, sh.A);
  const size_t b = lds_bytes<T>(Y, false);
  hipError_t e = set_lds(wave_sweep_kernel<T>, b);
  if (e != hipSuccess) return e;
  wave_sweep_kernel<T><<<(unsigned)E.W, 64, b, st>>>((const DevSys<T>*)St, E, Y, A, partial, trace);
  return hipGetLastError();
}

template <class T>
hipError_t wave_dd(const void* St, EnsView<T> E, WaveShape sh, SweepArgs A, double* partial, double* trace,
                   hipStream_t st) {
  WaveLayout Y = layout(sh.N, sh.nu, sh.nd, sh.A);
  const size_t b = lds_bytes<T>(Y, false);
  hipError_t e = set_lds(wave_dd_kernel<T>, b);
  if (e != hipSuccess) return e;
  wave_dd_kernel<T><<<(unsigned)E.W, 64, b, st>>>((const DevSys<T>*)St, E, Y, A, partial, trace);
  return hipGetLastError();
}

template <class T>
hipError_t wave_ae(const void* Sd, const void* St, EnsView<T> E, WaveShape sh, SweepArgs A, double* partial,
                   double* trace, hipStream_t st) {
  WaveLayout Y = layout(sh.N, sh.nu, sh.nd, sh.A);
  const size_t b = lds_bytes<T>(Y, true);  // the FP64 scratch's 3N doubles hold the proposal
  hipError_t e = set_lds(wave_ae_kernel<T>, b);
  if (e != hipSuccess) return e;
  wave_ae_kernel<T><<<(unsigned)E.W, 64, b, st>>>((const DevSys<double>*)Sd, (const DevSys<T>*)St, E, Y, A, partial,
                                                  trace);
  return hipGetLastError();
}

template <class T>
hipError_t wave_eval(const void* Sd, const void* St, int64_t W, const double* pos, WaveShape sh, EvalOut O,
                     int* sing, hipStream_t st) {
  WaveLayout Y = layout(sh.N, sh.nu, sh.nd, sh.A);
  const size_t b = lds_bytes<T>(Y, true);
  hipError_t e = set_lds(wave_eval_kernel<T>, b);
  if (e != hipSuccess) return e;
  wave_eval_kernel<T><<<(unsigned)W, 64, b, st>>>((const DevSys<double>*)Sd, (const DevSys<T>*)St, W, pos, Y, O,
                                                  sing);
  return hipGetLastError();
}

template hipError_t wave_build<double>(const void*, const void*, EnsView<double>, WaveShape, int, int*, hipStream_t);
template hipError_t wave_build<float>(const void*, const void*, EnsView<float>, WaveShape, int, int*, hipStream_t);
template hipError_t wave_sweep<double>(const void*, EnsView<double>, WaveShape, SweepArgs, double*, double*,
                                       hipStream_t);
template hipError_t wave_sweep<float>(const void*, EnsView<float>, WaveShape, SweepArgs, double*, double*,
                                      hipStream_t);
template hipError_t wave_dd<double>(const void*, EnsView<double>, WaveShape, SweepArgs, double*, double*,
                                    hipStream_t);
template hipError_t wave_dd<float>(const void*, EnsView<float>, WaveShape, SweepArgs, double*, double*, hipStream_t);
template hipError_t wave_ae<double>(const void*, const void*, EnsView<double>, WaveShape, SweepArgs, double*,
                                    double*, hipStream_t);
template hipError_t wave_ae<float>(const void*, const void*, EnsView<float>, WaveShape, SweepArgs, double*, double*,
                                   hipStream_t);
template hipError_t wave_eval<double>(const void*, const void*, int64_t, const double*, WaveShape, EvalOut, int*,
                                      hipStream_t);
template hipError_t wave_eval<float>(const void*, const void*, int64_t, const double*, WaveShape, EvalOut, int*,
                                     hipStream_t);

}  // namespace qmc
