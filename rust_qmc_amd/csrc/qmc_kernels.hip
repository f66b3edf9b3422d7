// qmc_kernels.hip -- gfx950 kernels of the walker engine and their
// launchers.  Walker state lives in HBM between launches as
// structure-of-arrays with the walker index fastest (coalesced 64-lane
// loads), and in VGPRs inside a launch.
#include "qmc_kernels.hpp"

#include "qmc_engine.hpp"

namespace qmc {

// ---------------------------------------------------------------------------
// initial positions: electron e of walker w = nuc[init_nuc[e]] + sigma z,
// z from Philox counter (w, 0, e, INIT), Box-Muller in FP64
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(64) init_positions_kernel(const DevSys<double>* __restrict__ Sp, T* pos,
                                                            int64_t W, int64_t ld, uint64_t seed,
                                                            int64_t w0) {
  const DevSys<double>& S = *Sp;
  const int64_t w = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (w >= W) return;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int e = 0; e < S.n_elec; ++e) {
    double z[3];
    Rng<double>::normals3(philox((uint32_t)(w0 + w), 0u, (uint32_t)e, RNG_INIT, k0, k1), z);
    const int c = S.init_nuc[e];
#pragma unroll
    for (int q = 0; q < 3; ++q) pos[(e * 3 + q) * ld + w] = (T)(S.nuc[c][q] + S.init_sigma * z[q]);
  }
}

// ---------------------------------------------------------------------------
// state build / refresh: Slater matrices and inverses in FP64 from the
// current positions, ln|D|, sign, Jastrow exponent; optionally E_L.
// ---------------------------------------------------------------------------
template <int NU, int ND, class Bs, class T>
__global__ void __launch_bounds__(64) build_state_kernel(const DevSys<double>* __restrict__ Sd,
                                                         const DevSys<T>* __restrict__ St, EnsView<T> E,
                                                         int want_eloc, int* singular_count) {
  const int64_t w = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (w >= (E.n_live ? *E.n_live : E.W)) return;
  WalkerRegs<NU, ND, T> R;
  load_positions<NU, ND>(*St, E.pos, E.ld, w, R);
  double lu, su, ldn, sd;
  bool s1 = sector_inverse<NU, T, Bs>(*Sd, 0, R.ru, R.au, lu, su);
  bool s2 = sector_inverse<ND, T, Bs>(*Sd, 1, R.rd, R.ad, ldn, sd);
  if ((s1 || s2) && singular_count) atomicAdd(singular_count, 1);
  store_inverse<NU, ND>(E.ainv, E.ld, w, R);
  E.logdet[w] = lu + ldn;
  E.sign[w] = su * sd;
  E.ju[w] = jastrow_u<NU, ND, T, Bs>(*Sd, R.ru, R.rd);
  if (want_eloc) E.eloc[w] = local_energy<NU, ND, T, Bs>(*St, R.ru, R.au, R.rd, R.ad, nullptr, nullptr, 0);
}

// ---------------------------------------------------------------------------
// persistent single-electron sweeps (sweep_block in qmc_engine.hpp)
// ---------------------------------------------------------------------------
template <int NU, int ND, class Bs, class Js, class T>
__global__ void __launch_bounds__(64) sweep_kernel(const DevSys<T>* __restrict__ Sp, EnsView<T> E,
                                                   SweepArgs A, double* __restrict__ partial,
                                                   double* __restrict__ el_trace) {
  if constexpr (Bs::lds_mo) {
    static_assert(Bs::n_ao <= MO_TAB_AO, "too many AOs for the LDS MO table");
    fill_mo_tab<NU, ND>(*Sp);
  }
  sweep_block<NU, ND, Bs, Js, T>(*Sp, E, A, partial, el_trace, blockIdx.x, gridDim.x, threadIdx.x);
}

// drift-diffusion sweeps (dd_sweep_block in qmc_engine.hpp)
template <int NU, int ND, class Bs, class Js, class T>
__global__ void __launch_bounds__(64) dd_sweep_kernel(const DevSys<T>* __restrict__ Sp, EnsView<T> E,
                                                      SweepArgs A, double* __restrict__ partial,
                                                      double* __restrict__ el_trace) {
  dd_sweep_block<NU, ND, Bs, Js, T>(*Sp, E, A, partial, el_trace, blockIdx.x, gridDim.x, threadIdx.x);
}

// ---------------------------------------------------------------------------
// all-electron Metropolis (MCMCSimulation::metropolis_step, vmc.rs:95-119):
// every electron moved at once, full recomputation of ln|psi|, E_L
// recomputed on acceptance only and cached otherwise.
// ---------------------------------------------------------------------------
template <int NU, int ND, class Bs, class T>
__global__ void __launch_bounds__(64) all_electron_kernel(const DevSys<double>* __restrict__ Sd,
                                                          const DevSys<T>* __restrict__ Sp, EnsView<T> E,
                                                          SweepArgs A, double* __restrict__ partial,
                                                          double* __restrict__ el_trace) {
  const DevSys<T>& S = *Sp;
  const int64_t w = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const bool live = w < E.W;
  const int64_t wl = live ? w : E.W - 1;
  WalkerRegs<NU, ND, T> R, Q;
  load_positions<NU, ND>(S, E.pos, E.ld, wl, R);
  double ldet = E.logdet[wl], ju = E.ju[wl], sgn = E.sign[wl], el = E.eloc[wl];
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const uint32_t wid = (uint32_t)(A.walker0 + wl);
  const T sigma = (T)A.step_size;
  const uint32_t N = (uint32_t)S.n_elec;
  for (int s = 0; s < A.n_steps; ++s) {
    const uint32_t step = A.step0 + (uint32_t)s;
#pragma unroll
    for (int k = 0; k < NU; ++k) {
      T z[3];
      Rng<T>::normals3(philox(wid, step, (uint32_t)S.sector_elec[0][k], RNG_ALL, k0, k1), z);
#pragma unroll
      for (int c = 0; c < 3; ++c) Q.ru[k][c] = R.ru[k][c] + sigma * z[c];
    }
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      T z[3];
      Rng<T>::normals3(philox(wid, step, (uint32_t)S.sector_elec[1][k], RNG_ALL, k0, k1), z);
#pragma unroll
      for (int c = 0; c < 3; ++c) Q.rd[k][c] = R.rd[k][c] + sigma * z[c];
    }
    double lu, su, ldn, sd;
    sector_inverse<NU, T, Bs>(*Sd, 0, Q.ru, Q.au, lu, su);
    sector_inverse<ND, T, Bs>(*Sd, 1, Q.rd, Q.ad, ldn, sd);
    const double jun = jastrow_u<NU, ND, T, Bs>(*Sd, Q.ru, Q.rd);
    const double U = u01d(philox(wid, step, N, RNG_ALL, k0, k1).x);
    int acc = 0;
    if (U < exp(2.0 * ((lu + ldn + jun) - (ldet + ju)))) {
      // only the positions persist; the inverses serve this E_L and go
#pragma unroll
      for (int k = 0; k < NU; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) R.ru[k][c] = Q.ru[k][c];
#pragma unroll
      for (int k = 0; k < ND; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) R.rd[k][c] = Q.rd[k][c];
      ldet = lu + ldn;
      ju = jun;
      sgn = su * sd;
      el = local_energy<NU, ND, T, Bs>(S, Q.ru, Q.au, Q.rd, Q.ad, nullptr, nullptr, 0);
      acc = 1;
    }
    if (el_trace && live) el_trace[(int64_t)s * A.trace_ld + w] = el;
    double e1 = live ? el : 0.0, e2 = live ? el * el : 0.0, a = live ? (double)acc : 0.0;
    e1 = wave_sum(e1);
    e2 = wave_sum(e2);
    a = wave_sum(a);
    if (threadIdx.x == 0) {
      double* p = partial + ((int64_t)s * gridDim.x + blockIdx.x) * 3;
      p[0] = e1;
      p[1] = e2;
      p[2] = a;
    }
  }
  if (live) {
    store_positions<NU, ND>(S, E.pos, E.ld, w, R);
    E.logdet[w] = ldet;
    E.ju[w] = ju;
    E.sign[w] = sgn;
    E.eloc[w] = el;
  }
}

// ---------------------------------------------------------------------------
// batched per-configuration evaluation of the trait methods.  Input
// positions are [W][N][3] FP64 in global electron order (the reference's
// Vec<Vector3<f64>> layout); outputs are written walker-major.
// ---------------------------------------------------------------------------
template <int NU, int ND, class Bs, class T>
__global__ void __launch_bounds__(64) eval_kernel(const DevSys<double>* __restrict__ Sd,
                                                  const DevSys<T>* __restrict__ St, int64_t W,
                                                  const double* __restrict__ pos, EvalOut O, int* singular_count) {
  const int64_t w = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (w >= W) return;
  const DevSys<T>& S = *St;
  const int N = NU + ND;
  // the determinants, inverses and ln J come from the FP64 positions; the
  // local-energy terms then run in T
  WalkerRegs<NU, ND, double> D;
#pragma unroll
  for (int k = 0; k < NU; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) D.ru[k][c] = pos[(w * N + S.sector_elec[0][k]) * 3 + c];
#pragma unroll
  for (int k = 0; k < ND; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) D.rd[k][c] = pos[(w * N + S.sector_elec[1][k]) * 3 + c];
  double lu, su, ldn, sd;
  bool s1 = sector_inverse<NU, double, Bs>(*Sd, 0, D.ru, D.au, lu, su);
  bool s2 = sector_inverse<ND, double, Bs>(*Sd, 1, D.rd, D.ad, ldn, sd);
  if ((s1 || s2) && singular_count) atomicAdd(singular_count, 1);
  if (O.logabs) O.logabs[w] = lu + ldn + jastrow_u<NU, ND, double, Bs>(*Sd, D.ru, D.rd);
  if (O.sign) O.sign[w] = su * sd;
  WalkerRegs<NU, ND, T> R;
#pragma unroll
  for (int k = 0; k < Cap<NU>::v; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) R.ru[k][c] = (T)D.ru[k][c];
#pragma unroll
    for (int m = 0; m < Cap<NU>::v; ++m) R.au[k][m] = (T)D.au[k][m];
  }
#pragma unroll
  for (int k = 0; k < Cap<ND>::v; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) R.rd[k][c] = (T)D.rd[k][c];
#pragma unroll
    for (int m = 0; m < Cap<ND>::v; ++m) R.ad[k][m] = (T)D.ad[k][m];
  }
  if (O.eloc || O.grad || O.lap) {
    // per-electron terms land in private arrays in global electron order
    const bool per_e = O.grad || O.lap;
    double g[3 * (N > 0 ? N : 1)], l[N > 0 ? N : 1];
    double el = local_energy<NU, ND, T, Bs>(S, R.ru, R.au, R.rd, R.ad, per_e ? g : nullptr, per_e ? l : nullptr, 1);
    if (O.eloc) O.eloc[w] = el;
    if (O.grad)
#pragma unroll
      for (int i = 0; i < 3 * N; ++i) O.grad[w * 3 * N + i] = g[i];
    if (O.lap)
#pragma unroll
      for (int i = 0; i < N; ++i) O.lap[w * N + i] = l[i];
  }
}

// deterministic reduction of per-block partials: out[s][q] = sum_b partial[s][b][q]
__global__ void __launch_bounds__(256) reduce_partials_kernel(const double* __restrict__ partial, int nblk,
                                                              int nq, double* __restrict__ out) {
  __shared__ double sh[256];
  const int s = blockIdx.x;
  for (int q = 0; q < nq; ++q) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) v += partial[((int64_t)s * nblk + b) * nq + q];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(int64_t)s * nq + q] = sh[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblocks(int64_t W) { return (unsigned)((W + 63) / 64); }

template <class T>
static hipError_t launch_init(const void* Sd, EnsView<T> E, uint64_t seed, int64_t w0, hipStream_t st) {
  init_positions_kernel<T><<<nblocks(E.W), 64, 0, st>>>((const DevSys<double>*)Sd, E.pos, E.W, E.ld, seed, w0);
  return hipGetLastError();
}

template <int NU, int ND, class Bs, class T>
struct Launch {
  static hipError_t build(const void* Sd, const void* St, EnsView<T> E, int want_eloc, int* sing,
                          hipStream_t st) {
    build_state_kernel<NU, ND, Bs, T><<<nblocks(E.W), 64, 0, st>>>((const DevSys<double>*)Sd,
                                                                  (const DevSys<T>*)St, E, want_eloc, sing);
    return hipGetLastError();
  }
  template <class Js>
  static hipError_t sweep(const void* St, EnsView<T> E, SweepArgs A, double* partial, double* trace,
                          hipStream_t st) {
    sweep_kernel<NU, ND, Bs, Js, T><<<nblocks(E.W), 64, 0, st>>>((const DevSys<T>*)St, E, A, partial, trace);
    return hipGetLastError();
  }
  template <class Js>
  static hipError_t dd(const void* St, EnsView<T> E, SweepArgs A, double* partial, double* trace,
                       hipStream_t st) {
    dd_sweep_kernel<NU, ND, Bs, Js, T><<<nblocks(E.W), 64, 0, st>>>((const DevSys<T>*)St, E, A, partial, trace);
    return hipGetLastError();
  }
  static hipError_t all_electron(const void* Sd, const void* St, EnsView<T> E, SweepArgs A, double* partial,
                                 double* trace, hipStream_t st) {
    all_electron_kernel<NU, ND, Bs, T><<<nblocks(E.W), 64, 0, st>>>((const DevSys<double>*)Sd,
                                                                   (const DevSys<T>*)St, E, A, partial, trace);
    return hipGetLastError();
  }
  static hipError_t eval(const void* Sd, const void* St, int64_t W, const double* pos, EvalOut O, int* sing,
                         hipStream_t st) {
    eval_kernel<NU, ND, Bs, T><<<nblocks(W), 64, 0, st>>>((const DevSys<double>*)Sd, (const DevSys<T>*)St, W,
                                                         pos, O, sing);
    return hipGetLastError();
  }
};

template <class B>
struct SweepBasis32 {
  using type = B;
};
template <class Bd>
struct SweepBasis32<CtBasis<Bd>> {
  using type = CtBasisL<Bd>;
};

template <int NU, int ND, class Bs, class JsFast>
static EngineOps make_ops(const char* name, bool (*match)(const DevSys<double>&)) {
  EngineOps o;
  o.nu = NU;
  o.nd = ND;
  o.name = name;
  o.match = match;
  o.build64 = &Launch<NU, ND, Bs, double>::build;
  o.build32 = &Launch<NU, ND, Bs, float>::build;
  // FP32 sweeps of compiled shell structures without Pade series take the
  // MO coefficients from LDS (CtBasisL): 4 % faster on CH4 STO-3G; the FP64
  // kernel, whose coefficients take twice the registers, and the series
  // kernels, which then spill to scratch, were slower that way
  using Bs32 = typename SweepBasis32<Bs>::type;
  o.sweep64 = &Launch<NU, ND, Bs, double>::template sweep<JsRt>;
  o.sweep32 = &Launch<NU, ND, Bs, float>::template sweep<JsRt>;
  o.sweep64_ns = &Launch<NU, ND, Bs, double>::template sweep<JsFast>;
  o.sweep32_ns = &Launch<NU, ND, Bs32, float>::template sweep<JsFast>;
  // All-electron moves always run the run-time-loop kernels: the CH4/STO-3G
  // all-electron kernel specialised to its shell structure disagreed with
  // them (and with the oracle) from the first step while every other
  // specialised kernel agrees, and no cause has been found in the source.
  o.ae64 = &Launch<NU, ND, RtBasis, double>::all_electron;
  o.ae32 = &Launch<NU, ND, RtBasis, float>::all_electron;
  o.dd64 = &Launch<NU, ND, Bs, double>::template dd<JsRt>;
  o.dd32 = &Launch<NU, ND, Bs, float>::template dd<JsRt>;
  o.dd64_ns = &Launch<NU, ND, Bs, double>::template dd<JsFast>;
  o.dd32_ns = &Launch<NU, ND, Bs, float>::template dd<JsFast>;
  o.eval64 = &Launch<NU, ND, Bs, double>::eval;
  o.eval32 = &Launch<NU, ND, Bs, float>::eval;
  return o;
}

template <int NU, int ND>
static EngineOps generic_ops() {
  return make_ops<NU, ND, RtBasis, JsRt>("generic", nullptr);
}

template <class Bd>
static EngineOps compiled_ops() {
  return make_ops<Bd::nu, Bd::nd, CtBasis<Bd>, JsNone>(Bd::name, &basis_matches<Bd, double>);
}

// Shell structures compiled into the register engine (tried first), then
// the electron-count shapes with run-time shell loops.
static const EngineOps kOps[] = {
    compiled_ops<BdH2>(),       compiled_ops<BdLiSeeded>(), compiled_ops<BdLiRef>(),
    compiled_ops<BdCH4Sto3g>(), compiled_ops<BdCH4631g>(),  generic_ops<1, 0>(),
    generic_ops<1, 1>(),        generic_ops<2, 1>(),        generic_ops<2, 2>(),
    generic_ops<3, 2>(),        generic_ops<3, 3>(),        generic_ops<4, 4>(),
    generic_ops<5, 5>(),
};

const EngineOps* find_engine(const DevSys<double>& D, bool generic_only) {
  for (const auto& o : kOps) {
    if (o.nu != D.n_up || o.nd != D.n_dn) continue;
    if (o.match && (generic_only || !o.match(D))) continue;
    return &o;
  }
  return nullptr;
}

hipError_t launch_init64(const void* Sd, EnsView<double> E, uint64_t seed, int64_t w0, hipStream_t st) {
  return launch_init<double>(Sd, E, seed, w0, st);
}
hipError_t launch_init32(const void* Sd, EnsView<float> E, uint64_t seed, int64_t w0, hipStream_t st) {
  return launch_init<float>(Sd, E, seed, w0, st);
}

hipError_t launch_reduce(const double* partial, int nsteps, int nblk, int nq, double* out, hipStream_t st) {
  reduce_partials_kernel<<<nsteps, 256, 0, st>>>(partial, nblk, nq, out);
  return hipGetLastError();
}

}  // namespace qmc
