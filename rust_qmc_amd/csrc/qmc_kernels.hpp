// qmc_kernels.hpp -- host-visible views of device buffers and the launcher
// table of the register engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmc {

// Walker ensemble resident in HBM, structure-of-arrays with the walker
// index fastest; ld is the padded walker stride.
template <class T>
struct EnsView {
  int64_t W, ld;
  T* pos;          // [N*3][ld]
  T* ainv;         // [NU*NU + ND*ND][ld], inverse Slater matrices [mo][row]
  double* logdet;  // [ld] ln|D_up D_dn|
  double* sign;    // [ld]
  double* ju;      // [ld] Jastrow exponent ln J
  double* eloc;    // [ld] last local energy
  // live walker count in device memory (a branching population), or null
  // for W; read by the drift-diffusion sweep and the state build
  const int64_t* n_live = nullptr;
};

struct SweepArgs {
  int n_steps;
  uint32_t step0;
  uint64_t seed;
  int64_t walker0;  // Philox walker index of lane 0
  double step_size;  // Gaussian width (Metropolis) or time step tau (drift-diffusion)
  int64_t trace_ld;  // walker stride of the optional E_L trace
  double max_drift;  // drift-diffusion: cap of |tau F| per electron
  // drift-diffusion: when set, step0 is read from here at launch (a
  // replayed graph of DMC steps, whose step lives on the device)
  const uint32_t* step0_dev = nullptr;
};

// Batched evaluation outputs; any pointer may be null.
struct EvalOut {
  double* logabs;  // [W]
  double* sign;    // [W]
  double* grad;    // [W][N][3] grad ln psi
  double* lap;     // [W][N] lap psi / psi
  double* eloc;    // [W]
};

template <class T>
struct DevSys;

struct EngineOps {
  int nu, nd;
  const char* name;                          // "generic" or the compiled shell structure
  bool (*match)(const DevSys<double>&);      // null for the generic shapes
  hipError_t (*build64)(const void*, const void*, EnsView<double>, int, int*, hipStream_t);
  hipError_t (*build32)(const void*, const void*, EnsView<float>, int, int*, hipStream_t);
  hipError_t (*sweep64)(const void*, EnsView<double>, SweepArgs, double*, double*, hipStream_t);
  hipError_t (*sweep32)(const void*, EnsView<float>, SweepArgs, double*, double*, hipStream_t);
  // the same sweeps compiled for a Jastrow without Pade series terms
  hipError_t (*sweep64_ns)(const void*, EnsView<double>, SweepArgs, double*, double*, hipStream_t);
  hipError_t (*sweep32_ns)(const void*, EnsView<float>, SweepArgs, double*, double*, hipStream_t);
  hipError_t (*dd64)(const void*, EnsView<double>, SweepArgs, double*, double*, hipStream_t);
  hipError_t (*dd32)(const void*, EnsView<float>, SweepArgs, double*, double*, hipStream_t);
  // the drift-diffusion sweeps compiled for a Jastrow without Pade series terms
  hipError_t (*dd64_ns)(const void*, EnsView<double>, SweepArgs, double*, double*, hipStream_t);
  hipError_t (*dd32_ns)(const void*, EnsView<float>, SweepArgs, double*, double*, hipStream_t);
  hipError_t (*ae64)(const void*, const void*, EnsView<double>, SweepArgs, double*, double*, hipStream_t);
  hipError_t (*ae32)(const void*, const void*, EnsView<float>, SweepArgs, double*, double*, hipStream_t);
  hipError_t (*eval64)(const void*, const void*, int64_t, const double*, EvalOut, int*, hipStream_t);
  hipError_t (*eval32)(const void*, const void*, int64_t, const double*, EvalOut, int*, hipStream_t);
};

// the register-engine kernels for a system: a compiled shell structure when
// one matches (unless generic_only), else the generic kernels of its
// electron-count shape; null when neither exists
const EngineOps* find_engine(const DevSys<double>& D, bool generic_only);
hipError_t launch_init64(const void* Sd, EnsView<double> E, uint64_t seed, int64_t w0, hipStream_t st);
hipError_t launch_init32(const void* Sd, EnsView<float> E, uint64_t seed, int64_t w0, hipStream_t st);
hipError_t launch_reduce(const double* partial, int nsteps, int nblk, int nq, double* out, hipStream_t st);

}  // namespace qmc
