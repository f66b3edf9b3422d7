// qmc_wave.hpp -- launchers of the wave engine (one wavefront per walker,
// walker state in LDS; any electron count up to 32 per spin).
#pragma once

#include "qmc_kernels.hpp"

namespace qmc {

struct WaveShape {
  int N, nu, nd, A;
};

template <class T>
hipError_t wave_build(const void* Sd, const void* St, EnsView<T> E, WaveShape sh, int want_eloc, int* sing,
                      hipStream_t st);
template <class T>
hipError_t wave_sweep(const void* St, EnsView<T> E, WaveShape sh, SweepArgs A, double* partial, double* trace,
                      hipStream_t st);
// drift-diffusion sweeps; blocks past *E.n_live (when set) write zero partials
template <class T>
hipError_t wave_dd(const void* St, EnsView<T> E, WaveShape sh, SweepArgs A, double* partial, double* trace,
                   hipStream_t st);
// all-electron Metropolis steps (positions, ln|psi| and E_L persist; the
// inverses must be rebuilt before single-electron sweeps)
template <class T>
hipError_t wave_ae(const void* Sd, const void* St, EnsView<T> E, WaveShape sh, SweepArgs A, double* partial,
                   double* trace, hipStream_t st);
template <class T>
hipError_t wave_eval(const void* Sd, const void* St, int64_t W, const double* pos, WaveShape sh, EvalOut O,
                     int* sing, hipStream_t st);

}  // namespace qmc
