// qmc_derivs.hpp -- launcher table of the derivative / moment kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmc {

struct DerivArgs {
  int64_t W;
  int P, M, Qp;                 // Jastrow parameters, nuclei, padded feature count
  const double* pos_aos;        // [W][N][3] input positions, or null
  const double* pos64;          // ensemble SoA positions (FP64), or null
  const float* pos32;           // ensemble SoA positions (FP32), or null
  int64_t ld;                   // SoA walker stride
  double* Y;                    // [Qp][ldy] feature matrix
  int64_t ldy;
};

template <class T>
struct DevSys;

struct DerivOps {
  int nu, nd;
  bool (*match)(const DevSys<double>&);  // compiled shell structure, or null (generic)
  hipError_t (*derivs64)(const void*, const void*, DerivArgs, hipStream_t);
  hipError_t (*derivs32)(const void*, const void*, DerivArgs, hipStream_t);
};

// derivative kernels for a system: its compiled shell structure when one
// matches (unless generic_only), else the generic kernels of its shape
const DerivOps* find_derivs(const DevSys<double>& D, bool generic_only);
// G[Qp][Qp] (upper triangle valid) = Y Y^T over W walkers; partial holds
// ceil(W/chunk) * Qp * Qp doubles.
hipError_t launch_gram(const double* Y, int64_t ldy, int64_t W, int Qp, double* partial, int chunk, double* G,
                       hipStream_t st);

}  // namespace qmc
