// qmc_derivs.hpp -- launcher table of the derivative / moment kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmc {

// Dependence of MO coefficients on the nuclear positions (MOs that follow
// the geometry, such as benzene's C-H sigma bonds): pair p names the
// coefficient C[spin][m][ao], where AO ao is value q of shell sh, and its
// entries [e0, e1) give dC/dR_{I,c}.
struct MoDerivPair {
  int spin, m, sh, q, e0, e1;
};
struct MoDerivEntry {
  int I, c;
  double d;
};

struct DerivArgs {
  int64_t W;
  int P, M, Qp;                 // Jastrow parameters, nuclei, padded feature count
  const double* pos_aos;        // [W][N][3] input positions, or null
  const double* pos64;          // ensemble SoA positions (FP64), or null
  const float* pos32;           // ensemble SoA positions (FP32), or null
  int64_t ld;                   // SoA walker stride
  double* Y;                    // [Qp][ldy] feature matrix
  int64_t ldy;
  const MoDerivPair* mdp = nullptr;  // coefficient derivatives (wave kernel only):
  const MoDerivEntry* mde = nullptr; // pairs of spin s are [mdp0[s], mdp0[s+1])
  int mdp0[3] = {0, 0, 0};
};

struct WaveShape;

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
// The derivative matrix Y from one wavefront per walker (any shape the wave
// engine runs; positions from the same sources as the register kernels).
template <class T>
hipError_t wave_derivs(const void* Sd, const void* St, DerivArgs A, const WaveShape& sh, hipStream_t st);
// G[Qp][Qp] (upper triangle valid) = Y Y^T over W walkers; partial holds
// gram_blocks(W) * Qp * Qp doubles.  Y's columns W .. W rounded up to 16
// must exist (ldy >= that) and hold zeros.
int gram_blocks(int64_t W);
hipError_t launch_gram(const double* Y, int64_t ldy, int64_t W, int Qp, double* partial, double* G, hipStream_t st);

}  // namespace qmc
