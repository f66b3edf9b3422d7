// qmc_noalloc.hpp -- the fixed device scratch array behind the
// allocation-free entry point, and the kernel that unpacks a derivative
// feature matrix into the caller's walker-major outputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmc {

// Capacity of the scratch array in doubles (64 MiB per GPU).
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = 8ll << 20;

// Device address of the current device's scratch array.  The first call
// on a device resolves it with hipGetSymbolAddress (that device current)
// and keeps it; later calls return the kept address.
hipError_t noalloc_scratch(double** ptr);

// Outputs of the unpack kernel; any pointer may be null.
struct DerivUnpack {
  int64_t W;
  int P, M;
  const double* Y;  // [Qp][W] feature matrix, walker fastest
  double* eloc;     // [W]
  double* O;        // [W][P]
  double* dE;       // [W][P]
  double* F;        // [W][M][3], electron part + fnn
  double* Om;       // [W][M][3]
  double fnn[48];   // nucleus-nucleus Hellmann-Feynman force [M][3]
};

hipError_t launch_deriv_unpack(const DerivUnpack& U, hipStream_t st);

}  // namespace qmc
