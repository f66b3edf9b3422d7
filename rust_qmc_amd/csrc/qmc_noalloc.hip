// qmc_noalloc.hip -- device scratch for the allocation-free derivative
// entry point (qmc_derivatives_noalloc): a file-scope device array that
// every GPU running the library holds until the process ends, and the
// kernel that turns the feature matrix it holds into walker-major outputs.
#include <mutex>

#include "qmc_noalloc.hpp"

namespace qmc {

// The feature matrix Y of qmc_derivatives_noalloc.  In host code this name
// is a host-side stand-in, not a device address: kernels and copies take
// the address noalloc_scratch() resolved for the current device.
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(double** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = (double*)p;
  }
  *ptr = g_ptr[dev];
  return hipSuccess;
}

// One thread per walker; the feature rows follow the derivative kernels'
// column map: O_k at k, Omega_{I,c} at P + 3I + c, E_L at P + 3M,
// dE_L/dp_k at P + 3M + 2 + k, the electron part of F_HF at
// 2P + 3M + 2 + 3I + c.
__global__ void __launch_bounds__(256) deriv_unpack_kernel(DerivUnpack U) {
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (w >= U.W) return;
  const int64_t W = U.W;
  const int P = U.P, M3 = 3 * U.M;
  const double* Y = U.Y;
  if (U.eloc) U.eloc[w] = Y[(int64_t)(P + M3) * W + w];
  if (U.O)
    for (int k = 0; k < P; ++k) U.O[w * P + k] = Y[(int64_t)k * W + w];
  if (U.dE)
    for (int k = 0; k < P; ++k) U.dE[w * P + k] = Y[(int64_t)(P + M3 + 2 + k) * W + w];
  if (U.Om)
    for (int q = 0; q < M3; ++q) U.Om[w * M3 + q] = Y[(int64_t)(P + q) * W + w];
  if (U.F)
    for (int q = 0; q < M3; ++q) U.F[w * M3 + q] = Y[(int64_t)(2 * P + M3 + 2 + q) * W + w] + U.fnn[q];
}

hipError_t launch_deriv_unpack(const DerivUnpack& U, hipStream_t st) {
  if (U.W <= 0) return hipSuccess;
  deriv_unpack_kernel<<<(unsigned)((U.W + 255) / 256), 256, 0, st>>>(U);
  return hipGetLastError();
}

}  // namespace qmc
