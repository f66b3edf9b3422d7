// qmc_tile.hip -- the tile engine: one 256-thread workgroup (four
// wavefronts) owns a tile of 16 walkers, 16 lanes per walker.  Built for
// spin sectors too large for the register engine (benzene: 21 + 21
// electrons, 66 AOs), with FP32 (mixed precision) or FP64 walker state.
//
// Per walker, lane g of its 16-lane group keeps columns g and g + 16 of the
// moving sector's inverse Slater matrix in VGPRs (the other sector's rest
// in LDS); positions live in LDS.  All 16 walkers of the tile move the same
// electron together, so one single-electron move is
//   A  proposal from the sector's random numbers, drawn ahead into LDS
//   B  AO values of the 16 proposals, shells spread over the group's lanes,
//      into the LDS matrix X[walker][ao]
//   C  MO values Phi = X C for all 16 walkers at once with
//      v_mfma_f32_16x16x4f32: one wavefront per 16-MO column tile, the
//      coefficients held as the B operand in VGPRs
//   D  per walker: v_j = sum_m Phi_m Ainv[m][j] on the column owners, the
//      ratio R = v_k, the Jastrow difference over partners spread on the
//      lanes (16-lane butterfly sums), the Metropolis test and the
//      Sherman-Morrison update of the owned columns.
// The local energy after each sweep puts the electron of sector slot k on
// lane k mod 16, the owner of its inverse column.
#include <algorithm>

#include "qmc_tile_dev.hpp"

namespace qmc {

// FP32 holds two workgroups per CU (208 VGPRs, no spills; 8,192 walkers
// make exactly two tiles per CU); FP64, whose inverse columns take twice
// the registers, one.
template <class T>
struct TileOcc {
  static constexpr int v = sizeof(T) == 4 ? 2 : 1;
};

template <int NU, int ND, int KS, class Js, class T>
__global__ void __launch_bounds__(TILE_T, TileOcc<T>::v)
    tile_sweep_kernel(const DevSys<T>* __restrict__ Sp, EnsView<T> E, TileArgs TA, SweepArgs A,
                      double* __restrict__ partial, double* __restrict__ el_trace) {
  extern __shared__ double lds_d[];
  tile_sweep_block<NU, ND, KS, Js, T>(*Sp, E, TA, A, partial, el_trace, blockIdx.x, gridDim.x, threadIdx.x,
                                      (T*)lds_d);
}

// drift-diffusion sweeps of a tile: the Metropolis kernel's layout and
// MO contraction, with the moved electron's AO weights staged per walker
// in LDS and the quantum force at both ends of the move
template <class T>
struct TileDdOcc {
  static constexpr int v = sizeof(T) == 4 ? 2 : 1;
};

template <int NU, int ND, int KS, class Js, class T>
__global__ void __launch_bounds__(TILE_T, TileDdOcc<T>::v)
    tile_dd_kernel(const DevSys<T>* __restrict__ Sp, EnsView<T> E, TileArgs TA, SweepArgs A,
                   double* __restrict__ partial, double* __restrict__ el_trace) {
  extern __shared__ double lds_d[];
  tile_sweep_block<NU, ND, KS, Js, T, true>(*Sp, E, TA, A, partial, el_trace, blockIdx.x, gridDim.x, threadIdx.x,
                                            (T*)lds_d);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool tile_plan(const DevSys<float>& D, TilePlan& P) {
  P.ok = false;
  P.N = D.n_elec;
  P.nu = D.n_up;
  P.nd = D.n_dn;
  P.A = D.n_ao;
  P.n_nuc = D.n_nuc;
  P.no_series = D.k_ee == 0 && D.k_en == 0;
  // compiled shape: 21 + 21 electrons, up to 68 AOs (benzene)
  if (P.nu != 21 || P.nd != 21 || P.A > 68 || P.N > 16 * TILE_EE || P.n_nuc > 16 || D.n_shell > 16 * TILE_MAXSH)
    return false;
  for (int i = 0; i < D.n_shell; ++i)
    if (D.sh_nprim[i] > TILE_MAXP) return false;
  P.n_shell = D.n_shell;
  for (int g = 0; g < 16; ++g)
    for (int j = 0; j < TILE_MAXSH; ++j) P.lane_sh[g][j] = -1;
  // shells spread over the 16 lanes of a group by slots: the lanes
  // evaluate slot j together, so a slot costs its largest primitive count.
  // Slot j takes shells 16 j .. 16 j + 15 of the descending primitive-count
  // order (benzene 6-31G: six 6-primitive and ten 3-primitive shells, then
  // two 3-primitive and twelve 1-primitive ones: 6 + 3 primitive passes),
  // in alternating lane order so that the lanes' totals stay even.
  int order[MAXSH];
  for (int i = 0; i < D.n_shell; ++i) order[i] = i;
  for (int i = 1; i < D.n_shell; ++i)  // stable insertion sort, descending primitive count
    for (int j = i; j > 0 && D.sh_nprim[order[j]] > D.sh_nprim[order[j - 1]]; --j) {
      const int t = order[j];
      order[j] = order[j - 1];
      order[j - 1] = t;
    }
  for (int i = 0; i < D.n_shell; ++i) {
    const int slot = i / 16, r = i % 16;
    if (slot >= TILE_MAXSH) return false;
    P.lane_sh[slot & 1 ? 15 - r : r][slot] = order[i];
    P.slot_np[slot] = std::max(P.slot_np[slot], D.sh_nprim[order[i]]);
    P.n_slot = slot + 1;
  }
  P.gauss = true;
  for (int i = 0; i < D.n_shell; ++i)
    if (D.sh_kind[i] != SH_S && D.sh_kind[i] != SH_SP && D.sh_kind[i] != SH_P) P.gauss = false;
  P.ok = true;
  return true;
}

template <int NU, int ND, int KS, class Js, class T>
static hipError_t launch_tile(const void* St, EnsView<T> E, const TilePlan& P, SweepArgs A, double* partial,
                              double* trace, hipStream_t st) {
  const TileArgs TA = make_tile_args<KS>(P, (int)sizeof(T));
  const size_t bytes = sizeof(T) * (size_t)TA.lds_floats;
  auto k = tile_sweep_kernel<NU, ND, KS, Js, T>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  k<<<(unsigned)tile_blocks(E.W), TILE_T, bytes, st>>>((const DevSys<T>*)St, E, TA, A, partial, trace);
  return hipGetLastError();
}

template <class T>
hipError_t tile_sweep(const void* St, EnsView<T> E, const TilePlan& P, SweepArgs A, double* partial, double* trace,
                      hipStream_t st) {
  if (!P.ok) return hipErrorInvalidValue;
  if (P.no_series) return launch_tile<21, 21, 17, JsNone, T>(St, E, P, A, partial, trace, st);
  return launch_tile<21, 21, 17, JsRt, T>(St, E, P, A, partial, trace, st);
}

template <int NU, int ND, int KS, class Js, class T>
static hipError_t launch_tile_dd(const void* St, EnsView<T> E, const TilePlan& P, SweepArgs A, double* partial,
                                 double* trace, hipStream_t st) {
  const TileArgs TA = make_tile_args<KS>(P, (int)sizeof(T), true);
  const size_t bytes = sizeof(T) * (size_t)TA.lds_floats;
  auto k = tile_dd_kernel<NU, ND, KS, Js, T>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  k<<<(unsigned)tile_blocks(E.W), TILE_T, bytes, st>>>((const DevSys<T>*)St, E, TA, A, partial, trace);
  return hipGetLastError();
}

template <class T>
hipError_t tile_dd(const void* St, EnsView<T> E, const TilePlan& P, SweepArgs A, double* partial, double* trace,
                   hipStream_t st) {
  if (!P.ok || E.n_live || !(A.step_size > 0.0)) return hipErrorInvalidValue;
  if (P.no_series) return launch_tile_dd<21, 21, 17, JsNone, T>(St, E, P, A, partial, trace, st);
  return launch_tile_dd<21, 21, 17, JsRt, T>(St, E, P, A, partial, trace, st);
}

template hipError_t tile_sweep<float>(const void*, EnsView<float>, const TilePlan&, SweepArgs, double*, double*,
                                      hipStream_t);
template hipError_t tile_dd<float>(const void*, EnsView<float>, const TilePlan&, SweepArgs, double*, double*,
                                   hipStream_t);
template hipError_t tile_dd<double>(const void*, EnsView<double>, const TilePlan&, SweepArgs, double*, double*,
                                    hipStream_t);
template hipError_t tile_sweep<double>(const void*, EnsView<double>, const TilePlan&, SweepArgs, double*, double*,
                                       hipStream_t);

}  // namespace qmc
