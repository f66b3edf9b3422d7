// qmc_tile.hpp -- launchers of the tile engine (a 256-thread workgroup owns
// a tile of 16 walkers; MFMA MO contraction), compiled for benzene-size
// spin sectors in mixed precision.
#pragma once

#include "qmc_kernels.hpp"

namespace qmc {

constexpr int TILE_W = 16;   // walkers per tile (one workgroup)
constexpr int TILE_MAXSH = 4;  // shells per lane in the AO phase
// Jastrow partners of a moved electron on the 16 lanes of its walker:
// electron p = g + 16 t on lane g for t < TILE_EE, then nucleus g
constexpr int TILE_EE = 3;

template <class T>
struct DevSys;

// what a tile launch needs to know about the system, prepared on the host
struct TilePlan {
  bool ok = false;  // a tile kernel is compiled for this system
  int N = 0, nu = 0, nd = 0, A = 0, n_nuc = 0, n_shell = 0;
  bool no_series = true;
  int lane_sh[16][TILE_MAXSH];  // shells of group lane g (-1: none)
  // every shell a contracted Gaussian of kind s, p or sp: the AO values of
  // a move take the branch-free path, slot j looping over slot_np[j]
  // primitives (the slot's largest count; records are zero-padded)
  bool gauss = false;
  int n_slot = 0, slot_np[TILE_MAXSH] = {};
};

// fills P; false when no tile kernel fits the system (21 + 21 electrons,
// at most 68 AOs, at most 48 electrons and 16 nuclei, at most 64 shells of
// at most 6 primitives)
bool tile_plan(const DevSys<float>& D, TilePlan& P);

// single-electron sweeps of an ensemble with FP32 state (mixed precision,
// St the FP32 system) or FP64 state (St the FP64 system); the FP64 inverse
// refresh and the state build keep the wave engine's HBM layout
template <class T>
hipError_t tile_sweep(const void* St, EnsView<T> E, const TilePlan& P, SweepArgs A, double* partial, double* trace,
                      hipStream_t st);

// drift-diffusion sweeps (A.step_size the time step, A.max_drift the drift
// cap) of an ensemble without a branching population, as tile_sweep
template <class T>
hipError_t tile_dd(const void* St, EnsView<T> E, const TilePlan& P, SweepArgs A, double* partial, double* trace,
                   hipStream_t st);

inline int64_t tile_blocks(int64_t W) { return (W + TILE_W - 1) / TILE_W; }

}  // namespace qmc
