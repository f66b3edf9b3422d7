// qmc_refresh.hpp -- the FP64 state refresh of benzene-size ensembles by
// tiles of 16 walkers (Slater matrices on FP64 matrix cores, Gauss-Jordan
// with 16 lanes per walker).
#pragma once

#include "qmc_tile.hpp"

namespace qmc {

// Rebuilds, from the positions, every walker's inverse Slater matrices (in
// the ensemble's precision T), ln|D_up D_dn|, its sign and ln J, in FP64:
// what the wave engine's build does when no local energy is wanted.  Needs
// a tile plan (21 + 21 electrons, at most 68 AOs), no branching population
// (E.n_live null) and a work buffer of tile_refresh_work(E.W) doubles.
// singular_count (may be null) counts walkers with a singular sector; their
// inverses become the identity.
size_t tile_refresh_work(int64_t W);
template <class T>
hipError_t tile_refresh(const void* Sd, EnsView<T> E, const TilePlan& P, double* work, int* singular_count,
                        hipStream_t st);

}  // namespace qmc
