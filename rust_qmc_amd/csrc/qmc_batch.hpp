// qmc_batch.hpp -- one persistent launch over ensembles of different
// molecules (mixed precision): a work queue of tiles, each taken by one
// 256-thread workgroup, runs the tile engine (benzene-size systems) or four
// register-engine blocks of 64 walkers (CH4 STO-3G, H2) in the same kernel.
#pragma once

#include "qmc_tile.hpp"

namespace qmc {

// segment kinds compiled into the fused kernel
enum : int { SEG_TILE_21_21 = 0, SEG_REG_CH4_STO3G = 1, SEG_REG_H2 = 2 };
constexpr int BATCH_MAXSEG = 8;

struct BatchSeg {
  int kind;
  const void* S;      // DevSys<float> in device memory
  EnsView<float> E;
  SweepArgs A;
  double* partial;    // [n_steps][nblk][3]
  int64_t nblk;       // native blocks: tiles of 16 walkers (tile) or 64 walkers (register)
  int64_t items;      // queue items: tiles, or groups of four register blocks
};

struct BatchLaunch {
  int nseg;
  BatchSeg seg[BATCH_MAXSEG];  // in queue order
  int64_t start[BATCH_MAXSEG + 1];
  int64_t total;
  int* counter;              // queue head, zero at launch
  TilePlan tile;             // plan of the tile segment (at most one tile system per batch)
};

// LDS bytes per workgroup of the fused kernel
size_t batch_lds_bytes(const BatchLaunch& B);
// resident workgroups per CU of the fused kernel at that LDS size
hipError_t batch_occupancy(size_t lds_bytes, int* blocks_per_cu);
// enqueue the fused kernel on st with grid workgroups (the queue head must
// already be zero on st)
hipError_t batch_launch(const BatchLaunch& B, int grid, hipStream_t st);

}  // namespace qmc
