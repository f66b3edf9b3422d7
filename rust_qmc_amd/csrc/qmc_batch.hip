// qmc_batch.hip -- the fused persistent kernel of a heterogeneous batch.
// Workgroups take queue items with one atomic increment each until the
// queue is empty, so every wavefront reaches the exit after the last item.
// Items are ordered by the host (most expensive first); each item runs all
// the launch's sweeps of its walkers, exactly as the ensemble's own kernel
// would (same Philox counters, same partial-sum layout).
#include "qmc_batch.hpp"

#include "qmc_tile_dev.hpp"

namespace qmc {

struct BatchDev {
  int nseg;
  BatchSeg seg[BATCH_MAXSEG];
  int64_t start[BATCH_MAXSEG + 1];
  int64_t total;
  int* counter;
  TileArgs TA;
};

__global__ void __launch_bounds__(256, 2) batch_sweep_kernel(BatchDev B) {
  extern __shared__ float lds_f[];
  __shared__ int64_t item;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) item = (int64_t)atomicAdd(B.counter, 1);
    __syncthreads();
    const int64_t t = item;
    __syncthreads();  // every thread has read the item before thread 0 takes the next
    if (t >= B.total) break;
    int si = 0;
    while (si + 1 < B.nseg && t >= B.start[si + 1]) ++si;
    const BatchSeg& G = B.seg[si];
    const int64_t it = t - B.start[si];
    if (G.kind == SEG_TILE_21_21) {
      tile_sweep_block<21, 21, 17, JsNone, float>(*(const DevSys<float>*)G.S, G.E, B.TA, G.A, G.partial, nullptr,
                                                  it, G.nblk, tid, lds_f);
    } else {
      // the ensemble's MO coefficients to the LDS table (all four waves,
      // before any of them may skip), then the register engine's blocks,
      // as the FP32 sweep kernel runs them (CtBasisL)
      const int64_t blk = it * 4 + (tid >> 6);
      if (G.kind == SEG_REG_CH4_STO3G) {
        fill_mo_tab<5, 5>(*(const DevSys<float>*)G.S);
        if (blk < G.nblk)
          sweep_block<5, 5, CtBasisL<BdCH4Sto3g>, JsNone, float>(*(const DevSys<float>*)G.S, G.E, G.A, G.partial,
                                                                 nullptr, blk, G.nblk, tid & 63);
      } else {
        fill_mo_tab<1, 1>(*(const DevSys<float>*)G.S);
        if (blk < G.nblk)
          sweep_block<1, 1, CtBasisL<BdH2>, JsNone, float>(*(const DevSys<float>*)G.S, G.E, G.A, G.partial, nullptr,
                                                           blk, G.nblk, tid & 63);
      }
    }
  }
}

static BatchDev to_dev(const BatchLaunch& B) {
  BatchDev D;
  D.nseg = B.nseg;
  for (int i = 0; i < B.nseg; ++i) D.seg[i] = B.seg[i];
  for (int i = 0; i <= B.nseg; ++i) D.start[i] = B.start[i];
  D.total = B.total;
  D.counter = B.counter;
  D.TA = B.tile.ok ? make_tile_args<17>(B.tile) : TileArgs{};
  return D;
}

size_t batch_lds_bytes(const BatchLaunch& B) {
  return B.tile.ok ? sizeof(float) * (size_t)make_tile_args<17>(B.tile).lds_floats : 0;
}

hipError_t batch_occupancy(size_t lds_bytes, int* blocks_per_cu) {
  hipError_t e = hipFuncSetAttribute((const void*)batch_sweep_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(lds_bytes > 0 ? lds_bytes : 1));
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_sweep_kernel, 256, lds_bytes);
}

hipError_t batch_launch(const BatchLaunch& B, int grid, hipStream_t st) {
  const size_t bytes = batch_lds_bytes(B);
  hipError_t e = hipFuncSetAttribute((const void*)batch_sweep_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(bytes > 0 ? bytes : 1));
  if (e != hipSuccess) return e;
  batch_sweep_kernel<<<grid, 256, bytes, st>>>(to_dev(B));
  return hipGetLastError();
}

}  // namespace qmc
