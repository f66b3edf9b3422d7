// qmc_dmc.hpp -- device side of the importance-sampled DMC population
// (ImportanceSampledDMC, is_dmc.rs:176-607): branching, compaction of the
// surviving and cloned walkers into the other population buffer, and the
// per-step population statistics.  Everything a step needs stays on the
// device (live counts, E_ref, force sums), so steps are enqueued without a
// host round trip.
#pragma once

#include "qmc_kernels.hpp"
#include "qmc_sys.hpp"

namespace qmc {

constexpr int DMC_BLOCK = 256;        // walkers per branching workgroup
constexpr int DMC_MAX_CLONES = 3;     // MAX_CLONES, is_dmc.rs:43
constexpr int DMC_MAX_BLOCKS = 4096;  // the single-workgroup scan covers 4 sums per thread

// device-resident control block of one DMC run
struct DmcCtl {
  double e_ref;              // reference energy of the next branching
  int64_t n_live[2];         // live walkers of population buffer 0 / 1
  int64_t dropped;           // clones dropped at the capacity
  int64_t n_force;           // walkers summed into force_sum
  double force_sum[MAXN * 3];  // production sum of the electronic HF force
  uint32_t step;             // the step being run; each step's finalize advances it
  uint32_t pad;
};

// Step arguments that say "read the step from ctl->step" (a replayed graph
// of steps has no per-step arguments): the Philox step, the trace slot (=
// the step) and the force flag (step >= n_burnin).
constexpr uint32_t DMC_STEP_CTL = 0xFFFFFFFFu;
constexpr int64_t DMC_TS_CTL = -2;
constexpr int DMC_FORCE_CTL = -1;

struct DmcArgs {
  int n3, nainv, n_nuc;      // 3N position rows, inverse rows, nuclei
  int64_t cap;               // walkers per population buffer (multiple of DMC_BLOCK)
  double tau_b, damping, n_target;  // branching time (tau, or 2 tau matched to the diffusion)
  uint64_t seed;
  int64_t n_burnin;          // steps before the forces are summed (for DMC_FORCE_CTL)
};

struct DmcBufs {
  DmcCtl* ctl;
  int* copies;     // [cap]
  int* loff;       // [cap] in-block exclusive offsets
  int64_t* bsum;   // [cap / DMC_BLOCK] block totals, then block offsets
  double* part;    // [cap / DMC_BLOCK][1 + 3 * n_nuc]
  double* e_trace;    // [steps] mean E_L after branching
  double* pop_trace;  // [steps] population after branching
};

// Branch the population of buffer cur (post-sweep E_L in src.eloc, pre-sweep
// in eprev_src) into buffer nxt and update E_ref; step t is the global step
// (Philox counter), trace slot ts its row in the traces (-1: none), and
// with_force adds the new population's HF force to ctl->force_sum.  The
// sweep's per-block partials [sweep_nblk][3] (accepted moves in column 2,
// one row per sweep_bs walkers: 64 for the register engine, 1 for the wave
// engine) are summed into acc_out[ts].
template <class T>
hipError_t dmc_branch_step(const void* Sd, const DmcArgs& a, const DmcBufs& b, int cur, EnsView<T> src,
                           const double* eprev_src, EnsView<T> dst, double* eprev_dst, uint32_t t, int64_t ts,
                           int with_force, const double* sweep_part, int sweep_nblk, int sweep_bs, double* acc_out,
                           hipStream_t st);
// a fresh population of n0 walkers in buffer b0 (positions, state and E_L
// already built): live count, pre-sweep E_L = E_L, E_ref = mean E_L
hipError_t dmc_reset(DmcCtl* ctl, int b0, int64_t n0, const double* eloc, double* eprev, hipStream_t st);
// ctl->step = t (a kernel, so that it orders with the steps on st)
hipError_t dmc_set_step(DmcCtl* ctl, uint32_t t, hipStream_t st);
// per-block partial sums [nblk][3M] of the electronic HF force over the
// first W walkers of an ensemble
template <class T>
hipError_t hf_partials(const void* Sd, EnsView<T> E, int64_t W, int n3, int n_nuc, double* part, int* nblk,
                       hipStream_t st);

}  // namespace qmc
