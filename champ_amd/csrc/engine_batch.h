// engine_batch.h -- kernels of a VMC step over several systems (handles) in
// one launch per kernel: the systems' walker groups are laid end to end in
// the grid, and each block runs as the matching block of its own system's
// launch, with that system's arguments.  Systems of uneven electron and basis
// counts share a launch; the kernels' lane-group widths and e-e-n order are
// common to the batch.
#pragma once
#include "engine_kernels.h"

namespace qmcw {

struct BatchArgs {
  const KArgs *A;          // [n] the systems' arguments for this kernel
  const int *boff;         // [n + 1] first block of each system in the grid
  const int *nw;           // [n] walkers each system launches over
  long long *const *nacc;  // [n] acceptance counters of the moves
  int n;
};

// the system that owns block blk (uniform over the block)
__device__ __forceinline__ int batch_sys(const BatchArgs &B, int blk) {
  int s = 0;
  while (s + 1 < B.n && blk >= B.boff[s + 1]) s++;
  return s;
}

// One-electron VMC move (proposal, test and commit) of electron ie in every
// system that has it
template <typename T, int G, int ORD>
__global__ QMCW_HEAVY void k_ratio_1e_batch(BatchArgs B, int ie, double delta, double fbias) {
  const int s = batch_sys(B, blockIdx.x);
  const KArgs &A = B.A[s];
  if (ie >= A.S.N) return;
  ratio_1e_b<T, G, ORD>(A, blockIdx.x - B.boff[s], ie, 2, nullptr, delta, fbias, nullptr, B.nacc[s]);
}

// The whole sweep of each system in one launch: a block moves its
// walkers' electrons one after another (the barrier after each move makes
// its commit visible to the next move's reads), so the blocks of small
// systems finish and free their compute units while large ones go on
template <typename T, int G, int ORD>
__global__ QMCW_HEAVY void k_sweep_1e_batch(BatchArgs B, double delta, double fbias) {
  const int s = batch_sys(B, blockIdx.x);
  const KArgs &A = B.A[s];
  const int N = A.S.N;
  for (int ie = 0; ie < N; ie++) {
    ratio_1e_b<T, G, ORD>(A, blockIdx.x - B.boff[s], ie, 2, nullptr, delta, fbias, nullptr, B.nacc[s]);
    __syncthreads();
  }
}

template <typename T, int G, int ORD>
__global__ QMCW_HEAVY void k_full_jas_batch(BatchArgs B) {
  const int s = batch_sys(B, blockIdx.x);
  full_jas_b<T, G, ORD>(B.A[s], blockIdx.x - B.boff[s]);
}

template <typename T, int G, int ORD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : QMCW_WPE, 8)))
void k_full_nlpp_batch(BatchArgs B) {
  const int s = batch_sys(B, blockIdx.x);
  const KArgs &A = B.A[s];
  if (A.S.nloc == 0) return;
  full_nlpp_b<T, G, ORD>(A, blockIdx.x - B.boff[s]);
}

#ifndef QMCW_ORD   // launched from the host translation unit only
template <typename T, int G>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : QMCW_WPE, 8)))
void k_full_det_batch(BatchArgs B) {
  const int s = batch_sys(B, blockIdx.x);
  full_det_b<T, G>(B.A[s], blockIdx.x - B.boff[s]);
}

template <int G>
__global__ __launch_bounds__(64) void k_full_fin_batch(BatchArgs B) {
  const int s = batch_sys(B, blockIdx.x);
  full_fin_b<G>(B.A[s], blockIdx.x - B.boff[s]);
}

// rotqua of every walker of the systems with a pseudopotential (64 walkers
// per block)
__global__ __launch_bounds__(64) void k_rotqua_batch(BatchArgs B) {
  const int s = batch_sys(B, blockIdx.x);
  const KArgs &A = B.A[s];
  if (A.S.nloc == 0) return;
  rotqua_b(A, blockIdx.x - B.boff[s]);
}

// the pseudopotential projection of each system, its (row block, column
// block, spin) grid flattened: gx row blocks of 64 (walker, electron) rows,
// gy column blocks of 64 basis functions, two spins
__device__ __forceinline__ void gproj_grid(const KArgs &A, int nw, int &gx, int &gy) {
  gx = (int)(((int64_t)nw * max(A.S.ns[0], A.S.ns[1]) + 63) / 64);
  gy = (A.S.NB + 63) / 64;
}
__global__ __launch_bounds__(64) void k_gproj_mfma_batch(BatchArgs B) {
  const int s = batch_sys(B, blockIdx.x);
  const KArgs &A = B.A[s];
  if (!A.gproj_mfma) return;
  int gx, gy;
  gproj_grid(A, B.nw[s], gx, gy);
  const int local = blockIdx.x - B.boff[s], per = gx * gy;
  const int sp = local / per, r = local % per;
  gproj_mfma_b(A, B.nw[s], r % gx, r / gx, sp);
}
#endif

}  // namespace qmcw
