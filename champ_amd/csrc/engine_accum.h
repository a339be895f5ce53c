// engine_accum.h -- sums over walkers of the products the linear method
// averages (qmc/deriv_mod.f90): dpsi_dpsi_bld (:683-723), dpsi_dpsi_eloc_bld
// (:898-926), dpsi_deloc_bld (:1375-1412), dpsi_eloc_bld (:960-1000), and the
// plain sums of dpsi, deloc and E_L behind dpsi_av, deloc_av and eloc_av.
// Their consumers are ovlp_lin_bld and ham_lin_energy_bld
// (qmc/opt_lin_mod.f90:249-348, 404-620).
//
// Accumulator layout (doubles): [0] number of samples, [1] sum E_L,
// [2] sum E_L^2, then sum dpsi [P], sum E_L dpsi [P], sum deloc [P], then
// the P x P matrices (row-major [i][j]) sum dpsi_i dpsi_j,
// sum E_L dpsi_i dpsi_j and sum dpsi_i deloc_j.
//
// The matrices are [W x P]^T [W x P] products: FP64 MFMA
// (v_mfma_f64_16x16x4f64) tiles, each wavefront owning a 32 x 32 block of all
// three matrices over one chunk of walkers; the chunks' partial sums are
// added in chunk order, so the result does not depend on scheduling.
#pragma once
#include "engine_kernels.h"

namespace qmcw {

constexpr int kAccChunk = 256;   // walkers per partial sum
constexpr int kAccBlk = 32;      // output block of a wavefront (2 x 2 MFMA tiles)

typedef double acc_d4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline int64_t accum_doubles(int P) { return 3 + 3 * (int64_t)P + 3 * (int64_t)P * P; }

// der: [W][ldd] rows of dpsi[P] followed by deloc[P]; E_L of walker w at
// st[w * lst + eoff].  part: [nchunk][3][P][P].
__global__ __launch_bounds__(64) void k_accum_gemm(const double *__restrict__ der, int ldd, int P,
                                                   const double *__restrict__ st, int64_t lst,
                                                   int64_t eoff, int W, double *__restrict__ part) {
  const int nb = (P + kAccBlk - 1) / kAccBlk;
  const int bi = blockIdx.x / nb, bj = blockIdx.x % nb, c = blockIdx.y;
  // MFMA operand layout: lane l supplies A(row l % 16, k = l / 16) and
  // B(k = l / 16, column l % 16); it receives D rows l / 16 + 4 q, q < 4,
  // of column l % 16.  Here k runs over four consecutive walkers.
  const int lane = threadIdx.x, r = lane & 15, kq = lane >> 4;
  const int i0 = bi * kAccBlk, j0 = bj * kAccBlk;
  const int w0 = c * kAccChunk, w1 = min(W, w0 + kAccChunk);
  acc_d4 acc[3][2][2];
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int u = 0; u < 2; u++) acc[m][t][u] = acc_d4{0.0, 0.0, 0.0, 0.0};
  // loads of the next walkers are issued ahead of this group's MFMAs
#pragma unroll 4
  for (int w = w0; w < w1; w += 4) {
    const int ww = w + kq;
    const double ok = ww < w1 ? 1.0 : 0.0;             // walkers past the chunk add zero
    const int wc = min(ww, w1 - 1);                    // and read an in-range row
    const double *row = der + (int64_t)wc * ldd;
    const double e = ok * st[(int64_t)wc * lst + eoff];
    double a[2], b[2], l[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int ii = i0 + 16 * t + r, jj = j0 + 16 * t + r;
      a[t] = ii < P ? ok * row[ii] : 0.0;
      b[t] = jj < P ? ok * row[jj] : 0.0;
      l[t] = jj < P ? ok * row[P + jj] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int u = 0; u < 2; u++) {
        acc[0][t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[u], acc[0][t][u], 0, 0, 0);
        acc[1][t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], e * b[u], acc[1][t][u], 0, 0, 0);
        acc[2][t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], l[u], acc[2][t][u], 0, 0, 0);
      }
  }
  const int64_t PP = (int64_t)P * P;
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int row = i0 + 16 * t + kq + 4 * q, col = j0 + 16 * u + r;
          if (row < P && col < P) part[((int64_t)c * 3 + m) * PP + (int64_t)row * P + col] = acc[m][t][u][q];
        }
}

// Vector sums of one chunk: partv[c][3 + 3P] as the accumulator's head.
__global__ __launch_bounds__(64) void k_accum_vec(const double *__restrict__ der, int ldd, int P,
                                                  const double *__restrict__ st, int64_t lst,
                                                  int64_t eoff, int W, double *__restrict__ partv) {
  const int t = blockIdx.x * 64 + threadIdx.x, c = blockIdx.y;
  const int w0 = c * kAccChunk, w1 = min(W, w0 + kAccChunk);
  double *o = partv + (int64_t)c * (3 + 3 * (int64_t)P);
  if (t < P) {
    double sd = 0.0, sde = 0.0, sl = 0.0;
#pragma unroll 8
    for (int w = w0; w < w1; w++) {
      const double *row = der + (int64_t)w * ldd;
      const double d = row[t], e = st[(int64_t)w * lst + eoff];
      sd += d;
      sde += d * e;
      sl += row[P + t];
    }
    o[3 + t] = sd;
    o[3 + P + t] = sde;
    o[3 + 2 * P + t] = sl;
  } else if (t == P) {
    double sw = 0.0, se = 0.0, se2 = 0.0;
#pragma unroll 8
    for (int w = w0; w < w1; w++) {
      const double e = st[(int64_t)w * lst + eoff];
      sw += 1.0;
      se += e;
      se2 += e * e;
    }
    o[0] = sw;
    o[1] = se;
    o[2] = se2;
  }
}

// acc[e] += sum over chunks, in chunk order
__global__ __launch_bounds__(256) void k_accum_reduce(const double *__restrict__ part,
                                                      const double *__restrict__ partv, int nch, int P,
                                                      double *__restrict__ acc) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nv = 3 + 3 * (int64_t)P, n = accum_doubles(P);
  if (e >= n) return;
  double s = 0.0;
  if (e < nv) {
    for (int c = 0; c < nch; c++) s += partv[(int64_t)c * nv + e];
  } else {
    const int64_t m = e - nv, PP3 = 3 * (int64_t)P * P;
    for (int c = 0; c < nch; c++) s += part[(int64_t)c * PP3 + m];
  }
  acc[e] += s;
}

__global__ __launch_bounds__(256) void k_accum_zero(double *__restrict__ acc, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) acc[e] = 0.0;
}

}  // namespace qmcw
