// engine_kernels.h -- kernels of the walker-batched wavefunction engine.
// Every kernel runs 64-lane workgroups; group g of G lanes owns walker
// blockIdx.x*(64/G)+g.  The walker count is padded to a multiple of 64/G
// on the host, so every group has a valid walker and all groups reach the
// same barriers.
#pragma once
#include "engine_device.h"

namespace qmcw {

struct KArgs {
  DevSys S;
  WLayout L;
  SLayout SL;
  MLayout ML;
  DLayout DL;
  double *st;        // walker state  [Wp][L.stride]
  double *sc;        // walker scratch [Wp][SL.stride]
  int *rng;          // [Wp][4]
  int *flags;        // [Wp] status bits
  int Wp;
  int derivs;        // compute Jastrow/NLPP parameter derivative pieces
  int full_orb;      // evaluate all NO orbitals (orbital-rotation derivatives)
  int lds_phi;       // k_ratio_1e keeps the moved electron's basis values in LDS
  int lds_wst;       // T values per walker of that LDS region (at least NB * 5)
  int een_tab;       // k_ratio_1e: the moved electron's e-e-n tables in the same region
                     // during the Jastrow row (een_tables, jas_pair_tab)
  int lds_det;       // k_full_det keeps a chunk's basis values in LDS
  int keep_state;    // the full evaluation takes the rows the committed moves kept in the state
                     // (distances, Jastrow pairs and, without full_orb, the orbitals)
                     // instead of recomputing them
  int lds_y;         // doubles per group of the LDS staging of eval_y in the move kernels (0: none)
  int lds_y_off;     // its offset (doubles) in the dynamic LDS, after the basis values
  int lds_nlpp;      // doubles per lane of k_full_nlpp's point scratch in LDS (0: global)
  int lds_dnlpp;     // doubles per lane of k_deriv_nlpp's point scratch in LDS (0: global)
  int bf16;          // orbital-contraction inputs rounded to bfloat16
  int chunk;         // electrons per basis chunk in the full evaluation
  int pnl_mfma;      // k_deriv_nlpp leaves psp_nonloc_orb's contraction to k_pnl_mfma
  int orb_mfma;      // k_full_det (full_orb) leaves the contraction of all orbitals to
                     // k_orb_mfma, writing every electron's basis values to DL.phiall
  int gproj_mfma;    // k_full_det leaves the pseudopotential projection out (k_gproj_mfma
                     // forms it for large determinants; a re-inversion needs none)
  int lds_inv;       // k_full_det: doubles per group of LDS for A and the orbital values (0: global)
  int lds_inv_off;   // their offset (doubles) in the dynamic LDS
  int dmc_rot;       // DMC evaluation: pseudopotential grid rotated per pair in range
  const double *dmove;  // DMC: nonzero at dmove[w * dmove_stride] if the walker moved
  int dmove_stride;     // (null: every walker evaluates and draws)
  double *tq;           // DMC: per walker (tq_stride doubles) the quadrature-point ratios
  int64_t tq_stride;    // [N][tm_kmax][nq] and points [N][tm_kmax][nq][3] kept for T-moves
  double *dm;           // DMC records (k_ratio_1e mode 3)
  qmcw_dmc_params dmp;  // and the generation's parameters
  const uint8_t *acc_in;  // k_ratio_1e mode 4: the walkers whose given move is committed
  int imetro;           // VMC mover: 1 metrop_mov1, 6 metrop_slat_mov1
  double deltar, deltat;  // and the accelerated mover's radial and angular steps
};

// offsets in a walker's T-move record
__host__ __device__ inline int64_t tq_ratio(const DevSys &S, int i, int j, int iq) {
  return ((int64_t)i * S.tm_kmax + j) * S.nquad + iq;
}
__host__ __device__ inline int64_t tq_point(const DevSys &S, int i, int j, int iq) {
  return (int64_t)S.N * S.tm_kmax * S.nquad + 3 * tq_ratio(S, i, j, iq);
}

}  // namespace qmcw
#include "engine_dmc_dev.h"
#include "engine_slat_dev.h"
namespace qmcw {

// occupancy hint of the heavy per-walker kernels (waves per SIMD)
#ifndef QMCW_WPE
#define QMCW_WPE 2
#endif
#define QMCW_HEAVY __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(QMCW_WPE, 8)))
#ifndef QMCW_MOVE_WPE
#define QMCW_MOVE_WPE 2
#endif
#define QMCW_MOVE __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(QMCW_MOVE_WPE, 8)))
// offset (past ML.misc + ns + nocc) of the Jastrow sums handed between the move kernels
constexpr int kMvJas = 16;

// the group's lane, group and walker in block bid of the launch (a kernel's
// own blockIdx.x, or the block's index within its system in a batch launch)
#define QMCW_GROUP_SETUP_B(bid)                                  \
  const int lane = threadIdx.x % G;                              \
  const int grp = threadIdx.x / G;                               \
  const int w = (bid) * (kWave / G) + grp;                       \
  const DevSys &S = A.S;                                         \
  double *ws = A.st + (int64_t)w * A.L.stride;                   \
  double *ss = A.sc + (int64_t)w * A.SL.stride;                  \
  (void)ss;
#define QMCW_GROUP_SETUP                                         \
  const int lane = threadIdx.x % G;                              \
  const int grp = threadIdx.x / G;                               \
  const int w = blockIdx.x * (kWave / G) + grp;                  \
  const DevSys &S = A.S;                                         \
  double *ws = A.st + (int64_t)w * A.L.stride;                   \
  double *ss = A.sc + (int64_t)w * A.SL.stride;                  \
  (void)ss;

// The walker's current (o = 0) or candidate (o = 1) slot of a determinant
// block (WLayout: A^-1, then T / alpha inverses / their determinants, then
// Y, each with its own slot number in the state)
__device__ __forceinline__ int wslot(const double *ws, int64_t at) { return ws[at] != 0.0 ? 1 : 0; }
template <typename P>
__device__ __forceinline__ P *w_ai(const WLayout &L, P *ws, int s, int o = 0) {
  return ws + L.ai[s][wslot(ws, L.par + s) ^ o];
}
template <typename P>
__device__ __forceinline__ P *w_tocc(const WLayout &L, P *ws, int s, int o = 0) {
  return ws + L.tocc[s][wslot(ws, L.par + 2 + s) ^ o];
}
template <typename P>
__device__ __forceinline__ P *w_inv(const WLayout &L, P *ws, int s, int o = 0) {
  return ws + L.inv[s][wslot(ws, L.par + 2 + s) ^ o];
}
template <typename P>
__device__ __forceinline__ P *w_detv(const WLayout &L, P *ws, int s, int o = 0) {
  return ws + L.detv[s][wslot(ws, L.par + 2 + s) ^ o];
}
template <typename P>
__device__ __forceinline__ P *w_y(const WLayout &L, P *ws, int s, int o = 0) {
  return ws + L.y[s][wslot(ws, L.par + 4) ^ o];
}

// bfloat16 rounding (nearest, ties to even) of an FP32 value, kept in FP32
__device__ __forceinline__ float bf16r(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xFFFF0000u);
}

// orbital value, gradient and laplacian o[0..4] = sum_b coef[b * NO + lab] phi_b[0..4]
// (coef: coefT with NO = all orbitals, or an occupied-orbital table with NO =
// nocc; phi with stride 5), accumulated in FP64.  These rows form the Slater
// matrices, whose determinant can be orders of magnitude below the product
// of its rows near a node: bfloat16 inputs there gave local-energy errors of
// tens of Ha, so the BF16 mode does not round them (it rounds only the
// pseudopotential projection, group_gproj).
template <typename T>
__device__ __forceinline__ void contract5(cptr<double> coefT, int NO, int NB, int lab,
                                          const T *__restrict__ ph, double *o) {
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
#pragma unroll 4
  for (int b = 0; b < NB; b++) {
    const double cf = coefT[(int64_t)b * NO + lab];
    const T *p = ph + 5 * b;
    a0 += cf * (double)p[0]; a1 += cf * (double)p[1]; a2 += cf * (double)p[2];
    a3 += cf * (double)p[3]; a4 += cf * (double)p[4];
  }
  o[0] = a0; o[1] = a1; o[2] = a2; o[3] = a3; o[4] = a4;
}

// ---------------------------------------------------------------------------
// In-place inverse of the column-major n x n block `a` by a group (CHAMP's
// matinv pivoting, rows in parallel).  tmp: n*n scratch.  Returns det (all
// lanes).  ipiv: LDS array of the group, >= n entries.
template <int G>
__device__ double group_matinv(double *a, double *tmp, int n, int lane, int *ipiv) {
  for (int i = lane; i < n; i += G) ipiv[i] = i;
  __syncthreads();
  double det = 1.0;
  bool singular = false;
  for (int i = 0; i < n; i++) {
    const int col = i * n;
    // matinv scans k = i..n-1 keeping the first strictly larger |a(ipiv(k), i)|
    double am = -1.0, vm = 0.0;
    int km = n;
    for (int k = i + lane; k < n; k += G) {
      double v = a[ipiv[k] + col];
      if (fabs(v) > am) { am = fabs(v); vm = v; km = k; }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      double a2 = __shfl_xor(am, o, G);
      int k2 = __shfl_xor(km, o, G);
      double v2 = __shfl_xor(vm, o, G);
      if (a2 > am || (a2 == am && k2 < km)) { am = a2; km = k2; vm = v2; }
    }
    __syncthreads();
    if (!singular) {
      if (lane == 0 && km != i) { int t = ipiv[i]; ipiv[i] = ipiv[km]; ipiv[km] = t; }
      if (km != i) det = -det;
      det *= vm;
      if (vm == 0.0) singular = true;
    }
    __syncthreads();
    const int p = ipiv[i];
    const double inv = singular ? 0.0 : 1.0 / vm;
    if (!singular && lane == 0) a[p + col] = 1.0;
    __syncthreads();
    if (!singular)
      for (int q = lane; q < n; q += G) a[p + q * n] *= inv;
    __syncthreads();
    if (!singular)
      for (int j = lane; j < n; j += G) {
        if (j == p) continue;
        double t = -a[j + col];
        a[j + col] = 0.0;
        for (int q = 0; q < n; q++) a[j + q * n] += t * a[p + q * n];
      }
    __syncthreads();
  }
  // row permutation: new a(r, c) = a(ipiv[r], c); then columns: new
  // a(r, ipiv[c]) = a(r, c).  A singular matrix returns det 0 and zeros.
  for (int q = lane; q < n * n; q += G) tmp[q] = a[q];
  __syncthreads();
  for (int q = lane; q < n * n; q += G) {
    int r = q % n, c = q / n;
    a[r + c * n] = singular ? 0.0 : tmp[ipiv[r] + c * n];
  }
  __syncthreads();
  for (int q = lane; q < n * n; q += G) tmp[q] = a[q];
  __syncthreads();
  for (int q = lane; q < n * n; q += G) {
    int r = q % n, c = q / n;
    a[r + ipiv[c] * n] = tmp[r + c * n];
  }
  __syncthreads();
  return singular ? 0.0 : det;
}

// group_matinv's elimination on a matrix staged in LDS, without the final
// permutation: la (column-major n x n) ends as the eliminated matrix whose
// row ipiv[r] and column c hold the inverse's element (r, ipiv[c]).  The
// operations and their order are group_matinv's, so the values are the
// same.  Returns det (0 and *singular set for a singular matrix).
template <int G>
__device__ double group_gj_lds(double *la, int n, int lane, int *ipiv, bool *singular_out) {
  for (int i = lane; i < n; i += G) ipiv[i] = i;
  __syncthreads();
  double det = 1.0;
  bool singular = false;
  for (int i = 0; i < n; i++) {
    const int col = i * n;
    double am = -1.0, vm = 0.0;
    int km = n;
    for (int k = i + lane; k < n; k += G) {
      double v = la[ipiv[k] + col];
      if (fabs(v) > am) { am = fabs(v); vm = v; km = k; }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      double a2 = __shfl_xor(am, o, G);
      int k2 = __shfl_xor(km, o, G);
      double v2 = __shfl_xor(vm, o, G);
      if (a2 > am || (a2 == am && k2 < km)) { am = a2; km = k2; vm = v2; }
    }
    __syncthreads();
    if (!singular) {
      if (lane == 0 && km != i) { int t = ipiv[i]; ipiv[i] = ipiv[km]; ipiv[km] = t; }
      if (km != i) det = -det;
      det *= vm;
      if (vm == 0.0) singular = true;
    }
    __syncthreads();
    const int p = ipiv[i];
    const double inv = singular ? 0.0 : 1.0 / vm;
    if (!singular && lane == 0) la[p + col] = 1.0;
    __syncthreads();
    if (!singular)
      for (int q = lane; q < n; q += G) la[p + q * n] *= inv;
    __syncthreads();
    if (!singular)
      for (int j = lane; j < n; j += G) {
        if (j == p) continue;
        double t = -la[j + col];
        la[j + col] = 0.0;
        for (int q = 0; q < n; q++) la[j + q * n] += t * la[p + q * n];
      }
    __syncthreads();
  }
  *singular_out = singular;
  return singular ? 0.0 : det;
}

// Y entries of one spin: y[e] = chii sum_q cus[yent_u[q]] inv[yent_pos[q]]
// over e = lane, lane + G, ...; the inputs never alias the output, so the
// loads of an entry's list are issued together.
template <int G>
__device__ __forceinline__ void group_y_spin(const int *__restrict__ off, const int *__restrict__ yu,
                                             const int *__restrict__ yp, const double *__restrict__ cus,
                                             const double *__restrict__ inv, double *__restrict__ y,
                                             int n, double chii, int lane) {
  for (int e = lane; e < n; e += G) {
    double acc = 0.0;
    const int q1 = off[e + 1];
#pragma unroll 4
    for (int q = off[e]; q < q1; q++) acc += cus[yu[q]] * inv[yp[q]];
    y[e] = acc * chii;
  }
}

// chi and Y of one walker from the alpha inverses/dets of both spins
// (qmc/deriv_fast_mod.f90:545-598).  inv/detv per spin, y out per spin.
// The determinant coefficients carry the permutation signs (DevSys::cdd).
template <int G>
__device__ double group_eval_y(const DevSys &S, const double *const inv[2],
                               const double *const detv[2], double *const y[2],
                               double *cup, int lane) {
  // chi = sum_d cdet sgn detup detdn ; cup_s(u) = sum over dets with spin-s part u
  const double *__restrict__ dv0 = detv[0];
  const double *__restrict__ dv1 = detv[1];
  double chi = 0.0;
#pragma unroll 4
  for (int d = lane; d < S.ndet; d += G) chi += cst(S.cdd)[d] * dv0[cst(S.iwdet[0])[d]] * dv1[cst(S.iwdet[1])[d]];
  chi = gsum<G>(chi);
  const double chii = 1.0 / chi;
  for (int t = lane; t < S.ndetu[0] + S.ndetu[1]; t += G) {
    const int s = t < S.ndetu[0] ? 0 : 1, u = s ? t - S.ndetu[0] : t;
    const int *__restrict__ iu = S.uiu[s];
    const int *__restrict__ id = S.uid[s];
    const double *__restrict__ cf = S.ucf[s];
    double cu = 0.0;
    const int q1 = cst(S.udet_off[s])[u + 1];
#pragma unroll 4
    for (int q = cst(S.udet_off[s])[u]; q < q1; q++) cu += cf[q] * dv0[iu[q]] * dv1[id[q]];
    cup[t] = cu;
  }
  __syncthreads();
  for (int s = 0; s < 2; s++)
    group_y_spin<G>(S.yent_off[s], S.yent_u[s], S.yent_pos[s], cup + (s ? S.ndetu[0] : 0), inv[s],
                    y[s], S.nocc[s] * S.ns[s], chii, lane);
  __syncthreads();
  return chi;
}

// Occupied orbitals (value, gradient, laplacian) of every electron into the
// walker state (and all orbitals with full_orb), A.chunk electrons at a time
// through phi (basis values in the engine's real type).
template <typename T, int G>
__device__ __forceinline__ void full_orbitals(const KArgs &A, double *ws, double *ss, T *phi, int lane) {
  const DevSys &S = A.S;
  const WLayout &L = A.L;
  const int N = S.N, M = S.M, NO = S.NO, noccmax = S.noccmax;
  const double *rven = ws + L.rvec_en, *ren = ws + L.r_en;
  double *oc = ws + L.oc;
  double *orbf = ss + A.SL.orb;
  for (int i0 = 0; i0 < N; i0 += A.chunk) {
    const int nch = min(A.chunk, N - i0);
    __syncthreads();
    for (int t = lane; t < nch * S.nshell; t += G) {
      const int ii = t / S.nshell, s = t % S.nshell;
      const int i = i0 + ii, ic = cst(S.sh_cent)[s];
      const double *rv = rven + (i * M + ic) * 3;
      eval_shell<T>(S, s, (T)rv[0], (T)rv[1], (T)rv[2], (T)ren[i * M + ic], 1,
                    phi + (int64_t)ii * S.NB * 5);
    }
    __syncthreads();
    // occupied orbitals of each electron's spin into the resident state
    for (int t = lane; t < nch * noccmax; t += G) {
      const int ii = t / noccmax, j = t % noccmax;
      const int i = i0 + ii, s = (i < S.nup) ? 0 : 1;
      double *o = oc + ((int64_t)i * noccmax + j) * 5;
      if (j >= S.nocc[s]) { for (int c = 0; c < 5; c++) o[c] = 0.0; continue; }
      contract5<T>(cst(S.coef_occ[s]), S.nocc[s], S.NB, j, phi + (int64_t)ii * S.NB * 5, o);
    }
    if (A.full_orb && A.orb_mfma) {
      // the basis values of the chunk, for k_orb_mfma
      double *pa = ss + A.DL.phiall + (int64_t)i0 * S.NB * 5;
      for (int t = lane; t < nch * S.NB * 5; t += G) pa[t] = (double)phi[t];
    } else if (A.full_orb) {
      for (int t = lane; t < nch * NO; t += G) {
        const int ii = t / NO, lab = t % NO;
        const int i = i0 + ii;
        contract5<T>(cst(S.coefT), NO, S.NB, lab, phi + (int64_t)ii * S.NB * 5,
                     orbf + ((int64_t)i * NO + lab) * 5);
      }
    }
  }
}

// gp[i][b] = sum_j G(j, i) coef(b, occ_j) for every electron, or only for
// electron `only` (FP64; BF16 mode rounds the result), coef_occ holding the
// occupied columns contiguously
template <int G>
__device__ __forceinline__ void group_gproj(const DevSys &S, double *__restrict__ gp,
                                            const double *__restrict__ g0, const double *__restrict__ g1,
                                            int bf16, int lane, int only = -1) {
  const int NB = S.NB;
  const int t0 = only < 0 ? 0 : only * NB, t1 = only < 0 ? S.N * NB : (only + 1) * NB;
  for (int t = t0 + lane; t < t1; t += G) {
    const int i = t / NB, b = t % NB;
    const int s = i < S.nup ? 0 : 1, ie = s ? i - S.nup : i, ns = S.ns[s], nocc = S.nocc[s];
    const double *__restrict__ gc = (s ? g1 : g0) + ie;
    const double *__restrict__ co = S.coef_occ[s] + (int64_t)b * nocc;
    double acc = 0.0;
#pragma unroll 8
    for (int j = 0; j < nocc; j++) acc += gc[j * ns] * co[j];
    gp[t] = bf16 ? (double)bf16r((float)acc) : acc;
  }
}

// group_eval_y with the walker's determinant values and alpha inverses of
// both spins staged in LDS first (one batched copy), so that the index
// chains of the chi, cup and Y sums end in LDS reads.  lds: this group's
// region of y_lds_doubles(S) doubles (derived from the kernel's dynamic LDS
// so that the compiler addresses it as LDS).
__host__ __device__ inline int y_lds_doubles(int ndetu0, int ndetu1, int kmax2) {
  return (ndetu0 + ndetu1) * (2 + kmax2);
}
template <int G>
__device__ __forceinline__ double group_eval_y_lds(const DevSys &S, const double *const inv[2],
                                                   const double *const detv[2], double *const y[2],
                                                   double *lds, int lane) {
  const int nu0 = S.ndetu[0], nu1 = S.ndetu[1], k2 = S.kmax2;
  double *dv0 = lds, *dv1 = dv0 + nu0, *iv0 = dv1 + nu1, *iv1 = iv0 + nu0 * k2, *cup = iv1 + nu1 * k2;
  {
    // segments: detv[0], detv[1], inv[0], inv[1]; four loads in flight per lane
    const int st1 = nu0, st2 = st1 + nu1, st3 = st2 + nu0 * k2, n = st3 + nu1 * k2;
    for (int base = 0; base < n; base += 4 * G) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = base + lane + u * G;
        const double *src = e < st1 ? detv[0] + e : e < st2 ? detv[1] + (e - st1)
                          : e < st3 ? inv[0] + (e - st2) : inv[1] + (e - st3);
        v[u] = e < n ? *src : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = base + lane + u * G;
        if (e < n) lds[e] = v[u];
      }
    }
  }
  __syncthreads();
  double chi = 0.0;
#pragma unroll 4
  for (int d = lane; d < S.ndet; d += G)
    chi += cst(S.cdd)[d] * dv0[cst(S.iwdet[0])[d]] * dv1[cst(S.iwdet[1])[d]];
  chi = gsum<G>(chi);
  const double chii = 1.0 / chi;
  for (int t = lane; t < nu0 + nu1; t += G) {
    const int s = t < nu0 ? 0 : 1, u = s ? t - nu0 : t;
    const cptr<int> iu = cst(S.uiu[s]), id = cst(S.uid[s]);
    const cptr<double> cf = cst(S.ucf[s]);
    double cu = 0.0;
    const int q1 = cst(S.udet_off[s])[u + 1];
#pragma unroll 4
    for (int q = cst(S.udet_off[s])[u]; q < q1; q++) cu += cf[q] * dv0[iu[q]] * dv1[id[q]];
    cup[t] = cu;
  }
  __syncthreads();
  for (int s = 0; s < 2; s++) {
    const cptr<int> off = cst(S.yent_off[s]), yu = cst(S.yent_u[s]), yp = cst(S.yent_pos[s]);
    const double *cus = cup + (s ? nu0 : 0), *iv = s ? iv1 : iv0;
    double *ys = y[s];
    for (int e = lane; e < S.nocc[s] * S.ns[s]; e += G) {
      double acc = 0.0;
      const int q1 = off[e + 1];
#pragma unroll 4
      for (int q = off[e]; q < q1; q++) acc += cus[yu[q]] * iv[yp[q]];
      ys[e] = acc * chii;
    }
  }
  __syncthreads();
  return chi;
}

// G = A^-1 - T Y A^-1 on reference rows, Y A^-1 on the others
// (qmc/deriv_fast_mod.f90:600-625).  ai col-major ns x ns, tocc ns x nocc,
// y nocc x ns, g out nocc x ns; tmp >= nocc*ns.
template <int G>
__device__ void group_g(const DevSys &S, int s, const double *ai,
                        const double *tocc, const double *y, double *g,
                        double *tmp, int lane) {
  int ns = S.ns[s], nocc = S.nocc[s];
  if (ns == 0) return;
  if (S.ndetu[s] == 1) {
    // one determinant of this spin: Y = 0, so G is A^-1 in the rows of the
    // reference orbitals (the products below would add exact zeros)
    for (int e = lane; e < ns * ns; e += G) {
      int k = e / ns, i = e % ns;
      g[cst(S.iref[s])[k] * ns + i] = ai[k + ns * i];
    }
    __syncthreads();
    return;
  }
  for (int e = lane; e < nocc * ns; e += G) {
    int j = e / ns, i = e % ns;
    double acc = 0.0;
    for (int k = 0; k < ns; k++) acc += y[j * ns + k] * ai[k + ns * i];
    tmp[e] = acc;   // yai
  }
  __syncthreads();
  for (int e = lane; e < nocc * ns; e += G) {
    int j = e / ns, i = e % ns;
    g[e] = tmp[e];
    (void)j;
  }
  __syncthreads();
  for (int e = lane; e < ns * ns; e += G) {
    int k = e / ns, i = e % ns;
    double acc = 0.0;
    for (int j = 0; j < nocc; j++) acc += tocc[k * nocc + j] * tmp[j * ns + i];
    g[cst(S.iref[s])[k] * ns + i] = ai[k + ns * i] - acc;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Full evaluation (hpsi): distances, Jastrow, orbitals, determinants (AMF),
// pseudopotential, local energy.  Refreshes the walker's resident state.
// The full evaluation runs as four kernels (distances + Jastrow; orbitals +
// determinants; pseudopotential; local energy) so that each carries fewer
// live registers; scalars pass through the walker's reduction scratch:
// phs[50] pe_en, [51] pe_ee, [52] ln J, [53] lap ln J, [54] lap ln D,
// [55] local pseudopotential, [56] nonlocal pseudopotential (SLayout::phs).
template <typename T, int G, int ORD>
__device__ __forceinline__ void full_jas_b(const KArgs &A, int bid_) {
  QMCW_GROUP_SETUP_B(bid_)
  const int N = S.N, M = S.M, NO = S.NO;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const double *x = ws + L.x;
  double *rven = ws + L.rvec_en, *ren = ws + L.r_en, *rvee = ws + L.rvec_ee, *ree = ws + L.r_ee;
  double *fso = ws + L.fso, *fijo = ws + L.fijo, *lapj = ws + L.lapj, *fjo = ws + L.fjo;
  double *vj = ss + SL.vj, *divv = ss + SL.divv, *vd = ss + SL.vd;
  double *red = ss + SL.phs;
  (void)NO; (void)x; (void)rven; (void)ren; (void)rvee; (void)ree; (void)fso; (void)fijo;
  (void)lapj; (void)fjo; (void)vj; (void)divv; (void)vd; (void)red;
  double pe_en = 0.0, pe_ee = 0.0, fsum = 0.0;
  if (lane == 0) A.flags[w] = 0;   // status of this evaluation (set by k_full_nlpp, k_full_fin)
  if (A.keep_state) {
    // Distances and Jastrow pair arrays as the committed moves left them: a
    // move recomputes every pair of the moved electron, so they hold the
    // values at the current positions; only the sums are formed here.
    for (int p = lane; p < N * M; p += G)
      if (S.nloc == 0) pe_en -= cst(S.znuc)[cst(S.iwctype)[p % M]] / ren[p];
    for (int p = lane; p < N * N; p += G) {
      const int i = p / N, j = p % N;
      if (j < i) pe_ee += 1.0 / ree[p];
      if (j <= i) fsum += fso[p];
    }
    pe_en = gsum<G>(pe_en);
    pe_ee = gsum<G>(pe_ee);
    fsum = gsum<G>(fsum);
  } else {
    // ---- distances (qmc/vmc/distances.f90:1-233)
    for (int p = lane; p < N * M; p += G) {
      int i = p / M, ic = p % M;
      double s2 = 0.0;
      for (int q = 0; q < 3; q++) {
        double d = x[3 * i + q] - cst(S.cent)[3 * ic + q];
        rven[3 * p + q] = d;
        s2 += d * d;
      }
      double r = sqrt(s2);
      ren[p] = r;
      if (S.nloc == 0) pe_en -= cst(S.znuc)[cst(S.iwctype)[ic]] / r;
    }
    for (int p = lane; p < N * N; p += G) {
      int i = p / N, j = p % N;
      if (j >= i) continue;
      double s2 = 0.0;
      for (int q = 0; q < 3; q++) {
        double d = x[3 * i + q] - x[3 * j + q];
        rvee[3 * p + q] = d;
        s2 += d * d;
      }
      double r = sqrt(s2);
      ree[p] = r;
      pe_ee += 1.0 / r;
    }
    pe_en = gsum<G>(pe_en);
    pe_ee = gsum<G>(pe_ee);
    __syncthreads();
    // ---- Jastrow pair arrays (qmc/vmc/jastrow4.f90)
    for (int p = lane; p < N * N; p += G) {
      int i = p / N, j = p % N;
      if (j > i) continue;
      if (j == i) {
        T f, g[3], lp;
        jas_en<T>(S, ren + i * M, rven + i * M * 3, f, g, lp);
        fso[p] = f;
        for (int q = 0; q < 3; q++) fijo[3 * p + q] = g[q];
        lapj[p] = lp;
        fsum += f;
      } else {
        PairOut<T> o;
        T rvij[3];
        for (int q = 0; q < 3; q++) rvij[q] = (T)rvee[3 * p + q];
        bool par = (i < S.nup) || (j >= S.nup);
        jas_pair<T, ORD>(S, par, (T)ree[p], rvij, ren + i * M, ren + j * M, rven + i * M * 3,
                    rven + j * M * 3, o);
        fso[p] = o.f;
        for (int q = 0; q < 3; q++) {
          fijo[3 * (i * N + j) + q] = o.gi[q];
          fijo[3 * (j * N + i) + q] = o.gj[q];
        }
        lapj[i * N + j] = o.li;
        lapj[j * N + i] = o.lj;
        fsum += o.f;
      }
    }
    fsum = gsum<G>(fsum);
  }
  __syncthreads();
  double d2j = 0.0;
  for (int i = lane; i < N; i += G) {
    double gx = 0, gy = 0, gz = 0, dv = lapj[i * N + i], lt = lapj[i * N + i];
    for (int j = 0; j < N; j++) {
      gx += fijo[3 * (i * N + j)];
      gy += fijo[3 * (i * N + j) + 1];
      gz += fijo[3 * (i * N + j) + 2];
      if (j != i) {
        dv += 0.5 * (lapj[i * N + j] + lapj[j * N + i]);
        lt += lapj[i * N + j];
      }
    }
    fjo[3 * i] = gx; fjo[3 * i + 1] = gy; fjo[3 * i + 2] = gz;
    vj[3 * i] = gx; vj[3 * i + 1] = gy; vj[3 * i + 2] = gz;
    divv[i] = dv;    // CHAMP's div_vj (pair terms split evenly)
    d2j += lt;
  }
  d2j = gsum<G>(d2j);
  if (lane == 0) {
    ws[L.fsumo] = fsum;
    red[50] = pe_en; red[51] = pe_ee; red[52] = fsum; red[53] = d2j;
  }
}
template <typename T, int G, int ORD>
__global__ QMCW_HEAVY void k_full_jas(KArgs A) {
  full_jas_b<T, G, ORD>(A, blockIdx.x);
}

template <typename T, int G>
__device__ __forceinline__ void full_det_b(const KArgs &A, int bid_) {
  QMCW_GROUP_SETUP_B(bid_)
  const int N = S.N, M = S.M, NO = S.NO;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const double *x = ws + L.x;
  double *rven = ws + L.rvec_en, *ren = ws + L.r_en, *rvee = ws + L.rvec_ee, *ree = ws + L.r_ee;
  double *fso = ws + L.fso, *fijo = ws + L.fijo, *lapj = ws + L.lapj, *fjo = ws + L.fjo;
  double *vj = ss + SL.vj, *divv = ss + SL.divv, *vd = ss + SL.vd;
  double *red = ss + SL.phs;
  (void)NO; (void)x; (void)rven; (void)ren; (void)rvee; (void)ree; (void)fso; (void)fijo;
  (void)lapj; (void)fjo; (void)vj; (void)divv; (void)vd; (void)red;
  __shared__ int s_ipiv[kWave / G][128];
  __shared__ int s_ipinv[kWave / G][128];
  // ---- orbitals (basis_fns + orbitals_loc_ana), chunks of electrons; the
  // chunk's basis values in LDS when the launch reserved it
  extern __shared__ double dyn_lds[];
  double *oc = ws + L.oc;
  const int noccmax = S.noccmax;
  // The rows of the occupied orbitals (value, gradient, laplacian) of every
  // electron at its current position are already in the state when each
  // position since the last full evaluation came from a committed move,
  // whose orbitals are those of the new position: then they are kept.
  if (A.keep_state && !A.full_orb) {
  } else if (A.lds_det)
    full_orbitals<T, G>(A, ws, ss, reinterpret_cast<T *>(dyn_lds) + (int64_t)grp * A.chunk * S.NB * 5, lane);
  else
    full_orbitals<T, G>(A, ws, ss, reinterpret_cast<T *>(ss + SL.phi), lane);
  __syncthreads();
  // ---- determinants (t_bld, alpha_bld, y_bld, g_bld)
  double *tmp = ss + SL.red;  // >= max(ns^2, nocc*ns) doubles
  for (int s = 0; s < 2; s++) {
    int ns = S.ns[s], nocc = S.nocc[s], off = s ? S.nup : 0;
    double *ai = w_ai(L, ws, s);
    if (ns == 0) {
      if (lane == 0) { ws[L.deta[s]] = 1.0; w_detv(L, ws, s)[0] = 1.0; }
      __syncthreads();
      continue;
    }
    double *tocc = w_tocc(L, ws, s);
    if (A.lds_inv) {
      // A and the occupied-orbital values in LDS: the elimination and
      // T = A^-1 orb read them there (large determinants)
      double *la = dyn_lds + A.lds_inv_off + (int64_t)grp * A.lds_inv;
      double *lo = la + ns * ns;
      int *ipv = s_ipiv[grp], *ipinv = s_ipinv[grp];
      for (int e = lane; e < ns * ns; e += G) {
        int i = e % ns, k = e / ns;
        la[i + ns * k] = oc[((int64_t)(off + i) * noccmax + cst(S.iref[s])[k]) * 5];
      }
      for (int e = lane; e < ns * nocc; e += G) {
        int i = e % ns, j = e / ns;
        lo[e] = oc[((int64_t)(off + i) * noccmax + j) * 5];
      }
      __syncthreads();
      bool singular;
      double det = group_gj_lds<G>(la, ns, lane, ipv, &singular);
      if (lane == 0) ws[L.deta[s]] = det;
      for (int i = lane; i < ns; i += G) ipinv[ipv[i]] = i;
      __syncthreads();
      // the inverse: ai(r, ipiv[c]) = la(ipiv[r], c)
      for (int q = lane; q < ns * ns; q += G) {
        int r = q % ns, c = q / ns;
        ai[r + ipv[c] * ns] = singular ? 0.0 : la[ipv[r] + c * ns];
      }
      // T(k, occ_j) = sum_i Ainv(k, i) orb(i, occ_j), Ainv(k, i) = la(ipiv[k], ipinv[i])
      for (int e = lane; e < ns * nocc; e += G) {
        int k = e / nocc, j = e % nocc;
        double acc = 0.0;
        if (!singular)
          for (int i = 0; i < ns; i++) acc += la[ipv[k] + ns * ipinv[i]] * lo[i + ns * j];
        tocc[e] = acc;
      }
      __syncthreads();
    } else {
    // A(i, k) = orb(i, ref_k): position of ref_k in occ is iref[k]
    for (int e = lane; e < ns * ns; e += G) {
      int i = e % ns, k = e / ns;
      ai[i + ns * k] = oc[((int64_t)(off + i) * noccmax + cst(S.iref[s])[k]) * 5];
    }
    __syncthreads();
    double det = group_matinv<G>(ai, tmp, ns, lane, s_ipiv[grp]);
    if (lane == 0) ws[L.deta[s]] = det;
    // T(k, occ_j) = sum_i Ainv(k, i) orb(i, occ_j)
    for (int e = lane; e < ns * nocc; e += G) {
      int k = e / nocc, j = e % nocc;
      double acc = 0.0;
      for (int i = 0; i < ns; i++)
        acc += ai[k + ns * i] * oc[((int64_t)(off + i) * noccmax + j) * 5];
      tocc[e] = acc;
    }
    __syncthreads();
    }
    // alpha blocks of the unique dets
    double *inv = w_inv(L, ws, s), *dv = w_detv(L, ws, s);
    for (int u = lane; u < S.ndetu[s]; u += G) {
      int k = cst(S.ord[s])[u];
      double *blk = inv + u * S.kmax2;
      for (int ii = 0; ii < k; ii++)
        for (int jj = 0; jj < k; jj++)
          blk[jj + k * ii] = tocc[cst(S.rf[s])[u * ns + jj] * nocc + cst(S.exo[s])[u * ns + ii]];
      dv[u] = small_matinv(blk, k);
    }
    __syncthreads();
  }
  {
    const double *invs[2] = {w_inv(L, ws, 0), w_inv(L, ws, 1)};
    const double *dvs[2] = {w_detv(L, ws, 0), w_detv(L, ws, 1)};
    double *ys[2] = {w_y(L, ws, 0), w_y(L, ws, 1)};
    double chi = group_eval_y<G>(S, invs, dvs, ys, ss + SL.cup, lane);
    if (lane == 0) ws[L.chi] = chi;
    __syncthreads();
    for (int s = 0; s < 2; s++)
      group_g<G>(S, s, w_ai(L, ws, s), w_tocc(L, ws, s), w_y(L, ws, s), ss + SL.g[s], tmp, lane);
  }
  // ---- grad / laplacian of ln D (qmc/vmc/determinant.f90:99-152)
  double d2d = 0.0;
  for (int i = lane; i < N; i += G) {
    int s = i < S.nup ? 0 : 1, ie = s ? i - S.nup : i;
    int ns = S.ns[s], nocc = S.nocc[s];
    const double *g = ss + SL.g[s];
    double dx = 0, dy = 0, dz = 0, ek = 0;
    for (int j = 0; j < nocc; j++) {
      double gj = g[j * ns + ie];
      const double *o = oc + ((int64_t)i * noccmax + j) * 5;
      dx += gj * o[1]; dy += gj * o[2]; dz += gj * o[3]; ek += gj * o[4];
    }
    vd[3 * i] = dx; vd[3 * i + 1] = dy; vd[3 * i + 2] = dz;
    double dvd = ek - (dx * dx + dy * dy + dz * dz);
    divv[i] += dvd;
    d2d += dvd;
  }
  d2d = gsum<G>(d2d);
  if (lane == 0) red[54] = d2d;
  // projected coefficients gp_i = G(:,i)^T coef_occ of the pseudopotential
  // (k_gproj_mfma instead, after this kernel, for large determinants)
  if (S.nloc > 0 && !A.gproj_mfma) group_gproj<G>(S, ss + SL.gproj, ss + SL.g[0], ss + SL.g[1], A.bf16, lane);
}
template <typename T, int G>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : QMCW_WPE, 8)))
void k_full_det(KArgs A) {
  full_det_b<T, G>(A, blockIdx.x);
}

#ifndef QMCW_ORD   // launched from the host translation unit only
typedef double gp_d4 __attribute__((ext_vector_type(4)));

// Projected coefficients of the pseudopotential on FP64 MFMA:
// gp[i][b] = sum_j G(j, i) coef_occ(b, j) for every walker w < nw and
// electron i of spin blockIdx.z, as one GEMM [nw*ns x nocc] x [nocc x NB]
// whose rows are the (walker, electron) pairs in walker order: block
// blockIdx.x takes 64 consecutive rows (16 walkers of a 4-electron spin, or
// one walker and part of the next for 49 electrons) and blockIdx.y a block of
// 64 basis functions, in v_mfma_f64_16x16x4f64 tiles, 16 rows by 16
// functions, k over the occupied orbitals four at a time.  The coefficients
// (the B operand) are shared by every walker.  One wavefront per block.
// Each sum runs in the MFMA's order, not group_gproj's, and is the same
// whatever rows share its wavefront.
__device__ __forceinline__ void gproj_mfma_b(const KArgs &A, int nw, int bx, int cb, int sp) {
  const DevSys &S = A.S;
  const int ns = S.ns[sp], nocc = S.nocc[sp], NB = S.NB, off = sp ? S.nup : 0;
  const int64_t rows = (int64_t)nw * ns, row0 = (int64_t)bx * 64;
  if (ns == 0 || row0 >= rows) return;
  const int64_t stride = A.SL.stride;
  const double *__restrict__ co = S.coef_occ[sp];        // coef(b, occ_j) at co[b * nocc + j]
  // operand layout: lane l supplies A(row l % 16, k = l / 16) and B(k = l / 16,
  // column l % 16); it receives D rows l / 16 + 4 q (q < 4) of column l % 16
  const int lane = threadIdx.x, r = lane & 15, kq = lane >> 4;
  const int b0 = cb * 64;
  // the G column of each of this lane's four A rows: G(j, i) of walker w at
  // sc[w] + SL.g[sp] + j * ns + i
  const double *__restrict__ ga[4];
  bool va[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int64_t rho = row0 + 16 * t + r;
    va[t] = rho < rows;
    const int64_t w = va[t] ? rho / ns : 0, i = va[t] ? rho % ns : 0;
    ga[t] = A.sc + w * stride + A.SL.g[sp] + i;
  }
  gp_d4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int u = 0; u < 4; u++) acc[t][u] = gp_d4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < nocc; k0 += 4) {
    const int kk = k0 + kq;
    double a[4], bb[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      a[t] = (va[t] && kk < nocc) ? ga[t][(int64_t)kk * ns] : 0.0;
      const int b = b0 + 16 * t + r;
      bb[t] = (b < NB && kk < nocc) ? co[(int64_t)b * nocc + kk] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int u = 0; u < 4; u++)
        acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], bb[u], acc[t][u], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int64_t rho = row0 + 16 * t + kq + 4 * q;
      if (rho >= rows) continue;
      const int64_t w = rho / ns, i = rho % ns;
      double *__restrict__ gpr = A.sc + w * stride + A.SL.gproj + (off + i) * NB;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int b = b0 + 16 * u + r;
        if (b < NB) {
          const double v = acc[t][u][q];
          gpr[b] = A.bf16 ? (double)bf16r((float)v) : v;
        }
      }
    }
}
__global__ __launch_bounds__(64) void k_gproj_mfma(KArgs A, int nw) {
  gproj_mfma_b(A, nw, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Every orbital of every electron (value, gradient, laplacian) on FP64
// MFMA for the derivatives: orbf[i][q][c] = sum_b coef(b, q) phi_i[b][c] for
// every walker w < nw, one GEMM [nw*N*5 x NB] x [NB x NO] whose rows are the
// (walker, electron, component) triples in walker order and whose B operand
// (the coefficients) every walker shares; block blockIdx.x takes 64 rows
// and blockIdx.y 64 orbitals, in the tiles of k_gproj_mfma.  The sums run in
// the MFMA's order, not contract5's.
__global__ __launch_bounds__(64) void k_orb_mfma(KArgs A, int nw) {
  const DevSys &S = A.S;
  const int N = S.N, NB = S.NB, NO = S.NO;
  const int64_t rows = (int64_t)nw * N * 5, row0 = (int64_t)blockIdx.x * 64;
  if (row0 >= rows) return;
  const int64_t stride = A.SL.stride;
  const cptr<double> co = cst(S.coefT);                  // coef(b, q) at co[b * NO + q]
  const int lane = threadIdx.x, r = lane & 15, kq = lane >> 4;
  const int c0 = blockIdx.y * 64;
  // A(row, b) of the row (w, i, c) at phiall_w[(i * NB + b) * 5 + c]
  const double *__restrict__ ga[4];
  bool va[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int64_t rho = row0 + 16 * t + r;
    va[t] = rho < rows;
    const int64_t w = va[t] ? rho / (N * 5) : 0, ic = va[t] ? rho % (N * 5) : 0;
    ga[t] = A.sc + w * stride + A.DL.phiall + (ic / 5) * NB * 5 + ic % 5;
  }
  gp_d4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int u = 0; u < 4; u++) acc[t][u] = gp_d4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < NB; k0 += 4) {
    const int kk = k0 + kq;
    double a[4], bb[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      a[t] = (va[t] && kk < NB) ? ga[t][(int64_t)kk * 5] : 0.0;
      const int q = c0 + 16 * t + r;
      bb[t] = (q < NO && kk < NB) ? co[(int64_t)kk * NO + q] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int u = 0; u < 4; u++)
        acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], bb[u], acc[t][u], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int q4 = 0; q4 < 4; q4++) {
      const int64_t rho = row0 + 16 * t + kq + 4 * q4;
      if (rho >= rows) continue;
      const int64_t w = rho / (N * 5), ic = rho % (N * 5), i = ic / 5, c = ic % 5;
      double *__restrict__ o = A.sc + w * stride + A.SL.orb + i * NO * 5 + c;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int q = c0 + 16 * u + r;
        if (q < NO) o[(int64_t)q * 5] = acc[t][u][q4];
      }
    }
}

// psp_nonloc_orb of the derivatives on FP64 MFMA: pnl[i][q] = sum_b
// coef(b, q) pbc_i[b] for every walker w < nw and electron i, one GEMM
// [nw*N x NB] x [NB x NO] with rows = (walker, electron) pairs in walker
// order (pbc_i: the electron's weighted basis sums, in the row of centre 0
// of its pairs) and the coefficients shared by every walker; block
// blockIdx.x takes 64 rows and blockIdx.y 64 orbitals, in the tiles of
// k_gproj_mfma.
__global__ __launch_bounds__(64) void k_pnl_mfma(KArgs A, int nw) {
  const DevSys &S = A.S;
  const int N = S.N, M = S.M, NB = S.NB, NO = S.NO;
  const int64_t rows = (int64_t)nw * N, row0 = (int64_t)blockIdx.x * 64;
  if (row0 >= rows) return;
  const int64_t stride = A.SL.stride;
  const cptr<double> co = cst(S.coefT);                  // coef(b, q) at co[b * NO + q]
  const int lane = threadIdx.x, r = lane & 15, kq = lane >> 4;
  const int c0 = blockIdx.y * 64;
  const double *__restrict__ ga[4];
  bool va[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int64_t rho = row0 + 16 * t + r;
    va[t] = rho < rows;
    const int64_t w = va[t] ? rho / N : 0, i = va[t] ? rho % N : 0;
    ga[t] = A.sc + w * stride + A.DL.pbc + i * M * NB;
  }
  gp_d4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int u = 0; u < 4; u++) acc[t][u] = gp_d4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < NB; k0 += 4) {
    const int kk = k0 + kq;
    double a[4], bb[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      a[t] = (va[t] && kk < NB) ? ga[t][kk] : 0.0;
      const int q = c0 + 16 * t + r;
      bb[t] = (q < NO && kk < NB) ? co[(int64_t)kk * NO + q] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int u = 0; u < 4; u++)
        acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], bb[u], acc[t][u], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int q4 = 0; q4 < 4; q4++) {
      const int64_t rho = row0 + 16 * t + kq + 4 * q4;
      if (rho >= rows) continue;
      const int64_t w = rho / N, i = rho % N;
      double *__restrict__ o = A.sc + w * stride + A.DL.pnl + i * NO;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int q = c0 + 16 * u + r;
        if (q < NO) o[q] = acc[t][u][q4];
      }
    }
}
#endif

// rotqua (qmc/vmc/rotqua.f90): a random rotation drawn from the stream st4,
// as the axes (x1, x2, x3), (y1, y2, y3), (z1, z2, z3) that take the
// reference grid q0 to q = q0x x + q0y y + q0z z.
__device__ __forceinline__ void rot_axes(int *st4, double *ax) {
  double x1, x2, x3, xsum;
  do {
    x1 = 1.0 - 2.0 * rannyu(st4);
    x2 = 1.0 - 2.0 * rannyu(st4);
    xsum = x1 * x1 + x2 * x2;
  } while (xsum >= 1.0);
  double xsum2 = 2.0 * sqrt(fabs(1.0 - xsum));
  x1 *= xsum2;
  x2 *= xsum2;
  x3 = 1.0 - 2.0 * xsum;
  double theta = acos(x3), sthet = sin(theta), cfi, sfi;
  if (sthet < 1e-5) {
    cfi = 1; sfi = 0; x1 = 0; x2 = 0; x3 = 1; sthet = 0;
  } else {
    cfi = x1 / sthet;
    sfi = x2 / sthet;
  }
  double yy1 = x3 * cfi, yy2 = x3 * sfi, yy3 = -sthet;
  double zz1 = -sfi, zz2 = cfi, zz3 = 0;
  double u1, u2, usum;
  do {
    u1 = rannyu(st4) * 2.0 - 1.0;
    u2 = rannyu(st4) * 2.0 - 1.0;
    usum = u1 * u1 + u2 * u2;
  } while (usum >= 1.0);
  double uu = sqrt(usum);
  u1 /= uu;
  u2 /= uu;
  ax[0] = x1; ax[1] = x2; ax[2] = x3;
  ax[3] = yy1 * u1 + zz1 * u2; ax[4] = yy2 * u1 + zz2 * u2; ax[5] = yy3 * u1 + zz3 * u2;
  ax[6] = yy1 * u2 - zz1 * u1; ax[7] = yy2 * u2 - zz2 * u1; ax[8] = yy3 * u2 - zz3 * u1;
}

// Pseudopotential quadrature points of the in-range pairs plist[0..npair)
// taken by this lane: sum of w exp(dJ) v P_l (Psi_D ratio).  rnew: 4M
// doubles and uq: N values of per-lane scratch (LDS or global), stored
// lane-interleaved.
template <typename T, int G, int ORD, bool PAIR>
__device__ __forceinline__ double nlpp_points(const KArgs &A, const double *ws, const double *ss,
                                              int npair, const int *plist, LaneArr<double> rnew,
                                              LaneArr<T> uq, double *tqw, int lane) {
  const DevSys &S = A.S;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const int N = S.N, M = S.M;
  const double *x = ws + L.x, *rven = ws + L.rvec_en, *ren = ws + L.r_en, *fso = ws + L.fso;
  const double *vnl = ss + SL.vnl, *foldv = ss + SL.fold;
  const T *bt = reinterpret_cast<const T *>(ss + SL.bt);
  const double *gp = ss + SL.gproj;
  const double *quad = ws + L.quad;
  const int nq = S.nquad;
  const LaneArr<double> rvnew = rnew + M;
  double vpsp = 0.0;
  const double *prot = ss + SL.prot;
  for (int t = lane; t < npair * nq; t += G) {
    const int iq = t % nq, p = plist[t / nq], ic = p % M, i = p / M;
    const double *v = vnl + p * (QMCW_MAX_LPOT + 1);
    const int ict = cst(S.iwctype)[ic];
    const double rei = ren[p];
    double qx, qy, qz;
    if (A.dmc_rot) {
      // this pair's own rotation
      const double *ax = prot + 10 * (t / nq);
      const double a = cst(S.q0)[iq], b = cst(S.q0)[nq + iq], c = cst(S.q0)[2 * nq + iq];
      qx = a * ax[0] + b * ax[3] + c * ax[6];
      qy = a * ax[1] + b * ax[4] + c * ax[7];
      qz = a * ax[2] + b * ax[5] + c * ax[8];
    } else {
      qx = quad[iq]; qy = quad[nq + iq]; qz = quad[2 * nq + iq];
    }
    const double *rv0 = rven + 3 * p;
    const double costh = (rv0[0] * qx + rv0[1] * qy + rv0[2] * qz) / rei;
    const double xn[3] = {rei * qx + cst(S.cent)[3 * ic], rei * qy + cst(S.cent)[3 * ic + 1],
                          rei * qz + cst(S.cent)[3 * ic + 2]};
    for (int jc = 0; jc < M; jc++) {
      const double dx = xn[0] - cst(S.cent)[3 * jc], dy = xn[1] - cst(S.cent)[3 * jc + 1],
                   dz = xn[2] - cst(S.cent)[3 * jc + 2];
      rvnew[3 * jc] = dx; rvnew[3 * jc + 1] = dy; rvnew[3 * jc + 2] = dz;
      rnew[jc] = (jc == ic) ? rei : sqrt(dx * dx + dy * dy + dz * dz);
    }
    // determinant ratio of moving electron i to the point
    double rd = 0.0;
    const double *gpi = gp + (int64_t)i * S.NB;
    for (int sh = 0; sh < S.nshell; sh++) {
      const int jc = cst(S.sh_cent)[sh];
      rd += (double)shell_dot<T, double>(S, sh, (T)rvnew[3 * jc], (T)rvnew[3 * jc + 1],
                                         (T)rvnew[3 * jc + 2], (T)rnew[jc], gpi);
    }
    // Jastrow value change against the electron's current pair sum: kept
    // per electron for small N; summed here for large N, where the sum's
    // loads overlap the point's other work (measured faster on the chain)
    double fold;
    if (N <= 32) {
      fold = foldv[i];
    } else {
      fold = fso[i * N + i];
      for (int jj = 0; jj < N; jj++)
        if (jj != i) fold += fso[jj < i ? i * N + jj : jj * N + i];
    }
    const double delta = (double)jas_point_val<T, ORD, PAIR>(S, i, xn, x, rnew, bt, uq) - fold;
    double vpot = 0.0;
    for (int l = 1; l <= cst(S.npotd)[ict]; l++)
      if (l != cst(S.lpotp1)[ict]) vpot += v[l - 1] * legendre_w(l, costh);
    vpsp += cst(S.q0)[3 * nq + iq] * exp(delta) * vpot * rd;
    if (tqw) {
      // kept for T-moves: the point and the Psi ratio there (rank j of the
      // pair among electron i's pairs in range)
      const int j = (int)prot[10 * (t / nq) + 9];
      if (j >= 0) {
        tqw[tq_ratio(S, i, j, iq)] = rd * exp(delta);
        double *qp = tqw + tq_point(S, i, j, iq);
        qp[0] = qx; qp[1] = qy; qp[2] = qz;
      }
    }
  }
  return vpsp;
}

template <typename T, int G, int ORD, bool PAIR = false>
__device__ __forceinline__ void full_nlpp_b(const KArgs &A, int bid_) {
  QMCW_GROUP_SETUP_B(bid_)
  const int N = S.N, M = S.M, NO = S.NO;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const double *x = ws + L.x;
  double *rven = ws + L.rvec_en, *ren = ws + L.r_en, *rvee = ws + L.rvec_ee, *ree = ws + L.r_ee;
  double *fso = ws + L.fso, *fijo = ws + L.fijo, *lapj = ws + L.lapj, *fjo = ws + L.fjo;
  double *vj = ss + SL.vj, *divv = ss + SL.divv, *vd = ss + SL.vd;
  double *red = ss + SL.phs;
  (void)NO; (void)x; (void)rven; (void)ren; (void)rvee; (void)ree; (void)fso; (void)fijo;
  (void)lapj; (void)fjo; (void)vj; (void)divv; (void)vd; (void)red;
  // ---- pseudopotential (qmc/vmc/nonloc_pot.f90, nonloc.f90).
  // V_nl = sum_G psp_nonloc_orb is linear in the orbitals, so per quadrature
  // point only the determinant ratio sum_j G(j,i) phi_occj = sum_b gp_i[b]
  // phi_b is needed, with gp_i = G(:,i)^T coef_occ formed once per electron.
  double vpsp = 0.0, ploc = 0.0;
  if (S.nloc > 0) {
    double *vnl = ss + SL.vnl;        // [N*M][QMCW_MAX_LPOT + 1] (last: in range)
    T *bt = reinterpret_cast<T *>(ss + SL.bt);
    for (int p = lane; p < N * M; p += G) {
      int ic = p % M, ict = cst(S.iwctype)[ic];
      double r = fmax(1e-10, ren[p]);
      double *v = vnl + p * (QMCW_MAX_LPOT + 1);
      bool inr = false;
      for (int l = 1; l <= cst(S.npotd)[ict]; l++) {
        double vl = ecp_radial(S, r, l - 1, ict);
        if (l == cst(S.lpotp1)[ict]) { ploc += vl - cst(S.znuc)[ict] / r; vl = 0.0; }
        else if (fabs(vl) > 1e-4) inr = true;
        v[l - 1] = vl;
      }
      v[QMCW_MAX_LPOT] = inr ? 1.0 : 0.0;
      // e-e-n variable of the other electrons at their current positions,
      // centre by centre (bt[ic * N + j])
      T bu, t1, t2;
      scaled<T>(S.isc, (T)S.scalek, (T)ren[p], bu, t1, t2);
      bt[ic * N + p / M] = T(1) - (T)S.scalek * bu;
    }
    ploc = gsum<G>(ploc);
    // the pair sum of the Jastrow of each electron at its current position
    // (what a quadrature point's value is compared with): one pass over the
    // lower triangle of the pair values, row by row with the lanes along
    // the row, so every load is contiguous.  Lane jj % G owns foldv[jj]:
    // it adds the column terms as the rows pass and the row total of row
    // jj (summed over the lanes) when its row comes.
    double *foldv = ss + SL.fold;
    for (int jj = lane; jj < N; jj += G) foldv[jj] = fso[jj * N + jj];
    for (int i = 1; i < N && N <= 32; i++) {
      double rs = 0.0;
      for (int jj = lane; jj < i; jj += G) {
        const double v = fso[i * N + jj];
        foldv[jj] += v;
        rs += v;
      }
      rs = gsum<G>(rs);
      if (lane == i % G) foldv[i] += rs;
    }
    __syncthreads();
    // quadrature points only of the pairs in range, spread evenly over the lanes
    int *plist = reinterpret_cast<int *>(ss + SL.plist);
    const int npair = group_compact<G>(vnl + QMCW_MAX_LPOT, QMCW_MAX_LPOT + 1, N * M, plist, lane);
    // DMC (qmc/vmc/nonloc.f90:89): every pair in range rotates the grid anew,
    // drawing from the walker's stream in pair order.  rotqua's rejection
    // loops make the number of draws vary, so one lane walks the stream and
    // keeps each pair's axes; walkers that moved no electron are not
    // evaluated by the reference and draw nothing.
    const bool mover = A.dmc_rot && (!A.dmove || A.dmove[(int64_t)w * A.dmove_stride] > 0.0);
    if (mover && lane == 0) {
      int *rs = A.rng + 4 * w;
      double *pax = ss + SL.prot;
      int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
      int prev = -1, j = 0;
      for (int k = 0; k < npair; k++) {
        rot_axes(st4, pax + 10 * k);
        const int i = plist[k] / M;
        j = (i == prev) ? j + 1 : 0;
        prev = i;
        if (j >= S.tm_kmax) A.flags[w] |= 4;     // more pairs than the T-move record holds
        pax[10 * k + 9] = j < S.tm_kmax ? (double)j : -1.0;
      }
      rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
    }
    __syncthreads();
    double *tqw = (mover && A.tq) ? A.tq + (int64_t)w * A.tq_stride : nullptr;
    // per-lane distances of the point and U of the other electrons: LDS
    // when the launch reserved it.  In DMC a walker that moved no electron
    // drew no rotations (its local energy is the one kept from the start of
    // the generation), so its points are not evaluated.
    if (A.dmc_rot && !mover) {
      vpsp = 0.0;
    } else if (A.lds_nlpp) {
      // the block's 64 lanes interleaved: 4M doubles, then N values
      extern __shared__ double dyn_lds[];
      const LaneArr<double> rq{dyn_lds + threadIdx.x, kWave};
      const LaneArr<T> uq{reinterpret_cast<T *>(dyn_lds + 4 * M * kWave) + threadIdx.x, kWave};
      vpsp = nlpp_points<T, G, ORD, PAIR>(A, ws, ss, npair, plist, rq, uq, tqw, lane);
    } else {
      // the walker's G lanes interleaved in its scratch
      const LaneArr<double> rq{ss + SL.qd + lane, G};
      const LaneArr<T> uq{reinterpret_cast<T *>(ss + SL.uq) + lane, G};
      vpsp = nlpp_points<T, G, ORD, PAIR>(A, ws, ss, npair, plist, rq, uq, tqw, lane);
    }
    vpsp = gsum<G>(vpsp);
  }
  if (lane == 0) { red[55] = ploc; red[56] = vpsp; }
}
// FP32 instances are held to 128 VGPRs (four waves per SIMD); PAIR: the
// point Jastrow two centres at a time (systems of many centres)
template <typename T, int G, int ORD, bool PAIR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : QMCW_WPE, 8)))
void k_full_nlpp(KArgs A) {
  full_nlpp_b<T, G, ORD, PAIR>(A, blockIdx.x);
}

// T-moves: Psi ratios of moving electron ie to points of the walker's
// current state, the linear form of the pseudopotential evaluation
// (sum_b gp_i[b] phi_b(x) times the Jastrow factor; qmc/dmc/dmc_elec/
// tmove.f90:207-255 computes the same ratio by a full one-electron update).
// A walker's G matrices and the e-e-n variables bt are rebuilt from its
// state first when its record is marked stale (after a committed T-move, or
// after branching moved it to another scratch slot); the projection gp of
// electron ie is formed on every call.
// pts/out: column c of walker w at c * cw + w (3 and 8 doubles);
// cnt[w]: the walker's number of points (0: nothing to do).
template <typename T, int G, int ORD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(QMCW_WPE, 8)))
void k_tm_ratios(KArgs A, int ie, int cw, const double *pts, const int *cnt, double *out,
                 double *stale) {
  QMCW_GROUP_SETUP
  const int N = S.N, M = S.M, nq = S.nquad;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const int n = cnt[w];
  if (n == 0) return;                  // the whole group leaves together
  const double *ren = ws + L.r_en, *rven = ws + L.rvec_en, *fso = ws + L.fso, *x = ws + L.x;
  T *bt = reinterpret_cast<T *>(ss + SL.bt);
  if (stale[w] != 0.0) {
    for (int sp = 0; sp < 2; sp++)
      group_g<G>(S, sp, w_ai(L, ws, sp), w_tocc(L, ws, sp), w_y(L, ws, sp), ss + SL.g[sp], ss + SL.red, lane);
    for (int p = lane; p < N * M; p += G) {
      T bu, t1, t2;
      scaled<T>(S.isc, (T)S.scalek, (T)ren[p], bu, t1, t2);
      bt[(p % M) * N + p / M] = T(1) - (T)S.scalek * bu;
    }
    __syncthreads();
    if (lane == 0) stale[w] = 0.0;
  }
  // the projected coefficients of electron ie only (the other rows are not
  // read before the next full evaluation rebuilds them all)
  group_gproj<G>(S, ss + SL.gproj, ss + SL.g[0], ss + SL.g[1], A.bf16, lane, ie);
  __syncthreads();
  const LaneArr<double> rnew{ss + SL.qd + lane, G}, rvnew = rnew + M;
  const LaneArr<T> uq{reinterpret_cast<T *>(ss + SL.uq) + lane, G};
  const double *gpi = ss + SL.gproj + (int64_t)ie * S.NB;
  double fold = fso[ie * N + ie];
  for (int jj = 0; jj < N; jj++)
    if (jj != ie) fold += fso[jj < ie ? ie * N + jj : jj * N + ie];
  for (int c = lane; c < n; c += G) {
    const double *xn = pts + 3 * ((int64_t)c * cw + w);
    for (int jc = 0; jc < M; jc++) {
      const double dx = xn[0] - cst(S.cent)[3 * jc], dy = xn[1] - cst(S.cent)[3 * jc + 1],
                   dz = xn[2] - cst(S.cent)[3 * jc + 2];
      rvnew[3 * jc] = dx; rvnew[3 * jc + 1] = dy; rvnew[3 * jc + 2] = dz;
      rnew[jc] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    double rd = 0.0;
    for (int sh = 0; sh < S.nshell; sh++) {
      const int jc = cst(S.sh_cent)[sh];
      rd += (double)shell_dot<T, double>(S, sh, (T)rvnew[3 * jc], (T)rvnew[3 * jc + 1],
                                         (T)rvnew[3 * jc + 2], (T)rnew[jc], gpi);
    }
    const double delta = (double)jas_point_val<T, ORD>(S, ie, xn, x, rnew, bt, uq) - fold;
    out[8 * ((int64_t)c * cw + w)] = rd * exp(delta);
  }
  (void)nq; (void)rven;
}

template <int G>
__device__ __forceinline__ void full_fin_b(const KArgs &A, int bid_) {
  QMCW_GROUP_SETUP_B(bid_)
  const int N = S.N, M = S.M, NO = S.NO;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const double *x = ws + L.x;
  double *rven = ws + L.rvec_en, *ren = ws + L.r_en, *rvee = ws + L.rvec_ee, *ree = ws + L.r_ee;
  double *fso = ws + L.fso, *fijo = ws + L.fijo, *lapj = ws + L.lapj, *fjo = ws + L.fjo;
  double *vj = ss + SL.vj, *divv = ss + SL.divv, *vd = ss + SL.vd;
  double *red = ss + SL.phs;
  (void)NO; (void)x; (void)rven; (void)ren; (void)rvee; (void)ree; (void)fso; (void)fijo;
  (void)lapj; (void)fjo; (void)vj; (void)divv; (void)vd; (void)red;
  // ---- local energy (qmc/vmc/hpsi.f90:336-344)
  const double pe_ee = red[51], fsum = red[52], d2j = red[53], d2d = red[54];
  const double vpsp = S.nloc > 0 ? red[56] : 0.0;
  const double pe = S.pecent + red[50] + pe_ee + (S.nloc > 0 ? red[55] + vpsp : 0.0);
  double v2 = 0.0;
  for (int i = lane; i < N; i += G)
    for (int q = 0; q < 3; q++) {
      double v = vj[3 * i + q] + vd[3 * i + q];
      v2 += v * v;
    }
  v2 = gsum<G>(v2);
  double d2lnpsi = d2d + d2j;
  double eloc = pe - S.hb * d2lnpsi - S.hb * v2;
  if (lane == 0) {
    double chi = ws[L.chi], du = ws[L.deta[0]], dd = ws[L.deta[1]];
    double psid = chi * du * dd;
    double lp = log(fabs(chi)) + log(fabs(du)) + log(fabs(dd));
    ws[L.logpsid] = lp;
    ws[L.sgn] = psid < 0 ? -1.0 : 1.0;
    ws[L.psij] = fsum;
    ws[L.psi2o] = 2.0 * (lp + fsum);
    ws[L.eloc] = eloc;
    ws[L.epot] = pe;
    red[60] = d2lnpsi;
    red[61] = vpsp;
    red[62] = pe_ee;
    int fl = 0;
    if (psid == 0.0) fl |= 1;
    if (!isfinite(eloc)) fl |= 2;
    A.flags[w] |= fl;           // k_full_jas cleared the word; k_full_nlpp may have set bit 4
  }
}
template <int G>
__global__ __launch_bounds__(64) void k_full_fin(KArgs A) {
  full_fin_b<G>(A, blockIdx.x);
}

// ---------------------------------------------------------------------------
// gradient of ln Psi of electron ie in the current state (eval_grad +
// Jastrow), qmc/dete_mod.f90:249-292 and qmc/psi_utils_mod.f90:457-483.
// Also the laplacian (lap Psi)/Psi when lap != nullptr.  All lanes return.
template <int G>
__device__ void group_grad_1e(const KArgs &A, const double *ws, int ie,
                              const double *orbv, const double *ai,
                              const double *tocc, const double *y,
                              const double *fj, double lapj, double *grad,
                              double *lapout, int lane) {
  const DevSys &S = A.S;
  int s = ie < S.nup ? 0 : 1, ns = S.ns[s], nocc = S.nocc[s];
  int e = s ? ie - S.nup : ie;
  // grad(c) = sum_k ai(k,e) d_c(ref_k) + sum_i sum_j y(j,i) ai(i,e) v_c(j)
  //   v_c(j) = d_c(occ_j) - sum_k d_c(ref_k) T(k, occ_j)
  double acc[4] = {0, 0, 0, 0};
  const int *iref = S.iref[s];
  for (int k = lane; k < ns; k += G) {
    double a = ai[k + ns * e];
    const double *o = orbv + iref[k] * 5;
    for (int c = 0; c < 4; c++) acc[c] += a * o[1 + c];
  }
  // w_i = ai(i, e) ; z_j = sum_i y(j,i) w_i ; the four components of v(j)
  // in one pass over the reference orbitals (each keeps its own order).
  // A single determinant has Y = 0: every z_j is zero and the loop adds
  // nothing.
  const double *aie = ai + ns * e;
  for (int j = S.single_det ? nocc : lane; j < nocc; j += G) {
    double z = 0.0;
#pragma unroll 4
    for (int i = 0; i < ns; i++) z += y[j * ns + i] * aie[i];
    if (z == 0.0) continue;
    double v0 = orbv[j * 5 + 1], v1 = orbv[j * 5 + 2], v2 = orbv[j * 5 + 3], v3 = orbv[j * 5 + 4];
#pragma unroll 4
    for (int k = 0; k < ns; k++) {
      const double *o = orbv + iref[k] * 5;
      const double t = tocc[k * nocc + j];
      v0 -= o[1] * t; v1 -= o[2] * t; v2 -= o[3] * t; v3 -= o[4] * t;
    }
    acc[0] += z * v0; acc[1] += z * v1; acc[2] += z * v2; acc[3] += z * v3;
  }
  for (int c = 0; c < 4; c++) acc[c] = gsum<G>(acc[c]);
  for (int c = 0; c < 3; c++) grad[c] = acc[c] + fj[c];
  if (lapout)
    *lapout = acc[3] + 2.0 * (fj[0] * acc[0] + fj[1] * acc[1] + fj[2] * acc[2]) + lapj;
}

__device__ __forceinline__ int metrop_accept(const KArgs &A, const double *ws, const double *mv,
                                             int w, int ie, int ns, int nocc, double delta,
                                             double fbias);
__device__ __forceinline__ void vmc_propose_w(const KArgs &A, int w, int ie, const double *ws,
                                              const double *vold, double delta, double fbias,
                                              double *xn, double *mv);
template <int G>
__device__ void group_commit(const KArgs &A, double *ws, const double *mv, int ie, int lane);

// Sherman-Morrison vectors of dete_update (qmc/dete_mod.f90:114-214):
// vi1(i) = sum_k orb(ref_k) A^-1(k, i) - delta(i, e) and
// vi2(j) = sum_k orb(ref_k) T(k, j) - orb(occ_j).  Outputs never alias inputs.
template <int G>
__device__ __forceinline__ void sm_vectors(double *__restrict__ vi1, double *__restrict__ vi2,
                                           const double *__restrict__ orbe, const double *__restrict__ ai,
                                           const double *__restrict__ tocc, const int *__restrict__ iref,
                                           int ns, int nocc, int e, int lane, bool no_t = false) {
  for (int i = lane; i < ns; i += G) {
    double acc = 0.0;
#pragma unroll 4
    for (int k = 0; k < ns; k++) acc += orbe[5 * iref[k]] * ai[k + ns * i];
    vi1[i] = acc - (i == e ? 1.0 : 0.0);
  }
  if (no_t) return;
  for (int j = lane; j < nocc; j += G) {
    double acc = 0.0;
#pragma unroll 4
    for (int k = 0; k < ns; k++) acc += orbe[5 * iref[k]] * tocc[k * nocc + j];
    vi2[j] = acc - orbe[5 * j];
  }
}

// rank-1 updates A^-1_new = A^-1 + u vi1^T, T_new = T + u vi2^T with
// u(k) = -A^-1(k, e) / ratio
template <int G>
__device__ __forceinline__ void sm_rank1(double *__restrict__ ain, double *__restrict__ tn,
                                         const double *__restrict__ ai, const double *__restrict__ tocc,
                                         const double *__restrict__ vi1, const double *__restrict__ vi2,
                                         int ns, int nocc, int e, double rat, int lane, bool no_t = false) {
#pragma unroll 4
  for (int q = lane; q < ns * ns; q += G) {
    const int k = q % ns, i = q / ns;
    ain[q] = ai[q] + (-ai[k + ns * e] / rat) * vi1[i];
  }
  if (no_t) return;
#pragma unroll 4
  for (int q = lane; q < ns * nocc; q += G) {
    const int k = q / nocc, j = q % nocc;
    tn[q] = tocc[q] + (-ai[k + ns * e] / rat) * vi2[j];
  }
}

// basis values (value, gradient, laplacian) of one position into phi, then
// the occupied orbitals of spin s into orbe[5 * nocc]
template <typename T, int G>
__device__ __forceinline__ void orbitals_1e(const KArgs &A, int s, int nocc, const double *ren,
                                            const double *rven, T *phi, double *orbe, int lane) {
  const DevSys &S = A.S;
  for (int sh = lane; sh < S.nshell; sh += G) {
    const int ic = cst(S.sh_cent)[sh];
    eval_shell<T>(S, sh, (T)rven[3 * ic], (T)rven[3 * ic + 1], (T)rven[3 * ic + 2], (T)ren[ic], 1, phi);
  }
  __syncthreads();
  for (int j = lane; j < nocc; j += G)
    contract5<T>(cst(S.coef_occ[s]), nocc, S.NB, j, phi, orbe + 5 * j);
}

// ---------------------------------------------------------------------------
// One-electron move of electron ie: propose (mode 1, metrop_mov1 with the
// walker's rannyu stream) or take xnew (mode 0), then distancese +
// jastrow4e + orbitals + dete_update + eval_grad into the pending buffers.
// out[w][8]: ratio, ln|psid_new|, sign, psij_new, grad(3), lap
template <typename T, int G, int ORD>
__device__ __forceinline__ void ratio_1e_b(const KArgs &A, int bid_, int ie, int mode,
                                           const double *xnew_in, double delta, double fbias,
                                           double *out, long long *nacc) {
  QMCW_GROUP_SETUP_B(bid_)
  // mode 4: commit the given move of the walkers that accepted it; the
  // others have nothing to do (a group leaves as a whole)
  if (mode == 4 && !A.acc_in[w]) return;
  const int N = S.N, M = S.M, noccmax = S.noccmax;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const MLayout &ML = A.ML;
  double *mv = ss + SL.mv;
  const int s = ie < S.nup ? 0 : 1, ns = S.ns[s], nocc = S.nocc[s];
  const int e = s ? ie - S.nup : ie;
  const double *x = ws + L.x;
  const double *fso = ws + L.fso, *fijo = ws + L.fijo, *lapj = ws + L.lapj, *fjo = ws + L.fjo;
  double *xn = mv + ML.xn;
  // ---- proposal (mode 1: ratio only; mode 2: ratio, test and commit;
  // mode 3: DMC proposal, test and commit, with A.dm and A.dmp; modes 0
  // and 4 take the position given)
  if (mode >= 1 && mode <= 3) {
    double lapo_dummy;
    double vold[3];
    double lapjo = 0.0;  // not needed for the proposal
    group_grad_1e<G>(A, ws, ie, ws + L.oc + (int64_t)ie * noccmax * 5, w_ai(L, ws, s),
                     w_tocc(L, ws, s), w_y(L, ws, s), fjo + 3 * ie, lapjo, vold, nullptr, lane);
    (void)lapo_dummy;
    if (mode == 3) {
      if (lane == 0) dmc_propose_w(A, w, ie, vold, A.dm + (int64_t)w * kDmStride, A.dmp, xn);
    } else if (lane == 0) {
      vmc_propose_w(A, w, ie, ws, vold, delta, fbias, xn, mv);
    }
  } else {
    if (lane < 3) xn[lane] = xnew_in[3 * w + lane];
  }
  __syncthreads();
  // ---- new distances of electron ie
  double *ren = mv + ML.ren, *rven = mv + ML.rven, *ree = mv + ML.ree, *rvee = mv + ML.rvee;
  for (int ic = lane; ic < M; ic += G) {
    double s2 = 0.0;
    for (int q = 0; q < 3; q++) {
      double d = xn[q] - cst(S.cent)[3 * ic + q];
      rven[3 * ic + q] = d;
      s2 += d * d;
    }
    ren[ic] = sqrt(s2);
  }
  for (int j = lane; j < N; j += G) {
    if (j == ie) continue;
    int hi = j < ie ? ie : j, lo = j < ie ? j : ie;
    double s2 = 0.0;
    for (int q = 0; q < 3; q++) {
      double xi = (hi == ie) ? xn[q] : x[3 * hi + q];
      double xj = (lo == ie) ? xn[q] : x[3 * lo + q];
      rvee[3 * j + q] = xi - xj;
      s2 += (xi - xj) * (xi - xj);
    }
    ree[j] = sqrt(s2);
  }
  __syncthreads();
  // ---- jastrow4e (+ per-electron laplacian)
  double *fsn = mv + ML.fsn, *fijn_i = mv + ML.fijn_i, *fijn_j = mv + ML.fijn_j,
         *lapn = mv + ML.lapn_i, *fjn = mv + ML.fjn;
  double dsum = 0.0, gix = 0, giy = 0, giz = 0, lapi = 0;
  // the e-e-n tables of the new position, one per centre, in the LDS region
  // the basis values take after the Jastrow row
  extern __shared__ double dyn_lds[];
  T *etab = reinterpret_cast<T *>(dyn_lds) + (int64_t)grp * A.lds_wst;
  if (A.een_tab) {
    een_tables<T, ORD, G>(S, ren, rven, etab, lane);
    __syncthreads();
  }
  for (int j = lane; j < N; j += G) {
    if (j == ie) {  // e-n diagonal term, on the lane the pair list leaves free
      T f, g[3], lp;
      jas_en<T>(S, ren, rven, f, g, lp);
      fsn[ie] = f;
      for (int q = 0; q < 3; q++) fijn_i[3 * ie + q] = g[q];
      lapn[ie] = lp;
      dsum += (double)f - fso[ie * N + ie];
      gix += (double)g[0] - fijo[3 * (ie * N + ie)];
      giy += (double)g[1] - fijo[3 * (ie * N + ie) + 1];
      giz += (double)g[2] - fijo[3 * (ie * N + ie) + 2];
      lapi += lp;
      continue;
    }
    int hi = j < ie ? ie : j, lo = j < ie ? j : ie;
    T rvij[3];
    for (int q = 0; q < 3; q++) rvij[q] = (T)rvee[3 * j + q];
    const double *rI = ren, *rvI = rven;
    PairOut<T> po;
    bool par = (hi < S.nup) || (lo >= S.nup);
    // moved electron's arrays go to whichever slot (hi or lo) it occupies
    // the other electron's distances formed from its position (see EnFromX)
    const EnRows pI{rI, rvI};
    const EnFromX pJ{x[3 * j], x[3 * j + 1], x[3 * j + 2]};
    if (A.een_tab) {
      const T rvmo[3] = {hi == ie ? rvij[0] : -rvij[0], hi == ie ? rvij[1] : -rvij[1],
                         hi == ie ? rvij[2] : -rvij[2]};
      jas_pair_tab<T, ORD>(S, par, (T)ree[j], rvmo, etab, pJ, hi == ie, po);
    } else if (hi == ie) {
      jas_pair_g<T, ORD>(S, par, (T)ree[j], rvij, pI, pJ, po);
    } else {
      jas_pair_g<T, ORD>(S, par, (T)ree[j], rvij, pJ, pI, po);
    }
    // old pair values before the stores to the pending buffers below
    const double f_old = fso[hi * N + lo];
    const double gi_old[3] = {fijo[3 * (ie * N + j)], fijo[3 * (ie * N + j) + 1], fijo[3 * (ie * N + j) + 2]};
    const double gj_old[3] = {fijo[3 * (j * N + ie)], fijo[3 * (j * N + ie) + 1], fijo[3 * (j * N + ie) + 2]};
    const double fj_old[3] = {fjo[3 * j], fjo[3 * j + 1], fjo[3 * j + 2]};
    // per-component selects keep po in registers
    const bool mi = hi == ie;
    const T g_ie[3] = {mi ? po.gi[0] : po.gj[0], mi ? po.gi[1] : po.gj[1], mi ? po.gi[2] : po.gj[2]};
    const T g_j[3] = {mi ? po.gj[0] : po.gi[0], mi ? po.gj[1] : po.gi[1], mi ? po.gj[2] : po.gi[2]};
    T l_ie = mi ? po.li : po.lj;
    T l_j = mi ? po.lj : po.li;
    fsn[j] = po.f;
    for (int q = 0; q < 3; q++) {
      fijn_i[3 * j + q] = g_ie[q];
      fijn_j[3 * j + q] = g_j[q];
    }
    lapn[j] = l_ie;
    mv[ML.lapn_i + N + j] = l_j;
    dsum += (double)po.f - f_old;
    gix += (double)g_ie[0] - gi_old[0];
    giy += (double)g_ie[1] - gi_old[1];
    giz += (double)g_ie[2] - gi_old[2];
    lapi += l_ie;
    // other electron's new gradient
    for (int q = 0; q < 3; q++)
      fjn[3 * j + q] = fj_old[q] + (double)g_j[q] - gj_old[q];
  }
  dsum = gsum<G>(dsum);
  gix = gsum<G>(gix); giy = gsum<G>(giy); giz = gsum<G>(giz);
  lapi = gsum<G>(lapi);
  const double fsumn = ws[L.fsumo] + dsum;
  double fji[3] = {fjo[3 * ie] + gix, fjo[3 * ie + 1] + giy, fjo[3 * ie + 2] + giz};
  if (lane == 0) { fjn[3 * ie] = fji[0]; fjn[3 * ie + 1] = fji[1]; fjn[3 * ie + 2] = fji[2]; }
  double lapjn = lapi + fji[0] * fji[0] + fji[1] * fji[1] + fji[2] * fji[2];
  // ---- orbitals of the new position (occupied set of the spin)
  // basis values of the new position: LDS when the launch reserved it (the
  // two calls keep the LDS pointer visible as such to the compiler)
  double *orbe = mv + ML.orbe;
  if (A.een_tab) __syncthreads();      // the tables' region takes the basis values
  if (A.lds_phi)
    orbitals_1e<T, G>(A, s, nocc, ren, rven, etab, orbe, lane);
  else
    orbitals_1e<T, G>(A, s, nocc, ren, rven, reinterpret_cast<T *>(ss + SL.phi), orbe, lane);
  __syncthreads();
  // ---- dete_update (qmc/dete_mod.f90:114-214)
  const double *ai = w_ai(L, ws, s);
  const double *tocc = w_tocc(L, ws, s);
  // candidates into the other slots of the state (the commit flips them)
  double *ain = w_ai(L, ws, s, 1), *tn = w_tocc(L, ws, s, 1);
  double rat = 0.0;
  for (int k = lane; k < ns; k += G) rat += orbe[5 * cst(S.iref[s])[k]] * ai[k + ns * e];
  rat = gsum<G>(rat);
  // vi1(i) = sum_k orbe(ref_k) ai(k, i) - delta(i, e); ui(k) = -ai(k,e)/rat
  double *vi1 = mv + ML.misc + 8;            // [ns]
  double *vi2 = vi1 + ns;                    // [nocc]
  // a single determinant needs only A^-1: T stays the fixed permutation
  // and the excited-determinant blocks, Y and chi do not change
  const bool single = S.single_det;
  sm_vectors<G>(vi1, vi2, orbe, ai, tocc, S.iref[s], ns, nocc, e, lane, single);
  __syncthreads();
  sm_rank1<G>(ain, tn, ai, tocc, vi1, vi2, ns, nocc, e, rat, lane, single);
  __syncthreads();
  double *invn = w_inv(L, ws, s, 1), *detn = w_detv(L, ws, s, 1);
  for (int u = single ? S.ndetu[s] : lane; u < S.ndetu[s]; u += G) {
    int k = cst(S.ord[s])[u];
    double *blk = invn + u * S.kmax2;
    for (int ii = 0; ii < k; ii++)
      for (int jj = 0; jj < k; jj++)
        blk[jj + k * ii] = tn[cst(S.rf[s])[u * ns + jj] * nocc + cst(S.exo[s])[u * ns + ii]];
    detn[u] = small_matinv(blk, k);
  }
  if (!single) __syncthreads();
  double chin;
  if (single) {
    // chi = cdd * 1 * 1 (the empty alpha blocks have determinant 1), as eval_y forms it
    chin = cst(S.cdd)[0] * w_detv(L, ws, 0)[0] * w_detv(L, ws, 1)[0];
  } else {
    const double *invs[2], *dvs[2];
    double *ys[2];
    invs[s] = invn; dvs[s] = detn; ys[s] = w_y(L, ws, s, 1);
    invs[1 - s] = w_inv(L, ws, 1 - s); dvs[1 - s] = w_detv(L, ws, 1 - s);
    ys[1 - s] = w_y(L, ws, 1 - s, 1);
    if (A.lds_y) {
      extern __shared__ double dyn_lds[];
      chin = group_eval_y_lds<G>(S, invs, dvs, ys, dyn_lds + A.lds_y_off + (int64_t)grp * A.lds_y, lane);
    } else {
      chin = group_eval_y<G>(S, invs, dvs, ys, ss + SL.cup, lane);
    }
  }
  if (!single) __syncthreads();
  const double *yn = w_y(L, ws, s, 1);
  double deta_n = ws[L.deta[s]] * rat;
  double gnew[3], lapn_tot;
  group_grad_1e<G>(A, ws, ie, orbe, ain, tn, yn, fji, lapjn, gnew, &lapn_tot, lane);
  if (lane == 0) {
    double psio = ws[L.chi] * ws[L.deta[0]] * ws[L.deta[1]];
    double psin = chin * deta_n * ws[L.deta[1 - s]];
    double logn = log(fabs(chin)) + log(fabs(deta_n)) + log(fabs(ws[L.deta[1 - s]]));
    mv[ML.misc + 0] = rat;
    mv[ML.misc + 5] = chin;
    mv[ML.misc + 6] = deta_n;
    mv[ML.misc + 7] = fsumn;
    mv[ML.misc + 8 + ns + nocc] = logn;
    double *o = out ? out + 8 * w : nullptr;
    double ratio = (psin / psio) * exp(fsumn - ws[L.fsumo]);
    mv[ML.misc + 9 + ns + nocc] = ratio;
    for (int q = 0; q < 3; q++) mv[ML.misc + 10 + ns + nocc + q] = gnew[q];
    if (o) {
      o[0] = ratio; o[1] = logn; o[2] = psin < 0 ? -1.0 : 1.0; o[3] = fsumn;
      o[4] = gnew[0]; o[5] = gnew[1]; o[6] = gnew[2]; o[7] = lapn_tot;
    }
  }
  // VMC sweep (mode 2), DMC (mode 3) or a given move (mode 4): test and
  // commit in the same launch
  if (mode >= 2) {
    __shared__ int s_acc[kWave / G];
    __syncthreads();
    if (lane == 0) {
      const int acc = mode == 4   ? 1
                      : mode == 3 ? dmc_decide_w(A, w, ie, mv, A.dm + (int64_t)w * kDmStride, A.dmp)
                                  : metrop_accept(A, ws, mv, w, ie, ns, nocc, delta, fbias);
      s_acc[grp] = acc;
      if (acc && nacc) nacc[w] += 1;
    }
    __syncthreads();
    if (s_acc[grp]) group_commit<G>(A, ws, mv, ie, lane);
  }
}
template <typename T, int G, int ORD>
__global__ QMCW_HEAVY void k_ratio_1e(KArgs A, int ie, int mode, const double *xnew_in,
                                      double delta, double fbias, double *out, long long *nacc) {
  ratio_1e_b<T, G, ORD>(A, blockIdx.x, ie, mode, xnew_in, delta, fbias, out, nacc);
}

// ---------------------------------------------------------------------------
// The same one-electron move as k_ratio_1e in three launches, each with its
// own register budget: (a) proposal, distances and Jastrow row; (b) orbitals
// of the new position and dete_update; (c) eval_y, the new gradient, the
// ratio and (mode 2) the Metropolis test and commit.  The pending buffers
// carry everything between them; the Jastrow sums that (c) needs go to
// ML.misc + kMvJas + {0..3} (fji, lap ln J).
template <typename T, int G, int ORD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : QMCW_MOVE_WPE, 8)))
void k_move_a(KArgs A, int ie, int mode, const double *xnew_in, double delta, double fbias) {
  QMCW_GROUP_SETUP
  const int N = S.N, M = S.M, noccmax = S.noccmax;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const MLayout &ML = A.ML;
  double *mv = ss + SL.mv;
  const int s = ie < S.nup ? 0 : 1, ns = S.ns[s], nocc = S.nocc[s];
  const int e = s ? ie - S.nup : ie;
  const double *x = ws + L.x;
  const double *fso = ws + L.fso, *fijo = ws + L.fijo, *lapj = ws + L.lapj, *fjo = ws + L.fjo;
  double *xn = mv + ML.xn;
  // ---- proposal (mode 1: ratio only; mode 2: ratio, test and commit)
  if (mode >= 1) {
    double lapo_dummy;
    double vold[3];
    double lapjo = 0.0;  // not needed for the proposal
    group_grad_1e<G>(A, ws, ie, ws + L.oc + (int64_t)ie * noccmax * 5, w_ai(L, ws, s),
                     w_tocc(L, ws, s), w_y(L, ws, s), fjo + 3 * ie, lapjo, vold, nullptr, lane);
    (void)lapo_dummy;
    if (lane == 0) vmc_propose_w(A, w, ie, ws, vold, delta, fbias, xn, mv);
  } else {
    if (lane < 3) xn[lane] = xnew_in[3 * w + lane];
  }
  __syncthreads();
  // ---- new distances of electron ie
  double *ren = mv + ML.ren, *rven = mv + ML.rven, *ree = mv + ML.ree, *rvee = mv + ML.rvee;
  for (int ic = lane; ic < M; ic += G) {
    double s2 = 0.0;
    for (int q = 0; q < 3; q++) {
      double d = xn[q] - cst(S.cent)[3 * ic + q];
      rven[3 * ic + q] = d;
      s2 += d * d;
    }
    ren[ic] = sqrt(s2);
  }
  for (int j = lane; j < N; j += G) {
    if (j == ie) continue;
    int hi = j < ie ? ie : j, lo = j < ie ? j : ie;
    double s2 = 0.0;
    for (int q = 0; q < 3; q++) {
      double xi = (hi == ie) ? xn[q] : x[3 * hi + q];
      double xj = (lo == ie) ? xn[q] : x[3 * lo + q];
      rvee[3 * j + q] = xi - xj;
      s2 += (xi - xj) * (xi - xj);
    }
    ree[j] = sqrt(s2);
  }
  __syncthreads();
  // ---- jastrow4e (+ per-electron laplacian)
  double *fsn = mv + ML.fsn, *fijn_i = mv + ML.fijn_i, *fijn_j = mv + ML.fijn_j,
         *lapn = mv + ML.lapn_i, *fjn = mv + ML.fjn;
  double dsum = 0.0, gix = 0, giy = 0, giz = 0, lapi = 0;
  for (int j = lane; j < N; j += G) {
    if (j == ie) {  // e-n diagonal term, on the lane the pair list leaves free
      T f, g[3], lp;
      jas_en<T>(S, ren, rven, f, g, lp);
      fsn[ie] = f;
      for (int q = 0; q < 3; q++) fijn_i[3 * ie + q] = g[q];
      lapn[ie] = lp;
      dsum += (double)f - fso[ie * N + ie];
      gix += (double)g[0] - fijo[3 * (ie * N + ie)];
      giy += (double)g[1] - fijo[3 * (ie * N + ie) + 1];
      giz += (double)g[2] - fijo[3 * (ie * N + ie) + 2];
      lapi += lp;
      continue;
    }
    int hi = j < ie ? ie : j, lo = j < ie ? j : ie;
    T rvij[3];
    for (int q = 0; q < 3; q++) rvij[q] = (T)rvee[3 * j + q];
    const double *rI = ren, *rvI = rven;
    PairOut<T> po;
    bool par = (hi < S.nup) || (lo >= S.nup);
    // moved electron's arrays go to whichever slot (hi or lo) it occupies
    // the other electron's distances formed from its position (see EnFromX)
    const EnRows pI{rI, rvI};
    const EnFromX pJ{x[3 * j], x[3 * j + 1], x[3 * j + 2]};
    if (hi == ie) jas_pair_g<T, ORD>(S, par, (T)ree[j], rvij, pI, pJ, po);
    else jas_pair_g<T, ORD>(S, par, (T)ree[j], rvij, pJ, pI, po);
    // old pair values before the stores to the pending buffers below
    const double f_old = fso[hi * N + lo];
    const double gi_old[3] = {fijo[3 * (ie * N + j)], fijo[3 * (ie * N + j) + 1], fijo[3 * (ie * N + j) + 2]};
    const double gj_old[3] = {fijo[3 * (j * N + ie)], fijo[3 * (j * N + ie) + 1], fijo[3 * (j * N + ie) + 2]};
    const double fj_old[3] = {fjo[3 * j], fjo[3 * j + 1], fjo[3 * j + 2]};
    // per-component selects keep po in registers
    const bool mi = hi == ie;
    const T g_ie[3] = {mi ? po.gi[0] : po.gj[0], mi ? po.gi[1] : po.gj[1], mi ? po.gi[2] : po.gj[2]};
    const T g_j[3] = {mi ? po.gj[0] : po.gi[0], mi ? po.gj[1] : po.gi[1], mi ? po.gj[2] : po.gi[2]};
    T l_ie = mi ? po.li : po.lj;
    T l_j = mi ? po.lj : po.li;
    fsn[j] = po.f;
    for (int q = 0; q < 3; q++) {
      fijn_i[3 * j + q] = g_ie[q];
      fijn_j[3 * j + q] = g_j[q];
    }
    lapn[j] = l_ie;
    mv[ML.lapn_i + N + j] = l_j;
    dsum += (double)po.f - f_old;
    gix += (double)g_ie[0] - gi_old[0];
    giy += (double)g_ie[1] - gi_old[1];
    giz += (double)g_ie[2] - gi_old[2];
    lapi += l_ie;
    // other electron's new gradient
    for (int q = 0; q < 3; q++)
      fjn[3 * j + q] = fj_old[q] + (double)g_j[q] - gj_old[q];
  }
  dsum = gsum<G>(dsum);
  gix = gsum<G>(gix); giy = gsum<G>(giy); giz = gsum<G>(giz);
  lapi = gsum<G>(lapi);
  const double fsumn = ws[L.fsumo] + dsum;
  double fji[3] = {fjo[3 * ie] + gix, fjo[3 * ie + 1] + giy, fjo[3 * ie + 2] + giz};
  if (lane == 0) { fjn[3 * ie] = fji[0]; fjn[3 * ie + 1] = fji[1]; fjn[3 * ie + 2] = fji[2]; }
  double lapjn = lapi + fji[0] * fji[0] + fji[1] * fji[1] + fji[2] * fji[2];
  if (lane == 0) {
    const int o = ML.misc + kMvJas + ns + nocc;
    mv[o] = fji[0]; mv[o + 1] = fji[1]; mv[o + 2] = fji[2]; mv[o + 3] = lapjn;
    mv[ML.misc + 7] = fsumn;
  }
}

template <typename T, int G>
__global__ QMCW_MOVE void k_move_b(KArgs A, int ie) {
  QMCW_GROUP_SETUP
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const MLayout &ML = A.ML;
  double *mv = ss + SL.mv;
  const int s = ie < S.nup ? 0 : 1, ns = S.ns[s], nocc = S.nocc[s];
  const int e = s ? ie - S.nup : ie;
  const double *ren = mv + ML.ren, *rven = mv + ML.rven;
  // ---- orbitals of the new position (occupied set of the spin)
  // basis values of the new position: LDS when the launch reserved it (the
  // two calls keep the LDS pointer visible as such to the compiler)
  extern __shared__ double dyn_lds[];
  double *orbe = mv + ML.orbe;
  if (A.lds_phi)
    orbitals_1e<T, G>(A, s, nocc, ren, rven, reinterpret_cast<T *>(dyn_lds) + (int64_t)grp * S.NB * 5,
                      orbe, lane);
  else
    orbitals_1e<T, G>(A, s, nocc, ren, rven, reinterpret_cast<T *>(ss + SL.phi), orbe, lane);
  __syncthreads();
  // ---- dete_update (qmc/dete_mod.f90:114-214)
  const double *ai = w_ai(L, ws, s);
  const double *tocc = w_tocc(L, ws, s);
  // candidates into the other slots of the state (the commit flips them)
  double *ain = w_ai(L, ws, s, 1), *tn = w_tocc(L, ws, s, 1);
  double rat = 0.0;
  for (int k = lane; k < ns; k += G) rat += orbe[5 * cst(S.iref[s])[k]] * ai[k + ns * e];
  rat = gsum<G>(rat);
  // vi1(i) = sum_k orbe(ref_k) ai(k, i) - delta(i, e); ui(k) = -ai(k,e)/rat
  double *vi1 = mv + ML.misc + 8;            // [ns]
  double *vi2 = vi1 + ns;                    // [nocc]
  sm_vectors<G>(vi1, vi2, orbe, ai, tocc, S.iref[s], ns, nocc, e, lane);
  __syncthreads();
  sm_rank1<G>(ain, tn, ai, tocc, vi1, vi2, ns, nocc, e, rat, lane);
  __syncthreads();
  double *invn = w_inv(L, ws, s, 1), *detn = w_detv(L, ws, s, 1);
  for (int u = lane; u < S.ndetu[s]; u += G) {
    int k = cst(S.ord[s])[u];
    double *blk = invn + u * S.kmax2;
    for (int ii = 0; ii < k; ii++)
      for (int jj = 0; jj < k; jj++)
        blk[jj + k * ii] = tn[cst(S.rf[s])[u * ns + jj] * nocc + cst(S.exo[s])[u * ns + ii]];
    detn[u] = small_matinv(blk, k);
  }
  __syncthreads();
  if (lane == 0) mv[ML.misc + 0] = rat;
}

template <int G>
__global__ QMCW_MOVE void k_move_c(KArgs A, int ie, int mode, double delta, double fbias,
                                   double *out, long long *nacc) {
  QMCW_GROUP_SETUP
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const MLayout &ML = A.ML;
  double *mv = ss + SL.mv;
  const int s = ie < S.nup ? 0 : 1, ns = S.ns[s], nocc = S.nocc[s];
  double *orbe = mv + ML.orbe, *ain = w_ai(L, ws, s, 1), *tn = w_tocc(L, ws, s, 1);
  double *invn = w_inv(L, ws, s, 1), *detn = w_detv(L, ws, s, 1);
  const double rat = mv[ML.misc + 0], fsumn = mv[ML.misc + 7];
  const int oj = ML.misc + kMvJas + ns + nocc;
  const double fji[3] = {mv[oj], mv[oj + 1], mv[oj + 2]};
  const double lapjn = mv[oj + 3];
  double chin;
  {
    const double *invs[2], *dvs[2];
    double *ys[2];
    invs[s] = invn; dvs[s] = detn; ys[s] = w_y(L, ws, s, 1);
    invs[1 - s] = w_inv(L, ws, 1 - s); dvs[1 - s] = w_detv(L, ws, 1 - s);
    ys[1 - s] = w_y(L, ws, 1 - s, 1);
    if (A.lds_y) {
      extern __shared__ double dyn_lds[];
      chin = group_eval_y_lds<G>(S, invs, dvs, ys, dyn_lds + A.lds_y_off + (int64_t)grp * A.lds_y, lane);
    } else {
      chin = group_eval_y<G>(S, invs, dvs, ys, ss + SL.cup, lane);
    }
  }
  __syncthreads();
  const double *yn = w_y(L, ws, s, 1);
  double deta_n = ws[L.deta[s]] * rat;
  double gnew[3], lapn_tot;
  group_grad_1e<G>(A, ws, ie, orbe, ain, tn, yn, fji, lapjn, gnew, &lapn_tot, lane);
  if (lane == 0) {
    double psio = ws[L.chi] * ws[L.deta[0]] * ws[L.deta[1]];
    double psin = chin * deta_n * ws[L.deta[1 - s]];
    double logn = log(fabs(chin)) + log(fabs(deta_n)) + log(fabs(ws[L.deta[1 - s]]));
    mv[ML.misc + 0] = rat;
    mv[ML.misc + 5] = chin;
    mv[ML.misc + 6] = deta_n;
    mv[ML.misc + 7] = fsumn;
    mv[ML.misc + 8 + ns + nocc] = logn;
    double *o = out ? out + 8 * w : nullptr;
    double ratio = (psin / psio) * exp(fsumn - ws[L.fsumo]);
    mv[ML.misc + 9 + ns + nocc] = ratio;
    for (int q = 0; q < 3; q++) mv[ML.misc + 10 + ns + nocc + q] = gnew[q];
    if (o) {
      o[0] = ratio; o[1] = logn; o[2] = psin < 0 ? -1.0 : 1.0; o[3] = fsumn;
      o[4] = gnew[0]; o[5] = gnew[1]; o[6] = gnew[2]; o[7] = lapn_tot;
    }
  }
  // VMC sweep (mode 2): Metropolis test and commit in the same launch
  if (mode == 2) {
    __shared__ int s_acc[kWave / G];
    __syncthreads();
    if (lane == 0) {
      const int acc = metrop_accept(A, ws, mv, w, ie, ns, nocc, delta, fbias);
      s_acc[grp] = acc;
      if (acc && nacc) nacc[w] += 1;
    }
    __syncthreads();
    if (s_acc[grp]) group_commit<G>(A, ws, mv, ie, lane);
  }
}

// Metropolis test of metrop_mov1 (qmc/vmc/metrop_mov1.f90:90-99, 226-250)
// for the pending move of electron ie, with the walker's rannyu (lane 0).
// One walker's VMC proposal of electron ie (lane 0) from its current
// velocity vold: metrop_mov1's box proposal with force bias (imetro 1,
// qmc/vmc/metrop_mov1.f90:63-99) or the accelerated mover (imetro 6,
// qmc/vmc/metrop_slat_mov1.f90:78-289).  The pending move buffer keeps what
// the acceptance needs.
__device__ __forceinline__ void vmc_propose_w(const KArgs &A, int w, int ie, const double *ws,
                                              const double *vold, double delta, double fbias,
                                              double *xn, double *mv) {
  const MLayout &ML = A.ML;
  int *rs = A.rng + 4 * w;
  int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
  if (A.imetro == 6) {
    slat_propose(A.S, A.deltar, A.deltat, ws + A.L.rvec_en + (int64_t)ie * A.S.M * 3, vold, fbias, st4, xn,
                 mv + ML.slat);
  } else {
    const double *x = ws + A.L.x;
    const double deltfi = 2.0 * fbias / delta;
    double fxop = 1.0;
    for (int k = 0; k < 3; k++) {
      double fx = fmin(fabs(vold[k]), deltfi);
      double ftest = fx * delta * 0.5;
      fx = copysign(fx, vold[k]);
      double dx;
      if (ftest <= rannyu(st4)) {
        dx = (rannyu(st4) - 0.5) * delta;
      } else {
        double r1 = rannyu(st4), r2 = rannyu(st4);
        dx = (fmax(r1, r2) - 0.5) * delta;
        if (vold[k] < 0) dx = -dx;
      }
      fxop *= (1.0 + fx * dx);
      xn[k] = x[3 * ie + k] + dx;
      mv[ML.misc + 2 + k] = dx;
    }
    mv[ML.misc + 1] = fxop;
  }
  rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
}

__device__ __forceinline__ int metrop_accept(const KArgs &A, const double *ws, const double *mv,
                                             int w, int ie, int ns, int nocc, double delta,
                                             double fbias) {
  const WLayout &L = A.L;
  const MLayout &ML = A.ML;
  int *rs = A.rng + 4 * w;
  int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
  const double logn = mv[ML.misc + 8 + ns + nocc];
  const double psi2n = 2.0 * (logn + mv[ML.misc + 7]);
  double p;
  if (A.imetro == 6) {
    // reverse move from the proposed point with its velocity
    // (metrop_slat_mov1.f90:299-336, 376-481); 0 for a geometrical rejection
    const double *rec = mv + ML.slat;
    double fx, area, rto;
    p = 0.0;
    if (slat_density(A.S, A.deltar, A.deltat, mv + ML.rven, mv + ML.misc + 10 + ns + nocc,
                     ws + L.rvec_en + (int64_t)ie * A.S.M * 3, fbias, fx, area, rto)) {
      const double rr = rec[2] / rto;
      p = rr * rr * exp(psi2n - ws[L.psi2o]) * fabs((fx * rec[1]) / (rec[0] * area));
      p = p >= 0 ? fmin(1.0, p) : 0.0;     // NaN (an unsampled proposal): reject
    }
  } else {
    const double deltfi = 2.0 * fbias / delta;
    double fxnp = 1.0;
    for (int k = 0; k < 3; k++) {
      const double vn = mv[ML.misc + 10 + ns + nocc + k];
      fxnp *= (1.0 - mv[ML.misc + 2 + k] * copysign(fmin(fabs(vn), deltfi), vn));
    }
    p = exp(psi2n - ws[L.psi2o]) * fxnp / mv[ML.misc + 1];
    p = fmin(1.0, p);
  }
  const int acc = rannyu(st4) < p;
  rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
  return acc;
}

// Commit the pending move of electron ie into the walker's state: jassav +
// detsav + dete_save (qmc/vmc/metrop_slat_mov1.f90:544-546).  All lanes of
// the group; no barriers.  The contiguous blocks (x, distance and orbital
// rows, gradients, A^-1, T, alpha inverses and determinants, Y) are copied
// as one flat list of segments, up to kCommitBatch loads in flight per lane
// before the stores; the pair-array rows follow, one electron per lane.
constexpr int kCommitSeg = 5;
constexpr int kCommitBatch = 16;

template <int G>
__device__ __forceinline__ void commit_pairs(const DevSys &S, const WLayout &L, const MLayout &ML,
                                             double *__restrict__ ws, const double *__restrict__ mv,
                                             int ie, int lane) {
  const int N = S.N;
  for (int j = lane; j < N; j += G) {
    if (j == ie) {
      const double f = mv[ML.fsn + ie], lp = mv[ML.lapn_i + ie];
      const double g0 = mv[ML.fijn_i + 3 * ie], g1 = mv[ML.fijn_i + 3 * ie + 1], g2 = mv[ML.fijn_i + 3 * ie + 2];
      ws[L.fso + ie * N + ie] = f;
      ws[L.fijo + 3 * (ie * N + ie)] = g0;
      ws[L.fijo + 3 * (ie * N + ie) + 1] = g1;
      ws[L.fijo + 3 * (ie * N + ie) + 2] = g2;
      ws[L.lapj + ie * N + ie] = lp;
      continue;
    }
    const int hi = j < ie ? ie : j, lo = j < ie ? j : ie;
    double v[13];
    v[0] = mv[ML.ree + j];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      v[1 + q] = mv[ML.rvee + 3 * j + q];
      v[4 + q] = mv[ML.fijn_i + 3 * j + q];
      v[7 + q] = mv[ML.fijn_j + 3 * j + q];
    }
    v[10] = mv[ML.fsn + j];
    v[11] = mv[ML.lapn_i + j];
    v[12] = mv[ML.lapn_i + N + j];
    ws[L.r_ee + hi * N + lo] = v[0];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      ws[L.rvec_ee + 3 * (hi * N + lo) + q] = v[1 + q];
      ws[L.fijo + 3 * (ie * N + j) + q] = v[4 + q];
      ws[L.fijo + 3 * (j * N + ie) + q] = v[7 + q];
    }
    ws[L.fso + hi * N + lo] = v[10];
    ws[L.lapj + ie * N + j] = v[11];
    ws[L.lapj + j * N + ie] = v[12];
  }
}

template <int G>
__device__ __forceinline__ void commit_blocks(const int (&st)[kCommitSeg + 1], const int (&dso)[kCommitSeg],
                                              const int (&sso)[kCommitSeg], double *__restrict__ ws,
                                              const double *__restrict__ mv, int lane) {
  const int n = st[kCommitSeg];
  int srel[kCommitSeg], drel[kCommitSeg];
#pragma unroll
  for (int k = 0; k < kCommitSeg; k++) { srel[k] = sso[k] - st[k]; drel[k] = dso[k] - st[k]; }
  for (int base = 0; base < n; base += G * kCommitBatch) {
    double v[kCommitBatch];
    int dst[kCommitBatch];
#pragma unroll
    for (int u = 0; u < kCommitBatch; u++) {
      const int e = base + lane + u * G;
      // segments are consecutive in e: the last one starting at or before e holds it
      int so = srel[0], d = drel[0];
#pragma unroll
      for (int k = 1; k < kCommitSeg; k++)
        if (e >= st[k]) { so = srel[k]; d = drel[k]; }
      dst[u] = e < n ? d + e : -1;
      v[u] = e < n ? mv[so + e] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kCommitBatch; u++)
      if (dst[u] >= 0) ws[dst[u]] = v[u];
  }
}

template <int G>
__device__ void group_commit(const KArgs &A, double *ws, const double *mv, int ie, int lane) {
  const DevSys &S = A.S;
  const int N = S.N, M = S.M, noccmax = S.noccmax;
  const WLayout &L = A.L;
  const MLayout &ML = A.ML;
  const int s = ie < S.nup ? 0 : 1, ns = S.ns[s], nocc = S.nocc[s];
  const int len[kCommitSeg] = {3, M, 3 * M, 3 * N, nocc * 5};
  const int dso[kCommitSeg] = {(int)L.x + 3 * ie, (int)L.r_en + ie * M, (int)L.rvec_en + ie * M * 3,
                               (int)L.fjo, (int)L.oc + ie * noccmax * 5};
  const int sso[kCommitSeg] = {(int)ML.xn, (int)ML.ren, (int)ML.rven, (int)ML.fjn, (int)ML.orbe};
  int st[kCommitSeg + 1];
  st[0] = 0;
#pragma unroll
  for (int k = 0; k < kCommitSeg; k++) st[k + 1] = st[k] + len[k];
  commit_blocks<G>(st, dso, sso, ws, mv, lane);
  commit_pairs<G>(S, L, ML, ws, mv, ie, lane);
  if (lane == 0) {
    // the candidate slots become the current ones: A^-1 of the spin, and
    // unless the system has a single determinant (whose move leaves them
    // unchanged) its T, alpha blocks and determinants and Y of both spins
    ws[L.par + s] = wslot(ws, L.par + s) ? 0.0 : 1.0;
    if (!S.single_det) {
      ws[L.par + 2 + s] = wslot(ws, L.par + 2 + s) ? 0.0 : 1.0;
      ws[L.par + 4] = wslot(ws, L.par + 4) ? 0.0 : 1.0;
    }
    const double chi = mv[ML.misc + 5], deta = mv[ML.misc + 6], fs = mv[ML.misc + 7];
    const double logn = mv[ML.misc + 8 + ns + nocc];
    ws[L.chi] = chi;
    ws[L.deta[s]] = deta;
    ws[L.fsumo] = fs;
    ws[L.psij] = fs;
    ws[L.logpsid] = logn;
    ws[L.psi2o] = 2.0 * (logn + fs);
  }
}

// Commit or drop the pending move of electron ie.  mode 1: Metropolis test
// with the walker's rannyu; mode 0: accept[w] given.  naccept[w] incremented
// on acceptance.
template <int G>
__global__ __launch_bounds__(64) void k_accept_1e(KArgs A, int ie, int mode,
                                                  const uint8_t *accept_in,
                                                  double delta, double fbias,
                                                  long long *naccept) {
  QMCW_GROUP_SETUP
  const SLayout &SL = A.SL;
  const double *mv = ss + SL.mv;
  const int s = ie < S.nup ? 0 : 1, ns = S.ns[s], nocc = S.nocc[s];
  __shared__ int s_acc[kWave / G];
  if (lane == 0) {
    const int acc = (mode == 1) ? metrop_accept(A, ws, mv, w, ie, ns, nocc, delta, fbias)
                                : (accept_in[w] != 0);
    s_acc[grp] = acc;
    if (acc && naccept) naccept[w] += 1;
  }
  __syncthreads();
  if (!s_acc[grp]) return;   // uniform within the group; no barriers follow
  group_commit<G>(A, ws, mv, ie, lane);
}

// grad/lap of electron ie in the current state: out[w][4]
template <int G>
__global__ __launch_bounds__(64) void k_grad_1e(KArgs A, int ie, double *out) {
  QMCW_GROUP_SETUP
  const WLayout &L = A.L;
  const int N = S.N, s = ie < S.nup ? 0 : 1;
  const double *lapj = ws + L.lapj, *fjo = ws + L.fjo;
  double lj = 0.0;
  for (int j = lane; j < N; j += G) lj += lapj[ie * N + j];
  lj = gsum<G>(lj);
  lj += fjo[3 * ie] * fjo[3 * ie] + fjo[3 * ie + 1] * fjo[3 * ie + 1] + fjo[3 * ie + 2] * fjo[3 * ie + 2];
  double g[3], lap;
  group_grad_1e<G>(A, ws, ie, ws + L.oc + (int64_t)ie * S.noccmax * 5, w_ai(L, ws, s),
                   w_tocc(L, ws, s), w_y(L, ws, s), fjo + 3 * ie, lj, g, &lap, lane);
  if (lane == 0) {
    out[4 * w] = g[0]; out[4 * w + 1] = g[1]; out[4 * w + 2] = g[2]; out[4 * w + 3] = lap;
  }
}

// random rotation of each walker's quadrature grid (qmc/vmc/rotqua.f90:621-682)
#ifndef QMCW_ORD   // launched from the host translation unit only
__device__ __forceinline__ void rotqua_b(const KArgs &A, int bid_) {
  int w = bid_ * kWave + threadIdx.x;
  if (w >= A.Wp) return;
  const DevSys &S = A.S;
  int *rs = A.rng + 4 * w;
  int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
  double ax[9];
  rot_axes(st4, ax);
  int nq = S.nquad;
  double *quad = A.st + (int64_t)w * A.L.stride + A.L.quad;
  for (int iq = 0; iq < nq; iq++) {
    double a = cst(S.q0)[iq], b = cst(S.q0)[nq + iq], c = cst(S.q0)[2 * nq + iq];
    quad[iq] = a * ax[0] + b * ax[3] + c * ax[6];
    quad[nq + iq] = a * ax[1] + b * ax[4] + c * ax[7];
    quad[2 * nq + iq] = a * ax[2] + b * ax[5] + c * ax[8];
  }
  rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
}
__global__ __launch_bounds__(64) void k_rotqua(KArgs A) { rotqua_b(A, blockIdx.x); }
#endif

}  // namespace qmcw
