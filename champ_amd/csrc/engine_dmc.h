// engine_dmc.h -- kernels of one DMC generation (qmc/rundmc_mod.f90:97-330):
// T-moves (qmc/dmc/dmc_elec/tmove.f90) when asked for, the drift-diffusion
// proposal with the averaged velocity (qmc/green_functions_mod.f90:70-80,
// 598-633; with the nucleus cusp for all-electron systems, :83-145),
// Metropolis test with the forward and backward Green's functions (:257-323,
// rundmc_mod.f90:142-196), and reweighting with ene_int_rewt
// (green_functions_mod.f90:382-472).  The one-electron ratios and commits
// reuse k_ratio_1e (given position) and k_accept_1e (given decision).
// Every walker draws from its own rannyu stream.
#pragma once
#include "engine_kernels.h"

namespace qmcw {

__global__ void k_dmc_gen_begin(KArgs A, double *dm) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  const double *ws = A.st + (int64_t)w * A.L.stride;
  double *d = dm + (int64_t)w * kDmStride;
  d[kDmEloc] = ws[A.L.eloc];
  d[kDmEpot] = ws[A.L.epot];
  for (int q = kDmSdd; q <= kDmNcross; q++) d[q] = 0.0;
  d[kDmTacc] = 0.0;
}

// proposal for electron ie; g[w][4] holds its current gradient
__global__ void k_dmc_propose(KArgs A, int ie, const double *g, double *xin, double *dm,
                              qmcw_dmc_params p) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  dmc_propose_w(A, w, ie, g + 4 * w, dm + (int64_t)w * kDmStride, p, xin + 3 * w);
}

// Metropolis decision for the pending move of electron ie (k_ratio_1e has
// left the ratio and the new gradient in the walker's move buffer)
__global__ void k_dmc_decide(KArgs A, int ie, double *dm, uint8_t *acc, qmcw_dmc_params p) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  const double *mv = A.sc + (int64_t)w * A.SL.stride + A.SL.mv;
  acc[w] = (uint8_t)dmc_decide_w(A, w, ie, mv, dm + (int64_t)w * kDmStride, p);
}

// after the full evaluation: walkers that moved no electron keep the local
// energy of the start of the generation (CHAMP does not re-evaluate them);
// effective time step, reweighting (ene_int_v9 or no_ene_int)
__global__ void k_dmc_finish(KArgs A, double *dm, qmcw_dmc_params p) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  const DevSys &S = A.S;
  double *ws = A.st + (int64_t)w * A.L.stride;
  const double *ss = A.sc + (int64_t)w * A.SL.stride;
  double *d = dm + (int64_t)w * kDmStride;
  if (d[kDmAcc] == 0.0) {
    ws[A.L.eloc] = d[kDmEloc];
    ws[A.L.epot] = d[kDmEpot];
  }
  const double *vj = ss + A.SL.vj, *vd = ss + A.SL.vd;
  double g2 = 0.0;
  for (int q = 0; q < 3 * S.N; q++) {
    const double v = vj[q] + vd[q];
    g2 += v * v;
  }
  double taunow = p.tau;
  if (p.itau_eff >= 1) taunow = p.tau * (p.itau_eff == 2 ? d[kDmSad] : d[kDmSpd]) / d[kDmSdd];
  double tauint = p.tau;
  if (p.itau_integ >= 1) tauint = (p.itau_eff == 3) ? p.taueff : taunow;
  double wt = d[kDmW];
  const double en = ws[A.L.eloc];
  if (p.idmc > 0) {
    const double pi = 3.141592653589793;
    const double eo = d[kDmEloc];
    double ewto, ewtn;
    if (p.ene_int == 1) {
      ewto = eo;
      ewtn = en;
    } else {
      const double nf = pow((double)S.N, p.pow_rewt);
      const double facto = fmax(sqrt(d[kDmG2]) / nf * tauint * p.c_rewt, 1e-9);
      const double factn = fmax(sqrt(g2) / nf * tauint * p.c_rewt, 1e-9);
      ewto = p.eest + (eo - p.eest) * (sqrt(pi) / 2.0) * erf(facto) / facto;
      ewtn = p.eest + (en - p.eest) * (sqrt(pi) / 2.0) * erf(factn) / factn;
    }
    double eint = (ewto + ewtn) / 2.0 - p.etrial;
    const double emin = p.eest - p.etrial - p.limit_rewt * p.esigma / sqrt(sqrt(p.tau));
    if (eint < emin) eint = emin;
    const double enetau = eint * ((p.itau_eff == 3) ? p.taueff : taunow);
    wt = wt * exp(-enetau) * p.pop_ctrl;
  }
  d[kDmW] = wt;
  d[kDmTau] = taunow;
  d[kDmG2] = g2;
}

// |grad ln Psi|^2 of the current state into the records (start of a run)
__global__ void k_dmc_init(KArgs A, double *dm, const double *weight, int nact) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  const double *ws = A.st + (int64_t)w * A.L.stride;
  const double *ss = A.sc + (int64_t)w * A.SL.stride;
  double *d = dm + (int64_t)w * kDmStride;
  double g2 = 0.0;
  for (int q = 0; q < 3 * A.S.N; q++) {
    const double v = ss[A.SL.vj + q] + ss[A.SL.vd + q];
    g2 += v * v;
  }
  for (int q = 0; q < kDmStride; q++) d[q] = 0.0;
  d[kDmG2] = g2;
  d[kDmEloc] = ws[A.L.eloc];
  d[kDmW] = w < nact ? weight[w] : 0.0;
}

// population sums of one generation: one workgroup, fixed-order tree
// out[8]: w, w E, w E^2, w taunow, pacc, nacc, ncross, ntry
__global__ __launch_bounds__(256) void k_dmc_sums(KArgs A, const double *dm, int nact,
                                                  double *out) {
  __shared__ double sh[8][256];
  const int t = threadIdx.x;
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int w = t; w < nact; w += 256) {
    const double *d = dm + (int64_t)w * kDmStride;
    const double e = A.st[(int64_t)w * A.L.stride + A.L.eloc];
    const double wt = d[kDmW];
    a[0] += wt; a[1] += wt * e; a[2] += wt * e * e; a[3] += wt * d[kDmTau];
    a[4] += d[kDmPacc]; a[5] += d[kDmNacc]; a[6] += d[kDmNcross]; a[7] += A.S.N;
  }
  for (int q = 0; q < 8; q++) sh[q][t] = a[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int q = 0; q < 8; q++) sh[q][t] += sh[q][t + o];
    __syncthreads();
  }
  if (t < 8) out[t] = sh[t][0];
}

// weights of the population (for branching on the host)
__global__ void k_dmc_weights(const double *dm, int nact, double *out) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nact) out[w] = dm[(int64_t)w * kDmStride + kDmW];
}

// population after branching: record d is a copy of record origin[d]
// nl_dev (may be null): the number of records, read on the device (the
// device split_join leaves it there); else nl
__global__ void k_dmc_gather(const double *st_src, double *st_dst, int64_t stride,
                             const double *dm_src, double *dm_dst, const int *origin,
                             const double *wnew, int nl, const double *tq_src, double *tq_dst,
                             int64_t tq_stride, const int *nl_dev) {
  const int d = blockIdx.x;
  if (d >= (nl_dev ? nl_dev[3] : nl)) return;
  const int o = origin[d];
  const double *src = st_src + (int64_t)o * stride;
  double *dst = st_dst + (int64_t)d * stride;
  for (int64_t q = threadIdx.x; q < stride; q += blockDim.x) dst[q] = src[q];
  if (tq_src)
    for (int64_t q = threadIdx.x; q < tq_stride; q += blockDim.x)
      tq_dst[(int64_t)d * tq_stride + q] = tq_src[(int64_t)o * tq_stride + q];
  if (threadIdx.x < kDmStride) {
    double v = dm_src[(int64_t)o * kDmStride + threadIdx.x];
    if (threadIdx.x == kDmW) v = wnew[d];
    dm_dst[(int64_t)d * kDmStride + threadIdx.x] = v;
  }
}

// ---------------------------------------------------------------------------
// split_join (qmc/dmc/dmc_elec/split_join_mod.f90:3-219) on the device, one
// workgroup: the same loops and draws as the host routine qmcw_split_join
// (joins in walker order with one rannyu per joined pair; the split
// threshold from the sorted weights; splits into the holes last-first, each
// copy with four 12-bit limbs drawn from the branching stream; compaction).
// The joins, splits and compaction are sequential in the walker index and
// the stream, so one lane runs them; the sort of the weights runs on every
// lane (bitonic, over the next power of two, padded with +inf).
// w[roundup(mwalk, 8)]: in, the n weights; out, the new population's
// weights, padded with 0 to a multiple of 8.  rng[mwalk][4]: the walkers' streams, moved
// with them.  origin[mwalk]: the old walker each new one copies.  sortbuf:
// npow2 doubles; under: n ints.  res: {walkers, branching events, status
// (0 ok, 1 more walkers than mwalk), walkers padded to a multiple of 8}.
__global__ __launch_bounds__(1024) void k_split_join(double *w, int n, int mwalk, int *brng, int *rng,
                                                     int *origin, double *sortbuf, int npow2, int *under,
                                                     int *res) {
  const int t = threadIdx.x, nt = blockDim.x;
  __shared__ int s_nunder, s_nbr;
  for (int i = t; i < mwalk; i += nt) origin[i] = i;
  if (t == 0) {
    int st4[4] = {brng[0], brng[1], brng[2], brng[3]};
    int nu = 0, ipair = 0, iw2 = 0, nbr = 0;
    for (int iw = 0; iw < n; iw++) {
      const double wi = w[iw];
      if (!(wi < 0.5)) continue;
      if (wi == 0.0) {
        nbr++;
        under[nu++] = iw;
      } else if (ipair == 0) {
        ipair = 1;
        iw2 = iw;
      } else {
        nbr++;
        ipair = 0;
        const double wt = wi + w[iw2];
        if (rannyu(st4) > wi / wt) {
          w[iw2] = wt;
          under[nu++] = iw;
        } else {
          w[iw] = wt;
          under[nu++] = iw2;
        }
      }
    }
    brng[0] = st4[0]; brng[1] = st4[1]; brng[2] = st4[2]; brng[3] = st4[3];
    s_nunder = nu;
    s_nbr = nbr;
  }
  __syncthreads();
  for (int i = t; i < npow2; i += nt) sortbuf[i] = i < n ? w[i] : __builtin_inf();
  __syncthreads();
  for (int k = 2; k <= npow2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < npow2; i += nt) {
        const int l = i ^ j;
        if (l > i) {
          const double a = sortbuf[i], b = sortbuf[l];
          const bool up = (i & k) == 0;
          if ((a > b) == up) { sortbuf[i] = b; sortbuf[l] = a; }
        }
      }
      __syncthreads();
    }
  if (t == 0) {
    int nu = s_nunder, nbr = s_nbr;
    const int nafter = n - nu, nmax_dupl = mwalk - nafter;
    auto at1 = [&](int i) { return n > 0 ? sortbuf[max(1, i) - 1] : 0.0; };   // 1-based, clamped below
    double wt_split = fmax(2.0, at1(n - nmax_dupl));
    int nmore = 0;
    for (int iw = max(1, n - nmax_dupl); iw <= n; iw++) {
      const double r = sortbuf[iw - 1] / wt_split;
      if (r >= 8.0) nmore += 3;
      else if (r >= 4.0) nmore += 2;
      else if (r >= 2.0) nmore += 1;
    }
    wt_split = (n - nmax_dupl + nmore < n) ? fmax(2.0, at1(n - nmax_dupl + nmore)) : 1e99;
    int st4[4] = {brng[0], brng[1], brng[2], brng[3]};
    int n2 = n, status = 0;
    for (int iw = 0; iw < n && !status; iw++) {
      if (!(w[iw] > wt_split)) continue;
      nbr++;
      int dst;
      if (nu > 0) {
        dst = under[--nu];
      } else {
        if (n2 >= mwalk) { status = 1; break; }
        dst = n2++;
      }
      w[iw] *= 0.5;
      w[dst] = w[iw];
      origin[dst] = origin[iw];
      for (int k = 0; k < 4; k++) rng[4 * dst + k] = (int)(rannyu(st4) * 4096.0) & 4095;
      rng[4 * dst + 3] |= 1;
    }
    while (!status && nu > 0) {
      const int dst = under[--nu], src = n2 - 1;
      n2--;
      if (dst != src) {
        w[dst] = w[src];
        origin[dst] = origin[src];
        for (int k = 0; k < 4; k++) rng[4 * dst + k] = rng[4 * src + k];
      }
    }
    brng[0] = st4[0]; brng[1] = st4[1]; brng[2] = st4[2]; brng[3] = st4[3];
    const int nl = (n2 + 7) / 8 * 8;     // the buffers hold mwalk rounded up to 8
    for (int d = n2; d < nl; d++) {       // padding records: copies of the first walker, weight 0
      origin[d] = origin[0];
      w[d] = 0.0;
      for (int k = 0; k < 4; k++) rng[4 * d + k] = rng[k];
    }
    res[0] = n2; res[1] = nbr; res[2] = status; res[3] = nl;
  }
}

// ---------------------------------------------------------------------------
// T-moves (qmc/dmc/dmc_elec/tmove.f90:1-182), one electron per step, one
// thread per walker.  Point c of walker w sits at column c * cw + w of the
// point buffers (cw: the column stride, at least the launch's walkers), so
// that each column is one k_ratio_1e launch over all walkers.

// nonlocal channel values of electron-centre distance r and whether one is
// in range (the criterion of the pseudopotential evaluation)
__device__ __forceinline__ bool tm_channels(const DevSys &S, double r_en, int ict, double *vps) {
  const double r = fmax(1e-10, r_en);
  bool inr = false;
  for (int l = 1; l <= cst(S.npotd)[ict]; l++) {
    vps[l - 1] = ecp_radial(S, r, l - 1, ict);
    if (l != cst(S.lpotp1)[ict] && fabs(vps[l - 1]) > 1e-4) inr = true;
  }
  return inr;
}

// weight max(0, sum_l w_q (exp(-tau v_l) - 1) P_l(cos) ratio) of one point
__device__ __forceinline__ double tm_weight(const DevSys &S, int ict, const double *vps, int iq,
                                            double costh, double ratio, double tau) {
  double g = 0.0;
  const double wq = cst(S.q0)[3 * S.nquad + iq];
  for (int l = 1; l <= cst(S.npotd)[ict]; l++)
    if (l != cst(S.lpotp1)[ict]) g += wq * (exp(-tau * vps[l - 1]) - 1.0) * legendre_w(l, costh) * ratio;
  return fmax(0.0, g);
}

// step 1 (tmove.f90:50-98): the points of electron ie's pairs in range, in
// pair order: those kept by the walker's last evaluation while it has
// accepted no T-move in this generation, else those of a fresh rotation
// drawn per pair.  tmq: unit vectors; tmx: positions (the electron's own
// position past the walker's points, so every column is a valid move).
__global__ void k_tm_points(KArgs A, int ie, int cw, double *tmq, double *tmx, double *dm,
                            int *need) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  const DevSys &S = A.S;
  const int M = S.M, nq = S.nquad, ncol = S.tm_kmax * nq;
  const double *ws = A.st + (int64_t)w * A.L.stride;
  double *d = dm + (int64_t)w * kDmStride;
  const double *tqw = A.tq + (int64_t)w * A.tq_stride;
  const double *ren = ws + A.L.r_en + ie * M;
  const double *xe = ws + A.L.x + 3 * ie;
  const bool fresh = d[kDmTacc] != 0.0;
  int *rs = A.rng + 4 * w;
  int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
  double vps[QMCW_MAX_LPOT];
  int j = 0;
  for (int ic = 0; ic < M && j < S.tm_kmax; ic++) {
    if (!tm_channels(S, ren[ic], cst(S.iwctype)[ic], vps)) continue;
    double ax[9];
    if (fresh) rot_axes(st4, ax);
    for (int iq = 0; iq < nq; iq++) {
      double q[3];
      if (fresh) {
        const double a = cst(S.q0)[iq], b = cst(S.q0)[nq + iq], c = cst(S.q0)[2 * nq + iq];
        q[0] = a * ax[0] + b * ax[3] + c * ax[6];
        q[1] = a * ax[1] + b * ax[4] + c * ax[7];
        q[2] = a * ax[2] + b * ax[5] + c * ax[8];
      } else {
        const double *qp = tqw + tq_point(S, ie, j, iq);
        q[0] = qp[0]; q[1] = qp[1]; q[2] = qp[2];
      }
      const int64_t col = (int64_t)(j * nq + iq) * cw + w;
      for (int k = 0; k < 3; k++) {
        tmq[3 * col + k] = q[k];
        tmx[3 * col + k] = ren[ic] * q[k] + cst(S.cent)[3 * ic + k];
      }
    }
    j++;
  }
  for (int c = j * nq; c < ncol; c++)
    for (int k = 0; k < 3; k++) {
      tmq[3 * ((int64_t)c * cw + w) + k] = 0.0;
      tmx[3 * ((int64_t)c * cw + w) + k] = xe[k];
    }
  d[kDmTmN] = j;
  need[w] = fresh ? j * nq : 0;     // fresh points that need ratios
  rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
}

// step 2 (tmove.f90:100-150, discrete_rand :184-205): the weights of the
// points (kept ratios, or rcol: those of the fresh points), the choice
// among staying and the points, and for det_balance1 the backward points of
// a fresh rotation about the chosen point's centre (tmb, tmbq; the
// electron's own position when nothing was chosen); nback[w]: backward
// points to evaluate, sel[w]: a move was chosen
__global__ void k_tm_select(KArgs A, int ie, int cw, const double *tmq, const double *tmx,
                            const double *rcol, double *tmb, double *tmbq, double *dm,
                            int *nback, int *sel, qmcw_dmc_params p) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  const DevSys &S = A.S;
  const int M = S.M, nq = S.nquad;
  const double *ws = A.st + (int64_t)w * A.L.stride;
  double *d = dm + (int64_t)w * kDmStride;
  const double *tqw = A.tq + (int64_t)w * A.tq_stride;
  const double *ren = ws + A.L.r_en + ie * M, *rven = ws + A.L.rvec_en + (int64_t)ie * M * 3;
  const double *xe = ws + A.L.x + 3 * ie;
  const bool fresh = d[kDmTacc] != 0.0;
  const int npair = (int)d[kDmTmN];
  int *rs = A.rng + 4 * w;
  int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
  double vps[QMCW_MAX_LPOT];
  // two passes over the same weights: their sum, then the cumulative choice
  double tot = 1.0, rnd = 0.0, psum = 1.0;
  int which = -1, wic = -1;
  double wpr = 0.0;
  int64_t wcol = 0;
  for (int pass = 0; pass < 2 && which < 0; pass++) {
    if (pass == 1) {
      rnd = tot * rannyu(st4);
      if (psum > rnd) { which = 0; break; }
    }
    int j = 0, last = -1, lic = -1;
    double lpr = 0.0;
    for (int ic = 0; ic < M && j < npair; ic++) {
      const int ict = cst(S.iwctype)[ic];
      if (!tm_channels(S, ren[ic], ict, vps)) continue;
      for (int iq = 0; iq < nq; iq++) {
        const int64_t col = (int64_t)(j * nq + iq) * cw + w;
        const double *q = tmq + 3 * col, *rv = rven + 3 * ic;
        const double costh = (rv[0] * q[0] + rv[1] * q[1] + rv[2] * q[2]) / ren[ic];
        const double ratio = fresh ? rcol[8 * col] : tqw[tq_ratio(S, ie, j, iq)];
        const double g = tm_weight(S, ict, vps, iq, costh, ratio, p.tau);
        if (pass == 0) {
          tot += g;
        } else {
          psum += g;
          last = j * nq + iq; lic = ic; lpr = ratio;
          if (psum > rnd) { which = last + 1; wic = ic; wpr = ratio; wcol = col; break; }
        }
      }
      if (which >= 0) break;
      j++;
    }
    if (pass == 1 && which < 0 && last >= 0) {   // rounding: the last point
      which = last + 1; wic = lic; wpr = lpr; wcol = (int64_t)last * cw + w;
    }
    if (pass == 1 && which < 0) which = 0;
  }
  d[kDmTmSel] = which > 0 ? 1.0 : 0.0;
  sel[w] = which > 0 ? 1 : 0;
  nback[w] = (which > 0 && p.tmoves == 1) ? nq : 0;
  if (which > 0) {
    for (int k = 0; k < 3; k++) d[kDmTmX + k] = tmx[3 * wcol + k];
    d[kDmTmPr] = wpr;
    d[kDmTmTot] = tot;
    d[kDmTmIc] = wic;
  }
  double ax[9];
  const bool back = which > 0 && p.tmoves == 1;
  if (back) rot_axes(st4, ax);
  for (int jq = 0; jq < nq; jq++) {
    const int64_t col = (int64_t)jq * cw + w;
    if (back) {
      const double a = cst(S.q0)[jq], b = cst(S.q0)[nq + jq], c = cst(S.q0)[2 * nq + jq];
      const double q[3] = {a * ax[0] + b * ax[3] + c * ax[6], a * ax[1] + b * ax[4] + c * ax[7],
                           a * ax[2] + b * ax[5] + c * ax[8]};
      for (int k = 0; k < 3; k++) {
        tmbq[3 * col + k] = q[k];
        tmb[3 * col + k] = ren[wic] * q[k] + cst(S.cent)[3 * wic + k];
      }
    } else {
      for (int k = 0; k < 3; k++) { tmbq[3 * col + k] = 0.0; tmb[3 * col + k] = xe[k]; }
    }
  }
  rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
}

// step 3 (tmove.f90:150-180): det_balance1 accepts with min(1, forward sum /
// backward sum) (bcol: ratios at the backward points, relative to the
// current state), casula1 always; the decision and the position to commit
// (the current one for walkers that keep the electron)
__global__ void k_tm_decide(KArgs A, int ie, int cw, const double *tmbq, const double *bcol,
                            double *xin, uint8_t *acc, double *dm, double *stale,
                            qmcw_dmc_params p) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  const DevSys &S = A.S;
  const int M = S.M, nq = S.nquad;
  const double *ws = A.st + (int64_t)w * A.L.stride;
  double *d = dm + (int64_t)w * kDmStride;
  const double *xe = ws + A.L.x + 3 * ie;
  int a = 0;
  if (d[kDmTmSel] != 0.0) {
    double pacc = 1.0;
    if (p.tmoves == 1) {
      const int ic = (int)d[kDmTmIc], ict = cst(S.iwctype)[ic];
      const double rei = ws[A.L.r_en + ie * M + ic];
      double vps[QMCW_MAX_LPOT];
      tm_channels(S, rei, ict, vps);
      double xen[3];
      for (int k = 0; k < 3; k++) xen[k] = d[kDmTmX + k] - cst(S.cent)[3 * ic + k];
      double bwd = 1.0;
      for (int jq = 0; jq < nq; jq++) {
        const int64_t col = (int64_t)jq * cw + w;
        const double *q = tmbq + 3 * col;
        const double costh = (xen[0] * q[0] + xen[1] * q[1] + xen[2] * q[2]) / rei;
        bwd += tm_weight(S, ict, vps, jq, costh, bcol[8 * col] / d[kDmTmPr], p.tau);
      }
      pacc = fmin(1.0, d[kDmTmTot] / bwd);
    }
    int *rs = A.rng + 4 * w;
    int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
    a = (rannyu(st4) < pacc) || p.tmoves == 2;
    rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
    if (a) { d[kDmTacc] = 1.0; d[kDmAcc] = 1.0; stale[w] = 1.0; }
  }
  acc[w] = (uint8_t)a;
  for (int k = 0; k < 3; k++) xin[3 * w + k] = a ? d[kDmTmX + k] : xe[k];
}

__global__ void k_fill_d(double *a, double v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

}  // namespace qmcw
