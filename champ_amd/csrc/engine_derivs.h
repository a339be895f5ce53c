// engine_derivs.h -- parameter derivatives of ln Psi and of the local energy
// (the optimisation sample): Jastrow parameter derivatives, the
// pseudopotential pieces (psp_nonloc_orb over all orbitals, dpe), the
// orbital-resolved local energy and the CSF / Jastrow / orbital-rotation
// blocks.  Everything here runs in FP64.
#pragma once
#include "engine_kernels.h"

namespace qmcw {

// Per-slot derivative pieces go to base[slot * rs + c], c = 0 value,
// 1..3 gradient, 4 laplacian.
struct SlotAcc {
  double *base;
  int64_t rs;
  __device__ __forceinline__ void add(int slot, int c, double v) const { base[slot * rs + c] += v; }
  // all five pieces of a slot: the old values are read together before the
  // stores, one memory round trip instead of five
  __device__ __forceinline__ void add5(int slot, double v, double g0, double g1, double g2,
                                       double l) const {
    double *p = base + slot * rs;
    const double o0 = p[0], o1 = p[1], o2 = p[2], o3 = p[3], o4 = p[4];
    p[0] = o0 + v; p[1] = o1 + g0; p[2] = o2 + g1; p[3] = o3 + g2; p[4] = o4 + l;
  }
};

// a-coefficient pieces of the e-n terms of one electron (distances r[M],
// vectors rv[3M]), qmc/vmc/deriv_jastrow4.f90:880-1090.  cv scales the value
// pieces (0: none); FULL adds the gradient and the laplacian.
template <bool FULL, typename RA, typename RVA>
__device__ void en_slots(const DevSys &S, RA r, RVA rv, double cv, const SlotAcc &acc) {
  const double k = S.scalek, ar = S.asymp_r;
  const int na = S.norda + 1;
  for (int ic = 0; ic < S.M; ic++) {
    const int it = cst(S.iwctype)[ic];
    const double ri = r[ic];
    double rr, d7, d9;
    scaled<double>(S.isc, k, ri, rr, d7, d9);
    const double *a = S.a4 + it * na;
    double rp = rr, arp = ar, rpm1 = 1.0, rpm2 = 0.0;   // rr^iord, ar^iord, rr^(iord-1), rr^(iord-2)
    for (int iw = 1; iw <= na; iw++) {
      double gen, geni, genii;
      if (iw == 1) {
        double bt = 1.0 + a[1] * rr, bta = 1.0 + a[1] * ar;
        gen = rr / bt - ar / bta;
        geni = 1.0 / (bt * bt);
        genii = -2.0 * a[1] * geni / bt;
      } else if (iw == 2) {
        double t = -a[0] * rr * rr, ti = -2.0 * a[0] * rr, tii = -2.0 * a[0];
        double bot0 = 1.0 + a[1] * rr, bt = bot0 * bot0, bt2 = bt * bt;
        double bta = 1.0 + a[1] * ar;
        gen = t / bt + a[0] * ar * ar / (bta * bta);
        geni = ti / (bt * bot0);
        genii = tii * (1.0 - 2.0 * a[1] * rr) / bt2;
      } else {
        // iord = iw - 1 >= 2: advance the powers to iord
        rpm2 = rpm1;
        rpm1 = rp;
        rp *= rr;
        arp *= ar;
        const int iord = iw - 1;
        gen = rp - arp;
        geni = iord * rpm1;
        genii = iord * (iord - 1) * rpm2;
      }
      const int slot = it * na + iw - 1;
      if (FULL) {
        genii = genii * d7 * d7 + geni * d9;
        geni = geni * d7 / ri;
        acc.add5(slot, cv != 0.0 ? cv * gen : 0.0, geni * rv[3 * ic], geni * rv[3 * ic + 1],
                 geni * rv[3 * ic + 2], genii + 2.0 * geni);
      } else if (cv != 0.0) {
        acc.add(slot, 0, cv * gen);
      }
    }
  }
}

// e-e-n term pieces of one (pair, centre): term values U^k (A^(l+m) B^m +
// A^m B^(l+m)) and, with FULL, their gradient wrt r_i and the part of the
// laplacian that belongs to electron i (qmc/vmc/deriv_jastrow4.f90:330-720).
template <bool FULL, int ORD>
__device__ __forceinline__ void een_slots(double U, double A, double B, double Ud1, double Ud2,
                                          double Ad1, double Ad2, double rij, double ri,
                                          const double *rvij, const double *rvi, double u2pst,
                                          double cv, const SlotAcc &acc, int slot0) {
  double up[ORD + 1], ap[ORD + 1], bp[ORD + 1];
  up[0] = ap[0] = bp[0] = 1.0;
#pragma unroll
  for (int q = 1; q <= ORD; q++) { up[q] = up[q - 1] * U; ap[q] = ap[q - 1] * A; bp[q] = bp[q - 1] * B; }
  int idx = 0;
#pragma unroll
  for (int k = 0; k <= ORD; k++)
#pragma unroll
    for (int m = 0; 2 * m <= ORD - k; m++)
#pragma unroll
      for (int l = 0; l <= ORD - k - 2 * m; l++)
        if (QMCW_EEN_TERM(k, l, m)) {
          const int lm = l + m;
          const int slot = slot0 + idx;
          idx++;
          const double P = ap[lm] * bp[m] + ap[m] * bp[lm];
          if (!FULL && cv != 0.0) acc.add(slot, 0, cv * up[k] * P);
          if (FULL) {
            double PA = (lm ? lm * ap[lm > 0 ? lm - 1 : 0] * bp[m] : 0.0) +
                        (m ? m * ap[m > 0 ? m - 1 : 0] * bp[lm] : 0.0);
            double PAA = (lm > 1 ? lm * (lm - 1) * ap[lm > 1 ? lm - 2 : 0] * bp[m] : 0.0) +
                         (m > 1 ? m * (m - 1) * ap[m > 1 ? m - 2 : 0] * bp[lm] : 0.0);
            double Uk = up[k];
            double Uk1 = k ? k * up[k > 0 ? k - 1 : 0] : 0.0;
            double Uk2 = k > 1 ? k * (k - 1) * up[k > 1 ? k - 2 : 0] : 0.0;
            double gu = Uk1 * P, guu = Uk2 * P, gi = Uk * PA, gii = Uk * PAA, gui = Uk1 * PA;
            guu = guu * Ud1 * Ud1 + gu * Ud2;
            gu = gu * Ud1 / rij;
            gui = gui * Ud1 * Ad1;
            gii = gii * Ad1 * Ad1 + gi * Ad2;
            gi = gi * Ad1 / ri;
            acc.add5(slot, cv != 0.0 ? cv * up[k] * P : 0.0, gi * rvi[0] + gu * rvij[0],
                     gi * rvi[1] + gu * rvij[1], gi * rvi[2] + gu * rvij[2],
                     2.0 * (gu + gi) + guu + gii + gui * u2pst / (ri * rij));
          }
        }
}

// b-coefficient and e-e-n term pieces of the pair (i, j) seen from electron
// i: rij, rvij = r_i - r_j, the distances of i and j to the centres and i's
// vectors to them (qmc/vmc/deriv_jastrow4.f90:95-720).  cv scales the value
// pieces (0: none); FULL adds the gradient wrt r_i and i's laplacian part.
template <bool FULL>
__device__ void b_slots(const DevSys &S, bool par, double rij, const double *rvij, double cv,
                        const SlotAcc &acc) {
  const double k = S.scalek, ar = S.asymp_r;
  double sspinn = 1.0;
  int isb = 0;
  if (par) {
    if (S.nspin2b == 2) isb = 1;
    else if (S.nocuspb == 0) sspinn = 0.5;
  }
  const int nb = S.nordb + 1;
  const double *b = S.b + isb * nb;
  double u, dd1, dd2;
  scaled<double>(S.isc, k, rij, u, dd1, dd2);
  double up_ = u, arp = ar, upm1 = 1.0, upm2 = 0.0;
  for (int iw = 1; iw <= nb; iw++) {
    double gee, geeu, geeuu;
    if (iw == 1) {
      double bt = 1.0 + b[1] * u, bta = 1.0 + b[1] * ar;
      gee = sspinn * u / bt - sspinn * ar / bta;
      geeu = sspinn / (bt * bt);
      geeuu = -2.0 * b[1] * geeu / bt;
    } else if (iw == 2) {
      double t = -sspinn * b[0] * u * u, tu = -sspinn * 2.0 * b[0] * u, tuu = -sspinn * 2.0 * b[0];
      double bot0 = 1.0 + b[1] * u, bt = bot0 * bot0, bt2 = bt * bt;
      double bta = 1.0 + b[1] * ar;
      gee = t / bt + sspinn * b[0] * ar * ar / (bta * bta);
      geeu = tu / (bt * bot0);
      geeuu = tuu * (1.0 - 2.0 * b[1] * u) / bt2;
    } else {
      upm2 = upm1;
      upm1 = up_;
      up_ *= u;
      arp *= ar;
      const int iord = iw - 1;
      gee = up_ - arp;
      geeu = iord * upm1;
      geeuu = iord * (iord - 1) * upm2;
    }
    const int slot = S.slot_b0 + isb * nb + iw - 1;
    if (FULL) {
      geeuu = geeuu * dd1 * dd1 + geeu * dd2;
      geeu = geeu * dd1 / rij;
      acc.add5(slot, cv != 0.0 ? cv * gee : 0.0, geeu * rvij[0], geeu * rvij[1], geeu * rvij[2],
               geeuu + 2.0 * geeu);
    } else if (cv != 0.0) {
      acc.add(slot, 0, cv * gee);
    }
  }
}

template <bool FULL, int ORD>
__device__ void pair_slots(const DevSys &S, bool par, double rij, const double *rvij,
                           const double *ri_v, const double *rj_v, const double *rvi, double cv,
                           const SlotAcc &acc) {
  const double k = S.scalek;
  b_slots<FULL>(S, par, rij, rvij, cv, acc);
  if (S.nordc <= 1) return;
  double U, Ud1, Ud2;
  scaled<double>(S.isc, k, rij, U, Ud1, Ud2);
  U = 1.0 - k * U; Ud1 = -k * Ud1; Ud2 = -k * Ud2;
  for (int ic = 0; ic < S.M; ic++) {
    const int it = cst(S.iwctype)[ic];
    const double ri = ri_v[ic], rj = rj_v[ic];
    double A, Ad1, Ad2, B, Bd1, Bd2;
    scaled<double>(S.isc, k, ri, A, Ad1, Ad2);
    scaled<double>(S.isc, k, rj, B, Bd1, Bd2);
    A = 1.0 - k * A; Ad1 = -k * Ad1; Ad2 = -k * Ad2;
    B = 1.0 - k * B;
    (void)Bd1; (void)Bd2;
    const double u2pst = rij * rij + ri * ri - rj * rj;
    een_slots<FULL, ORD>(U, A, B, Ud1, Ud2, Ad1, Ad2, rij, ri, rvij,
                         FULL ? rvi + 3 * ic : nullptr, u2pst, cv, acc,
                         S.slot_c0 + it * S.nterms);
  }
}

// Value pieces, weighted by cv, of the b-coefficient and e-e-n terms of all
// pairs (i, j), electron i at xn (distances ri_v to the centres) and every
// other electron where it is (btd[j * M + ic] = 1 - k scaled(r_j,ic);
// uqd: per-lane scratch of N values).  Summed over j, the term values of a
// centre are A^(l+m) S(k, m) + A^m S(k, l+m) with S(k, n) = sum_j U_j^k B_j^n,
// so each slot takes one addition per call.
template <int ORD, typename RA, typename UA>
__device__ void point_pair_slots(const DevSys &S, int i, const double *xn, const double *x,
                                 RA ri_v, const double *btd, UA uqd, double cv, const SlotAcc &acc) {
  const int N = S.N, M = S.M;
  const double k = S.scalek, ar = S.asymp_r;
  // b-coefficient pieces summed over the other electrons in registers
  // (b_slots' value terms), one update per slot afterwards
  const int nb = S.nordb + 1;
  double ab0[QMCW_MAX_ORD + 1], ab1[QMCW_MAX_ORD + 1];
#pragma unroll
  for (int q = 0; q <= QMCW_MAX_ORD; q++) ab0[q] = ab1[q] = 0.0;
  for (int jj = 0; jj < N; jj++) {
    if (jj == i) continue;
    const double dx = xn[0] - x[3 * jj], dy = xn[1] - x[3 * jj + 1], dz = xn[2] - x[3 * jj + 2];
    const double rij = sqrt(dx * dx + dy * dy + dz * dz);
    double u, d1, d2;
    scaled<double>(S.isc, k, rij, u, d1, d2);
    uqd[jj] = 1.0 - k * u;
    double sspinn = 1.0;
    int isb = 0;
    if ((i < S.nup) == (jj < S.nup)) {
      if (S.nspin2b == 2) isb = 1;
      else if (S.nocuspb == 0) sspinn = 0.5;
    }
    const cptr<double> b = cst(S.b) + isb * nb;
    double up_ = u, arp = ar;
#pragma unroll
    for (int iw = 1; iw <= QMCW_MAX_ORD + 1; iw++) {
      if (iw > nb) continue;
      double gee;
      if (iw == 1) {
        const double bt = 1.0 + b[1] * u, bta = 1.0 + b[1] * ar;
        gee = sspinn * u / bt - sspinn * ar / bta;
      } else if (iw == 2) {
        const double t = -sspinn * b[0] * u * u;
        const double bot0 = 1.0 + b[1] * u, bt = bot0 * bot0, bta = 1.0 + b[1] * ar;
        gee = t / bt + sspinn * b[0] * ar * ar / (bta * bta);
      } else {
        up_ *= u;
        arp *= ar;
        gee = up_ - arp;
      }
      if (isb) ab1[iw - 1] += cv * gee;
      else ab0[iw - 1] += cv * gee;
    }
  }
  {
    double *bs = acc.base + (int64_t)S.slot_b0 * acc.rs;
    const int nsb = S.nspin2b == 2 ? 2 : 1;
    double o0[QMCW_MAX_ORD + 1], o1[QMCW_MAX_ORD + 1];
#pragma unroll
    for (int q = 0; q <= QMCW_MAX_ORD; q++) {
      o0[q] = q < nb ? bs[(int64_t)q * acc.rs] : 0.0;
      o1[q] = (nsb == 2 && q < nb) ? bs[(int64_t)(nb + q) * acc.rs] : 0.0;
    }
#pragma unroll
    for (int q = 0; q <= QMCW_MAX_ORD; q++) {
      if (q < nb) bs[(int64_t)q * acc.rs] = o0[q] + ab0[q];
      if (nsb == 2 && q < nb) bs[(int64_t)(nb + q) * acc.rs] = o1[q] + ab1[q];
    }
  }
  if (S.nordc <= 1) return;
  constexpr int NG = een_gi<ORD>(ORD, 0) + 1;
  for (int ic = 0; ic < M; ic++) {
    double sg[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) sg[q] = 0.0;
    for (int jj = 0; jj < N; jj++) {
      if (jj == i) continue;
      const double U = uqd[jj], B = btd[jj * M + ic];
      double uk = 1.0;
#pragma unroll
      for (int kk = 0; kk <= ORD; kk++) {
        double t = uk;
#pragma unroll
        for (int n = 0; n <= ORD - kk; n++) {
          sg[een_gi<ORD>(kk, n)] += t;
          t *= B;
        }
        uk *= U;
      }
    }
    double A, t1, t2;
    scaled<double>(S.isc, k, ri_v[ic], A, t1, t2);
    A = 1.0 - k * A;
    double ap[ORD + 1];
    ap[0] = 1.0;
#pragma unroll
    for (int q = 1; q <= ORD; q++) ap[q] = ap[q - 1] * A;
    const int slot0 = S.slot_c0 + cst(S.iwctype)[ic] * S.nterms;
    int idx = 0;
#pragma unroll
    for (int kk = 0; kk <= ORD; kk++)
#pragma unroll
      for (int m = 0; 2 * m <= ORD - kk; m++)
#pragma unroll
        for (int l = 0; l <= ORD - kk - 2 * m; l++)
          if (QMCW_EEN_TERM(kk, l, m)) {
            acc.add(slot0 + idx, 0, cv * (ap[l + m] * sg[een_gi<ORD>(kk, m)] +
                                          ap[m] * sg[een_gi<ORD>(kk, l + m)]));
            idx++;
          }
  }
}

// ---------------------------------------------------------------------------
// Jastrow parameter derivatives of the current configuration: gvalue(p),
// g(p, i) = d grad_i ln J / dp and d2g(p) = d lap ln J / dp
// (qmc/vmc/deriv_jastrow4.f90).  Each lane takes whole electrons: it visits
// every pair (i, j) from i's side, so its rows of the slot pieces are its own.
template <int G, int ORD>
__global__ QMCW_HEAVY void k_deriv_jas(KArgs A) {
  QMCW_GROUP_SETUP
  const int N = S.N, M = S.M, nsl = S.nslot;
  const WLayout &L = A.L;
  const DLayout &DL = A.DL;
  const double *ren = ws + L.r_en, *rven = ws + L.rvec_en, *ree = ws + L.r_ee, *rvee = ws + L.rvec_ee;
  double *raw = ss + DL.raw;
  for (int t = lane; t < nsl * N * 5; t += G) raw[t] = 0.0;
  __syncthreads();
  for (int i = lane; i < N; i += G) {
    const SlotAcc acc{raw + 5 * i, (int64_t)N * 5};
    en_slots<true>(S, ren + i * M, rven + i * M * 3, 1.0, acc);
    for (int j = 0; j < N; j++) {
      if (j == i) continue;
      const int hi = i > j ? i : j, lo = i > j ? j : i;
      const double sg = i > j ? 1.0 : -1.0;   // rvec_ee(hi, lo) = r_hi - r_lo
      const double *rv = rvee + 3 * (hi * N + lo);
      const double rvij[3] = {sg * rv[0], sg * rv[1], sg * rv[2]};
      const bool par = (hi < S.nup) || (lo >= S.nup);
      pair_slots<true, ORD>(S, par, ree[hi * N + lo], rvij, ren + i * M, ren + j * M,
                       rven + i * M * 3, j < i ? 1.0 : 0.0, acc);
    }
  }
  __syncthreads();
  double *gv = ss + DL.gvalue, *gj = ss + DL.gjas, *d2g = ss + DL.d2g;
  for (int t = lane; t < S.npj * N; t += G) {
    const int p = t / N, i = t % N;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int q = cst(S.pj_off)[p]; q < cst(S.pj_off)[p + 1]; q++) {
      const double w = cst(S.pj_w)[q];
      const double *r = raw + ((int64_t)cst(S.pj_slot)[q] * N + i) * 5;
      g0 += w * r[1]; g1 += w * r[2]; g2 += w * r[3];
    }
    gj[3 * t] = g0; gj[3 * t + 1] = g1; gj[3 * t + 2] = g2;
  }
  for (int p = lane; p < S.npj; p += G) {
    double v = 0.0, d = 0.0;
    for (int q = cst(S.pj_off)[p]; q < cst(S.pj_off)[p + 1]; q++) {
      const double w = cst(S.pj_w)[q];
      const double *r = raw + (int64_t)cst(S.pj_slot)[q] * N * 5;
      double sv = 0.0, sd = 0.0;
      for (int i = 0; i < N; i++) { sv += r[5 * i]; sd += r[5 * i + 4]; }
      v += w * sv;
      d += w * sd;
    }
    gv[p] = v;
    d2g[p] = d;
  }
}

// The (electron, centre) pairs plist[kp] taken by this lane, whole: the
// weighted basis sums (pbc rows) and the slot pieces of dpe (dpe_raw rows)
// over the pair's quadrature points.  rnew: 4M doubles, uq: N values of the
// kernel's T, uqd: N doubles of per-lane scratch (LDS or global), stored
// lane-interleaved.
template <typename T, int G, int ORD>
__device__ __forceinline__ void deriv_nlpp_pairs(const KArgs &A, const double *ws, double *ss, int npair,
                                                 const int *plist, LaneArr<double> rnew, LaneArr<T> uq,
                                                 LaneArr<double> uqd, int lane) {
  const DevSys &S = A.S;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const DLayout &DL = A.DL;
  const int N = S.N, M = S.M, NB = S.NB, nsl = S.nslot, nq = S.nquad;
  const double *x = ws + L.x, *ren = ws + L.r_en, *rven = ws + L.rvec_en, *fso = ws + L.fso;
  const double *vnl = ss + SL.vnl, *quad = ws + L.quad, *gp = ss + SL.gproj, *btd = ss + DL.btd;
  const T *bt = reinterpret_cast<const T *>(ss + SL.bt);    // filled by k_full_nlpp
  double *pbc = ss + DL.pbc, *draw = ss + DL.dpe_raw;
  const LaneArr<double> rvnew = rnew + M;
  for (int kp = lane; kp < npair; kp += G) {
    const int p = plist[kp];
    double *pbrow = pbc + (int64_t)p * NB;
    double *drow = draw + (int64_t)p * nsl;
    const double *v = vnl + p * (QMCW_MAX_LPOT + 1);
    for (int b = 0; b < NB; b++) pbrow[b] = 0.0;
    const int i = p / M, ic = p % M, ict = cst(S.iwctype)[ic];
    const double *gpi = gp + (int64_t)i * NB;
    const double rei = ren[p];
    const SlotAcc acc{drow, 1};
    double fold = fso[i * N + i];
    for (int jj = 0; jj < N; jj++)
      if (jj != i) fold += fso[jj < i ? i * N + jj : jj * N + i];
    double wsum = 0.0;
    for (int iq = 0; iq < nq; iq++) {
      const double qx = quad[iq], qy = quad[nq + iq], qz = quad[2 * nq + iq];
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
      // Jastrow factor and channel sum first, then one pass over the shells
      // forms both the determinant ratio and the weighted basis sums
      const double delta = (double)jas_point_val<T, ORD>(S, i, xn, x, rnew, bt, uq) - fold;
      double vpot = 0.0;
      for (int l = 1; l <= cst(S.npotd)[ict]; l++)
        if (l != cst(S.lpotp1)[ict]) vpot += v[l - 1] * legendre_w(l, costh);
      const double fac = cst(S.q0)[3 * nq + iq] * exp(delta) * vpot;
      double rat = 0.0;
      for (int sh = 0; sh < S.nshell; sh++) {
        const int jc = cst(S.sh_cent)[sh];
        shell_val_axpy<T>(S, sh, (T)rvnew[3 * jc], (T)rvnew[3 * jc + 1], (T)rvnew[3 * jc + 2],
                          (T)rnew[jc], gpi, fac, pbrow, rat);
      }
      const double wr = fac * rat;
      wsum += wr;
      en_slots<false>(S, rnew, rvnew, wr, acc);
      point_pair_slots<ORD>(S, i, xn, x, rnew, btd, uqd, wr, acc);
    }
    // the same terms at the current position, weighted by the summed weight
    en_slots<false>(S, ren + i * M, rven + i * M * 3, -wsum, acc);
    point_pair_slots<ORD>(S, i, x + 3 * i, x, ren + i * M, btd, uqd, -wsum, acc);
  }
}

// ---------------------------------------------------------------------------
// Pseudopotential pieces of the derivatives (qmc/vmc/nonloc.f90:184-197,
// qmc/fit/deriv_nonloc_pot.f90:59-64): psp_nonloc_orb(i, q) over all
// orbitals and dpe(p) = sum over quadrature points of
// w exp(dJ) v P_l (Psi_D ratio) d(dJ)/dp.  Each lane takes whole
// (electron, centre) pairs and owns their rows; k_full_nlpp has filled the
// channel values (SLayout::vnl) of the current configuration.
// psp_nonloc_orb is linear in the orbitals, psp_nonloc_orb(i, q) =
// sum_b coef(b, q) sum_points fac phi_b(point), so per point only the
// weighted basis values are accumulated (NB terms) and the contraction with
// the coefficients runs once per electron; the determinant ratio at a point
// is sum_b gproj_i[b] phi_b(point).
template <typename T, int G, int ORD>
__global__ QMCW_HEAVY void k_deriv_nlpp(KArgs A) {
  QMCW_GROUP_SETUP
  const int N = S.N, M = S.M, NO = S.NO, NB = S.NB, nsl = S.nslot;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const DLayout &DL = A.DL;
  const double *ren = ws + L.r_en;
  const double *vnl = ss + SL.vnl;
  double *pbc = ss + DL.pbc, *draw = ss + DL.dpe_raw;
  double *btd = ss + DL.btd;
  for (int p = lane; p < N * M; p += G) {
    double bu, t1, t2;
    scaled<double>(S.isc, S.scalek, ren[p], bu, t1, t2);
    btd[p] = 1.0 - S.scalek * bu;
    double *drow = draw + (int64_t)p * nsl;
    for (int q = 0; q < nsl; q++) drow[q] = 0.0;
  }
  // whole pairs in range per lane
  int *plist = reinterpret_cast<int *>(ss + SL.plist);
  const int npair = group_compact<G>(vnl + QMCW_MAX_LPOT, QMCW_MAX_LPOT + 1, N * M, plist, lane);
  __syncthreads();
  // per-lane distances of the points and e-e-n variables of the other
  // electrons: LDS when the launch reserved it
  if (A.lds_dnlpp) {
    // the block's 64 lanes interleaved: 4M doubles, N values of T, N doubles
    extern __shared__ double dyn_lds[];
    const LaneArr<double> rq{dyn_lds + threadIdx.x, kWave};
    const LaneArr<T> uq{reinterpret_cast<T *>(dyn_lds + 4 * M * kWave) + threadIdx.x, kWave};
    const LaneArr<double> uqd{dyn_lds + (int64_t)(A.lds_dnlpp - N) * kWave + threadIdx.x, kWave};
    deriv_nlpp_pairs<T, G, ORD>(A, ws, ss, npair, plist, rq, uq, uqd, lane);
  } else {
    // the walker's G lanes interleaved in its scratch
    const LaneArr<double> rq{ss + SL.qd + lane, G};
    const LaneArr<T> uq{reinterpret_cast<T *>(ss + SL.uq) + lane, G};
    const LaneArr<double> uqd{ss + DL.uqd + lane, G};
    deriv_nlpp_pairs<T, G, ORD>(A, ws, ss, npair, plist, rq, uq, uqd, lane);
  }
  __syncthreads();
  // per-electron basis sums over the centres in range, kept in the row of
  // centre 0 (each lane reads and writes only its own (i, b) column)
  for (int t = lane; t < N * NB; t += G) {
    const int i = t / NB, b = t % NB;
    double a = 0.0;
    for (int ic = 0; ic < M; ic++)
      if (vnl[(i * M + ic) * (QMCW_MAX_LPOT + 1) + QMCW_MAX_LPOT] != 0.0)
        a += pbc[(int64_t)(i * M + ic) * NB + b];
    pbc[(int64_t)i * M * NB + b] = a;
  }
  double *dsl = ss + DL.dpe_slot;
  for (int sl = lane; sl < nsl; sl += G) {
    double a = 0.0;
    for (int p = 0; p < N * M; p++) a += draw[(int64_t)p * nsl + sl];
    dsl[sl] = a;
  }
  __syncthreads();
  double *pnl = ss + DL.pnl;
  // (k_pnl_mfma forms pnl after this kernel when A.pnl_mfma)
  for (int t = A.pnl_mfma ? N * NO : lane; t < N * NO; t += G) {
    const int i = t / NO, q = t % NO;
    const double *row = pbc + (int64_t)i * M * NB;
    double a = 0.0;
    for (int b = 0; b < NB; b++) a += cst(S.coefT)[(int64_t)b * NO + q] * row[b];
    pnl[t] = a;
  }
  double *dpe = ss + DL.dpe;
  for (int p = lane; p < S.npj; p += G) {
    double a = 0.0;
    for (int q = cst(S.pj_off)[p]; q < cst(S.pj_off)[p + 1]; q++) a += cst(S.pj_w)[q] * dsl[cst(S.pj_slot)[q]];
    dpe[p] = a;
  }
}

// ---------------------------------------------------------------------------
// Assembly of d ln Psi / dp and d E_L / dp in the order CSF, Jastrow,
// orbitals (qmc/deriv_mod.f90:219-239): elocal_orb_bld, dt_bld,
// alpha_deloc_bld, dy_bld, dg_bld, csf_derivs_bld
// (qmc/deriv_fast_mod.f90:174-823), dpsi_orb_bld / deloc_orb_bld
// (qmc/deriv_orb_mod.f90:1486-1630) and the Jastrow denergy of
// qmc/vmc/hpsi.f90:366-374.  out[w * 2P + p] = dpsi, out[w * 2P + P + p] = deloc.
template <int G>
__global__ QMCW_HEAVY void k_derivs(KArgs A, double *out) {
  QMCW_GROUP_SETUP
  const int N = S.N, NO = S.NO, P = S.nparm;
  const WLayout &L = A.L;
  const SLayout &SL = A.SL;
  const DLayout &DL = A.DL;
  const double hb = S.hb;
  const double *orb = ss + SL.orb;
  const double *vj = ss + SL.vj, *vd = ss + SL.vd, *red = ss + SL.phs;
  const double *pnl = ss + DL.pnl;
  double *B = ss + DL.B;
  const double vpsp = S.nloc > 0 ? red[56] : 0.0;
  const double epl = ws[L.epot] - vpsp;          // local part of the potential
  double vj2 = 0.0;
  for (int q = lane; q < 3 * N; q += G) vj2 += vj[q] * vj[q];
  vj2 = gsum<G>(vj2);
  const double ddjas = (red[53] + vj2) / N;
  // ---- elocal_orb_bld
  for (int t = lane; t < N * NO; t += G) {
    const int i = t / NO;
    const double *o = orb + (int64_t)t * 5;
    const double dv = vj[3 * i] * o[1] + vj[3 * i + 1] * o[2] + vj[3 * i + 2] * o[3];
    double b = -hb * (o[4] + 2.0 * dv + ddjas * o[0]) + epl * o[0] / N;
    if (S.nloc > 0) b += pnl[t];
    B[t] = b;
  }
  __syncthreads();
  // ---- full T, dt_bld, alpha_deloc_bld and the k x k inverse products
  for (int s = 0; s < 2; s++) {
    const int ns = S.ns[s], off = s ? S.nup : 0;
    if (ns == 0) continue;
    const double *ai = w_ai(L, ws, s);
    double *tf = ss + DL.tfull[s], *dt = ss + DL.dt[s], *aib = ss + DL.aib;
    for (int e = lane; e < ns * NO; e += G) {
      const int k = e / NO, q = e % NO;
      double a1 = 0.0, a2 = 0.0;
      for (int i = 0; i < ns; i++) {
        const double aki = ai[k + ns * i];
        a1 += aki * orb[((int64_t)(off + i) * NO + q) * 5];
        a2 += aki * B[(off + i) * NO + q];
      }
      tf[e] = a1;
      aib[e] = a2;
    }
    __syncthreads();
    for (int e = lane; e < ns * NO; e += G) {
      const int k = e / NO, q = e % NO;
      double a = 0.0;
      for (int r = 0; r < ns; r++) a += aib[k * NO + cst(S.ref[s])[r]] * tf[r * NO + q];
      dt[e] = aib[e] - a;
    }
    __syncthreads();
    const double *inv = w_inv(L, ws, s);
    double *del = ss + DL.del[s], *idti = ss + DL.idti[s];
    for (int u = lane; u < S.ndetu[s]; u += G) {
      const int k = cst(S.ord[s])[u];
      const double *iv = inv + u * S.kmax2;
      const int *rf = S.rf[s] + u * ns, *le = S.exl[s] + u * ns;
      double d = 0.0;
      for (int j = 0; j < k; j++)
        for (int i = 0; i < k; i++) d += iv[j + k * i] * dt[rf[i] * NO + le[j]];
      del[u] = d;
      // idti = inv (fdt inv), fdt(j, i) = dt(rf_j, le_i)
      double *id = idti + u * S.kmax2;
      for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) {
          double a = 0.0;
          for (int kk = 0; kk < k; kk++) {
            double dti = 0.0;
            for (int k2 = 0; k2 < k; k2++) dti += dt[rf[kk] * NO + le[k2]] * iv[k2 + k * i];
            a += iv[j + k * kk] * dti;
          }
          id[j + k * i] = a;
        }
    }
    __syncthreads();
  }
  // ---- coefficient sums per unique det and esum
  const double chii = 1.0 / ws[L.chi];
  const double *dv0 = w_detv(L, ws, 0), *dv1 = w_detv(L, ws, 1);
  const double *del0 = ss + DL.del[0], *del1 = ss + DL.del[1];
  // over the dets that contain unique det u (the CSR lists of eval_y, in
  // ascending det order)
  for (int t = lane; t < S.ndetu[0] + S.ndetu[1]; t += G) {
    const int s = t < S.ndetu[0] ? 0 : 1, u = s ? t - S.ndetu[0] : t;
    const cptr<int> uiu = cst(S.uiu[s]), uid = cst(S.uid[s]);
    const cptr<double> ucf = cst(S.ucf[s]);
    double c0 = 0.0, c1 = 0.0;
    const int q1 = cst(S.udet_off[s])[u + 1];
#pragma unroll 4
    for (int q = cst(S.udet_off[s])[u]; q < q1; q++) {
      const int iu = uiu[q], id = uid[q];
      const double cf = ucf[q] * dv0[iu] * dv1[id];
      c0 += cf;
      c1 += cf * ((S.ns[0] ? del0[iu] : 0.0) + (S.ns[1] ? del1[id] : 0.0));
    }
    (ss + DL.c0[s])[u] = c0;
    (ss + DL.c1[s])[u] = c1;
  }
  double es = 0.0;
  for (int s = 0; s < 2; s++) {
    const int ns = S.ns[s], nocc = S.nocc[s];
    const double *y = w_y(L, ws, s), *dt = ss + DL.dt[s];
    for (int e = lane; e < nocc * ns; e += G) {
      const int j = e / ns, i = e % ns;
      es += y[e] * dt[i * NO + cst(S.occ[s])[j]];
    }
  }
  const double esum = gsum<G>(es);
  __syncthreads();
  // ---- dy_bld
  for (int s = 0; s < 2; s++) {
    const int ns = S.ns[s], nocc = S.nocc[s];
    const double *inv = w_inv(L, ws, s), *idti = ss + DL.idti[s], *y = w_y(L, ws, s);
    const double *c0 = ss + DL.c0[s], *c1 = ss + DL.c1[s];
    double *dy = ss + DL.dy[s];
    // contributions to entry e (the Y-entry CSR lists of eval_y, ascending u)
    const cptr<int> yoff = cst(S.yent_off[s]), yu = cst(S.yent_u[s]), yp = cst(S.yent_pos[s]);
    for (int e = lane; e < nocc * ns; e += G) {
      double a = 0.0;
      const int q1 = yoff[e + 1];
#pragma unroll 4
      for (int q = yoff[e]; q < q1; q++) {
        const int u = yu[q], o = yp[q];
        a += c1[u] * inv[o] - c0[u] * idti[o];
      }
      dy[e] = a * chii - esum * y[e];
    }
  }
  __syncthreads();
  // ---- dg_bld
  for (int s = 0; s < 2; s++) {
    const int ns = S.ns[s], nocc = S.nocc[s], off = s ? S.nup : 0;
    if (ns == 0) continue;
    const double *tf = ss + DL.tfull[s], *dt = ss + DL.dt[s], *dy = ss + DL.dy[s];
    const double *y = w_y(L, ws, s), *g = ss + SL.g[s], *ai = w_ai(L, ws, s);
    double *aux = ss + DL.aux4, *dg = ss + DL.dg[s];
    for (int e = lane; e < nocc * ns; e += G) aux[e] = 0.0;
    __syncthreads();
    for (int e = lane; e < ns * ns; e += G) {
      const int k = e / ns, i = e % ns;
      double a = 0.0;
      for (int j = 0; j < nocc; j++) {
        const int oj = cst(S.occ[s])[j];
        a += tf[k * NO + oj] * dy[j * ns + i] + dt[k * NO + oj] * y[j * ns + i];
      }
      aux[cst(S.iref[s])[k] * ns + i] = a;
    }
    __syncthreads();
    for (int e = lane; e < nocc * ns; e += G) {
      const int j = e / ns, i = e % ns;
      double a3 = 0.0;
      for (int q = 0; q < ns; q++) a3 += g[j * ns + q] * B[(off + q) * NO + cst(S.ref[s])[i]];
      aux[e] = dy[e] - a3 - aux[e];
    }
    __syncthreads();
    for (int e = lane; e < nocc * ns; e += G) {
      const int j = e / ns, i = e % ns;
      double a = 0.0;
      for (int q = 0; q < ns; q++) a += aux[j * ns + q] * ai[q + ns * i];
      dg[e] = a;
    }
    __syncthreads();
  }
  // ---- outputs
  double *o1 = out + (int64_t)w * 2 * P, *o2 = o1 + P;
  for (int ip = lane; ip < S.npcsf; ip += G) {
    const int icsf = cst(S.pcsf)[ip];
    double dd = 0.0, de = 0.0;
    for (int q = cst(S.csf_det_off)[icsf]; q < cst(S.csf_det_off)[icsf + 1]; q++) {
      const int d = cst(S.csf_det)[q];
      const int iu = cst(S.iwdet[0])[d], id = cst(S.iwdet[1])[d];
      const double cf = cst(S.dsgn[0])[iu] * cst(S.dsgn[1])[id] * cst(S.csf_det_coef)[q] * dv0[iu] * dv1[id] * chii;
      const double dl = (S.ns[0] ? del0[iu] : 0.0) + (S.ns[1] ? del1[id] : 0.0);
      dd += cf;
      de += cf * (dl - esum);
    }
    o1[ip] = dd;
    o2[ip] = de;
  }
  const double *gv = ss + DL.gvalue, *gj = ss + DL.gjas, *d2g = ss + DL.d2g, *dpe = ss + DL.dpe;
  for (int ip = lane; ip < S.npj; ip += G) {
    double de = -hb * d2g[ip] + (S.nloc > 0 ? dpe[ip] : 0.0);
    for (int i = 0; i < N; i++)
      for (int k = 0; k < 3; k++)
        de -= 2.0 * hb * gj[(ip * N + i) * 3 + k] * (vj[3 * i + k] + vd[3 * i + k]);
    o1[S.npcsf + ip] = gv[ip];
    o2[S.npcsf + ip] = de;
  }
  for (int ip = lane; ip < S.nporb; ip += G) {
    const int oi = cst(S.porb_i)[ip], oj = cst(S.porb_j)[ip];
    double dp = 0.0, de = 0.0;
    for (int pass = 0; pass < (cst(S.porb_rev)[ip] ? 2 : 1); pass++) {
      const int a = pass ? oj : oi, b = pass ? oi : oj;
      const double sg = pass ? -1.0 : 1.0;
      for (int s = 0; s < 2; s++) {
        const int ns = S.ns[s], off = s ? S.nup : 0;
        const int pa = cst(S.iocc_lab[s])[a];
        if (ns == 0 || pa < 0) continue;
        const double *g = ss + SL.g[s], *dg = ss + DL.dg[s];
        for (int k = 0; k < ns; k++) {
          const double ob = orb[((int64_t)(off + k) * NO + b) * 5];
          dp += sg * g[pa * ns + k] * ob;
          de += sg * (g[pa * ns + k] * B[(off + k) * NO + b] + dg[pa * ns + k] * ob);
        }
      }
    }
    o1[S.npcsf + S.npj + ip] = dp;
    o2[S.npcsf + S.npj + ip] = de;
  }
}

}  // namespace qmcw
