// engine_dmc_dev.h -- device pieces of a DMC move shared by the DMC kernels
// and by k_ratio_1e's fused DMC mode: the per-walker DMC record, the
// averaged velocity and Green's functions (qmc/green_functions_mod.f90),
// the proposal of one walker and its Metropolis decision.  Included by
// engine_kernels.h after the kernel argument block.
#pragma once

namespace qmcw {

// per-walker DMC record, kDmStride doubles
enum {
  kDmEloc = 0,    // E_L at the start of the generation
  kDmG2 = 1,      // |grad ln Psi|^2 (all electrons) of the current state
  kDmW = 2,       // weight
  kDmSdd = 3,     // sum of diffusion displacements squared
  kDmSpd = 4,     // sum pacc * dd
  kDmSad = 5,     // sum accepted * dd
  kDmAcc = 6,     // any move accepted
  kDmPacc = 7,    // sum of acceptance probabilities
  kDmNacc = 8,    // accepted moves
  kDmNcross = 9,  // node-crossing proposals
  kDmTau = 10,    // taunow of the generation
  kDmV = 11,      // averaged velocity of the current proposal (3)
  kDmEpot = 14,   // potential energy at the start of the generation
  kDmDd = 15,     // dd of the current proposal
  kDmPo = 16,     // all-electron mover: forward probability of the current proposal
  kDmTacc = 17,   // a T-move was accepted in this generation
  kDmTmSel = 18,  // T-move of the current electron proposed
  kDmTmPr = 19,   // its Psi ratio
  kDmTmTot = 20,  // forward sum of the T-move weights (with 1 for staying)
  kDmTmIc = 21,   // centre of the proposed point
  kDmTmX = 22,    // proposed position (3)
  kDmTmN = 25,    // pairs in range of the current electron
  kDmStride = 28
};

// averaged_1e_velocity with adrift = adrift_default
__device__ __forceinline__ void dmc_vav(double tau, double adrift, int nd, const double *g,
                                        double *v) {
  const double vn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  double vav;
  if (adrift * vn * vn * tau <= 0.0) vav = vn;
  else vav = (-1.0 + pow(1.0 + nd * adrift * vn * vn * tau, 1.0 / nd)) / (adrift * vn * tau);
  for (int k = 0; k < 3; k++) v[k] = vav * g[k] / vn;
}

// drift-diffusion Green's function of the move xo -> xn with velocity vav
__device__ __forceinline__ double dmc_prob(double tau, const double *vav, const double *xo,
                                           const double *xn) {
  const double pi = 3.141592653589793;
  double d2 = 0.0;
  for (int k = 0; k < 3; k++) {
    const double d = xn[k] - xo[k] - vav[k] * tau;
    d2 += d * d;
  }
  const double nrm = sqrt(2.0 * pi * tau);
  return exp(-d2 / (2.0 * tau)) / (nrm * nrm * nrm);
}

// CHAMP's gauss (qmc/dmc/gauss.f90): cos(2 pi u1) sqrt(-2 ln u2)
__device__ __forceinline__ double dmc_gauss(int *st4) {
  const double pi = 3.141592653589793;
  const double c = cos(2.0 * pi * rannyu(st4));
  return c * sqrt(-2.0 * log(rannyu(st4)));
}

// All-electron mover (nloc == 0): drift-diffusion with the nucleus cusp,
// qmc/green_functions_mod.f90:83-145 (sample) and :268-323 (probability),
// with the nucleus term of averaged_1e_velocity (:616-622).  The geometry
// depends on the starting point: e-n vectors rv[3M] (electron minus centre)
// and distances r[M] there, and the gradient g of ln Psi there.
struct CuspGeo {
  double z[3], x[3], xdrift[3], c[3];  // axes, bent drift point, nearest centre
  double pgaus, zeta;
};
__device__ __forceinline__ void dmc_cusp_geo(const DevSys &S, double tau, int ndrift, const double *g,
                                             const double *rv, const double *r, CuspGeo &cg) {
  const double eps = 1e-10, adrift0 = 0.1;
  int iw = 0;
  for (int ic = 1; ic < S.M; ic++)
    if (r[ic] < r[iw]) iw = ic;                       // minloc: the first minimum
  const double *xe = rv + 3 * iw;
  const double zn = sqrt(xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2]);
  for (int k = 0; k < 3; k++) cg.z[k] = xe[k] / zn;
  const double gz = g[0] * cg.z[0] + g[1] * cg.z[1] + g[2] * cg.z[2];
  for (int k = 0; k < 3; k++) cg.x[k] = g[k] - gz * cg.z[k];
  if (sqrt(cg.x[0] * cg.x[0] + cg.x[1] * cg.x[1] + cg.x[2] * cg.x[2]) < eps) {
    cg.x[0] = eps * (1.0 - cg.z[0] * cg.z[0]);
    cg.x[1] = eps * (-cg.z[0] * cg.z[1]);
    cg.x[2] = eps * (-cg.z[0] * cg.z[2]);
  }
  const double xnrm = sqrt(cg.x[0] * cg.x[0] + cg.x[1] * cg.x[1] + cg.x[2] * cg.x[2]);
  for (int k = 0; k < 3; k++) cg.x[k] /= xnrm;
  // averaged velocity with adrift from the nearest nucleus
  const double Z = cst(S.znuc)[cst(S.iwctype)[iw]];
  const double vn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  const double rmin = zn, hafzr2 = (Z * rmin / 2.0) * (Z * rmin / 2.0);
  const double adrift = (1.0 + eps + (xe[0] * g[0] + xe[1] * g[1] + xe[2] * g[2]) / (vn * rmin)) / 2.0 +
                        adrift0 * hafzr2 / (1.0 + hafzr2);
  double v[3];
  dmc_vav(tau, adrift, ndrift, g, v);
  const double zi = r[iw];
  double zf = zi + (v[0] * cg.z[0] + v[1] * cg.z[1] + v[2] * cg.z[2]) * tau;
  cg.pgaus = zf >= 0.0 ? 1.0 - erfc(zf / sqrt(2.0 * tau)) / 2.0 : erfc(-zf / sqrt(2.0 * tau)) / 2.0;
  zf = fmax(0.0, zf);                                 // no drift past the nucleus
  const double xf = (v[0] * cg.x[0] + v[1] * cg.x[1] + v[2] * cg.x[2]) * tau * fmin(1.0, zf / ((zf + zi) / 2.0));
  for (int k = 0; k < 3; k++) {
    cg.c[k] = cst(S.cent)[3 * iw + k];
    cg.xdrift[k] = cg.c[k] + xf * cg.x[k] + zf * cg.z[k];
  }
  cg.zeta = sqrt(1.0 / tau + Z * Z);
}

// probability density of proposing xn from the geometry cg
__device__ __forceinline__ double dmc_cusp_prob(const CuspGeo &cg, double tau, const double *xn) {
  const double pi = 3.141592653589793;
  double d2 = 0.0, rn2 = 0.0;
  for (int k = 0; k < 3; k++) {
    d2 += (xn[k] - cg.xdrift[k]) * (xn[k] - cg.xdrift[k]);
    rn2 += (xn[k] - cg.c[k]) * (xn[k] - cg.c[k]);
  }
  const double nrm = sqrt(2.0 * pi * tau);
  const double ddp = exp(-d2 / (2.0 * tau)) / (nrm * nrm * nrm);
  return cg.pgaus * ddp + (1.0 - cg.pgaus) * cg.zeta * cg.zeta * cg.zeta * exp(-2.0 * cg.zeta * sqrt(rn2)) / pi;
}

// One walker's proposal for electron ie (the body of sample_1e_move): g is
// the gradient of ln Psi of the electron at its current position; the
// proposed position goes to xn, the record d receives what the decision
// needs.
__device__ __forceinline__ void dmc_propose_w(const KArgs &A, int w, int ie, const double *g,
                                              double *d, const qmcw_dmc_params &p, double *xn) {
  const double *ws = A.st + (int64_t)w * A.L.stride;
  int *rs = A.rng + 4 * w;
  int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
  const double rt = sqrt(p.tau);
  double d2 = 0.0;
  if (A.S.nloc == 0) {
    // all-electron: gaussian about the bent drift point with probability
    // pgaus, otherwise exponential about the nearest nucleus
    const int M = A.S.M;
    CuspGeo cg;
    dmc_cusp_geo(A.S, p.tau, p.n_drift, g, ws + A.L.rvec_en + ie * M * 3, ws + A.L.r_en + ie * M, cg);
    double x[3];
    if (rannyu(st4) < cg.pgaus) {
      for (int k = 0; k < 3; k++) x[k] = cg.xdrift[k] + rt * dmc_gauss(st4);
    } else {
      const double pi = 3.141592653589793;
      const double u1 = rannyu(st4), u2 = rannyu(st4), u3 = rannyu(st4);
      const double r = (-0.5 / cg.zeta) * log(u1 * u2 * u3);
      const double costht = 2.0 * (rannyu(st4) - 0.5);
      const double sintht = sqrt(1.0 - costht * costht);
      const double phi = 2.0 * pi * rannyu(st4);
      x[0] = cg.c[0] + r * sintht * cos(phi);
      x[1] = cg.c[1] + r * sintht * sin(phi);
      x[2] = cg.c[2] + r * costht;
    }
    for (int k = 0; k < 3; k++) {
      xn[k] = x[k];
      d2 += (x[k] - cg.xdrift[k]) * (x[k] - cg.xdrift[k]);
    }
    d[kDmPo] = dmc_cusp_prob(cg, p.tau, x);
  } else {
    double v[3];
    dmc_vav(p.tau, p.adrift, p.n_drift, g, v);
    for (int k = 0; k < 3; k++) {
      const double xd = rt * dmc_gauss(st4);
      xn[k] = ws[A.L.x + 3 * ie + k] + v[k] * p.tau + xd;
      d2 += xd * xd;
    }
    for (int k = 0; k < 3; k++) d[kDmV + k] = v[k];
  }
  rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
  d[kDmDd] = d2;
  d[kDmSdd] += d2;
}

// One walker's Metropolis decision on the pending move of electron ie
// (rundmc_mod.f90:158-196): the move buffer mv holds the ratio and the new
// gradient; returns 1 to accept.
__device__ __forceinline__ int dmc_decide_w(const KArgs &A, int w, int ie, const double *mv, double *d,
                                            const qmcw_dmc_params &p) {
  const DevSys &S = A.S;
  const double *ws = A.st + (int64_t)w * A.L.stride;
  const int s = ie < S.nup ? 0 : 1, ns = S.ns[s], nocc = S.nocc[s];
  const double ratio = mv[A.ML.misc + 9 + ns + nocc];
  const double *gn = mv + A.ML.misc + 10 + ns + nocc;
  const double *xn = mv + A.ML.xn;
  const double *xo = ws + A.L.x + 3 * ie;
  double po, pn;
  if (S.nloc == 0) {
    // backward move from the proposed point, with its geometry
    CuspGeo cg;
    dmc_cusp_geo(S, p.tau, p.n_drift, gn, mv + A.ML.rven, mv + A.ML.ren, cg);
    po = d[kDmPo];
    pn = dmc_cusp_prob(cg, p.tau, xo);
  } else {
    double vn[3];
    dmc_vav(p.tau, p.adrift, p.n_drift, gn, vn);
    po = dmc_prob(p.tau, d + kDmV, xo, xn);
    pn = dmc_prob(p.tau, vn, xn, xo);
  }
  double pa = fmin(1.0, ratio * ratio * pn / po);
  if (p.iacc_rej == 0) pa = 1.0;
  if (ratio <= 0.0) {
    if (p.icross <= 0) pa = 0.0;
    d[kDmNcross] += 1.0;
  }
  int *rs = A.rng + 4 * w;
  int st4[4] = {rs[0], rs[1], rs[2], rs[3]};
  const int a = rannyu(st4) < pa;
  rs[0] = st4[0]; rs[1] = st4[1]; rs[2] = st4[2]; rs[3] = st4[3];
  const double dd = d[kDmDd];
  d[kDmSpd] += pa * dd;
  d[kDmSad] += a * dd;
  d[kDmPacc] += pa;
  d[kDmNacc] += a;
  if (a) d[kDmAcc] = 1.0;
  return a;
}

}  // namespace qmcw
