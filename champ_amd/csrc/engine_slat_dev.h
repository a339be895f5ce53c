// engine_slat_dev.h -- CHAMP's accelerated Metropolis mover (imetro = 6,
// qmc/vmc/metrop_slat_mov1.f90:78-546) for one walker on one lane: the
// proposal in the nearest nucleus' local (r, theta, phi) coordinates and
// the forward and reverse transition densities of the acceptance.  Included
// by engine_kernels.h after the kernel argument block.
#pragma once

namespace qmcw {

constexpr int kSlatMaxTry = 1 << 20;    // tries of each rejection loop of a proposal

// incomplete gamma function for a = 5/2 with xae = x^a e^-x given; returns
// -gammac (iflag = -1) for x >= a + 4 (qmc/vmc/gammai.f90:1-98; the loop
// counts are the reference's fixed ones)
__device__ __forceinline__ double slat_gammai(double a, double x, double xae, int &iflag) {
  if (x < a + 4) {
    iflag = 1;
    double ap = a, sum = 1 / a, del = sum;
    const int iter = (int)(21 + 3 * (x - a));
    for (int n = 1; n <= iter; n++) {
      ap = ap + 1;
      del = del * x / ap;
      sum = sum + del;
    }
    return sum * xae;
  }
  iflag = -1;
  double a0 = 1, a1 = x, b0 = 0, b1 = 1, fac = 1, g = 0;
  int iter = 20 + (int)llround(a - x);
  if (iter < 7) iter = 7;
  for (int n = 1; n <= iter; n++) {
    const double an = (double)n, ana = an - a;
    a0 = (a1 + a0 * ana) * fac;
    b0 = (b1 + b0 * ana) * fac;
    const double anf = an * fac;
    a1 = x * a0 + anf * a1;
    b1 = x * b0 + anf * b1;
    fac = 1 / a1;
    g = b1 * fac;
  }
  return -(g * xae);
}

// zeta and co of the radial function sqrt(r)|1 + co r|exp(-zeta r) for the
// radial velocity vr at distance r (metrop_slat_mov1.f90:125-140)
__device__ __forceinline__ void slat_zeta_co(double zcusp, double vr, double r, double &zeta, double &co) {
  const double eps = 1e-10;
  double root = (zcusp + vr) * (zcusp + vr - 4.0 / r), z;
  if (root >= 0) {
    root = sqrt(root);
    z = 0.5 * (zcusp - vr - root);
    if (z <= 0) z = 0.5 * (zcusp - vr + root);
    if (z <= 0) z = eps;
  } else {
    z = vr < 0 ? -vr - eps : 1.0;
  }
  zeta = z;
  co = (z + vr) / (1.0 - (z + vr) * r);
}

// integral of the radial function over [rbot, rtop] divided by bot
// (metrop_slat_mov1.f90:198-232)
__device__ __forceinline__ double slat_area(double zeta, double co, double rbot, double rtop, double bot) {
  const double d5b2 = 2.5, d2b3 = .666666666666667, g5b2 = 1.329340388179137;
  int fb, ft, fz;
  const double rzero = -1.0 / co;
  const double zrbot = zeta * rbot, zrtop = zeta * rtop;
  const double zebot = sqrt(zrbot * zrbot * zrbot) * exp(-zrbot);
  const double zetop = sqrt(zrtop * zrtop * zrtop) * exp(-zrtop);
  const double g52bot = slat_gammai(d5b2, zrbot, zrbot * zebot, fb);
  const double g52top = slat_gammai(d5b2, zrtop, zrtop * zetop, ft);
  if (rzero < rbot || rzero > rtop) {
    const double g52dif = (fb * ft == 1) ? g52top - g52bot : g52top - g52bot + g5b2;
    const double g32dif = d2b3 * (g52dif + zetop - zebot);
    return fabs(g32dif + co * g52dif / zeta) / (bot * sqrt(zeta * zeta * zeta));
  }
  const double zrzer = zeta * rzero;
  const double zezer = sqrt(zrzer * zrzer * zrzer) * exp(-zrzer);
  const double g52zer = slat_gammai(d5b2, zrzer, zrzer * zezer, fz);
  const double g52dif1 = (fb * fz == 1) ? g52zer - g52bot : g52zer - g52bot + g5b2;
  const double g32dif1 = d2b3 * (g52dif1 + zezer - zebot);
  const double g52dif2 = (ft * fz == 1) ? g52top - g52zer : g52top - g52zer + g5b2;
  const double g32dif2 = d2b3 * (g52dif2 + zetop - zezer);
  return (fabs(g32dif1 + co * g52dif1 / zeta) + fabs(g32dif2 + co * g52dif2 / zeta)) /
         (bot * sqrt(zeta * zeta * zeta));
}

// nearest nucleus from the e-n vectors rv[M][3] (squared distances, the first minimum)
__device__ __forceinline__ int slat_nearest(int M, const double *rv) {
  double best = 99.e9;
  int ic = 0;
  for (int j = 0; j < M; j++) {
    const double d = rv[3 * j] * rv[3 * j] + rv[3 * j + 1] * rv[3 * j + 1] + rv[3 * j + 2] * rv[3 * j + 2];
    if (d < best) { best = d; ic = j; }
  }
  return ic;
}

// Proposal from the position with e-n vectors rv[M][3] and velocity v:
// xn, and rec = {fxop, areao, r of xn from the nearest nucleus}
// (metrop_slat_mov1.f90:78-289).  st4: the walker's rannyu stream.
__device__ __forceinline__ void slat_propose(const DevSys &S, double deltar, double deltat, const double *rv,
                                             const double *v, double fbias, int *st4, double *xn,
                                             double *rec) {
  const double eps = 1e-10, pi = 3.141592653589793, deltri = 1.0 / deltar;
  const int nearo = slat_nearest(S.M, rv);
  const double *rvm = rv + 3 * nearo;
  const double zn = cst(S.znuc)[cst(S.iwctype)[nearo]];
  const double zcusp = S.nloc == 0 ? zn : 0.0;
  const double rmino = sqrt(rvm[0] * rvm[0] + rvm[1] * rvm[1] + rvm[2] * rvm[2]);
  const double rbot = rmino * deltri, rtop = rmino * deltar;
  double vr = 0, vp = 0, xa[3], ya[3], za[3];
  for (int k = 0; k < 3; k++) vr += v[k] * rvm[k];
  vr /= rmino;
  for (int k = 0; k < 3; k++) {
    za[k] = rvm[k] / rmino;
    xa[k] = v[k] - vr * za[k];
    vp += xa[k] * xa[k];
  }
  vp = sqrt(vp);
  if (vp < eps) {
    xa[0] = eps * (1 - za[0] * za[0]);
    xa[1] = eps * (-za[0] * za[1]);
    xa[2] = eps * (-za[0] * za[2]);
    vp = eps * sqrt(1 - za[0] * za[0]);
  }
  for (int k = 0; k < 3; k++) xa[k] /= vp;
  vr = fmax(vr, -2 * zn);
  ya[0] = za[1] * xa[2] - za[2] * xa[1];
  ya[1] = za[2] * xa[0] - za[0] * xa[2];
  ya[2] = za[0] * xa[1] - za[1] * xa[0];
  vp *= fbias;
  double zeta, co;
  slat_zeta_co(zcusp, vr, rmino, zeta, co);
  const double root = sqrt((1.5 * co - zeta) * (1.5 * co - zeta) + 2 * zeta * co);
  double rmax1 = co != 0 ? (1.5 * co - zeta + root) / (2 * zeta * co) : rtop;
  rmax1 = fmin(fmax(rmax1, rbot), rtop);
  double gmax = sqrt(rmax1) * fabs(1 + co * rmax1) * exp(-zeta * rmax1);
  if (co < 0) {
    double rmax2 = (1.5 * co - zeta - root) / (2 * zeta * co);
    rmax2 = fmin(fmax(rmax2, rbot), rtop);
    gmax = fmax(gmax, sqrt(rmax2) * fabs(1 + co * rmax2) * exp(-zeta * rmax2));
  }
  const double bot = sqrt(rmino) * fabs(1 + co * rmino) * exp(-zeta * rmino);
  double rtry, top;
  // rejection sampling of r; a proposal still unsampled after kSlatMaxTry
  // tries of either loop (only non-finite input gets there) gets fxop = NaN,
  // which the acceptance rejects
  bool ok = false;
  for (int t = 0; t < kSlatMaxTry && !ok; t++) {
    rtry = ((deltar - deltri) * rannyu(st4) + deltri) * rmino;
    top = sqrt(rtry) * fabs(1 + co * rtry) * exp(-zeta * rtry);
    ok = top / gmax > rannyu(st4);
  }
  double fxop = ok ? top / bot : __builtin_nan("");
  double areao = slat_area(zeta, co, rbot, rtop, bot);
  const double raver = 0.5 * (rmino + rtry);
  const double deltt = deltat + (2.0 - deltat) / (1.0 + (zn * raver) * (zn * raver));
  const double costht = 1 - deltt * rannyu(st4);
  const double zprime = rtry * costht;
  const double sintht = sqrt(1 - costht * costht);
  areao *= deltt;
  const double term = fmin(vp * raver * sintht, 1.0);
  const double fm = 1 + term;
  double phitry = 0.0;
  ok = false;
  for (int t = 0; t < kSlatMaxTry && !ok; t++) {    // accepts with probability >= 1/2 per try
    phitry = pi * rannyu(st4);
    top = 1 + term * cos(phitry);
    ok = top > rannyu(st4) * fm;
  }
  fxop = ok ? fxop * top : __builtin_nan("");
  const double xprime = rtry * sintht * cos(phitry);
  double yprime = sqrt(fmax(0.0, rtry * rtry - xprime * xprime - zprime * zprime));
  if (rannyu(st4) < 0.5) yprime = -yprime;
  for (int k = 0; k < 3; k++)
    xn[k] = (xa[k] * xprime + ya[k] * yprime + za[k] * zprime) + cst(S.cent)[3 * nearo + k];
  rec[0] = fxop;
  rec[1] = areao;
  rec[2] = rtry;
}

// Density factors of the move from the point with e-n vectors rvf[M][3] and
// velocity vf to the point with e-n vectors rvt[M][3]
// (metrop_slat_mov1.f90:299-336, 376-475): the density is fx / (area 2 pi
// rto^2).  Returns 0 when the destination is outside what the start's
// sampler reaches (the reference's geometrical rejection).
__device__ __forceinline__ int slat_density(const DevSys &S, double deltar, double deltat, const double *rvf,
                                            const double *vf, const double *rvt, double fbias, double &fx,
                                            double &area, double &rto) {
  const double eps = 1e-10;
  const int nearn = slat_nearest(S.M, rvf);
  const double *rn = rvf + 3 * nearn, *ro = rvt + 3 * nearn;
  const double zn = cst(S.znuc)[cst(S.iwctype)[nearn]];
  double rminn = 0, rminon = 0, dot = 0;
  for (int k = 0; k < 3; k++) rminn += rn[k] * rn[k];
  rminn = sqrt(rminn);
  for (int k = 0; k < 3; k++) {
    rminon += ro[k] * ro[k];
    dot += rn[k] * ro[k];
  }
  rminon = sqrt(rminon);
  dot = dot / (rminn * rminon);
  const double costht = dot, sintht = sqrt(1 - costht * costht);
  const double ravern = 0.5 * (rminn + rminon);
  const double delttn = deltat + (2.0 - deltat) / (1.0 + (zn * ravern) * (zn * ravern));
  rto = rminon;
  if (rminon > rminn * deltar || dot < 1 - delttn) return 0;
  double vr = 0, vp = 0, xa[3];
  for (int k = 0; k < 3; k++) vr += vf[k] * rn[k];
  vr /= rminn;
  for (int k = 0; k < 3; k++) {
    xa[k] = vf[k] - vr * rn[k] / rminn;
    vp += xa[k] * xa[k];
  }
  vp = sqrt(vp);
  if (vp < eps) {
    xa[0] = eps * (1 - rn[0] * rn[0]);
    xa[1] = eps * (-rn[0] * rn[1]);
    xa[2] = eps * (-rn[0] * rn[2]);
    vp = eps * sqrt(1 - rn[0] * rn[0] * (2 - rminn * rminn));
  }
  for (int k = 0; k < 3; k++) xa[k] /= vp;
  const double zcusp = S.nloc == 0 ? zn : 0.0;
  vr = fmax(vr, -2 * zn);
  vp *= fbias;
  double zeta, co;
  slat_zeta_co(zcusp, vr, rminn, zeta, co);
  const double bot = sqrt(rminn) * fabs(1 + co * rminn) * exp(-zeta * rminn);
  const double top = sqrt(rminon) * fabs(1 + co * rminon) * exp(-zeta * rminon);
  double f = top / bot;
  area = slat_area(zeta, co, rminn / deltar, rminn * deltar, bot) * delttn;
  const double term = fmin(vp * ravern * sintht, 1.0);
  double cosphi = 0, rnorm = 0;
  for (int k = 0; k < 3; k++) {
    const double t2 = ro[k] / rminon - costht * rn[k] / rminn;
    rnorm += t2 * t2;
    cosphi += t2 * xa[k];
  }
  cosphi = cosphi / sqrt(rnorm);
  f *= (1 + term * cosphi);
  fx = f;
  return 1;
}

}  // namespace qmcw
