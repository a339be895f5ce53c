// engine_device.h -- device functions of the wavefunction engine.
//
// A walker is handled by a group of G lanes of one wavefront (G in
// {8,16,32,64}); a 64-lane workgroup holds 64/G walkers.  Group-parallel
// loops stride over lanes, reductions use wavefront shuffles of width G.
// Orbital and Jastrow arithmetic is templated on the real type T (double for
// the FP64 engine, float for the mixed engine); determinants, inverses and
// energies are always double.
#pragma once
#include "engine.h"

namespace qmcw {

// The system tables (DevSys) are written only by the host between launches.
// Reading them through the constant address space lets the compiler turn a
// load at a uniform index into a scalar load through the scalar cache
// instead of a vector load with its own wait.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> cst(const T *p) {
  return (cptr<T>)p;
}

template <int G>
__device__ __forceinline__ double gsum(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}
template <int G>
__device__ __forceinline__ float gsumf(float v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}
// inclusive prefix sum over the G lanes of a group (lane = position in it)
template <int G>
__device__ __forceinline__ int gscan_incl(int v, int lane) {
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    const int t = __shfl_up(v, d, G);
    if (lane >= d) v += t;
  }
  return v;
}

// Ordered list of the entries of flags[0..n) (stride fs) that are nonzero,
// built by the G lanes of a group: each lane takes a contiguous chunk, a
// prefix sum places the chunks.  Returns the count (all lanes).
template <int G>
__device__ __forceinline__ int group_compact(const double *flags, int fs, int n, int *list, int lane) {
  const int chunk = (n + G - 1) / G, b0 = lane * chunk, b1 = min(n, b0 + chunk);
  int c = 0;
  for (int q = b0; q < b1; q++) c += flags[(int64_t)q * fs] != 0.0;
  const int incl = gscan_incl<G>(c, lane);
  int o = incl - c;
  for (int q = b0; q < b1; q++)
    if (flags[(int64_t)q * fs] != 0.0) list[o++] = q;
  return __shfl(incl, G - 1, G);
}

template <int G>
__device__ __forceinline__ int gsumi(int v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}
// argmax of |v| over the group (first index wins among equal values, as a
// sequential scan from the lowest index would)
template <int G>
__device__ __forceinline__ void gargmax(double v, int idx, double &bv, int &bi) {
  double a = fabs(v);
  bi = idx;
  bv = v;
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    double a2 = __shfl_xor(a, o, G);
    double v2 = __shfl_xor(bv, o, G);
    int i2 = __shfl_xor(bi, o, G);
    bool take = (a2 > a) || (a2 == a && i2 < bi);
    if (take) { a = a2; bv = v2; bi = i2; }
  }
}

// ---------------------------------------------------------------------------
// scaled distances (isc = 2: (1-exp(-k r))/k ; isc = 4: r/(1+k r)) with first
// and second derivative; the e-e-n variant is 1 - k * rr.
template <typename T>
__device__ __forceinline__ void scaled(int isc, T k, T r, T &rr, T &d1, T &d2) {
  if (isc == 2) {
    T e = exp(-k * r);
    rr = (T(1) - e) / k;
    d1 = e;
    d2 = -k * e;
  } else {
    T deni = T(1) / (T(1) + k * r);
    rr = r * deni;
    d1 = deni * deni;
    d2 = T(-2) * k * deni * d1;
  }
}

// e-e-n polynomial sum_t c_t U^k (A^(l+m) B^m + A^m B^(l+m)) with the terms
// visited in a compile-time order (k, then m, then l); the coefficients are
// stored in that order (DevSys::c).  Every power index is a constant after
// unrolling, so the power tables live in registers.
#define QMCW_EEN_TERM(k, l, m) (((k) >= 1 || (m) >= 1) && ((l) + 2 * (m) >= 1) && ((k) + (l) + 2 * (m) >= 2))

template <typename T, int ORD>
__device__ __forceinline__ T een_value(T U, T A, T B, cptr<double> c) {
  T up[ORD + 1], ap[ORD + 1], bp[ORD + 1];
  up[0] = ap[0] = bp[0] = T(1);
#pragma unroll
  for (int q = 1; q <= ORD; q++) { up[q] = up[q - 1] * U; ap[q] = ap[q - 1] * A; bp[q] = bp[q - 1] * B; }
  T f = 0;
  int idx = 0;
#pragma unroll
  for (int k = 0; k <= ORD; k++)
#pragma unroll
    for (int m = 0; 2 * m <= ORD - k; m++)
#pragma unroll
      for (int l = 0; l <= ORD - k - 2 * m; l++)
        if (QMCW_EEN_TERM(k, l, m)) {
          f += T(c[idx]) * up[k] * (ap[l + m] * bp[m] + ap[m] * bp[l + m]);
          idx++;
        }
  return f;
}

// value and the partial derivatives wrt U, A, B (first and second, and the
// mixed U-A, U-B ones) of the e-e-n polynomial
template <typename T>
struct EenD { T f, fU, fUU, fA, fAA, fB, fBB, fUA, fUB; };

template <typename T, int ORD>
__device__ __forceinline__ void een_derivs(T U, T A, T B, cptr<double> c, EenD<T> &o) {
  T up[ORD + 1], ap[ORD + 1], bp[ORD + 1];
  up[0] = ap[0] = bp[0] = T(1);
#pragma unroll
  for (int q = 1; q <= ORD; q++) { up[q] = up[q - 1] * U; ap[q] = ap[q - 1] * A; bp[q] = bp[q - 1] * B; }
  o.f = o.fU = o.fUU = o.fA = o.fAA = o.fB = o.fBB = o.fUA = o.fUB = T(0);
  int idx = 0;
#pragma unroll
  for (int k = 0; k <= ORD; k++)
#pragma unroll
    for (int m = 0; 2 * m <= ORD - k; m++)
#pragma unroll
      for (int l = 0; l <= ORD - k - 2 * m; l++)
        if (QMCW_EEN_TERM(k, l, m)) {
          const int lm = l + m;
          T cl = T(c[idx]);
          idx++;
          T P = ap[lm] * bp[m] + ap[m] * bp[lm];
          T PA = (lm ? T(lm) * ap[lm > 0 ? lm - 1 : 0] * bp[m] : T(0)) +
                 (m ? T(m) * ap[m > 0 ? m - 1 : 0] * bp[lm] : T(0));
          T PB = (lm ? T(lm) * bp[lm > 0 ? lm - 1 : 0] * ap[m] : T(0)) +
                 (m ? T(m) * bp[m > 0 ? m - 1 : 0] * ap[lm] : T(0));
          T PAA = (lm > 1 ? T(lm * (lm - 1)) * ap[lm > 1 ? lm - 2 : 0] * bp[m] : T(0)) +
                  (m > 1 ? T(m * (m - 1)) * ap[m > 1 ? m - 2 : 0] * bp[lm] : T(0));
          T PBB = (lm > 1 ? T(lm * (lm - 1)) * bp[lm > 1 ? lm - 2 : 0] * ap[m] : T(0)) +
                  (m > 1 ? T(m * (m - 1)) * bp[m > 1 ? m - 2 : 0] * ap[lm] : T(0));
          T Uk = up[k];
          T Uk1 = k ? T(k) * up[k > 0 ? k - 1 : 0] : T(0);
          T Uk2 = k > 1 ? T(k * (k - 1)) * up[k > 1 ? k - 2 : 0] : T(0);
          o.f += cl * Uk * P;
          o.fU += cl * Uk1 * P;
          o.fUU += cl * Uk2 * P;
          o.fA += cl * Uk * PA;
          o.fAA += cl * Uk * PAA;
          o.fB += cl * Uk * PB;
          o.fBB += cl * Uk * PBB;
          o.fUA += cl * Uk1 * PA;
          o.fUB += cl * Uk1 * PB;
        }
}

// Pair term of electrons (i, j), i > j: e-e Pade + polynomial and e-e-n
// polynomial.  Returns value; gi/gj gradient wrt r_i and r_j; li/lj the
// laplacian wrt r_i and r_j.
template <typename T>
struct PairOut {
  T f, gi[3], gj[3], li, lj;
};

// An electron's distances to the centres for jas_pair: read from its rows
// of the state (EnRows) or formed from its position the way the moves and
// the full evaluation form the rows (EnFromX: the same differences and the
// same square root, so the same values), which spares the lanes of a group
// a strided read of each other electron's rows.
struct EnRows {
  const double *r, *rv;
  __device__ __forceinline__ void get(const DevSys &, int ic, double &ri, double *rvi) const {
    ri = r[ic];
    rvi[0] = rv[3 * ic]; rvi[1] = rv[3 * ic + 1]; rvi[2] = rv[3 * ic + 2];
  }
};
struct EnFromX {
  double x0, x1, x2;
  __device__ __forceinline__ void get(const DevSys &S, int ic, double &ri, double *rvi) const {
    const double p[3] = {x0, x1, x2};
    double s2 = 0.0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const double d = p[q] - cst(S.cent)[3 * ic + q];
      rvi[q] = d;
      s2 += d * d;
    }
    ri = sqrt(s2);
  }
};

template <typename T, int ORD, typename PI, typename PJ>
__device__ void jas_pair_g(const DevSys &S, bool parallel_spin, T rij, const T *rvij, const PI &pi,
                           const PJ &pj, PairOut<T> &o) {
  const T k = T(S.scalek);
  T sspinn = T(1);
  int isb = 0;
  double asymp_b = S.asymp_jasb0;
  if (parallel_spin) {
    if (S.nspin2b == 2) isb = 1;
    else if (S.nocuspb == 0) sspinn = T(0.5);
    asymp_b = S.asymp_jasb1;
  }
  const double *b = S.b + isb * (S.nordb + 1);
  T u, d1, d2;
  scaled<T>(S.isc, k, rij, u, d1, d2);
  T bot = T(1) + T(b[1]) * u;
  T f = sspinn * T(b[0]) * u / bot;
  T fu = sspinn * T(b[0]) / (bot * bot);
  T fuu = T(-2) * fu * T(b[1]) / bot;
  f -= T(asymp_b);
  T up = u;  // u^(iord-1)
  T upm = T(1);
  for (int iord = 2; iord <= S.nordb; iord++) {
    T bb = T(b[iord]);
    fuu += bb * T(iord * (iord - 1)) * upm;
    fu += bb * T(iord) * up;
    upm = up;
    up *= u;
    f += bb * up;
  }
  fuu = fuu * d1 * d1 + fu * d2;
  fu = fu * d1 / rij;
  o.f = f;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    o.gi[q] = fu * rvij[q];
    o.gj[q] = -fu * rvij[q];
  }
  o.li = fuu + T(2) * fu;
  o.lj = o.li;
  if (S.nordc <= 1) return;
  // e-e-n: variables 1 - k rr of r_ij, r_i, r_j
  T U, Ud1, Ud2;
  scaled<T>(S.isc, k, rij, U, Ud1, Ud2);
  U = T(1) - k * U;
  Ud1 = -k * Ud1;
  Ud2 = -k * Ud2;
  for (int ic = 0; ic < S.M; ic++) {
    int it = cst(S.iwctype)[ic];
    double rid, rjd, rvi[3], rvj[3];
    pi.get(S, ic, rid, rvi);
    pj.get(S, ic, rjd, rvj);
    T ri = (T)rid, rj = (T)rjd;
    T A, Ad1, Ad2, B, Bd1, Bd2;
    scaled<T>(S.isc, k, ri, A, Ad1, Ad2);
    scaled<T>(S.isc, k, rj, B, Bd1, Bd2);
    A = T(1) - k * A; Ad1 = -k * Ad1; Ad2 = -k * Ad2;
    B = T(1) - k * B; Bd1 = -k * Bd1; Bd2 = -k * Bd2;
    EenD<T> e;
    een_derivs<T, ORD>(U, A, B, cst(S.c) + it * S.nterms, e);
    T fc = e.f, fU = e.fU, fUU = e.fUU, fA = e.fA, fAA = e.fAA, fB = e.fB, fBB = e.fBB,
      fUA = e.fUA, fUB = e.fUB;
    // chain rule to unscaled distances
    T gU = fU * Ud1 / rij;
    T gUU = fUU * Ud1 * Ud1 + fU * Ud2;
    T gA = fA * Ad1 / ri, gAA = fAA * Ad1 * Ad1 + fA * Ad2;
    T gB = fB * Bd1 / rj, gBB = fBB * Bd1 * Bd1 + fB * Bd2;
    T gUA = fUA * Ud1 * Ad1, gUB = fUB * Ud1 * Bd1;
    o.f += fc;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      o.gi[q] += gA * (T)rvi[q] + gU * rvij[q];
      o.gj[q] += gB * (T)rvj[q] - gU * rvij[q];
    }
    // cos of the angle between r_ij and r_i: (rij^2 + ri^2 - rj^2)/(2 rij ri)
    T cI = (rij * rij + ri * ri - rj * rj) / (rij * ri);
    T cJ = (rij * rij + rj * rj - ri * ri) / (rij * rj);
    o.li += T(2) * (gU + gA) + gUU + gAA + gUA * cI;
    o.lj += T(2) * (gU + gB) + gUU + gBB + gUB * cJ;
  }
}

template <typename T, int ORD>
__device__ __forceinline__ void jas_pair(const DevSys &S, bool parallel_spin, T rij, const T *rvij,
                                         const double *ri_v, const double *rj_v, const double *rvi,
                                         const double *rvj, PairOut<T> &o) {
  jas_pair_g<T, ORD>(S, parallel_spin, rij, rvij, EnRows{ri_v, rvi}, EnRows{rj_v, rvj}, o);
}

// A per-lane array stored lane-interleaved: element e of this lane at
// p[e * s] (s = the lanes sharing the region), so that the lanes' accesses
// to one element are contiguous
template <typename V>
struct LaneArr {
  V *p;
  int s;
  __device__ __forceinline__ V &operator[](int e) const { return p[(int64_t)e * s]; }
  __device__ __forceinline__ LaneArr<V> operator+(int e) const { return {p + (int64_t)e * s, s}; }
};

// ri_v: a pointer or a LaneArr of the M distances
template <typename T, typename RA>
__device__ T jas_en_val(const DevSys &S, RA ri_v) {
  const T k = T(S.scalek);
  T f = 0;
  for (int ic = 0; ic < S.M; ic++) {
    int it = cst(S.iwctype)[ic];
    const cptr<double> a = cst(S.a4) + it * (S.norda + 1);
    T u, d1, d2;
    scaled<T>(S.isc, k, (T)ri_v[ic], u, d1, d2);
    T fe = T(a[0]) * u / (T(1) + T(a[1]) * u) - T(cst(S.asymp_jasa)[it]);
    T up = u;
    for (int iord = 2; iord <= S.norda; iord++) {
      up *= u;
      fe += T(a[iord]) * up;
    }
    f += fe;
  }
  return f;
}

// e-e-n polynomial of a fixed A as a bivariate polynomial in (U, B):
// sum_t c_t U^k (A^(l+m) B^m + A^m B^(l+m)) = sum_k U^k sum_n gam(k, n) B^n,
// n <= ORD - k, rows of gam stored one after the other.  With A fixed for
// a point and a centre, each other electron then costs one Horner
// evaluation instead of the term-by-term sum.
template <int ORD>
__host__ __device__ constexpr int een_gi(int k, int n) { return k * (ORD + 1) - k * (k - 1) / 2 + n; }

template <typename T, int ORD>
__device__ __forceinline__ void een_gamma(T A, cptr<double> c, T *gam) {
  T ap[ORD + 1];
  ap[0] = T(1);
#pragma unroll
  for (int q = 1; q <= ORD; q++) ap[q] = ap[q - 1] * A;
#pragma unroll
  for (int q = 0; q < een_gi<ORD>(ORD, 0) + 1; q++) gam[q] = T(0);
  int idx = 0;
#pragma unroll
  for (int k = 0; k <= ORD; k++)
#pragma unroll
    for (int m = 0; 2 * m <= ORD - k; m++)
#pragma unroll
      for (int l = 0; l <= ORD - k - 2 * m; l++)
        if (QMCW_EEN_TERM(k, l, m)) {
          const T ct = T(c[idx]);
          idx++;
          gam[een_gi<ORD>(k, m)] += ct * ap[l + m];
          gam[een_gi<ORD>(k, l + m)] += ct * ap[m];
        }
}

template <typename T, int ORD>
__device__ __forceinline__ T een_horner(const T *gam, T U, T B) {
  T f = T(0);
#pragma unroll
  for (int k = ORD; k >= 0; k--) {
    T row = T(0);
#pragma unroll
    for (int n = ORD - k; n >= 0; n--) row = row * B + gam[een_gi<ORD>(k, n)];
    f = f * U + row;
  }
  return f;
}

// The e-e-n polynomial of a fixed A (the moved electron at a centre) with
// its first and second A derivatives, each as a bivariate polynomial in
// (U, B) laid out as een_gamma's table: g[0..NG) the value, g[NG..2NG) d/dA,
// g[2NG..3NG) d2/dA2 (NG = een_gi<ORD>(ORD, 0) + 1).
template <typename T, int ORD>
__device__ __forceinline__ void een_gamma3(T A, cptr<double> c, T *g) {
  constexpr int NG = een_gi<ORD>(ORD, 0) + 1;
  T ap[ORD + 1];
  ap[0] = T(1);
#pragma unroll
  for (int q = 1; q <= ORD; q++) ap[q] = ap[q - 1] * A;
  T r[3 * NG];
#pragma unroll
  for (int q = 0; q < 3 * NG; q++) r[q] = T(0);
  int idx = 0;
#pragma unroll
  for (int k = 0; k <= ORD; k++)
#pragma unroll
    for (int m = 0; 2 * m <= ORD - k; m++)
#pragma unroll
      for (int l = 0; l <= ORD - k - 2 * m; l++)
        if (QMCW_EEN_TERM(k, l, m)) {
          const T ct = T(c[idx]);
          idx++;
          const int lm = l + m, a = een_gi<ORD>(k, m), b = een_gi<ORD>(k, lm);
          r[a] += ct * ap[lm];
          r[b] += ct * ap[m];
          if (lm >= 1) r[NG + a] += ct * T(lm) * ap[lm >= 1 ? lm - 1 : 0];
          if (m >= 1) r[NG + b] += ct * T(m) * ap[m >= 1 ? m - 1 : 0];
          if (lm >= 2) r[2 * NG + a] += ct * T(lm * (lm - 1)) * ap[lm >= 2 ? lm - 2 : 0];
          if (m >= 2) r[2 * NG + b] += ct * T(m * (m - 1)) * ap[m >= 2 ? m - 2 : 0];
        }
#pragma unroll
  for (int q = 0; q < 3 * NG; q++) g[q] = r[q];
}

// T values per centre of the move's e-e-n tables: the three tables of
// een_gamma3, then dA/dr, d2A/dr2, r, 1/r and the vector to the centre of
// the moved electron
template <int ORD>
__host__ __device__ constexpr int een_tab_stride() { return 3 * (een_gi<ORD>(ORD, 0) + 1) + 8; }

// The moved electron's tables at its new position (distances rn and vectors
// rvn to the centres), one centre per lane
template <typename T, int ORD, int G>
__device__ __forceinline__ void een_tables(const DevSys &S, const double *rn, const double *rvn, T *tab,
                                           int lane) {
  constexpr int NG = een_gi<ORD>(ORD, 0) + 1, TS = een_tab_stride<ORD>();
  const T k = T(S.scalek);
  for (int ic = lane; ic < S.M; ic += G) {
    T A, Ad1, Ad2;
    scaled<T>(S.isc, k, (T)rn[ic], A, Ad1, Ad2);
    A = T(1) - k * A;
    T *tb = tab + ic * TS;
    een_gamma3<T, ORD>(A, cst(S.c) + cst(S.iwctype)[ic] * S.nterms, tb);
    tb[3 * NG] = -k * Ad1;
    tb[3 * NG + 1] = -k * Ad2;
    tb[3 * NG + 2] = (T)rn[ic];
    tb[3 * NG + 3] = T(1) / (T)rn[ic];
#pragma unroll
    for (int q = 0; q < 3; q++) tb[3 * NG + 4 + q] = (T)rvn[3 * ic + q];
  }
}

// jas_pair for a pair of the moved electron m (at the position of the
// tables tab) and another electron o
// (distances from po): the e-e-n polynomial and its derivatives come from
// Horner sums over the tables instead of the term-by-term sum.  rvmo = r_m
// - r_o.  Output in jas_pair's convention with electron i = m when m_hi,
// else i = o.
template <typename T, int ORD, typename PO>
__device__ void jas_pair_tab(const DevSys &S, bool parallel_spin, T rij, const T *rvmo, const T *tab,
                             const PO &po, bool m_hi, PairOut<T> &out) {
  constexpr int NG = een_gi<ORD>(ORD, 0) + 1, TS = een_tab_stride<ORD>();
  const T k = T(S.scalek);
  T sspinn = T(1);
  int isb = 0;
  double asymp_b = S.asymp_jasb0;
  if (parallel_spin) {
    if (S.nspin2b == 2) isb = 1;
    else if (S.nocuspb == 0) sspinn = T(0.5);
    asymp_b = S.asymp_jasb1;
  }
  const double *b = S.b + isb * (S.nordb + 1);
  T u, d1, d2;
  scaled<T>(S.isc, k, rij, u, d1, d2);
  T bot = T(1) + T(b[1]) * u;
  T f = sspinn * T(b[0]) * u / bot;
  T fu = sspinn * T(b[0]) / (bot * bot);
  T fuu = T(-2) * fu * T(b[1]) / bot;
  f -= T(asymp_b);
  T up = u, upm = T(1);
  for (int iord = 2; iord <= S.nordb; iord++) {
    T bb = T(b[iord]);
    fuu += bb * T(iord * (iord - 1)) * upm;
    fu += bb * T(iord) * up;
    upm = up;
    up *= u;
    f += bb * up;
  }
  fuu = fuu * d1 * d1 + fu * d2;
  fu = fu * d1 / rij;
  T gm[3], go[3];
#pragma unroll
  for (int q = 0; q < 3; q++) { gm[q] = fu * rvmo[q]; go[q] = -fu * rvmo[q]; }
  T lm = fuu + T(2) * fu, lo = lm;
  if (S.nordc > 1) {
    T U, Ud1, Ud2;
    scaled<T>(S.isc, k, rij, U, Ud1, Ud2);
    U = T(1) - k * U;
    Ud1 = -k * Ud1;
    Ud2 = -k * Ud2;
    T uq[ORD + 1];
    uq[0] = T(1);
#pragma unroll
    for (int q = 1; q <= ORD; q++) uq[q] = uq[q - 1] * U;
    const T irij = T(1) / rij, rij2 = rij * rij;
    for (int ic = 0; ic < S.M; ic++) {
      const T *tb = tab + ic * TS;
      double rod, rvo[3];
      po.get(S, ic, rod, rvo);
      const T ro = (T)rod;
      T B, Bd1, Bd2;
      scaled<T>(S.isc, k, ro, B, Bd1, Bd2);
      B = T(1) - k * B; Bd1 = -k * Bd1; Bd2 = -k * Bd2;
      // rows of the tables in B (with the first two B derivatives of the
      // value table), then the sums over the powers of U
      T F = 0, FU = 0, FUU = 0, FB = 0, FBB = 0, FUB = 0, FA = 0, FAA = 0, FUA = 0;
#pragma unroll
      for (int kk = 0; kk <= ORD; kk++) {
        T r = 0, rb = 0, rbb = 0, ra = 0, raa = 0;
#pragma unroll
        for (int n = ORD - kk; n >= 0; n--) {
          rbb = rbb * B + T(2) * rb;
          rb = rb * B + r;
          r = r * B + tb[een_gi<ORD>(kk, n)];
          ra = ra * B + tb[NG + een_gi<ORD>(kk, n)];
          raa = raa * B + tb[2 * NG + een_gi<ORD>(kk, n)];
        }
        F += r * uq[kk];
        FB += rb * uq[kk];
        FBB += rbb * uq[kk];
        FA += ra * uq[kk];
        FAA += raa * uq[kk];
        if (kk >= 1) {
          FU += T(kk) * r * uq[kk - 1];
          FUB += T(kk) * rb * uq[kk - 1];
          FUA += T(kk) * ra * uq[kk - 1];
        }
        if (kk >= 2) FUU += T(kk * (kk - 1)) * r * uq[kk - 2];
      }
      const T Ad1 = tb[3 * NG], Ad2 = tb[3 * NG + 1], rm = tb[3 * NG + 2], irm = tb[3 * NG + 3];
      const T iro = T(1) / ro;
      // chain rule to unscaled distances (A: the moved electron, B: the other)
      const T gU = FU * Ud1 * irij, gUU = FUU * Ud1 * Ud1 + FU * Ud2;
      const T gA = FA * Ad1 * irm, gAA = FAA * Ad1 * Ad1 + FA * Ad2;
      const T gB = FB * Bd1 * iro, gBB = FBB * Bd1 * Bd1 + FB * Bd2;
      const T gUA = FUA * Ud1 * Ad1, gUB = FUB * Ud1 * Bd1;
      f += F;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        gm[q] += gA * tb[3 * NG + 4 + q] + gU * rvmo[q];
        go[q] += gB * (T)rvo[q] - gU * rvmo[q];
      }
      const T cM = (rij2 + rm * rm - ro * ro) * (irij * irm);
      const T cO = (rij2 + ro * ro - rm * rm) * (irij * iro);
      lm += T(2) * (gU + gA) + gUU + gAA + gUA * cM;
      lo += T(2) * (gU + gB) + gUU + gBB + gUB * cO;
    }
  }
  out.f = f;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    out.gi[q] = m_hi ? gm[q] : go[q];
    out.gj[q] = m_hi ? go[q] : gm[q];
  }
  out.li = m_hi ? lm : lo;
  out.lj = m_hi ? lo : lm;
}

// e-e Pade + polynomial of a pair at scaled distance u
template <typename T>
__device__ __forceinline__ T jas_b_val(const DevSys &S, bool parallel_spin, T u) {
  T sspinn = T(1);
  int isb = 0;
  double asymp_b = S.asymp_jasb0;
  if (parallel_spin) {
    if (S.nspin2b == 2) isb = 1;
    else if (S.nocuspb == 0) sspinn = T(0.5);
    asymp_b = S.asymp_jasb1;
  }
  const double *b = S.b + isb * (S.nordb + 1);
  T f = sspinn * T(b[0]) * u / (T(1) + T(b[1]) * u) - T(asymp_b);
  T up = u;
  for (int iord = 2; iord <= S.nordb; iord++) {
    up *= u;
    f += T(b[iord]) * up;
  }
  return f;
}

// Jastrow value of electron i placed at xn (distances rnew to the centres),
// with every other electron where it is: its e-n terms plus the e-e and
// e-e-n terms of all pairs (i, j).  bt[ic * N + j] = 1 - k scaled(r_j,ic)
// of the current configuration; uq is per-lane scratch of N values.  With
// PAIR the centres go two at a time, so that each U of the other electrons
// is read once for both (each centre's sum keeps its own order; a kernel
// instance of its own, for many centres: faster on the 50-centre chain,
// slower on the 2-centre dimer, whose registers it crowds).
template <typename T, int ORD, bool PAIR = false, typename RA, typename UA>
__device__ __forceinline__ T jas_point_val(const DevSys &S, int i, const double *xn, const double *x,
                                           RA rnew, const T *bt, UA uq) {
  const int N = S.N, M = S.M;
  const T k = T(S.scalek);
  T f = jas_en_val<T>(S, rnew);
  for (int jj = 0; jj < N; jj++) {
    if (jj == i) continue;
    const double dx = xn[0] - x[3 * jj], dy = xn[1] - x[3 * jj + 1], dz = xn[2] - x[3 * jj + 2];
    const T rij = (T)sqrt(dx * dx + dy * dy + dz * dz);
    const bool par = (i < S.nup) == (jj < S.nup);
    T u, d1, d2;
    scaled<T>(S.isc, k, rij, u, d1, d2);
    f += jas_b_val<T>(S, par, u);
    uq[jj] = T(1) - k * u;
  }
  if (S.nordc <= 1) return f;
  constexpr int NG = een_gi<ORD>(ORD, 0) + 1;
  int ic = 0;
  for (; PAIR && ic + 1 < M; ic += 2) {
    T A0, A1, t1, t2;
    scaled<T>(S.isc, k, (T)rnew[ic], A0, t1, t2);
    scaled<T>(S.isc, k, (T)rnew[ic + 1], A1, t1, t2);
    T gam0[NG], gam1[NG];
    een_gamma<T, ORD>(T(1) - k * A0, cst(S.c) + cst(S.iwctype)[ic] * S.nterms, gam0);
    een_gamma<T, ORD>(T(1) - k * A1, cst(S.c) + cst(S.iwctype)[ic + 1] * S.nterms, gam1);
    const T *b0 = bt + ic * N, *b1 = b0 + N;
    T fc0 = T(0), fc1 = T(0);
    for (int jj = 0; jj < N; jj++) {
      if (jj == i) continue;
      const T u = uq[jj];
      fc0 += een_horner<T, ORD>(gam0, u, b0[jj]);
      fc1 += een_horner<T, ORD>(gam1, u, b1[jj]);
    }
    f += fc0;
    f += fc1;
  }
  for (; ic < M; ic++) {
    T A, t1, t2;
    scaled<T>(S.isc, k, (T)rnew[ic], A, t1, t2);
    T gam[NG];
    een_gamma<T, ORD>(T(1) - k * A, cst(S.c) + cst(S.iwctype)[ic] * S.nterms, gam);
    const T *b0 = bt + ic * N;
    T fc = T(0);
    for (int jj = 0; jj < N; jj++) {
      if (jj == i) continue;
      fc += een_horner<T, ORD>(gam, uq[jj], b0[jj]);
    }
    f += fc;
  }
  return f;
}

// e-n Pade + polynomial of one electron summed over centres
template <typename T>
__device__ void jas_en(const DevSys &S, const double *ri_v, const double *rvi, T &f,
                       T *g, T &lap) {
  const T k = T(S.scalek);
  f = 0;
  g[0] = g[1] = g[2] = 0;
  lap = 0;
  for (int ic = 0; ic < S.M; ic++) {
    int it = cst(S.iwctype)[ic];
    const cptr<double> a = cst(S.a4) + it * (S.norda + 1);
    T r = (T)ri_v[ic], u, d1, d2;
    scaled<T>(S.isc, k, r, u, d1, d2);
    T bot = T(1) + T(a[1]) * u;
    T fe = T(a[0]) * u / bot - T(cst(S.asymp_jasa)[it]);
    T fu = T(a[0]) / (bot * bot);
    T fuu = T(-2) * fu * T(a[1]) / bot;
    T up = u, upm = T(1);
    for (int iord = 2; iord <= S.norda; iord++) {
      T aa = T(a[iord]);
      fuu += aa * T(iord * (iord - 1)) * upm;
      fu += aa * T(iord) * up;
      upm = up;
      up *= u;
      fe += aa * up;
    }
    fuu = fuu * d1 * d1 + fu * d2;
    fu = fu * d1 / r;
    f += fe;
#pragma unroll
    for (int q = 0; q < 3; q++) g[q] += fu * (T)rvi[3 * ic + q];
    lap += fuu + T(2) * fu;
  }
}

// ---------------------------------------------------------------------------
// basis shell: r^a sum_p d_p exp(-z_p r^q) times real solid harmonics S_lm.
// Writes phi[b*5 + {val, dx, dy, dz, lap}] (lap only when want_d).
template <typename T>
__device__ __forceinline__ void solid_harm(int l, T x, T y, T z, T *Sv, T (*Sg)[3]) {
  // real solid harmonics with <Y^2> = 4 pi/(2l+1) normalisation, m order
  // 0, 1, -1, 2, -2, ... ; stored at index (m>0 ? 2m-1 : (m<0 ? -2m : 0))
  const T x2 = x * x, y2 = y * y, z2 = z * z;
  switch (l) {
    case 0:
      Sv[0] = T(1); Sg[0][0] = Sg[0][1] = Sg[0][2] = T(0);
      break;
    case 1:
      Sv[0] = z; Sg[0][0] = 0; Sg[0][1] = 0; Sg[0][2] = 1;
      Sv[1] = x; Sg[1][0] = 1; Sg[1][1] = 0; Sg[1][2] = 0;
      Sv[2] = y; Sg[2][0] = 0; Sg[2][1] = 1; Sg[2][2] = 0;
      break;
    case 2: {
      const T c1 = T(1.7320508075688772), c2 = T(0.8660254037844386);
      Sv[0] = T(0.5) * (T(2) * z2 - x2 - y2);
      Sg[0][0] = -x; Sg[0][1] = -y; Sg[0][2] = T(2) * z;
      Sv[1] = c1 * x * z; Sg[1][0] = c1 * z; Sg[1][1] = 0; Sg[1][2] = c1 * x;
      Sv[2] = c1 * y * z; Sg[2][0] = 0; Sg[2][1] = c1 * z; Sg[2][2] = c1 * y;
      Sv[3] = c2 * (x2 - y2); Sg[3][0] = T(2) * c2 * x; Sg[3][1] = T(-2) * c2 * y; Sg[3][2] = 0;
      Sv[4] = T(2) * c2 * x * y; Sg[4][0] = T(2) * c2 * y; Sg[4][1] = T(2) * c2 * x; Sg[4][2] = 0;
      break;
    }
    case 3: {
      const T c0 = T(0.5), c1 = T(0.6123724356957945), c2 = T(1.9364916731037085),
              c3 = T(0.7905694150420949);
      T rho2 = x2 + y2;
      Sv[0] = c0 * z * (T(2) * z2 - T(3) * rho2);
      Sg[0][0] = T(-6) * c0 * z * x; Sg[0][1] = T(-6) * c0 * z * y;
      Sg[0][2] = c0 * (T(6) * z2 - T(3) * rho2);
      T w = T(4) * z2 - rho2;
      Sv[1] = c1 * w * x;
      Sg[1][0] = c1 * (w - T(2) * x2); Sg[1][1] = T(-2) * c1 * x * y; Sg[1][2] = T(8) * c1 * z * x;
      Sv[2] = c1 * w * y;
      Sg[2][0] = T(-2) * c1 * x * y; Sg[2][1] = c1 * (w - T(2) * y2); Sg[2][2] = T(8) * c1 * z * y;
      Sv[3] = c2 * z * (x2 - y2);
      Sg[3][0] = T(2) * c2 * z * x; Sg[3][1] = T(-2) * c2 * z * y; Sg[3][2] = c2 * (x2 - y2);
      Sv[4] = T(2) * c2 * z * x * y;
      Sg[4][0] = T(2) * c2 * z * y; Sg[4][1] = T(2) * c2 * z * x; Sg[4][2] = T(2) * c2 * x * y;
      Sv[5] = c3 * (x2 * x - T(3) * x * y2);
      Sg[5][0] = c3 * (T(3) * x2 - T(3) * y2); Sg[5][1] = T(-6) * c3 * x * y; Sg[5][2] = 0;
      Sv[6] = c3 * (T(3) * x2 * y - y2 * y);
      Sg[6][0] = T(6) * c3 * x * y; Sg[6][1] = c3 * (T(3) * x2 - T(3) * y2); Sg[6][2] = 0;
      break;
    }
    case 4: {
      const T c0 = T(0.125), c1 = T(0.7905694150420949), c2 = T(0.5590169943749475),
              c3 = T(2.0916500663351889), c4 = T(0.7395099728874520);
      T r2 = x2 + y2 + z2;
      Sv[0] = c0 * (T(35) * z2 * z2 - T(30) * r2 * z2 + T(3) * r2 * r2);
      // d/dx: -60 x z2 + 12 r2 x ; d/dz: 140 z^3 - 60 z r2 - 60 z^3 + 12 r2 z
      Sg[0][0] = c0 * (T(-60) * x * z2 + T(12) * r2 * x);
      Sg[0][1] = c0 * (T(-60) * y * z2 + T(12) * r2 * y);
      Sg[0][2] = c0 * (T(80) * z2 * z - T(48) * r2 * z);
      T w1 = T(7) * z2 - T(3) * r2;   // d/dx = -6x, d/dz = 8z
      Sv[1] = c1 * z * w1 * x;
      Sg[1][0] = c1 * z * (w1 - T(6) * x2); Sg[1][1] = c1 * z * (T(-6) * x * y);
      Sg[1][2] = c1 * x * (w1 + T(8) * z2);
      Sv[2] = c1 * z * w1 * y;
      Sg[2][0] = c1 * z * (T(-6) * x * y); Sg[2][1] = c1 * z * (w1 - T(6) * y2);
      Sg[2][2] = c1 * y * (w1 + T(8) * z2);
      T w2 = T(7) * z2 - r2;          // d/dx = -2x, d/dz = 12z
      T p2 = x2 - y2, q2 = T(2) * x * y;
      Sv[3] = c2 * w2 * p2;
      Sg[3][0] = c2 * (T(-2) * x * p2 + w2 * T(2) * x);
      Sg[3][1] = c2 * (T(-2) * y * p2 - w2 * T(2) * y);
      Sg[3][2] = c2 * T(12) * z * p2;
      Sv[4] = c2 * w2 * q2;
      Sg[4][0] = c2 * (T(-2) * x * q2 + w2 * T(2) * y);
      Sg[4][1] = c2 * (T(-2) * y * q2 + w2 * T(2) * x);
      Sg[4][2] = c2 * T(12) * z * q2;
      T p3 = x2 * x - T(3) * x * y2, q3 = T(3) * x2 * y - y2 * y;
      Sv[5] = c3 * z * p3;
      Sg[5][0] = c3 * z * (T(3) * x2 - T(3) * y2); Sg[5][1] = c3 * z * (T(-6) * x * y); Sg[5][2] = c3 * p3;
      Sv[6] = c3 * z * q3;
      Sg[6][0] = c3 * z * (T(6) * x * y); Sg[6][1] = c3 * z * (T(3) * x2 - T(3) * y2); Sg[6][2] = c3 * q3;
      T p4 = x2 * x2 - T(6) * x2 * y2 + y2 * y2, q4 = T(4) * x * y * (x2 - y2);
      Sv[7] = c4 * p4;
      Sg[7][0] = c4 * (T(4) * x2 * x - T(12) * x * y2); Sg[7][1] = c4 * (T(-12) * x2 * y + T(4) * y2 * y);
      Sg[7][2] = 0;
      Sv[8] = c4 * q4;
      Sg[8][0] = c4 * (T(12) * x2 * y - T(4) * y2 * y); Sg[8][1] = c4 * (T(4) * x2 * x - T(12) * x * y2);
      Sg[8][2] = 0;
      break;
    }
    default: {  // l = 5: S_5m = f_m(z, rho) {A_m, B_m}, A_m + i B_m = (x + i y)^|m|
      const T rho = x2 + y2 + z2;
      // f_m = N_m Pi_5^m with its partials in z and rho; N_m Pi_5^m's
      // constants: sqrt(15)/8, sqrt(105)/4, sqrt(70)/16, 3 sqrt(35)/8, 3 sqrt(14)/16
      const T c1 = T(0.4841229182759271), c2 = T(2.5617376914898995), c3 = T(0.5229125165837972),
              c4 = T(2.2185299186623560), c5 = T(0.7015607600201140);
      T f[6], fz[6], fr[6];
      f[0] = T(0.125) * z * (T(63) * z2 * z2 - T(70) * rho * z2 + T(15) * rho * rho);
      fz[0] = T(0.125) * (T(315) * z2 * z2 - T(210) * rho * z2 + T(15) * rho * rho);
      fr[0] = T(0.125) * z * (T(-70) * z2 + T(30) * rho);
      f[1] = c1 * (T(21) * z2 * z2 - T(14) * rho * z2 + rho * rho);
      fz[1] = c1 * (T(84) * z2 * z - T(28) * rho * z);
      fr[1] = c1 * (T(-14) * z2 + T(2) * rho);
      f[2] = c2 * z * (T(3) * z2 - rho);
      fz[2] = c2 * (T(9) * z2 - rho);
      fr[2] = -c2 * z;
      f[3] = c3 * (T(9) * z2 - rho);
      fz[3] = c3 * T(18) * z;
      fr[3] = -c3;
      f[4] = c4 * z; fz[4] = c4; fr[4] = T(0);
      f[5] = c5; fz[5] = T(0); fr[5] = T(0);
      T A[6], B[6];
      A[0] = T(1); B[0] = T(0);
#pragma unroll
      for (int m = 1; m <= 5; m++) {
        A[m] = A[m - 1] * x - B[m - 1] * y;
        B[m] = A[m - 1] * y + B[m - 1] * x;
      }
      Sv[0] = f[0];
      Sg[0][0] = T(2) * x * fr[0]; Sg[0][1] = T(2) * y * fr[0]; Sg[0][2] = fz[0] + T(2) * z * fr[0];
#pragma unroll
      for (int m = 1; m <= 5; m++) {
        // d(x+iy)^m/dx = m (x+iy)^(m-1), d/dy = i m (x+iy)^(m-1)
        const T gx = T(2) * x * fr[m], gy = T(2) * y * fr[m], gz = fz[m] + T(2) * z * fr[m];
        Sv[2 * m - 1] = f[m] * A[m];
        Sg[2 * m - 1][0] = gx * A[m] + f[m] * T(m) * A[m - 1];
        Sg[2 * m - 1][1] = gy * A[m] - f[m] * T(m) * B[m - 1];
        Sg[2 * m - 1][2] = gz * A[m];
        Sv[2 * m] = f[m] * B[m];
        Sg[2 * m][0] = gx * B[m] + f[m] * T(m) * B[m - 1];
        Sg[2 * m][1] = gy * B[m] + f[m] * T(m) * A[m - 1];
        Sg[2 * m][2] = gz * B[m];
      }
      break;
    }
  }
}

__device__ __forceinline__ int m_slot(int m) { return m > 0 ? 2 * m - 1 : (m < 0 ? -2 * m : 0); }

template <typename T, int L, int ST>
__device__ __forceinline__ void shell_write(T x, T y, T z, T f, T fr, T lapfac, int want,
                                            cptr<int> bas, T *phi) {
  T Sv[2 * L + 1], Sg[2 * L + 1][3];
  solid_harm<T>(L, x, y, z, Sv, Sg);
#pragma unroll
  for (int q = 0; q < 2 * L + 1; q++) {
    int b = bas[q];
    if (b < 0) continue;
    T *o = phi + ST * b;
    o[0] = Sv[q] * f;
    if (ST == 5 && want) {
      o[1] = f * Sg[q][0] + Sv[q] * fr * x;
      o[2] = f * Sg[q][1] + Sv[q] * fr * y;
      o[3] = f * Sg[q][2] + Sv[q] * fr * z;
      o[4] = Sv[q] * lapfac;
    }
  }
}

// Tabulated radial function k at r: R, and with want R' and R''
// (qmc/vmc/splfit.f90:13-76).  xr is the 1-based grid coordinate; below the
// first point the polynomial ce, at or beyond the last point the tail
// a0 exp(-a1 r), else the cubic spline between points jx and jx+1, whose
// positions are read from the table.  jx = int(xr) lies in [1, nr-1] on
// the spline branch, so both points are inside the function's table.
template <typename T>
__device__ __forceinline__ void num_radial(const DevSys &S, int k, T r, int want, T &R, T &R1,
                                           T &R2) {
  const cptr<double> p = cst(S.num_par) + k * kNumPar;
  const int igrid = (int)p[0], nr = (int)p[1];
  const T r0 = T(p[2]), ih = T(p[3]);
  T xr;
  if (igrid == 1)
    xr = (r - r0) * ih + T(1);
  else if (igrid == 2)
    xr = log(r / r0) * ih + T(1);
  else
    xr = log((r + r0) / r0) * ih + T(1);
  if (!(xr >= T(1))) {
    const T c1 = T(p[5]), c2 = T(p[6]), c3 = T(p[7]), c4 = T(p[8]);
    R = T(p[4]) + r * (c1 + r * (c2 + r * (c3 + r * c4)));
    R1 = c1 + r * (T(2) * c2 + r * (T(3) * c3 + r * T(4) * c4));
    R2 = T(2) * c2 + r * (T(6) * c3 + r * T(12) * c4);
    return;
  }
  if (xr >= T(nr)) {
    const T a1 = T(p[10]);
    R = T(p[9]) * exp(-a1 * r);
    R1 = -a1 * R;
    R2 = a1 * a1 * R;
    return;
  }
  const int jx = (int)xr;
  const cptr<double> t = cst(S.num_tab) + 4 * ((int64_t)p[11] + jx - 1);
  const T ref0 = T(t[0]), y0 = T(t[1]), d0 = T(t[2]);
  const T ref1 = T(t[4]), y1 = T(t[5]), d1 = T(t[6]);
  const T delh = ref1 - ref0, idh = T(1) / delh;
  const T bb = (r - ref0) * idh, aa = (ref1 - r) * idh;
  const T h6 = delh * delh * T(1.0 / 6.0);
  R = aa * y0 + bb * y1 + (aa * (aa * aa - T(1)) * d0 + bb * (bb * bb - T(1)) * d1) * h6;
  if (want) {
    R1 = (y1 - y0) * idh + ((T(3) * bb * bb - T(1)) * d1 - (T(3) * aa * aa - T(1)) * d0) * delh *
                               T(1.0 / 6.0);
    R2 = aa * d0 + bb * d1;
  }
}

// Radial part f = R_k(r) r^-l of a tabulated shell (value only).
template <typename T>
__device__ __forceinline__ T num_radial_f(const DevSys &S, int k, int l, T r) {
  T R, R1, R2;
  num_radial<T>(S, k, r, 0, R, R1, R2);
  const T ri = T(1) / r;
  for (int q = 0; q < l; q++) R *= ri;
  return R;
}

// Evaluate shell s at displacement (x, y, z) from its centre, distance r.
// want: 0 value only, 1 value + gradient + laplacian.  ST is the stride of
// phi per basis function (5: value, gradient, laplacian; 1: value only).
template <typename T, int ST = 5>
__device__ void eval_shell(const DevSys &S, int s, T x, T y, T z, T r, int want,
                           T *phi) {
  int l = cst(S.sh_l)[s], a = cst(S.sh_a)[s];
  if (r == T(0)) { r = T(1e-30); x = T(1e-30); }
  const int kn = S.has_num ? cst(S.sh_num)[s] : -1;
  if (kn >= 0) {
    // f = R r^-l: f' = (R' - l R / r) r^-l,
    // f'' = (R'' - 2 l R' / r + l (l + 1) R / r^2) r^-l
    T R, R1, R2;
    num_radial<T>(S, kn, r, want, R, R1, R2);
    const T ri = T(1) / r;
    T ril = T(1);
    for (int q = 0; q < l; q++) ril *= ri;
    const T f = R * ril;
    T fr = 0, lapfac = 0;
    if (want) {
      const T f1 = (R1 - T(l) * R * ri) * ril;
      const T f2 = (R2 - T(2 * l) * R1 * ri + T(l * (l + 1)) * R * ri * ri) * ril;
      fr = f1 * ri;
      lapfac = f2 + T(2 * (l + 1)) * fr;
    }
    const cptr<int> bas = cst(S.sh_bas) + s * kMaxShellM;
    switch (l) {
      case 0: shell_write<T, 0, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
      case 1: shell_write<T, 1, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
      case 2: shell_write<T, 2, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
      case 3: shell_write<T, 3, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
      case 4: shell_write<T, 4, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
    default: shell_write<T, 5, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
    }
    return;
  }
  // radial f = r^a g(r), g = sum d exp(-z r^q); need f, f'/r, f''
  T g = 0, g1 = 0, g2 = 0;  // g, dg/dr, d2g/dr2
  const bool gauss = (S.basis_type == QMCW_BASIS_GAUSSIAN);
  const cptr<double> pz = cst(S.prim_zeta), pc = cst(S.prim_coef);
  for (int p = cst(S.sh_prim_off)[s]; p < cst(S.sh_prim_off)[s + 1]; p++) {
    T zeta = T(pz[p]), d = T(pc[p]);
    if (gauss) {
      T e = d * exp(-zeta * r * r);
      g += e;
      g1 += T(-2) * zeta * r * e;
      g2 += (T(4) * zeta * zeta * r * r - T(2) * zeta) * e;
    } else {
      T e = d * exp(-zeta * r);
      g += e;
      g1 += -zeta * e;
      g2 += zeta * zeta * e;
    }
  }
  T f, fr, f2;  // f, f'/r, f''
  if (a == 0) {
    f = g;
    fr = g1 / r;
    f2 = g2;
  } else {
    T ra1 = T(1);
    for (int q = 0; q < a - 1; q++) ra1 *= r;   // r^(a-1)
    T ra = ra1 * r;
    T ra2 = (a >= 2) ? ra1 / r : T(0);          // r^(a-2) (only used when a >= 2)
    f = ra * g;
    fr = (T(a) * ra1 * g + ra * g1) / r;
    f2 = (a >= 2 ? T(a * (a - 1)) * ra2 * g : T(0)) + T(2 * a) * ra1 * g1 + ra * g2;
  }
  // laplacian of S f(r): S (f'' + 2 (l+1) f'/r)
  T lapfac = f2 + T(2 * (l + 1)) * fr;
  const cptr<int> bas = cst(S.sh_bas) + s * kMaxShellM;
  switch (l) {
    case 0: shell_write<T, 0, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
    case 1: shell_write<T, 1, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
    case 2: shell_write<T, 2, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
    case 3: shell_write<T, 3, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
    case 4: shell_write<T, 4, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
    default: shell_write<T, 5, ST>(x, y, z, f, fr, lapfac, want, bas, phi); break;
  }
}

// sum over the shell's components of gp[basis] * phi_basis at (x, y, z).
// Shells are complete: slot q holds m = 0, 1, -1, 2, -2, ...; an absent
// component has basis index -1.
template <typename T, int L, typename GP>
__device__ __forceinline__ T shell_dot_l(T x, T y, T z, cptr<int> bas, const GP *gp) {
  T Sv[2 * L + 1], Sg[2 * L + 1][3];
  solid_harm<T>(L, x, y, z, Sv, Sg);
  T acc = 0;
#pragma unroll
  for (int q = 0; q < 2 * L + 1; q++) {
    int b = bas[q];
    if (b >= 0) acc += T(gp[b]) * Sv[q];   // gp already rounded in the BF16 engine
  }
  return acc;
}
template <typename T, typename GP = T>
__device__ T shell_dot(const DevSys &S, int s, T x, T y, T z, T r, const GP *gp) {
  int l = cst(S.sh_l)[s], a = cst(S.sh_a)[s];
  const bool gauss = (S.basis_type == QMCW_BASIS_GAUSSIAN);
  T g = 0;
  const int kn = S.has_num ? cst(S.sh_num)[s] : -1;
  if (kn >= 0) {
    g = num_radial_f<T>(S, kn, l, r);
  } else {
    const cptr<double> pz = cst(S.prim_zeta), pc = cst(S.prim_coef);
    for (int p = cst(S.sh_prim_off)[s]; p < cst(S.sh_prim_off)[s + 1]; p++) {
      T zeta = T(pz[p]), d = T(pc[p]);
      g += d * exp(gauss ? -zeta * r * r : -zeta * r);
    }
    for (int q = 0; q < a; q++) g *= r;
  }
  const cptr<int> bas = cst(S.sh_bas) + s * kMaxShellM;
  T acc;
  switch (l) {
    case 0: acc = shell_dot_l<T, 0, GP>(x, y, z, bas, gp); break;
    case 1: acc = shell_dot_l<T, 1, GP>(x, y, z, bas, gp); break;
    case 2: acc = shell_dot_l<T, 2, GP>(x, y, z, bas, gp); break;
    case 3: acc = shell_dot_l<T, 3, GP>(x, y, z, bas, gp); break;
    case 4: acc = shell_dot_l<T, 4, GP>(x, y, z, bas, gp); break;
    default: acc = shell_dot_l<T, 5, GP>(x, y, z, bas, gp); break;
  }
  return acc * g;
}

// Values of shell s's basis functions at (x, y, z), r, accumulated in FP64:
// rat += sum gp[b] phi_b and pb[b] += fac phi_b (the pseudopotential's
// determinant ratio and weighted basis sums in one pass, no stored values).
template <typename T, int L>
__device__ __forceinline__ void shell_axpy_l(T x, T y, T z, T g, cptr<int> bas, const double *gp,
                                             double fac, double *pb, double &rat) {
  T Sv[2 * L + 1], Sg[2 * L + 1][3];
  solid_harm<T>(L, x, y, z, Sv, Sg);
  // the shell's sums first (its components are distinct), then the stores
  double old[2 * L + 1];
#pragma unroll
  for (int q = 0; q < 2 * L + 1; q++) old[q] = bas[q] >= 0 ? pb[bas[q]] : 0.0;
#pragma unroll
  for (int q = 0; q < 2 * L + 1; q++) {
    const int b = bas[q];
    if (b >= 0) {
      const double v = (double)(Sv[q] * g);
      rat += gp[b] * v;
      pb[b] = old[q] + fac * v;
    }
  }
}
template <typename T>
__device__ __forceinline__ void shell_val_axpy(const DevSys &S, int s, T x, T y, T z, T r,
                                               const double *gp, double fac, double *pb, double &rat) {
  const int l = cst(S.sh_l)[s], a = cst(S.sh_a)[s];
  const bool gauss = (S.basis_type == QMCW_BASIS_GAUSSIAN);
  T g = 0;
  const int kn = S.has_num ? cst(S.sh_num)[s] : -1;
  if (kn >= 0) {
    g = num_radial_f<T>(S, kn, l, r);
  } else {
    const cptr<double> pz = cst(S.prim_zeta), pc = cst(S.prim_coef);
    for (int p = cst(S.sh_prim_off)[s]; p < cst(S.sh_prim_off)[s + 1]; p++) {
      const T zeta = T(pz[p]), d = T(pc[p]);
      g += d * exp(gauss ? -zeta * r * r : -zeta * r);
    }
    for (int q = 0; q < a; q++) g *= r;
  }
  const cptr<int> bas = cst(S.sh_bas) + s * kMaxShellM;
  switch (l) {
    case 0: shell_axpy_l<T, 0>(x, y, z, g, bas, gp, fac, pb, rat); break;
    case 1: shell_axpy_l<T, 1>(x, y, z, g, bas, gp, fac, pb, rat); break;
    case 2: shell_axpy_l<T, 2>(x, y, z, g, bas, gp, fac, pb, rat); break;
    case 3: shell_axpy_l<T, 3>(x, y, z, g, bas, gp, fac, pb, rat); break;
    case 4: shell_axpy_l<T, 4>(x, y, z, g, bas, gp, fac, pb, rat); break;
    default: shell_axpy_l<T, 5>(x, y, z, g, bas, gp, fac, pb, rat); break;
  }
}

// ---------------------------------------------------------------------------
// CHAMP's pivoted Gauss-Jordan (qmc/vmc/matinv.f90) on a column-major k x k
// block held by one lane.  Singular -> det = 0, inverse zeroed.
__device__ inline double small_matinv(double *a, int n) {
  if (n == 0) return 1.0;
  if (n == 1) {   // the general path gives 1 * (1 / a) and det 1 * a (0 and 0 when singular)
    const double a0 = a[0];
    a[0] = a0 == 0.0 ? 0.0 : 1.0 / a0;
    return a0 == 0.0 ? 0.0 : a0;
  }
  int ipiv[kMaxKmax];
  double atemp[kMaxKmax];
  for (int i = 0; i < n; i++) ipiv[i] = i;
  double det = 1.0;
  for (int i = 0; i < n; i++) {
    int col = i * n;
    double adiag = a[ipiv[i] + col];
    int idiag = i;
    for (int k = i; k < n; k++)
      if (fabs(a[ipiv[k] + col]) > fabs(adiag)) { adiag = a[ipiv[k] + col]; idiag = k; }
    if (idiag != i) {
      det = -det;
      int t = ipiv[i]; ipiv[i] = ipiv[idiag]; ipiv[idiag] = t;
    }
    det *= adiag;
    int p = ipiv[i];
    a[p + col] = 1.0;
    if (adiag == 0.0) {
      for (int q = 0; q < n * n; q++) a[q] = 0.0;
      return 0.0;
    }
    double inv = 1.0 / adiag;
    for (int q = 0; q < n; q++) a[p + q * n] *= inv;
    for (int j = 0; j < n; j++)
      if (j != p) {
        double t = -a[j + col];
        a[j + col] = 0.0;
        for (int q = 0; q < n; q++) a[j + q * n] += t * a[p + q * n];
      }
  }
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < n; i++) atemp[i] = a[i + j * n];
    for (int i = 0; i < n; i++) a[i + j * n] = atemp[ipiv[i]];
  }
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) atemp[j] = a[i + j * n];
    for (int j = 0; j < n; j++) a[i + ipiv[j] * n] = atemp[j];
  }
  return det;
}

// (2l+1) P_l(cos) with l given as l+1 (qmc/vmc/nonloc.f90:249-272)
__device__ __forceinline__ double legendre_w(int l1, double c) {
  switch (l1) {
    case 1: return 1.0;
    case 2: return 3.0 * c;
    case 3: return 2.5 * (3 * c * c - 1);
    case 4: return 3.5 * c * (5 * c * c - 3);
    default: return 1.125 * (35 * c * c * c * c - 30 * c * c + 3);
  }
}

__device__ __forceinline__ double ecp_radial(const DevSys &S, double r, int l0, int ict) {
  double v = 0, rsq = r * r;
  int b = cst(S.ecp_off)[ict * QMCW_MAX_LPOT + l0], e = cst(S.ecp_off)[ict * QMCW_MAX_LPOT + l0 + 1];
  for (int q = b; q < e; q++) {
    int pw = cst(S.ecp_power)[q];
    double p = 1.0;
    if (pw != 2) p = pow(r, (double)(pw - 2));
    v += p * cst(S.ecp_coef)[q] * exp(-cst(S.ecp_expo)[q] * rsq);
  }
  return v;
}

// CHAMP's rannyu on a 4 x 12-bit limb state (qmc/vmc/rannyu.f90:53-74)
__host__ __device__ __forceinline__ double rannyu(int *l) {
  const int m1 = 502, m2 = 1521, m3 = 4071, m4 = 2107;
  const double two12i = 2.44140625e-4;
  int l1 = l[0], l2 = l[1], l3 = l[2], l4 = l[3];
  int i1 = l1 * m4 + l2 * m3 + l3 * m2 + l4 * m1;
  int i2 = l2 * m4 + l3 * m3 + l4 * m2;
  int i3 = l3 * m4 + l4 * m3;
  int i4 = l4 * m4;
  l4 = i4 & 4095;
  i3 = i3 + (i4 >> 12);
  l3 = i3 & 4095;
  i2 = i2 + (i3 >> 12);
  l2 = i2 & 4095;
  l1 = (i1 + (i2 >> 12)) & 4095;
  l[0] = l1; l[1] = l2; l[2] = l3; l[3] = l4;
  return two12i * ((double)l1 + two12i * ((double)l2 + two12i * ((double)l3 + two12i * (double)l4)));
}

}  // namespace qmcw
