/*
 * oracle.c -- FP64 CPU restatement of CHAMP's wavefunction hot path.
 *
 * TEST INFRASTRUCTURE (the checker): only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this.  Each function names the
 * CHAMP routine (file:line) whose algebra it restates; loop orders follow
 * the Fortran so that the arithmetic matches it term by term.
 *
 * Parity pins (see tests/test_oracle_*.py): closed forms, central finite
 * differences, brute-force determinant expansions, the reference's own
 * rannyu/matinv/rotqua compiled under oracle/_ref, and the VMC energies of
 * CHAMP's Be transcript (examples/be_vb1_jcas24_vmc_j_opt_lin.out).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXT 64       /* max e-e-n terms (nordc <= 7) */
#define MAXEQ 16
#define MAXPJ 512

static __thread char g_err[256];
const char *orc_error(void) { return g_err; }

/* ------------------------------------------------------------------ */
/* system                                                              */
/* ------------------------------------------------------------------ */
struct orc_sys {
  int nelec, nup, ndn, ncent, nctype, nbasis, norb, nterms;
  int basis_type, nrad;
  double *cent, *znuc, pecent;
  int *iwctype;
  int *rad_n, *rad_prim_off, *ctype_bas_off, *bas_rad, *bas_l, *bas_m;
  double *prim_zeta, *prim_coef;
  /* tabulated radial functions (rad_num[r] >= 0: table of radial fn r) */
  int *rad_num, nnum, *num_igrid, *num_nr, *num_off;
  double *num_h, *num_r0, *num_rwf, *num_d2, *num_ce, *num_ae;
  double *coef; /* [norb][nbasis] */
  /* dets */
  int ndet, ncsf;
  int *det_up, *det_dn;
  double *csf_coef, *csf_det_coef;
  int *csf_det_off, *csf_det;
  double *cdet_in_wf;
  int ndetup, ndetdn, *iwdetup, *iwdetdn, *iworbdup, *iworbddn;
  int noccup, noccdn, *occup, *occdn, *ioccup, *ioccdn;
  int *irefup, *irefdn, *ivirup, *ivirdn;
  int *ordup, *orddn, *sgnup, *sgndn;
  int *lab_excup, *lab_excdn, *refup, *refdn, *excup, *excdn; /* [ndet][nsig] */
  int kmax;
  /* Jastrow */
  int isc, nspin2b, nocuspb, norda, nordb, nordc;
  double scalek, asymp_r, asymp_jasa[64], asymp_jasb[2];
  double *a4, *b, *c;
  int neqs, iwc4[MAXEQ];
  double dmat[MAXEQ][MAXT];
  int nvdepend[MAXEQ][64];
  int iwdepend[MAXEQ][MAXT][64]; /* index into the type's varied-c list */
  double cdep[MAXEQ][MAXT][64];
  /* pseudopotential */
  int nloc, nquad;
  int npotd[64], lpotp1[64];
  int *ecp_off, *ecp_power;
  double *ecp_coef, *ecp_expo;
  double xq0[QMCW_MAX_NQUAD], yq0[QMCW_MAX_NQUAD], zq0[QMCW_MAX_NQUAD],
      wq[QMCW_MAX_NQUAD];
  /* parameters */
  int npa, npb, npc, npcsf, nporb, npj, nparm;
  int *pa_type, *pa_idx, *pb_spin, *pb_idx, *pc_type, *pc_idx;
  int *pcsf, *porb_i, *porb_j, *porb_rev;
  int nparmc_t[64];      /* varied c per type */
  int pc_first[64];      /* index of the type's first varied c in the list */
  double hb;
  /* precision emulation for error studies (orc_set_emulation): bit 0 rounds
     the inputs of the full orbital rows (value, gradient, laplacian) to
     bfloat16, bit 1 those of the value-only rows (quadrature points) */
  int emul;
  /* VMC mover: 1 metrop_mov1, 6 metrop_slat_mov1 (orc_set_mover) */
  int imetro;
  double deltar, deltat;
};

static void *dupa(const void *p, size_t n) {
  if (!p || n == 0) return calloc(1, 8);
  void *q = malloc(n);
  memcpy(q, p, n);
  return q;
}
#define DUPD(p, n) ((double *)dupa((p), sizeof(double) * (size_t)(n)))
#define DUPI(p, n) ((int *)dupa((p), sizeof(int) * (size_t)(n)))

/* qmc/vmc/jastrow4.f90:495-516 */
int32_t qmcw_nterms4(int32_t nord) {
  int i = 0;
  for (int n = 2; n <= nord; n++)
    for (int k = n - 1; k >= 0; k--) {
      int l_hi = (k == 0) ? n - k - 2 : n - k;
      for (int l = l_hi; l >= 0; l--) {
        int m = (n - k - l) / 2;
        if (2 * m == n - k - l) i++;
      }
    }
  return i;
}
int32_t orc_nterms(const orc_sys *o) { return o->nterms; }
int32_t orc_nparm(const orc_sys *o) { return o->nparm; }

/* qmc/vmc/scale_dist.f90:584-800 (scale_dist2) for isc = 2, 4; value only
   variant is scale_dist (:239-330). */
static void scale_dist2(const orc_sys *o, double r, double *rr, double *dd1,
                        double *dd2) {
  double k = o->scalek;
  if (k == 0.0) {
    *rr = r; *dd1 = 1; *dd2 = 0;
    return;
  }
  if (o->isc == 2) {
    *rr = (1.0 - exp(-k * r)) / k;
    *dd1 = 1.0 - k * *rr;
    *dd2 = -k * *dd1;
  } else { /* isc == 4 */
    double deni = 1.0 / (1.0 + k * r);
    *rr = r * deni;
    *dd1 = deni * deni;
    *dd2 = -2.0 * k * deni * *dd1;
  }
}
/* qmc/vmc/scale_dist.f90:857-885 (switch_scale2), isc < 8 */
static void switch_scale2(const orc_sys *o, double *rr, double *dd1,
                          double *dd2) {
  double k = o->scalek;
  *rr = 1.0 - k * *rr;
  *dd1 = -k * *dd1;
  *dd2 = -k * *dd2;
}

/* qmc/vmc/scale_dist.f90:1-237 (set_scale_dist), isc = 2, 4, ijas = 4 */
static void set_scale_dist(orc_sys *o) {
  double k = o->scalek;
  o->asymp_r = 1.0 / k;
  double ar = o->asymp_r;
  for (int it = 0; it < o->nctype; it++) {
    const double *a = o->a4 + it * (o->norda + 1);
    double v = a[0] * ar / (1.0 + a[1] * ar);
    for (int iord = 2; iord <= o->norda; iord++) v += a[iord] * pow(ar, iord);
    o->asymp_jasa[it] = v;
  }
  for (int i = 0; i < 2; i++) {
    double sspinn;
    int isp;
    if (i == 0) {
      sspinn = 1.0;
      isp = 0;
    } else {
      sspinn = (o->nspin2b == 1 && o->nocuspb == 0) ? 0.5 : 1.0;
      isp = o->nspin2b - 1;
    }
    const double *b = o->b + isp * (o->nordb + 1);
    double v = sspinn * b[0] * ar / (1.0 + b[1] * ar);
    for (int iord = 2; iord <= o->nordb; iord++) v += b[iord] * pow(ar, iord);
    o->asymp_jasb[i] = v;
  }
}
void orc_asymp(const orc_sys *o, double *aa, double *ab) {
  for (int it = 0; it < o->nctype; it++) aa[it] = o->asymp_jasa[it];
  ab[0] = o->asymp_jasb[0];
  ab[1] = o->asymp_jasb[1];
}

/* qmc/fit/cuspinit4.f90:1-60 (d matrix, iwc4) and checkdepend4 :62-156 */
static void cuspinit4(orc_sys *o) {
  int nordc = o->nordc;
  o->neqs = (nordc > 1) ? 2 * (nordc - 1) : 0;
  memset(o->dmat, 0, sizeof(o->dmat));
  memset(o->iwc4, 0, sizeof(o->iwc4));
  int i = 0;
  for (int n = 2; n <= nordc; n++)
    for (int k = n - 1; k >= 0; k--) {
      int l_hi = (k == 0) ? n - k - 2 : n - k;
      for (int l = l_hi; l >= 0; l--) {
        int m = (n - k - l) / 2;
        if (2 * m == n - k - l) {
          i++;
          if (k == 1 && o->iwc4[n - 2] == 0) o->iwc4[n - 2] = i;
          if (k == 0 && o->iwc4[n + nordc - 3] == 0) o->iwc4[n + nordc - 3] = i;
          int iord = l + 2 * m;
          o->dmat[iord - 1][i - 1] -= 2 * k;
          if (l + m > 0) {
            iord = nordc - 1 + k + m;
            o->dmat[iord - 1][i - 1] -= (l + m);
          }
          if (m > 0) {
            iord = nordc - 1 + k + l + m;
            o->dmat[iord - 1][i - 1] -= m;
          }
        }
      }
    }
  /* checkdepend4: dependent c's as linear functions of the varied ones */
  memset(o->nvdepend, 0, sizeof(o->nvdepend));
  memset(o->cdep, 0, sizeof(o->cdep));
  int neqs = o->neqs;
  for (int ie = 1; ie <= neqs; ie++) {
    if (ie == nordc) continue;
    for (int j = 1; j <= o->nterms; j++) {
      if (fabs(o->dmat[ie - 1][j - 1]) > 1e-10) {
        for (int it = 0; it < o->nctype; it++)
          for (int l = 0; l < o->nparmc_t[it]; l++) {
            int idx = o->pc_idx[o->pc_first[it] + l];
            if (j == idx) {
              int nv = o->nvdepend[ie - 1][it]++;
              o->iwdepend[ie - 1][nv][it] = l;
              o->cdep[ie - 1][nv][it] =
                  -o->dmat[ie - 1][j - 1] / o->dmat[ie - 1][o->iwc4[ie - 1] - 1];
            }
          }
      }
    }
  }
  if (nordc > 1) {
    for (int it = 0; it < o->nctype; it++)
      for (int l = 0; l < o->nparmc_t[it]; l++) {
        int nv = o->nvdepend[nordc - 1][it]++;
        o->iwdepend[nordc - 1][nv][it] = l;
      }
    for (int ie = 1; ie <= neqs; ie++) {
      if (ie == nordc) continue;
      double factor = -o->dmat[nordc - 1][o->iwc4[ie - 1] - 1] /
                      o->dmat[nordc - 1][o->iwc4[nordc - 1] - 1];
      for (int j = 1; j <= o->nterms; j++)
        for (int it = 0; it < o->nctype; it++)
          for (int l = 0; l < o->nparmc_t[it]; l++) {
            int idx = o->pc_idx[o->pc_first[it] + l];
            if (j == idx)
              o->cdep[nordc - 1][l][it] -=
                  factor * (o->dmat[ie - 1][j - 1] /
                            o->dmat[ie - 1][o->iwc4[ie - 1] - 1]);
          }
    }
  }
}
int32_t orc_cusp(const orc_sys *o, double *d, int32_t *iwc4) {
  for (int i = 0; i < o->neqs; i++) {
    iwc4[i] = o->iwc4[i];
    for (int j = 0; j < o->nterms; j++) d[i * o->nterms + j] = o->dmat[i][j];
  }
  return o->neqs;
}

/* qmc/fit/cuspexact4.f90:1-74 */
void orc_cuspexact(const orc_sys *o, double *c) {
  int nordc = o->nordc, nt = o->nterms;
  for (int it = 0; it < o->nctype; it++) {
    double *ct = c + it * nt;
    for (int i = 1; i <= nordc - 1; i++) {
      double sum = 0;
      for (int j = 1; j <= nt; j++)
        if (j != o->iwc4[i - 1]) sum += o->dmat[i - 1][j - 1] * ct[j - 1];
      ct[o->iwc4[i - 1] - 1] = -sum / o->dmat[i - 1][o->iwc4[i - 1] - 1];
    }
    for (int i = nordc + 1; i <= 2 * (nordc - 1); i++) {
      double sum = 0;
      for (int j = 1; j <= nt; j++)
        if (j != o->iwc4[i - 1]) sum += o->dmat[i - 1][j - 1] * ct[j - 1];
      ct[o->iwc4[i - 1] - 1] = -sum / o->dmat[i - 1][o->iwc4[i - 1] - 1];
    }
    if (nordc > 1) {
      int i = nordc;
      double sum = 0;
      for (int j = 1; j <= nt; j++)
        if (j != o->iwc4[i - 1]) sum += o->dmat[i - 1][j - 1] * ct[j - 1];
      ct[o->iwc4[i - 1] - 1] = -sum / o->dmat[i - 1][o->iwc4[i - 1] - 1];
    }
  }
}

/* qmc/deriv_fast_mod.f90:85-101 (xorbs): orbs1 entries absent from orbs2 */
static int xorbs(const int *o1, int n1, const int *o2, int n2, int *res) {
  int n = 0;
  for (int i = 0; i < n1; i++) {
    int found = 0;
    for (int j = 0; j < n2; j++)
      if (o2[j] == o1[i]) found = 1;
    if (!found) res[n++] = o1[i];
  }
  return n;
}
/* qmc/deriv_fast_mod.f90:103-139 (index_of, nswaps) */
static int nswaps(const int *exc, const int *ref, int len) {
  int n = 1, nswp = 0;
  for (int i = 1; i <= len; i++) {
    int ind = -1;
    for (int j = 0; j < len; j++)
      if (ref[j] == exc[i - 1]) { ind = j + 1; break; }
    if (ind != -1) {
      nswp += (i - n) + (ind - n);
      n++;
    }
  }
  return nswp;
}
static int cmp_int(const void *a, const void *b) {
  return (*(const int *)a) - (*(const int *)b);
}

/* qmc/vmc/determinant_up_dn.f90 (unique dets), qmc/deriv_fast_mod.f90:211-259
   (occ_bld) and :261-365 (excit_bld), :141-169 (cdet_in_wf_bld) */
static void setup_dets(orc_sys *o) {
  int nup = o->nup, ndn = o->ndn, ndet = o->ndet;
  o->iwdetup = calloc(ndet, sizeof(int));
  o->iwdetdn = calloc(ndet, sizeof(int));
  o->iworbdup = calloc((size_t)ndet * (nup + 1), sizeof(int));
  o->iworbddn = calloc((size_t)ndet * (ndn + 1), sizeof(int));
  int nu = 0, nd = 0;
  for (int id = 0; id < ndet; id++) {
    const int *du = o->det_up + id * nup;
    int f = -1;
    for (int u = 0; u < nu && f < 0; u++)
      if (!memcmp(o->iworbdup + u * nup, du, sizeof(int) * nup)) f = u;
    if (f < 0) {
      memcpy(o->iworbdup + nu * nup, du, sizeof(int) * nup);
      f = nu++;
    }
    o->iwdetup[id] = f;
    const int *dd = o->det_dn + id * ndn;
    f = -1;
    for (int u = 0; u < nd && f < 0; u++)
      if (!memcmp(o->iworbddn + u * ndn, dd, sizeof(int) * ndn)) f = u;
    if (f < 0) {
      memcpy(o->iworbddn + nd * ndn, dd, sizeof(int) * ndn);
      f = nd++;
    }
    o->iwdetdn[id] = f;
  }
  if (ndn == 0) nd = 1;
  o->ndetup = nu;
  o->ndetdn = nd;
  /* occupied orbitals (sorted union) */
  int *tmp = malloc(sizeof(int) * (o->norb + 1));
  int n = 0;
  for (int u = 0; u < nu; u++)
    for (int e = 0; e < nup; e++) {
      int orb = o->iworbdup[u * nup + e], f = 0;
      for (int q = 0; q < n; q++) if (tmp[q] == orb) f = 1;
      if (!f) tmp[n++] = orb;
    }
  qsort(tmp, n, sizeof(int), cmp_int);
  o->noccup = n;
  o->occup = DUPI(tmp, n);
  n = 0;
  for (int u = 0; u < o->ndetdn && ndn > 0; u++)
    for (int e = 0; e < ndn; e++) {
      int orb = o->iworbddn[u * ndn + e], f = 0;
      for (int q = 0; q < n; q++) if (tmp[q] == orb) f = 1;
      if (!f) tmp[n++] = orb;
    }
  qsort(tmp, n, sizeof(int), cmp_int);
  o->noccdn = n;
  o->occdn = DUPI(tmp, n);
  free(tmp);
  /* index maps */
  o->ioccup = malloc(sizeof(int) * o->norb);
  o->ioccdn = malloc(sizeof(int) * o->norb);
  for (int q = 0; q < o->norb; q++) o->ioccup[q] = o->ioccdn[q] = -1;
  for (int q = 0; q < o->noccup; q++) o->ioccup[o->occup[q]] = q;
  for (int q = 0; q < o->noccdn; q++) o->ioccdn[o->occdn[q]] = q;
  o->irefup = malloc(sizeof(int) * (nup + 1));
  o->irefdn = malloc(sizeof(int) * (ndn + 1));
  for (int k = 0; k < nup; k++) o->irefup[k] = o->ioccup[o->iworbdup[k]];
  for (int k = 0; k < ndn; k++) o->irefdn[k] = o->ioccdn[o->iworbddn[k]];
  o->ivirup = malloc(sizeof(int) * (o->noccup + 1));
  o->ivirdn = malloc(sizeof(int) * (o->noccdn + 1));
  {
    int nv = 0;
    for (int q = 0; q < o->noccup; q++) {
      int f = 0;
      for (int k = 0; k < nup; k++) if (o->irefup[k] == q) f = 1;
      if (!f) o->ivirup[nv++] = q;
    }
    nv = 0;
    for (int q = 0; q < o->noccdn; q++) {
      int f = 0;
      for (int k = 0; k < ndn; k++) if (o->irefdn[k] == q) f = 1;
      if (!f) o->ivirdn[nv++] = q;
    }
  }
  /* excitations relative to the first unique det of each spin */
  o->ordup = calloc(nu + 1, sizeof(int));
  o->orddn = calloc(o->ndetdn + 1, sizeof(int));
  o->sgnup = calloc(nu + 1, sizeof(int));
  o->sgndn = calloc(o->ndetdn + 1, sizeof(int));
  o->lab_excup = calloc((size_t)(nu + 1) * (nup + 1), sizeof(int));
  o->refup = calloc((size_t)(nu + 1) * (nup + 1), sizeof(int));
  o->excup = calloc((size_t)(nu + 1) * (nup + 1), sizeof(int));
  o->lab_excdn = calloc((size_t)(o->ndetdn + 1) * (ndn + 1), sizeof(int));
  o->refdn = calloc((size_t)(o->ndetdn + 1) * (ndn + 1), sizeof(int));
  o->excdn = calloc((size_t)(o->ndetdn + 1) * (ndn + 1), sizeof(int));
  int kmax = 0;
  int labref[256];
  for (int u = 0; u < nu; u++) {
    const int *ref = o->iworbdup, *det = o->iworbdup + u * nup;
    int ne = xorbs(ref, nup, det, nup, labref);
    xorbs(det, nup, ref, nup, o->lab_excup + u * nup);
    for (int j = 0; j < ne; j++) {
      for (int k = 0; k < nup; k++) if (ref[k] == labref[j]) o->refup[u * nup + j] = k;
      o->excup[u * nup + j] = o->ioccup[o->lab_excup[u * nup + j]];
    }
    o->sgnup[u] = 1 - 2 * (nswaps(det, ref, nup) % 2);
    o->ordup[u] = ne;
    if (ne > kmax) kmax = ne;
  }
  if (ndn > 0) {
    for (int u = 0; u < o->ndetdn; u++) {
      const int *ref = o->iworbddn, *det = o->iworbddn + u * ndn;
      int ne = xorbs(ref, ndn, det, ndn, labref);
      xorbs(det, ndn, ref, ndn, o->lab_excdn + u * ndn);
      for (int j = 0; j < ne; j++) {
        for (int k = 0; k < ndn; k++) if (ref[k] == labref[j]) o->refdn[u * ndn + j] = k;
        o->excdn[u * ndn + j] = o->ioccdn[o->lab_excdn[u * ndn + j]];
      }
      o->sgndn[u] = 1 - 2 * (nswaps(det, ref, ndn) % 2);
      o->orddn[u] = ne;
      if (ne > kmax) kmax = ne;
    }
  } else {
    o->sgndn[0] = 1;
    o->orddn[0] = 0;
  }
  o->kmax = kmax;
  o->cdet_in_wf = calloc(ndet, sizeof(double));
}

static void build_cdet(orc_sys *o) {
  memset(o->cdet_in_wf, 0, sizeof(double) * o->ndet);
  for (int icsf = 0; icsf < o->ncsf; icsf++) {
    double ccsf = o->csf_coef[icsf];
    for (int q = o->csf_det_off[icsf]; q < o->csf_det_off[icsf + 1]; q++)
      o->cdet_in_wf[o->csf_det[q]] += o->csf_det_coef[q] * ccsf;
  }
}

void orc_det_info(const orc_sys *o, int32_t *info) {
  info[0] = o->ndetup;
  info[1] = o->ndetdn;
  info[2] = o->noccup;
  info[3] = o->noccdn;
  info[4] = o->kmax;
}

/* qmc/vmc/rotqua.f90:396-428 (gesqua, octahedron), the 6-point rule:
   +-x, +-y, +-z with weights 1/6 */
static void gesqua6(double *xq, double *yq, double *zq, double *wq) {
  for (int j = 0; j < 6; j++) { xq[j] = yq[j] = zq[j] = 0; wq[j] = 1.0 / 6.0; }
  xq[0] = 1.0; xq[1] = -1.0;
  yq[2] = 1.0; yq[3] = -1.0;
  zq[4] = 1.0; zq[5] = -1.0;
}

/* qmc/vmc/rotqua.f90:300-500 (gesqua), 12-point icosahedral rule with the
   reference's value of pi (3.14159265359d0, rotqua.f90:325) */
static void gesqua12(double *xq, double *yq, double *zq, double *wq) {
  const double pi = 3.14159265359;
  for (int j = 0; j < 12; j++) { xq[j] = yq[j] = zq[j] = 0; wq[j] = 1.0 / 12.0; }
  zq[0] = 1.0;
  zq[1] = -1.0;
  double fi0 = pi / 5.0, s5 = sqrt(5.0);
  double cstha = 1.0 / s5, sntha = 2.0 / s5;
  for (int j = 1; j <= 5; j++) {
    double rk2 = 2.0 * (double)(j - 1);
    xq[j + 1] = sntha * cos(rk2 * fi0);
    yq[j + 1] = sntha * sin(rk2 * fi0);
    zq[j + 1] = cstha;
    double rk21 = rk2 + 1.0;
    xq[j + 6] = sntha * cos(rk21 * fi0);
    yq[j + 6] = sntha * sin(rk21 * fi0);
    zq[j + 6] = -cstha;
  }
}
void orc_quadrature(const orc_sys *o, double *xq, double *yq, double *zq,
                    double *wq) {
  for (int i = 0; i < o->nquad; i++) {
    xq[i] = o->xq0[i]; yq[i] = o->yq0[i]; zq[i] = o->zq0[i]; wq[i] = o->wq[i];
  }
}

orc_sys *orc_create(const qmcw_system *s) {
  if (s->basis_type != QMCW_BASIS_SLATER && s->basis_type != QMCW_BASIS_GAUSSIAN) {
    snprintf(g_err, sizeof g_err, "basis_type %d unsupported", s->basis_type);
    return NULL;
  }
  if (s->isc != 2 && s->isc != 4) {
    snprintf(g_err, sizeof g_err, "isc %d unsupported (2, 4)", s->isc);
    return NULL;
  }
  if (s->nctype > 64 || s->nordc > 7 || s->norda > QMCW_MAX_ORD ||
      s->nordb > QMCW_MAX_ORD) {
    snprintf(g_err, sizeof g_err, "system too large for the oracle");
    return NULL;
  }
  orc_sys *o = calloc(1, sizeof(orc_sys));
  o->nelec = s->nelec; o->nup = s->nup; o->ndn = s->nelec - s->nup;
  o->ncent = s->ncent; o->nctype = s->nctype;
  o->cent = DUPD(s->cent, 3 * s->ncent);
  o->iwctype = DUPI(s->iwctype, s->ncent);
  o->znuc = DUPD(s->znuc, s->nctype);
  o->pecent = 0;
  for (int i = 1; i < s->ncent; i++)
    for (int j = 0; j < i; j++) {
      double d2 = 0;
      for (int k = 0; k < 3; k++) {
        double d = o->cent[3 * i + k] - o->cent[3 * j + k];
        d2 += d * d;
      }
      o->pecent += o->znuc[o->iwctype[i]] * o->znuc[o->iwctype[j]] / sqrt(d2);
    }
  o->basis_type = s->basis_type; o->nrad = s->nrad;
  o->rad_n = DUPI(s->rad_n, s->nrad);
  o->rad_prim_off = DUPI(s->rad_prim_off, s->nrad + 1);
  int nprim = s->rad_prim_off[s->nrad];
  o->prim_zeta = DUPD(s->prim_zeta, nprim);
  o->prim_coef = DUPD(s->prim_coef, nprim);
  if (s->rad_num && s->nnum > 0) {
    o->rad_num = DUPI(s->rad_num, s->nrad);
    o->nnum = s->nnum;
    o->num_igrid = DUPI(s->num_igrid, s->nnum);
    o->num_nr = DUPI(s->num_nr, s->nnum);
    o->num_off = DUPI(s->num_off, s->nnum + 1);
    o->num_h = DUPD(s->num_h, s->nnum);
    o->num_r0 = DUPD(s->num_r0, s->nnum);
    o->num_rwf = DUPD(s->num_rwf, s->num_off[s->nnum]);
    o->num_d2 = DUPD(s->num_d2, s->num_off[s->nnum]);
    o->num_ce = DUPD(s->num_ce, 5 * s->nnum);
    o->num_ae = DUPD(s->num_ae, 2 * s->nnum);
  }
  o->ctype_bas_off = DUPI(s->ctype_bas_off, s->nctype + 1);
  int nbt = s->ctype_bas_off[s->nctype];
  o->bas_rad = DUPI(s->bas_rad, nbt);
  o->bas_l = DUPI(s->bas_l, nbt);
  o->bas_m = DUPI(s->bas_m, nbt);
  o->nbasis = s->nbasis; o->norb = s->norb;
  o->coef = DUPD(s->coef, (size_t)s->norb * s->nbasis);
  o->ndet = s->ndet; o->ncsf = s->ncsf;
  o->det_up = DUPI(s->det_up, s->ndet * s->nup);
  o->det_dn = DUPI(s->det_dn, s->ndet * (s->nelec - s->nup));
  o->csf_coef = DUPD(s->csf_coef, s->ncsf);
  o->csf_det_off = DUPI(s->csf_det_off, s->ncsf + 1);
  o->csf_det = DUPI(s->csf_det, s->csf_det_off[s->ncsf]);
  o->csf_det_coef = DUPD(s->csf_det_coef, s->csf_det_off[s->ncsf]);
  o->isc = s->isc; o->nspin2b = s->nspin2b; o->nocuspb = s->nocuspb;
  o->norda = s->norda; o->nordb = s->nordb; o->nordc = s->nordc;
  o->scalek = s->scalek;
  o->nterms = qmcw_nterms4(s->nordc);
  o->a4 = DUPD(s->a4, s->nctype * (s->norda + 1));
  o->b = DUPD(s->b, s->nspin2b * (s->nordb + 1));
  o->c = DUPD(s->c, s->nctype * o->nterms);
  o->nloc = s->nloc;
  o->nquad = s->nloc > 0 ? s->nquad : 0;
  if (s->nloc > 0) {
    if (s->nquad != 12 && s->nquad != 6) {
      snprintf(g_err, sizeof g_err, "only nquad = 6 or 12 is supported");
      free(o);
      return NULL;
    }
    for (int it = 0; it < s->nctype; it++) {
      o->npotd[it] = s->npotd[it];
      o->lpotp1[it] = s->lpotp1[it];
    }
    int ntermecp = s->ecp_off[s->nctype * QMCW_MAX_LPOT];
    o->ecp_off = DUPI(s->ecp_off, s->nctype * QMCW_MAX_LPOT + 1);
    o->ecp_coef = DUPD(s->ecp_coef, ntermecp);
    o->ecp_power = DUPI(s->ecp_power, ntermecp);
    o->ecp_expo = DUPD(s->ecp_expo, ntermecp);
    if (s->nquad == 6)
      gesqua6(o->xq0, o->yq0, o->zq0, o->wq);
    else
      gesqua12(o->xq0, o->yq0, o->zq0, o->wq);
  }
  o->npa = s->nparma_tot; o->npb = s->nparmb_tot; o->npc = s->nparmc_tot;
  o->pa_type = DUPI(s->jas_a_type, o->npa); o->pa_idx = DUPI(s->jas_a_idx, o->npa);
  o->pb_spin = DUPI(s->jas_b_spin, o->npb); o->pb_idx = DUPI(s->jas_b_idx, o->npb);
  o->pc_type = DUPI(s->jas_c_type, o->npc); o->pc_idx = DUPI(s->jas_c_idx, o->npc);
  for (int it = 0; it < s->nctype; it++) { o->nparmc_t[it] = 0; o->pc_first[it] = -1; }
  for (int p = 0; p < o->npc; p++) {
    int it = o->pc_type[p];
    if (o->pc_first[it] < 0) o->pc_first[it] = p;
    o->nparmc_t[it]++;
  }
  o->npcsf = s->nparmcsf; o->pcsf = DUPI(s->csf_parm, s->nparmcsf);
  o->nporb = s->nparmorb;
  o->porb_i = DUPI(s->orb_parm_i, s->nparmorb);
  o->porb_j = DUPI(s->orb_parm_j, s->nparmorb);
  o->porb_rev = DUPI(s->orb_parm_rev, s->nparmorb);
  o->npj = o->npa + o->npb + o->npc;
  o->nparm = o->npcsf + o->npj + o->nporb;
  o->hb = s->hb > 0 ? s->hb : 0.5;
  cuspinit4(o);
  set_scale_dist(o);
  setup_dets(o);
  build_cdet(o);
  return o;
}

void orc_set_params(orc_sys *o, const double *a4, const double *b,
                    const double *c, const double *csf_coef,
                    const double *coef) {
  if (a4) memcpy(o->a4, a4, sizeof(double) * o->nctype * (o->norda + 1));
  if (b) memcpy(o->b, b, sizeof(double) * o->nspin2b * (o->nordb + 1));
  if (c) memcpy(o->c, c, sizeof(double) * o->nctype * o->nterms);
  if (csf_coef) memcpy(o->csf_coef, csf_coef, sizeof(double) * o->ncsf);
  if (coef) memcpy(o->coef, coef, sizeof(double) * o->norb * o->nbasis);
  set_scale_dist(o);
  build_cdet(o);
}

void orc_destroy(orc_sys *o) {
  if (!o) return;
  void *ptrs[] = {o->cent, o->znuc, o->iwctype, o->rad_n, o->rad_prim_off,
                  o->ctype_bas_off, o->bas_rad, o->bas_l, o->bas_m,
                  o->prim_zeta, o->prim_coef, o->coef, o->det_up, o->det_dn,
                  o->csf_coef, o->csf_det_coef, o->csf_det_off, o->csf_det,
                  o->cdet_in_wf, o->iwdetup, o->iwdetdn, o->iworbdup,
                  o->iworbddn, o->occup, o->occdn, o->ioccup, o->ioccdn,
                  o->irefup, o->irefdn, o->ivirup, o->ivirdn, o->ordup,
                  o->orddn, o->sgnup, o->sgndn, o->lab_excup, o->lab_excdn,
                  o->refup, o->refdn, o->excup, o->excdn, o->a4, o->b, o->c,
                  o->ecp_off, o->ecp_power, o->ecp_coef, o->ecp_expo,
                  o->pa_type, o->pa_idx, o->pb_spin, o->pb_idx, o->pc_type,
                  o->pc_idx, o->pcsf, o->porb_i, o->porb_j, o->porb_rev,
                  o->rad_num, o->num_igrid, o->num_nr, o->num_off, o->num_h,
                  o->num_r0, o->num_rwf, o->num_d2, o->num_ce, o->num_ae};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  free(o);
}

/* ------------------------------------------------------------------ */
/* RNG                                                                 */
/* ------------------------------------------------------------------ */
/* qmc/vmc/rannyu.f90:11-21 (setrn) */
void orc_setrn(int32_t *l, const int32_t *seed) {
  for (int i = 0; i < 4; i++) l[i] = seed[i];
  l[3] = 2 * (l[3] / 2) + 1;
}
/* qmc/vmc/rannyu.f90:53-74: R <- 11^13 R mod 2^48 in 12-bit limbs */
double orc_rannyu(int32_t *l) {
  const int m1 = 502, m2 = 1521, m3 = 4071, m4 = 2107;
  const double two12i = 2.44140625e-4;
  int l1 = l[0], l2 = l[1], l3 = l[2], l4 = l[3];
  int i1 = l1 * m4 + l2 * m3 + l3 * m2 + l4 * m1;
  int i2 = l2 * m4 + l3 * m3 + l4 * m2;
  int i3 = l3 * m4 + l4 * m3;
  int i4 = l4 * m4;
  l4 = i4 % 4096;
  i3 = i3 + i4 / 4096;
  l3 = i3 % 4096;
  i2 = i2 + i3 / 4096;
  l2 = i2 % 4096;
  l1 = (i1 + i2 / 4096) % 4096;
  l[0] = l1; l[1] = l2; l[2] = l3; l[3] = l4;
  return two12i * ((double)l1 +
                   two12i * ((double)l2 +
                             two12i * ((double)l3 + two12i * ((double)l4))));
}
/* qmc/dmc/gauss.f90 (Box-Muller, cos branch; pi from const_mod is
   4*atan(1) in CHAMP's setup) */
double orc_gauss(int32_t *l) {
  const double pi = 3.141592653589793;
  double g = cos(2 * pi * orc_rannyu(l));
  return g * sqrt(-2 * log(orc_rannyu(l)));
}
/* qmc/vmc/rotqua.f90:621-682 */
void orc_rotqua(int32_t *st, const double *xq0, const double *yq0,
                const double *zq0, int32_t nq, double *xq, double *yq,
                double *zq) {
  double x1, x2, x3, xsum;
  do {
    x1 = 1.0 - 2.0 * orc_rannyu(st);
    x2 = 1.0 - 2.0 * orc_rannyu(st);
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
    u1 = orc_rannyu(st) * 2.0 - 1.0;
    u2 = orc_rannyu(st) * 2.0 - 1.0;
    usum = u1 * u1 + u2 * u2;
  } while (usum >= 1.0);
  double uu = sqrt(usum);
  u1 /= uu;
  u2 /= uu;
  double y1 = yy1 * u1 + zz1 * u2, y2 = yy2 * u1 + zz2 * u2, y3 = yy3 * u1 + zz3 * u2;
  double z1 = yy1 * u2 - zz1 * u1, z2 = yy2 * u2 - zz2 * u1, z3 = yy3 * u2 - zz3 * u1;
  for (int iq = 0; iq < nq; iq++) {
    xq[iq] = xq0[iq] * x1 + yq0[iq] * y1 + zq0[iq] * z1;
    yq[iq] = xq0[iq] * x2 + yq0[iq] * y2 + zq0[iq] * z2;
    zq[iq] = xq0[iq] * x3 + yq0[iq] * y3 + zq0[iq] * z3;
  }
}

/* ------------------------------------------------------------------ */
/* matinv: qmc/vmc/matinv.f90:1-129, pivoted Gauss-Jordan on a column-major
   n x n array; singular -> det = 0 and inverse = 0 (:60-69) */
double orc_matinv(double *a, int32_t n) {
  if (n == 0) return 1.0;
  int ipivot[512];
  double atemp[512];
  for (int i = 1; i <= n; i++) ipivot[i - 1] = i;
  double determ = 1.0;
  int iclm = -n;
#define A(idx) a[(idx)-1]
  for (int i = 1; i <= n; i++) {
    iclm += n;
    double adiag = A(ipivot[i - 1] + iclm);
    int idiag = i;
    for (int k = i; k <= n; k++)
      if (fabs(A(ipivot[k - 1] + iclm)) > fabs(adiag)) {
        adiag = A(ipivot[k - 1] + iclm);
        idiag = k;
      }
    if (idiag != i) {
      determ = -determ;
      int itemp = ipivot[i - 1];
      ipivot[i - 1] = ipivot[idiag - 1];
      ipivot[idiag - 1] = itemp;
    }
    determ = adiag * determ;
    A(ipivot[i - 1] + iclm) = 1.0;
    if (adiag == 0.0) {
      for (int q = 0; q < n * n; q++) a[q] = 0;
      return 0.0;
    }
    double adiagi = 1.0 / adiag;
    /* scal(n, adiagi, a(ipivot(i)), n) */
    for (int q = 0; q < n; q++) A(ipivot[i - 1] + q * n) *= adiagi;
    for (int j = 1; j <= n; j++)
      if (j != ipivot[i - 1]) {
        double t = -A(j + iclm);
        A(j + iclm) = 0.0;
        /* axpy(n, t, a(ipivot(i)), n, a(j), n) */
        for (int q = 0; q < n; q++) A(j + q * n) += t * A(ipivot[i - 1] + q * n);
      }
  }
  int jn = -n;
  for (int j = 1; j <= n; j++) {
    jn += n;
    for (int i = 1; i <= n; i++) atemp[i - 1] = A(i + jn);
    for (int i = 1; i <= n; i++) A(i + jn) = atemp[ipivot[i - 1] - 1];
  }
  for (int j = 1; j <= n; j++) ipivot[j - 1] = (ipivot[j - 1] - 1) * n;
  for (int i = 1; i <= n; i++) {
    jn = -n;
    for (int j = 1; j <= n; j++) {
      jn += n;
      atemp[j - 1] = A(i + jn);
    }
    for (int j = 1; j <= n; j++) A(i + ipivot[j - 1]) = atemp[j - 1];
  }
#undef A
  return determ;
}

/* ------------------------------------------------------------------ */
/* basis functions: qmc/vmc/basis_fns.f90:38-350 (explicit Y_lm to l=4)  */
/* ------------------------------------------------------------------ */
typedef struct { double th[5][5], dth[5][5][3], ph[9], dph[9][2]; } ylm_tab;
#define PH(m) ph[(m) + 4]
#define DPH(m, k) dph[(m) + 4][k]

static void ylm_fill(double xx, double yy, double zz, double r, ylm_tab *T) {
  double ri = 1 / r, ri2 = ri * ri, ri3 = ri2 * ri, ri4 = ri3 * ri, ri5 = ri4 * ri;
  double r2 = r * r, xx2 = xx * xx, yy2 = yy * yy, zz2 = zz * zz;
  double xhat = xx * ri, yhat = yy * ri, zhat = zz * ri;
  const double cd0 = 0.5, cd1 = sqrt(3.0), cd2 = sqrt(3.0) / 2;
  const double cf0 = 0.5, cf1 = sqrt(3.0 / 8.0), cf2 = sqrt(15.0) / 2, cf3 = sqrt(5.0 / 8.0);
  const double cg0 = 0.125, cg1 = sqrt(10.0) / 4, cg2 = sqrt(5.0) / 4,
               cg3 = sqrt(70.0) / 4, cg4 = sqrt(35.0) / 8;
  double *ph = T->ph;
  double(*dph)[2] = T->dph;
  PH(0) = 1; DPH(0, 0) = 0; DPH(0, 1) = 0;
  PH(1) = xx; DPH(1, 0) = 1; DPH(1, 1) = 0;
  PH(-1) = yy; DPH(-1, 0) = 0; DPH(-1, 1) = 1;
  PH(2) = xx2 - yy2; DPH(2, 0) = 2 * xx; DPH(2, 1) = -2 * yy;
  PH(-2) = 2 * xx * yy; DPH(-2, 0) = 2 * yy; DPH(-2, 1) = 2 * xx;
  PH(3) = PH(2) * PH(1) - PH(-2) * PH(-1); DPH(3, 0) = 3 * PH(2); DPH(3, 1) = -3 * PH(-2);
  PH(-3) = PH(-2) * PH(1) + PH(-1) * PH(2); DPH(-3, 0) = 3 * PH(-2); DPH(-3, 1) = 3 * PH(2);
  PH(4) = xx2 * xx2 - 6 * xx2 * yy2 + yy2 * yy2;
  DPH(4, 0) = 4 * xx * xx2 - 12 * xx * yy2; DPH(4, 1) = -12 * xx2 * yy + 4 * yy * yy2;
  PH(-4) = 4 * xx * yy * (xx2 - yy2);
  DPH(-4, 0) = 12 * xx2 * yy - 4 * yy * yy2; DPH(-4, 1) = 4 * xx * xx2 - 12 * xx * yy2;
  double(*th)[5] = T->th;
  double(*dth)[5][3] = T->dth;
  th[0][0] = 1; dth[0][0][0] = dth[0][0][1] = dth[0][0][2] = 0;
  th[1][0] = ri * zz;
  dth[1][0][0] = -ri3 * xx * zz; dth[1][0][1] = -ri3 * yy * zz; dth[1][0][2] = ri3 * (xx2 + yy2);
  th[1][1] = ri;
  dth[1][1][0] = -ri3 * xx; dth[1][1][1] = -ri3 * yy; dth[1][1][2] = -ri3 * zz;
  th[2][0] = cd0 * (3 * zhat * zhat - 1);
  dth[2][0][0] = -6 * cd0 * ri * zhat * zhat * xhat;
  dth[2][0][1] = -6 * cd0 * ri * zhat * zhat * yhat;
  dth[2][0][2] = 6 * cd0 * ri4 * (xx2 + yy2) * zz;
  th[2][1] = cd1 * ri2 * zz;
  dth[2][1][0] = -2 * cd1 * ri4 * zz * xx; dth[2][1][1] = -2 * cd1 * ri4 * zz * yy;
  dth[2][1][2] = cd1 * ri4 * (xx2 + yy2 - zz2);
  th[2][2] = cd2 * ri2;
  dth[2][2][0] = -2 * cd2 * ri4 * xx; dth[2][2][1] = -2 * cd2 * ri4 * yy; dth[2][2][2] = -2 * cd2 * ri4 * zz;
  th[3][0] = cf0 * ri3 * zz * (2 * zz2 - 3 * (xx2 + yy2));
  dth[3][0][0] = 3 * cf0 * ri5 * (xx2 + yy2 - 4 * zz2) * zz * xx;
  dth[3][0][1] = 3 * cf0 * ri5 * (xx2 + yy2 - 4 * zz2) * zz * yy;
  dth[3][0][2] = -3 * cf0 * ri5 * ((xx2 + yy2) * (xx2 + yy2) - 4 * zz2 * (xx2 + yy2));
  th[3][1] = cf1 * ri3 * (4 * zz2 - (xx2 + yy2));
  dth[3][1][0] = cf1 * ri5 * (xx2 + yy2 - 14 * zz2) * xx;
  dth[3][1][1] = cf1 * ri5 * (xx2 + yy2 - 14 * zz2) * yy;
  dth[3][1][2] = cf1 * ri5 * (11 * (xx2 + yy2) - 4 * zz2) * zz;
  th[3][2] = cf2 * ri3 * zz;
  dth[3][2][0] = -3 * cf2 * ri5 * zz * xx; dth[3][2][1] = -3 * cf2 * ri5 * zz * yy;
  dth[3][2][2] = cf2 * ri5 * (xx2 + yy2 - 2 * zz2);
  th[3][3] = cf3 * ri3;
  dth[3][3][0] = -3 * cf3 * ri5 * xx; dth[3][3][1] = -3 * cf3 * ri5 * yy; dth[3][3][2] = -3 * cf3 * ri5 * zz;
  th[4][0] = cg0 * (35 * zz2 * zz2 - 30 * r2 * zz2 + 3 * r2 * r2) * ri4;
  dth[4][0][0] = cg0 * (-60 * xx * zz2 + 12 * r2 * xx) * ri4 + th[4][0] * (-4 * ri2 * xx);
  dth[4][0][1] = cg0 * (-60 * yy * zz2 + 12 * r2 * yy) * ri4 + th[4][0] * (-4 * ri2 * yy);
  dth[4][0][2] = cg0 * (140 * zz2 * zz - (60 * zz * zz2 + 60 * r2 * zz) + 12 * r2 * zz) * ri4 +
                 th[4][0] * (-4 * ri2 * zz);
  th[4][1] = cg1 * zz * (7 * zz2 - 3 * r2) * ri4;
  dth[4][1][0] = cg1 * zz * (-6 * xx) * ri4 + th[4][1] * (-4 * ri2 * xx);
  dth[4][1][1] = cg1 * zz * (-6 * yy) * ri4 + th[4][1] * (-4 * ri2 * yy);
  dth[4][1][2] = cg1 * ((7 * zz2 - 3 * r2) + zz * (14 * zz - 6 * zz)) * ri4 + th[4][1] * (-4 * ri2 * zz);
  th[4][2] = cg2 * (7 * zz2 - r2) * ri4;
  dth[4][2][0] = cg2 * (-2 * xx) * ri4 + th[4][2] * (-4 * ri2 * xx);
  dth[4][2][1] = cg2 * (-2 * yy) * ri4 + th[4][2] * (-4 * ri2 * yy);
  dth[4][2][2] = cg2 * (14 * zz - 2 * zz) * ri4 + th[4][2] * (-4 * ri2 * zz);
  th[4][3] = cg3 * zz * ri4;
  dth[4][3][0] = th[4][3] * (-4 * ri2 * xx); dth[4][3][1] = th[4][3] * (-4 * ri2 * yy);
  dth[4][3][2] = cg3 * ri4 + th[4][3] * (-4 * ri2 * zz);
  th[4][4] = cg4 * ri4;
  dth[4][4][0] = th[4][4] * (-4 * ri2 * xx); dth[4][4][1] = th[4][4] * (-4 * ri2 * yy);
  dth[4][4][2] = th[4][4] * (-4 * ri2 * zz);
}

/* tabulated radial function k at r: value and first two derivatives
   (qmc/vmc/splfit.f90:13-76).  The grid index jx = int(xr) decides the
   piece: jx < 1 the small-r polynomial ce, jx >= nr the exponential tail
   ae0 exp(-ae1 r), else the cubic spline on the 1-based points jx, jx+1. */
static void num_radial(const orc_sys *o, int k, double r, double *w) {
  const int igrid = o->num_igrid[k], nr = o->num_nr[k];
  const double hb = o->num_h[k], r0 = o->num_r0[k];
  double xr;
  if (igrid == 1)
    xr = (r - r0) / hb + 1;
  else if (igrid == 2)
    xr = log(r / r0) / log(hb) + 1;
  else
    xr = log((r + r0) / r0) / log(hb) + 1.0;
  if (!(xr >= 1.0)) {
    const double *ce = o->num_ce + 5 * k;
    w[0] = ce[0];
    w[1] = w[2] = 0;
    for (int i = 2; i <= 5; i++) {
      w[0] += ce[i - 1] * pow(r, i - 1);
      w[1] += (i - 1) * ce[i - 1] * pow(r, i - 2);
      w[2] += (i - 1) * (i - 2) * ce[i - 1] * pow(r, i - 3);
    }
    return;
  }
  if (xr >= nr) {
    const double *ae = o->num_ae + 2 * k;
    w[0] = ae[0] * exp(-ae[1] * r);
    w[1] = -ae[1] * w[0];
    w[2] = -ae[1] * w[1];
    return;
  }
  const int jx = (int)xr;
  double ref0, ref1, delh;
  if (igrid == 1) {
    ref0 = r0 + hb * (jx - 1);
    ref1 = ref0 + hb;
    delh = hb;
  } else if (igrid == 2) {
    ref0 = r0 * pow(hb, jx - 1);
    ref1 = ref0 * hb;
    delh = ref1 - ref0;
  } else {
    ref0 = r0 * pow(hb, jx - 1) - r0;
    ref1 = (ref0 + r0) * hb - r0;
    delh = ref1 - ref0;
  }
  const double *y = o->num_rwf + o->num_off[k], *y2 = o->num_d2 + o->num_off[k];
  const double bb = (r - ref0) / delh, aa = (ref1 - r) / delh;
  const double cc = aa * (aa * aa - 1) * delh * delh / 6, dd = bb * (bb * bb - 1) * delh * delh / 6;
  w[0] = aa * y[jx - 1] + bb * y[jx] + cc * y2[jx - 1] + dd * y2[jx];
  w[1] = (y[jx] - y[jx - 1]) / delh +
         (-(3 * aa * aa - 1) * y2[jx - 1] + (3 * bb * bb - 1) * y2[jx]) * delh / 6;
  w[2] = aa * y2[jx - 1] + bb * y2[jx];
}

void orc_num_radial(const orc_sys *o, int32_t k, int32_t n, const double *r, double *out) {
  for (int i = 0; i < n; i++) num_radial(o, k, r[i], out + 3 * i);
}

/* radial value and first two derivatives: basis_fns.f90:118-157, summed
   over the primitives of a contracted function; a tabulated function
   through num_radial (basis_fns.f90:161-165) */
static void radial(const orc_sys *o, int irb, double r, double *w) {
  if (o->rad_num && o->rad_num[irb] >= 0) {
    num_radial(o, o->rad_num[irb], r, w);
    return;
  }
  int n = o->rad_n[irb];
  double rn = abs(n);
  double rm3 = pow(r, rn - 3), rm2 = rm3 * r, rm1 = rm2 * r;
  w[0] = w[1] = w[2] = 0;
  for (int p = o->rad_prim_off[irb]; p < o->rad_prim_off[irb + 1]; p++) {
    double zeta = o->prim_zeta[p], d = o->prim_coef[p], zr, ex, w3;
    if (o->basis_type == QMCW_BASIS_SLATER) {
      zr = zeta * r;
      ex = exp(-zr);
      w3 = rm3 * ((rn - 1) * (rn - 2 - zr) - (rn - 1 - zr) * zr) * ex;
    } else {
      zr = 2 * zeta * r * r;
      ex = exp(-0.5 * zr);
      w3 = rm3 * ((rn - 1) * (rn - 2) - (2 * rn - 1 - zr) * zr) * ex;
    }
    w[0] += d * rm1 * ex;
    w[1] += d * rm2 * ((rn - 1) - zr) * ex;
    w[2] += d * w3;
  }
}

/* Real solid harmonic S_lm(x, y, z) = r^l Y_lm with <Y_lm^2> = 4 pi/(2l+1)
   and its gradient, from the closed form (Helgaker, Jorgensen and Olsen,
   Molecular Electronic-Structure Theory, eqs. 6.4.47-6.4.50):
   S_lm = N_m Pi_l^|m|(z, r^2) {Re, Im}(x + i y)^|m| (Re for m >= 0),
   Pi_l^m = sum_k (-1)^k 2^-l C(l,k) C(2l-2k,l) (l-2k)!/(l-2k-m)! r^2k z^(l-2k-m),
   N_0 = 1, N_m = sqrt(2 (l-m)!/(l+m)!).  For l >= 5 CHAMP generates these
   by recursion (qmc/vmc/basis_fns.f90:1962-2444); for l <= 4 it has the
   explicit forms of ylm_fill, which agree with both. */
static double fact_d(int n) { double f = 1; for (int i = 2; i <= n; i++) f *= i; return f; }
static double binom_d(int n, int k) { return fact_d(n) / (fact_d(k) * fact_d(n - k)); }
static void solid_harmonic(int l, int m, double x, double y, double z, double *S, double *g) {
  const int am = abs(m);
  const double rho = x * x + y * y + z * z;
  /* Pi and its partials in z and rho */
  double pi = 0, piz = 0, pir = 0;
  for (int k = 0; 2 * k <= l - am; k++) {
    const int e = l - 2 * k - am;
    const double c = ((k & 1) ? -1.0 : 1.0) * binom_d(l, k) * binom_d(2 * l - 2 * k, l) *
                     fact_d(l - 2 * k) / (fact_d(e) * pow(2.0, l));
    pi += c * pow(rho, k) * pow(z, e);
    if (e > 0) piz += c * pow(rho, k) * e * pow(z, e - 1);
    if (k > 0) pir += c * k * pow(rho, k - 1) * pow(z, e);
  }
  /* (x + i y)^am and (x + i y)^(am-1) */
  double re = 1, im = 0, re1 = 0, im1 = 0;
  for (int q = 0; q < am; q++) {
    re1 = re; im1 = im;
    const double t = re * x - im * y;
    im = re * y + im * x;
    re = t;
  }
  double a, ax, ay;
  if (m >= 0) { a = re; ax = am * re1; ay = -am * im1; }
  else { a = im; ax = am * im1; ay = am * re1; }
  const double nm = am == 0 ? 1.0 : sqrt(2.0 * fact_d(l - am) / fact_d(l + am));
  *S = nm * pi * a;
  g[0] = nm * (2 * x * pir * a + pi * ax);
  g[1] = nm * (2 * y * pir * a + pi * ay);
  g[2] = nm * ((piz + 2 * z * pir) * a);
}

void orc_solid_harmonic(int32_t l, int32_t m, const double *x, double *out) {
  solid_harmonic(l, m, x[0], x[1], x[2], out, out + 1);
}

static void basis_at(const orc_sys *o, const double *rvec, const double *r_en,
                     double *phi, double *dphi, double *d2phi) {
  int ib = 0;
  double wfv[256][3];
  ylm_tab T;
  for (int ic = 0; ic < o->ncent; ic++) {
    int ict = o->iwctype[ic];
    double xc[3] = {rvec[3 * ic], rvec[3 * ic + 1], rvec[3 * ic + 2]};
    double r = r_en[ic];
    if (r == 0.0) { r = 1e-99; xc[0] = 1e-99; } /* basis_fns.f90:101-105 */
    double ri = 1 / r, ri2 = ri * ri;
    int b0 = o->ctype_bas_off[ict], b1 = o->ctype_bas_off[ict + 1];
    /* radial functions needed by this type */
    for (int q = b0; q < b1; q++) radial(o, o->bas_rad[q], r, wfv[q - b0]);
    ylm_fill(xc[0], xc[1], xc[2], r, &T);
    double xhat = xc[0] * ri, yhat = xc[1] * ri, zhat = xc[2] * ri;
    for (int q = b0; q < b1; q++, ib++) {
      int l = o->bas_l[q], m = o->bas_m[q], mabs = abs(m);
      const double *w = wfv[q - b0];
      if (l >= 5) {
        /* Y = S / r^l, grad Y = grad S / r^l - l S x / r^(l+2) */
        double Sv, gS[3];
        solid_harmonic(l, m, xc[0], xc[1], xc[2], &Sv, gS);
        const double ril = pow(ri, l), ylm = Sv * ril;
        phi[ib] = ylm * w[0];
        if (dphi) {
          const double xh[3] = {xhat, yhat, zhat};
          for (int k = 0; k < 3; k++)
            dphi[3 * ib + k] = ylm * xh[k] * w[1] + (gS[k] * ril - l * ylm * ri * xh[k]) * w[0];
          d2phi[ib] = ylm * (w[2] + 2 * ri * w[1] - l * (l + 1) * ri2 * w[0]);
        }
        continue;
      }
      double th = T.th[l][mabs], ph = T.ph[m + 4];
      double ylm = th * ph;
      phi[ib] = ylm * w[0];
      if (dphi) {
        dphi[3 * ib + 0] = ylm * xhat * w[1] + (th * T.dph[m + 4][0] + T.dth[l][mabs][0] * ph) * w[0];
        dphi[3 * ib + 1] = ylm * yhat * w[1] + (th * T.dph[m + 4][1] + T.dth[l][mabs][1] * ph) * w[0];
        dphi[3 * ib + 2] = ylm * zhat * w[1] + (T.dth[l][mabs][2] * ph) * w[0];
        d2phi[ib] = ylm * (w[2] + 2 * ri * w[1] - l * (l + 1) * ri2 * w[0]);
      }
    }
  }
}

void orc_basis(const orc_sys *o, const double *xe, double *phi, double *dphi,
               double *d2phi) {
  double rvec[3 * 256], r[256];
  for (int ic = 0; ic < o->ncent; ic++) {
    double s = 0;
    for (int k = 0; k < 3; k++) {
      rvec[3 * ic + k] = xe[k] - o->cent[3 * ic + k];
      s += rvec[3 * ic + k] * rvec[3 * ic + k];
    }
    r[ic] = sqrt(s);
  }
  basis_at(o, rvec, r, phi, dphi, d2phi);
}

/* bfloat16 rounding (nearest, ties to even) of a double through FP32 */
static double bf16_round(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

void orc_set_emulation(orc_sys *o, int32_t mode) { o->emul = mode; }

/* qmc/vmc/orbitals_loc_ana.f90:50-64 / _grade.f90:40-51 / _anae :76-102 */
static void orbitals_at(const orc_sys *o, const double *rvec, const double *r,
                        double *orb, double *dorb, double *ddorb,
                        double *work) {
  int nb = o->nbasis;
  double *phi = work, *dphi = work + nb, *d2phi = work + 4 * nb;
  basis_at(o, rvec, r, phi, dorb ? dphi : NULL, d2phi);
  const int rnd = dorb ? (o->emul & 1) : (o->emul & 2);
  if (rnd) {
    for (int m = 0; m < nb; m++) {
      phi[m] = bf16_round(phi[m]);
      if (dorb) {
        for (int k = 0; k < 3; k++) dphi[3 * m + k] = bf16_round(dphi[3 * m + k]);
        d2phi[m] = bf16_round(d2phi[m]);
      }
    }
  }
  for (int io = 0; io < o->norb; io++) {
    const double *cf = o->coef + (size_t)io * nb;
    double v = 0, g0 = 0, g1 = 0, g2 = 0, l = 0;
    for (int m = 0; m < nb; m++) {
      const double c = rnd ? bf16_round(cf[m]) : cf[m];
      v += c * phi[m];
      if (dorb) {
        g0 += c * dphi[3 * m];
        g1 += c * dphi[3 * m + 1];
        g2 += c * dphi[3 * m + 2];
        l += c * d2phi[m];
      }
    }
    orb[io] = v;
    if (dorb) {
      dorb[3 * io] = g0; dorb[3 * io + 1] = g1; dorb[3 * io + 2] = g2;
      ddorb[io] = l;
    }
  }
}
void orc_orbitals(const orc_sys *o, const double *xe, double *orb,
                  double *dorb, double *ddorb) {
  double rvec[3 * 256], r[256];
  for (int ic = 0; ic < o->ncent; ic++) {
    double s = 0;
    for (int k = 0; k < 3; k++) {
      rvec[3 * ic + k] = xe[k] - o->cent[3 * ic + k];
      s += rvec[3 * ic + k] * rvec[3 * ic + k];
    }
    r[ic] = sqrt(s);
  }
  double *work = malloc(sizeof(double) * 5 * o->nbasis);
  orbitals_at(o, rvec, r, orb, dorb, ddorb, work);
  free(work);
}

/* ------------------------------------------------------------------ */
/* Jastrow, ijas = 4                                                   */
/* ------------------------------------------------------------------ */
/* Output of one pair (i > j) or one e-n diagonal term.  Parameter
   derivative slots are indexed by the Jastrow parameter list. */
typedef struct {
  double f;          /* fso(i,j) */
  double gi[3], gj[3]; /* fijo(:,i,j), fijo(:,j,i) */
  double d2;         /* d2ijo(i,j) */
  double li, lj;     /* lapjijo(i,j), lapjijo(j,i) */
} pair_out;

typedef struct {
  int mode;          /* 0: none, 1: values go(p), 2: values + grad + d2 */
  double *gv;        /* [npj] accumulate go */
  double *gi, *gj;   /* [npj][3] gradient wrt electron i, j (mode 2) */
  double *d2;        /* [npj] (mode 2) */
} pderiv;

static inline void pd_add(pderiv *P, int p, double c, double gp, const double *vi,
                          const double *vj, double d2) {
  if (!P || P->mode == 0) return;
  P->gv[p] += c * gp;
  if (P->mode == 2) {
    for (int k = 0; k < 3; k++) {
      P->gi[3 * p + k] += c * vi[k];
      if (vj) P->gj[3 * p + k] += c * vj[k];
    }
    P->d2[p] += c * d2;
  }
}

/* e-e and e-e-n terms of one pair: qmc/vmc/jastrow4.f90:56-280 (value,
   gradient, laplacian), parameter derivatives as qmc/vmc/deriv_jastrow4.f90
   :95-330 (b) and :330-720 (c, incl. dependent c via cdep). */
static void jas_pair(const orc_sys *o, int i, int j, double rij,
                     const double *rvec_ij, const double *ri_v,
                     const double *rj_v, const double *rveci,
                     const double *rvecj, pair_out *out, pderiv *P) {
  double uu[QMCW_MAX_ORD + 4], ss[QMCW_MAX_ORD + 4], tt[QMCW_MAX_ORD + 4],
      rri[QMCW_MAX_ORD + 4], rrj[QMCW_MAX_ORD + 4];
  /* index offset 3 so that uu[-3..-1] = 0 */
#define U(k) uu[(k) + 3]
#define S(k) ss[(k) + 3]
#define TT(k) tt[(k) + 3]
#define RI(k) rri[(k) + 3]
#define RJ(k) rrj[(k) + 3]
  for (int q = -3; q <= -1; q++) { U(q) = 0; S(q) = 0; TT(q) = 0; RI(q) = 0; RJ(q) = 0; }
  U(0) = 1; S(0) = 2; TT(0) = 1; RI(0) = 1; RJ(0) = 1;
  memset(out, 0, sizeof(*out));
  double sspinn = 1;
  int isb = 0, ipar = 0;
  if (i < o->nup || j >= o->nup) {
    if (o->nspin2b == 2) isb = 1;
    else if (o->nocuspb == 0) sspinn = 0.5;
    ipar = 1;
  }
  double dd1, dd2;
  scale_dist2(o, rij, &U(1), &dd1, &dd2);
  const double *b = o->b + isb * (o->nordb + 1);
  {
    double top = sspinn * b[0] * U(1), topu = sspinn * b[0];
    double bot = 1 + b[1] * U(1), botu = b[1], bot2 = bot * bot;
    double fee = top / bot, feeu = topu / bot2, feeuu = -2 * feeu * botu / bot;
    fee -= o->asymp_jasb[ipar];
    for (int iord = 2; iord <= o->nordb; iord++) {
      U(iord) = U(1) * U(iord - 1);
      fee += b[iord] * U(iord);
      feeu += b[iord] * iord * U(iord - 1);
      feeuu += b[iord] * iord * (iord - 1) * U(iord - 2);
    }
    feeuu = feeuu * dd1 * dd1 + feeu * dd2;
    feeu = feeu * dd1 / rij;
    out->f += fee;
    for (int k = 0; k < 3; k++) {
      out->gi[k] += feeu * rvec_ij[k];
      out->gj[k] -= feeu * rvec_ij[k];
    }
    out->d2 += 2. * (feeuu + 2 * feeu);
    out->li += feeuu + 2 * feeu;
    out->lj += feeuu + 2 * feeu;
    if (P && P->mode) {
      double ar = o->asymp_r;
      for (int p = 0; p < o->npb; p++) {
        if (o->pb_spin[p] != isb) continue;
        int iw = o->pb_idx[p];
        double gee, geeu, geeuu;
        if (iw == 1) {
          double t = sspinn * U(1), tu = sspinn;
          double bt = 1 + b[1] * U(1), btu = b[1];
          double bta = 1 + b[1] * ar;
          gee = t / bt - sspinn * ar / bta;
          geeu = tu / (bt * bt);
          geeuu = -2 * btu * geeu / bt;
        } else if (iw == 2) {
          double t = -sspinn * b[0] * U(1) * U(1), tu = -sspinn * 2 * b[0] * U(1),
                 tuu = -sspinn * 2 * b[0];
          double bot0 = 1 + b[1] * U(1), bt = bot0 * bot0, bt2 = bt * bt;
          double bta = 1 + b[1] * ar;
          gee = t / bt + sspinn * b[0] * ar * ar / (bta * bta);
          geeu = tu / (bt * bot0);
          geeuu = tuu * (1 - 2 * b[1] * U(1)) / bt2;
        } else {
          int iord = iw - 1;
          gee = U(iord) - pow(ar, iord);
          geeu = iord * U(iord - 1);
          geeuu = iord * (iord - 1) * U(iord - 2);
        }
        geeuu = geeuu * dd1 * dd1 + geeu * dd2;
        geeu = geeu * dd1 / rij;
        double vi[3], vj[3];
        for (int k = 0; k < 3; k++) { vi[k] = geeu * rvec_ij[k]; vj[k] = -vi[k]; }
        pd_add(P, o->npa + p, 1.0, gee, vi, vj, 2. * (geeuu + 2 * geeu));
      }
    }
  }
  if (o->nordc <= 1) return;
  scale_dist2(o, rij, &U(1), &dd1, &dd2);
  switch_scale2(o, &U(1), &dd1, &dd2);
  for (int iord = 2; iord <= o->nordc; iord++) U(iord) = U(1) * U(iord - 1);
  for (int ic = 0; ic < o->ncent; ic++) {
    int it = o->iwctype[ic];
    double ri = ri_v[ic], rj = rj_v[ic];
    double dd7, dd9, dd8, dd10;
    scale_dist2(o, ri, &RI(1), &dd7, &dd9);
    scale_dist2(o, rj, &RJ(1), &dd8, &dd10);
    switch_scale2(o, &RI(1), &dd7, &dd9);
    switch_scale2(o, &RJ(1), &dd8, &dd10);
    double s = ri + rj, t = ri - rj;
    double u2pst = rij * rij + s * t, u2mst = rij * rij - s * t;
    for (int iord = 1; iord <= o->nordc; iord++) {
      RI(iord) = RI(1) * RI(iord - 1);
      RJ(iord) = RJ(1) * RJ(iord - 1);
      S(iord) = RI(iord) + RJ(iord);
      TT(iord) = RI(iord) * RJ(iord);
    }
    const double *c = o->c + it * o->nterms;
    double fc = 0, fu = 0, fuu = 0, fi = 0, fii = 0, fj = 0, fjj = 0, fui = 0, fuj = 0;
    int ll = 0;
    for (int n = 2; n <= o->nordc; n++)
      for (int k = n - 1; k >= 0; k--) {
        int l_hi = (k == 0) ? n - k - 2 : n - k;
        for (int l = l_hi; l >= 0; l--) {
          int m = (n - k - l) / 2;
          if (2 * m != n - k - l) continue;
          ll++;
          double p1 = U(k), p1u = k * U(k - 1), p1uu = k * (k - 1) * U(k - 2);
          double p2 = S(l) * TT(m);
          double p2i = (l + m) * RI(l + m - 1) * RJ(m) + m * RI(m - 1) * RJ(l + m);
          double p2ii = (l + m) * (l + m - 1) * RI(l + m - 2) * RJ(m) +
                        m * (m - 1) * RI(m - 2) * RJ(l + m);
          double p2j = (l + m) * RJ(l + m - 1) * RI(m) + m * RJ(m - 1) * RI(l + m);
          double p2jj = (l + m) * (l + m - 1) * RJ(l + m - 2) * RI(m) +
                        m * (m - 1) * RJ(m - 2) * RI(l + m);
          double cl = c[ll - 1];
          fc += cl * p1 * p2;
          fu += cl * p1u * p2;
          fuu += cl * p1uu * p2;
          fi += cl * p1 * p2i;
          fii += cl * p1 * p2ii;
          fj += cl * p1 * p2j;
          fjj += cl * p1 * p2jj;
          fui += cl * p1u * p2i;
          fuj += cl * p1u * p2j;
          if (P && P->mode && o->nparmc_t[it] + 0 >= 0) {
            /* which parameters does term ll feed? */
            int direct = -1;
            for (int q = 0; q < o->nparmc_t[it]; q++)
              if (o->pc_idx[o->pc_first[it] + q] == ll) direct = q;
            int dep_eq = -1;
            if (direct < 0)
              for (int id = 0; id < o->neqs; id++)
                if (ll == o->iwc4[id] && o->nvdepend[id][it] > 0) dep_eq = id;
            if (direct >= 0 || dep_eq >= 0) {
              double gp = p1 * p2, gu = p1u * p2, guu = p1uu * p2;
              double gi = p1 * p2i, gii = p1 * p2ii, gj = p1 * p2j, gjj = p1 * p2jj;
              double gui = p1u * p2i, guj = p1u * p2j;
              guu = guu * dd1 * dd1 + gu * dd2;
              gu = gu * dd1 / rij;
              gui = gui * dd1 * dd7;
              guj = guj * dd1 * dd8;
              gii = gii * dd7 * dd7 + gi * dd9;
              gjj = gjj * dd8 * dd8 + gj * dd10;
              gi = gi * dd7 / ri;
              gj = gj * dd8 / rj;
              double vi[3], vj[3];
              for (int kk = 0; kk < 3; kk++) {
                vi[kk] = gi * rveci[3 * ic + kk] + gu * rvec_ij[kk];
                vj[kk] = gj * rvecj[3 * ic + kk] - gu * rvec_ij[kk];
              }
              double d2 = 2 * (2 * gu + gi + gj) + 2 * guu + gii + gjj +
                          gui * u2pst / (ri * rij) + guj * u2mst / (rj * rij);
              if (direct >= 0) {
                pd_add(P, o->npa + o->npb + o->pc_first[it] + direct, 1.0, gp, vi, vj, d2);
              } else {
                for (int id = 0; id < o->nvdepend[dep_eq][it]; id++) {
                  int q = o->iwdepend[dep_eq][id][it];
                  pd_add(P, o->npa + o->npb + o->pc_first[it] + q,
                         o->cdep[dep_eq][id][it], gp, vi, vj, d2);
                }
              }
            }
          }
        }
      }
    fuu = fuu * dd1 * dd1 + fu * dd2;
    fu = fu * dd1 / rij;
    fui = fui * dd1 * dd7;
    fuj = fuj * dd1 * dd8;
    fii = fii * dd7 * dd7 + fi * dd9;
    fjj = fjj * dd8 * dd8 + fj * dd10;
    fi = fi * dd7 / ri;
    fj = fj * dd8 / rj;
    out->f += fc;
    for (int k = 0; k < 3; k++) {
      out->gi[k] += fi * rveci[3 * ic + k] + fu * rvec_ij[k];
      out->gj[k] += fj * rvecj[3 * ic + k] - fu * rvec_ij[k];
    }
    out->d2 += 2 * (2 * fu + fi + fj) + 2 * fuu + fii + fjj +
               fui * u2pst / (ri * rij) + fuj * u2mst / (rj * rij);
    out->li += 2 * (fu + fi) + fuu + fii + fui * u2pst / (ri * rij);
    out->lj += 2 * (fu + fj) + fuu + fjj + fuj * u2mst / (rj * rij);
  }
#undef U
#undef S
#undef TT
#undef RI
#undef RJ
}

/* e-n terms of electron i: qmc/vmc/jastrow4.f90:340-440; derivatives
   qmc/vmc/deriv_jastrow4.f90:880-1090 */
static void jas_en(const orc_sys *o, const double *ri_v, const double *rveci,
                   pair_out *out, pderiv *P) {
  memset(out, 0, sizeof(*out));
  double rri[QMCW_MAX_ORD + 2];
  double ar = o->asymp_r;
  for (int ic = 0; ic < o->ncent; ic++) {
    int it = o->iwctype[ic];
    double ri = ri_v[ic], dd7, dd9;
    scale_dist2(o, ri, &rri[1], &dd7, &dd9);
    rri[0] = 1;
    const double *a = o->a4 + it * (o->norda + 1);
    double top = a[0] * rri[1], topi = a[0];
    double bot = 1 + a[1] * rri[1], boti = a[1], bot2 = bot * bot;
    double fen = top / bot, feni = topi / bot2, fenii = -2 * feni * boti / bot;
    fen -= o->asymp_jasa[it];
    for (int iord = 2; iord <= o->norda; iord++) {
      rri[iord] = pow(rri[1], iord);
      fen += a[iord] * rri[iord];
      feni += a[iord] * iord * rri[iord - 1];
      fenii += a[iord] * iord * (iord - 1) * (iord >= 2 ? rri[iord - 2] : 0);
    }
    fenii = fenii * dd7 * dd7 + feni * dd9;
    feni = feni * dd7 / ri;
    out->f += fen;
    for (int k = 0; k < 3; k++) out->gi[k] += feni * rveci[3 * ic + k];
    out->d2 += fenii + 2 * feni;
    out->li += fenii + 2 * feni;
    if (P && P->mode) {
      for (int p = 0; p < o->npa; p++) {
        if (o->pa_type[p] != it) continue;
        int iw = o->pa_idx[p];
        double gen, geni, genii;
        if (iw == 1) {
          double bt = 1 + a[1] * rri[1], bta = 1 + a[1] * ar;
          gen = rri[1] / bt - ar / bta;
          geni = 1 / (bt * bt);
          genii = -2 * a[1] * geni / bt;
        } else if (iw == 2) {
          double t = -a[0] * rri[1] * rri[1], ti = -2 * a[0] * rri[1], tii = -2 * a[0];
          double bot0 = 1 + a[1] * rri[1], bt = bot0 * bot0, bt2 = bt * bt;
          double bta = 1 + a[1] * ar;
          gen = t / bt + a[0] * ar * ar / (bta * bta);
          geni = ti / (bt * bot0);
          genii = tii * (1 - 2 * a[1] * rri[1]) / bt2;
        } else {
          int iord = iw - 1;
          gen = pow(rri[1], iord) - pow(ar, iord);
          geni = iord * pow(rri[1], iord - 1);
          genii = iord * (iord - 1) * (iord >= 2 ? pow(rri[1], iord - 2) : 0);
        }
        genii = genii * dd7 * dd7 + geni * dd9;
        geni = geni * dd7 / ri;
        double vi[3];
        for (int k = 0; k < 3; k++) vi[k] = geni * rveci[3 * ic + k];
        pd_add(P, p, 1.0, gen, vi, NULL, genii + 2 * geni);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* walker state (psi_t, qmc/psi_type_mod.f90:5-39)                     */
/* ------------------------------------------------------------------ */
struct orc_walker {
  int N, M, NO, nup, ndn;
  double *x, *rvec_en, *r_en, *rvec_ee, *r_ee;
  double *fso, *fijo, *d2ijo, *lapjij, *fjo, fsumo, d2o;
  double *orb, *dorb, *ddorb;
  double *aiup, *aidn, deta_up, deta_dn, *tup, *tdn;
  double *invup, *invdn, *detup, *detdn, chi, *yup, *ydn, *gup, *gdn;
  /* last full evaluation */
  double logpsid, sgnpsid, psij, eloc, pe, pe_ee, d2j, d2lnpsi, vpsp, eloc_pot_loc;
  double *vj, *vd, *divvj, *divvd, *ekinen;
  double *psp_nonloc_orb;
  int have_derivs;
  double *gvalue, *g, *d2g, *dpe;
  double xq[QMCW_MAX_NQUAD], yq[QMCW_MAX_NQUAD], zq[QMCW_MAX_NQUAD];
  double psi2o;
  /* pending one-electron move */
  int pend_ie;
  double xnew[3], *r_en_n, *rvec_en_n, *r_ee_n, *rvec_ee_n;
  double *fsn, *fijn_i, *fijn_j, *lapn_i, *lapn_j, *fjn, fsumn, fsn_diag, fijn_diag[3], lapn_diag;
  double *orbe, *dorbe, *ddorbe;
  double *ain, detan, *tn, *invn, *detn, chin, *yupn, *ydnn;
  double *work, *scr_basis, *orbq, *gd, *gold;
  /* last DMC-mode evaluation, per electron: the quadrature points of its
     pairs in range (unit vectors, pair order) and the Psi ratios there
     (psi_t quadx / quadr, qmc/psi_type_mod.f90:34) */
  double *tqx, *tqr;
  int *tqn;
};

#define ALLOCD(n) ((double *)calloc((size_t)(n) + 1, sizeof(double)))

orc_walker *orc_walker_create(const orc_sys *o) {
  orc_walker *w = calloc(1, sizeof(orc_walker));
  int N = o->nelec, M = o->ncent, NO = o->norb, nup = o->nup, ndn = o->ndn;
  int k2 = o->kmax * o->kmax + 1;
  w->N = N; w->M = M; w->NO = NO; w->nup = nup; w->ndn = ndn;
  w->x = ALLOCD(3 * N);
  w->rvec_en = ALLOCD(3 * N * M); w->r_en = ALLOCD(N * M);
  w->rvec_ee = ALLOCD(3 * N * N); w->r_ee = ALLOCD(N * N);
  w->fso = ALLOCD(N * N); w->fijo = ALLOCD(3 * N * N); w->d2ijo = ALLOCD(N * N);
  w->lapjij = ALLOCD(N * N); w->fjo = ALLOCD(3 * N);
  w->orb = ALLOCD(N * NO); w->dorb = ALLOCD(3 * N * NO); w->ddorb = ALLOCD(N * NO);
  w->aiup = ALLOCD(nup * nup); w->aidn = ALLOCD(ndn * ndn);
  w->tup = ALLOCD(nup * NO); w->tdn = ALLOCD(ndn * NO);
  w->invup = ALLOCD(o->ndetup * k2); w->invdn = ALLOCD(o->ndetdn * k2);
  w->detup = ALLOCD(o->ndetup); w->detdn = ALLOCD(o->ndetdn);
  w->yup = ALLOCD(o->noccup * nup); w->ydn = ALLOCD(o->noccdn * ndn);
  w->gup = ALLOCD(o->noccup * nup); w->gdn = ALLOCD(o->noccdn * ndn);
  w->vj = ALLOCD(3 * N); w->vd = ALLOCD(3 * N); w->divvj = ALLOCD(N);
  w->divvd = ALLOCD(N); w->ekinen = ALLOCD(N);
  w->psp_nonloc_orb = ALLOCD(N * NO);
  int npj = o->npj;
  w->gvalue = ALLOCD(npj); w->g = ALLOCD(3 * N * npj); w->d2g = ALLOCD(npj);
  w->dpe = ALLOCD(npj);
  for (int q = 0; q < o->nquad; q++) {
    w->xq[q] = o->xq0[q]; w->yq[q] = o->yq0[q]; w->zq[q] = o->zq0[q];
  }
  w->pend_ie = -1;
  w->r_en_n = ALLOCD(M); w->rvec_en_n = ALLOCD(3 * M);
  w->r_ee_n = ALLOCD(N); w->rvec_ee_n = ALLOCD(3 * N);
  w->fsn = ALLOCD(N); w->fijn_i = ALLOCD(3 * N); w->fijn_j = ALLOCD(3 * N);
  w->lapn_i = ALLOCD(N); w->lapn_j = ALLOCD(N); w->fjn = ALLOCD(3 * N);
  w->orbe = ALLOCD(NO); w->dorbe = ALLOCD(3 * NO); w->ddorbe = ALLOCD(NO);
  int ns = nup > ndn ? nup : ndn, nocc = o->noccup > o->noccdn ? o->noccup : o->noccdn;
  int nd = o->ndetup > o->ndetdn ? o->ndetup : o->ndetdn;
  w->ain = ALLOCD(ns * ns); w->tn = ALLOCD(ns * nocc);
  w->invn = ALLOCD(nd * k2); w->detn = ALLOCD(nd);
  w->yupn = ALLOCD(o->noccup * nup); w->ydnn = ALLOCD(o->noccdn * ndn);
  int nb = o->nbasis;
  w->work = ALLOCD(8 * nb + 8 * NO * (N + 4) + 64 * N + 8 * ns * NO + 6 * npj + 4096);
  w->scr_basis = ALLOCD(5 * nb + 8);
  w->orbq = ALLOCD(NO);
  w->gd = ALLOCD(npj);
  w->gold = ALLOCD(npj);
  w->tqx = ALLOCD(3 * N * M * QMCW_MAX_NQUAD);
  w->tqr = ALLOCD(N * M * QMCW_MAX_NQUAD);
  w->tqn = calloc((size_t)N + 1, sizeof(int));
  return w;
}

void orc_walker_destroy(orc_walker *w) {
  if (!w) return;
  double *p[] = {w->x, w->rvec_en, w->r_en, w->rvec_ee, w->r_ee, w->fso, w->fijo,
                 w->d2ijo, w->lapjij, w->fjo, w->orb, w->dorb, w->ddorb, w->aiup,
                 w->aidn, w->tup, w->tdn, w->invup, w->invdn, w->detup, w->detdn,
                 w->yup, w->ydn, w->gup, w->gdn, w->vj, w->vd, w->divvj, w->divvd,
                 w->ekinen, w->psp_nonloc_orb, w->gvalue, w->g, w->d2g, w->dpe,
                 w->r_en_n, w->rvec_en_n, w->r_ee_n, w->rvec_ee_n, w->fsn, w->fijn_i,
                 w->fijn_j, w->lapn_i, w->lapn_j, w->fjn, w->orbe, w->dorbe, w->ddorbe,
                 w->ain, w->tn, w->invn, w->detn, w->yupn, w->ydnn, w->work,
                 w->scr_basis, w->orbq, w->gd, w->gold, w->tqx, w->tqr};
  for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++) free(p[i]);
  free(w->tqn);
  free(w);
}

void orc_walker_get(const orc_sys *o, const orc_walker *w, double *x,
                    double *psi2o) {
  if (x) memcpy(x, w->x, sizeof(double) * 3 * o->nelec);
  if (psi2o) *psi2o = w->psi2o;
}

/* qmc/vmc/distances.f90:1-233 (finite systems) */
static void distances(const orc_sys *o, orc_walker *w, double *pe,
                      double *pe_ee) {
  int N = o->nelec, M = o->ncent;
  double pe_en = 0;
  *pe = o->pecent;
  for (int ic = 0; ic < M; ic++)
    for (int i = 0; i < N; i++) {
      double s = 0;
      for (int k = 0; k < 3; k++) {
        double d = w->x[3 * i + k] - o->cent[3 * ic + k];
        w->rvec_en[(i * M + ic) * 3 + k] = d;
        s += d * d;
      }
      w->r_en[i * M + ic] = sqrt(s);
      if (o->nloc == 0) pe_en += -o->znuc[o->iwctype[ic]] / w->r_en[i * M + ic];
    }
  double pee = 0;
  for (int i = 1; i < N; i++)
    for (int j = 0; j < i; j++) {
      double s = 0;
      for (int k = 0; k < 3; k++) {
        double d = w->x[3 * i + k] - w->x[3 * j + k];
        w->rvec_ee[(i * N + j) * 3 + k] = d;
        s += d * d;
      }
      w->r_ee[i * N + j] = sqrt(s);
      pee += 1.0 / w->r_ee[i * N + j];
    }
  *pe += pe_en + pee;
  *pe_ee = pee;
}

/* full Jastrow into the walker: qmc/vmc/jastrow4.f90 / deriv_jastrow4.f90 */
static void jastrow_full_w(const orc_sys *o, orc_walker *w, int derivs) {
  int N = o->nelec, M = o->ncent, npj = o->npj;
  pderiv P = {0, 0, 0, 0, 0}, *PP = NULL;
  double *gvi = w->work, *gvj = w->work + 3 * npj;
  if (derivs && npj > 0) {
    memset(w->gvalue, 0, sizeof(double) * npj);
    memset(w->g, 0, sizeof(double) * 3 * N * npj);
    memset(w->d2g, 0, sizeof(double) * npj);
    P.mode = 2; P.gv = w->gvalue; P.gi = gvi; P.gj = gvj; P.d2 = w->d2g;
    PP = &P;
  }
  double fsum = 0, d2 = 0;
  memset(w->fjo, 0, sizeof(double) * 3 * N);
  memset(w->divvj, 0, sizeof(double) * N);
  double lapj_tmp[4096];
  (void)lapj_tmp;
  for (int i = 1; i < N; i++)
    for (int j = 0; j < i; j++) {
      pair_out po;
      if (PP) { memset(gvi, 0, sizeof(double) * 3 * npj); memset(gvj, 0, sizeof(double) * 3 * npj); }
      jas_pair(o, i, j, w->r_ee[i * N + j], w->rvec_ee + (i * N + j) * 3,
               w->r_en + i * M, w->r_en + j * M, w->rvec_en + i * M * 3,
               w->rvec_en + j * M * 3, &po, PP);
      if (PP)
        for (int p = 0; p < npj; p++)
          for (int k = 0; k < 3; k++) {
            w->g[(p * N + i) * 3 + k] += gvi[3 * p + k];
            w->g[(p * N + j) * 3 + k] += gvj[3 * p + k];
          }
      w->fso[i * N + j] = po.f;
      for (int k = 0; k < 3; k++) {
        w->fijo[(i * N + j) * 3 + k] = po.gi[k];
        w->fijo[(j * N + i) * 3 + k] = po.gj[k];
      }
      w->d2ijo[i * N + j] = po.d2;
      w->lapjij[i * N + j] = po.li;
      w->lapjij[j * N + i] = po.lj;
      fsum += po.f;
      for (int k = 0; k < 3; k++) {
        w->fjo[3 * i + k] += po.gi[k];
        w->fjo[3 * j + k] += po.gj[k];
      }
      w->divvj[i] += po.d2 / 2;
      w->divvj[j] += po.d2 / 2;
      d2 += po.d2;
    }
  for (int i = 0; i < N; i++) {
    pair_out po;
    if (PP) memset(gvi, 0, sizeof(double) * 3 * npj);
    jas_en(o, w->r_en + i * M, w->rvec_en + i * M * 3, &po, PP);
    if (PP)
      for (int p = 0; p < npj; p++)
        for (int k = 0; k < 3; k++) w->g[(p * N + i) * 3 + k] += gvi[3 * p + k];
    w->fso[i * N + i] = po.f;
    for (int k = 0; k < 3; k++) w->fijo[(i * N + i) * 3 + k] = po.gi[k];
    w->d2ijo[i * N + i] = po.d2;
    w->lapjij[i * N + i] = po.li;
    fsum += po.f;
    for (int k = 0; k < 3; k++) w->fjo[3 * i + k] += po.gi[k];
    w->divvj[i] += po.d2;
    d2 += po.d2;
  }
  w->fsumo = fsum;
  w->d2o = d2;
  w->psij = fsum;
  w->d2j = d2;
  memcpy(w->vj, w->fjo, sizeof(double) * 3 * N);
  w->have_derivs = (PP != NULL);
}

/* qmc/deriv_fast_mod.f90:545-598 (eval_y) */
static void eval_y(const orc_sys *o, const double *invup, const double *invdn,
                   const double *detup, const double *detdn, double *chi,
                   double *yup, double *ydn) {
  int k2 = o->kmax * o->kmax + 1;
  double cup[4096], cdn[4096];
  double ch = 0;
  for (int u = 0; u < o->ndetup; u++) cup[u] = 0;
  for (int u = 0; u < o->ndetdn; u++) cdn[u] = 0;
  for (int id = 0; id < o->ndet; id++) {
    double cdet = o->cdet_in_wf[id];
    int iup = o->iwdetup[id], idn = o->iwdetdn[id];
    int sgn = o->sgnup[iup] * o->sgndn[idn];
    double coef = sgn * cdet * detup[iup] * detdn[idn];
    ch += coef;
    cup[iup] += coef;
    cdn[idn] += coef;
  }
  int nup = o->nup, ndn = o->ndn;
  memset(yup, 0, sizeof(double) * o->noccup * nup);
  memset(ydn, 0, sizeof(double) * o->noccdn * ndn);
  for (int u = 0; u < o->ndetup; u++) {
    int k = o->ordup[u];
    for (int i = 0; i < k; i++)
      for (int j = 0; j < k; j++)
        yup[o->excup[u * nup + j] * nup + o->refup[u * nup + i]] +=
            cup[u] * invup[u * k2 + j + k * i];
  }
  for (int u = 0; u < o->ndetdn && ndn > 0; u++) {
    int k = o->orddn[u];
    for (int i = 0; i < k; i++)
      for (int j = 0; j < k; j++)
        ydn[o->excdn[u * ndn + j] * ndn + o->refdn[u * ndn + i]] +=
            cdn[u] * invdn[u * k2 + j + k * i];
  }
  double chii = 1.0 / ch;
  for (int q = 0; q < o->noccup * nup; q++) yup[q] *= chii;
  for (int q = 0; q < o->noccdn * ndn; q++) ydn[q] *= chii;
  *chi = ch;
}

/* qmc/deriv_fast_mod.f90:367-398 (t_bld), :427-466 (alpha_bld) for one spin */
static int t_alpha_spin(const orc_sys *o, orc_walker *w, int spin) {
  int N = o->nelec, NO = o->norb, nup = o->nup;
  int ns = spin ? o->ndn : nup, off = spin ? nup : 0;
  const int *ref = spin ? o->iworbddn : o->iworbdup;
  double *ai = spin ? w->aidn : w->aiup, *t = spin ? w->tdn : w->tup;
  if (ns == 0) { if (spin) w->deta_dn = 1; else w->deta_up = 1; return 0; }
  (void)N;
  /* A(i,k) = orb(i, ref_k), column-major: ai[i + ns*k] */
  for (int i = 0; i < ns; i++)
    for (int k = 0; k < ns; k++) ai[i + ns * k] = w->orb[(off + i) * NO + ref[k]];
  double det = orc_matinv(ai, ns);
  if (spin) w->deta_dn = det; else w->deta_up = det;
  /* T(k, o) = sum_i ainv(k, i) orb(i, o) */
  for (int k = 0; k < ns; k++)
    for (int q = 0; q < NO; q++) {
      double s = 0;
      for (int i = 0; i < ns; i++) s += ai[k + ns * i] * w->orb[(off + i) * NO + q];
      t[k * NO + q] = s;
    }
  int k2 = o->kmax * o->kmax + 1;
  int nd = spin ? o->ndetdn : o->ndetup;
  const int *ord = spin ? o->orddn : o->ordup, *rf = spin ? o->refdn : o->refup,
            *le = spin ? o->lab_excdn : o->lab_excup;
  double *inv = spin ? w->invdn : w->invup, *dets = spin ? w->detdn : w->detup;
  for (int u = 0; u < nd; u++) {
    int k = ord[u];
    for (int i = 0; i < k; i++)
      for (int j = 0; j < k; j++)
        inv[u * k2 + j + k * i] = t[rf[u * ns + j] * NO + le[u * ns + i]];
    dets[u] = orc_matinv(inv + u * k2, k);
  }
  return det == 0.0 ? QMCW_ERR_SINGULAR : 0;
}

/* qmc/deriv_fast_mod.f90:600-625 (g_bld) for one spin */
static void g_bld_spin(const orc_sys *o, orc_walker *w, int spin) {
  int NO = o->norb, ns = spin ? o->ndn : o->nup;
  int nocc = spin ? o->noccdn : o->noccup;
  const int *occ = spin ? o->occdn : o->occup, *iref = spin ? o->irefdn : o->irefup,
            *ivir = spin ? o->ivirdn : o->ivirup;
  const double *y = spin ? w->ydn : w->yup, *ai = spin ? w->aidn : w->aiup,
               *t = spin ? w->tdn : w->tup;
  double *g = spin ? w->gdn : w->gup;
  if (ns == 0) return;
  double *yai = w->work, *tyai = w->work + nocc * ns;
  for (int j = 0; j < nocc; j++)
    for (int i = 0; i < ns; i++) {
      double s = 0;
      for (int k = 0; k < ns; k++) s += y[j * ns + k] * ai[k + ns * i];
      yai[j * ns + i] = s;
    }
  for (int k = 0; k < ns; k++)
    for (int i = 0; i < ns; i++) {
      double s = 0;
      for (int j = 0; j < nocc; j++) s += t[k * NO + occ[j]] * yai[j * ns + i];
      tyai[k * ns + i] = s;
    }
  for (int k = 0; k < ns; k++)
    for (int i = 0; i < ns; i++) g[iref[k] * ns + i] = ai[k + ns * i] - tyai[k * ns + i];
  for (int v = 0; v < nocc - ns; v++)
    for (int i = 0; i < ns; i++) g[ivir[v] * ns + i] = yai[ivir[v] * ns + i];
}

/* qmc/vmc/determinant.f90:1-152 with t_bld/alpha_bld/y_bld/g_bld */
static int determinant_w(const orc_sys *o, orc_walker *w, double *d2lndet) {
  int N = o->nelec, M = o->ncent, NO = o->norb, nup = o->nup;
  for (int i = 0; i < N; i++)
    orbitals_at(o, w->rvec_en + i * M * 3, w->r_en + i * M, w->orb + i * NO,
                w->dorb + i * NO * 3, w->ddorb + i * NO, w->work);
  int st = t_alpha_spin(o, w, 0);
  int st2 = t_alpha_spin(o, w, 1);
  if (o->ndn == 0) { w->detdn[0] = 1; }
  eval_y(o, w->invup, w->invdn, w->detup, w->detdn, &w->chi, w->yup, w->ydn);
  g_bld_spin(o, w, 0);
  g_bld_spin(o, w, 1);
  double sgn = 1;
  double psid = w->chi * w->deta_up * w->deta_dn;
  if (psid < 0) sgn = -1;
  w->logpsid = log(fabs(w->chi)) + log(fabs(w->deta_up)) + log(fabs(w->deta_dn));
  w->sgnpsid = sgn;
  double d2 = 0;
  for (int i = 0; i < N; i++) {
    int up = i < nup;
    int ns = up ? nup : o->ndn, ie = up ? i : i - nup;
    int nocc = up ? o->noccup : o->noccdn;
    const int *occ = up ? o->occup : o->occdn;
    const double *g = up ? w->gup : w->gdn;
    double dd[3] = {0, 0, 0}, ek = 0;
    for (int j = 0; j < nocc; j++) {
      double gj = g[j * ns + ie];
      int jj = occ[j];
      for (int k = 0; k < 3; k++) dd[k] += gj * w->dorb[(i * NO + jj) * 3 + k];
      ek += gj * w->ddorb[i * NO + jj];
    }
    for (int k = 0; k < 3; k++) w->vd[3 * i + k] = dd[k];
    w->divvd[i] = ek - (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
    w->ekinen[i] = -0.5 * ek;
    d2 += w->divvd[i];
  }
  *d2lndet = d2;
  return st ? st : st2;
}

/* qmc/vmc/readps_gauss.f90:146-165 (gauss_pot) */
static double gauss_pot(const orc_sys *o, double r, int l, int ict) {
  double v = 0, rsq = r * r;
  int b = o->ecp_off[ict * QMCW_MAX_LPOT + l], e = o->ecp_off[ict * QMCW_MAX_LPOT + l + 1];
  for (int q = b; q < e; q++) {
    double p = (o->ecp_power[q] != 2) ? pow(r, o->ecp_power[q] - 2) : 1.0;
    v += p * o->ecp_coef[q] * exp(-o->ecp_expo[q] * rsq);
  }
  return v;
}
/* qmc/vmc/nonloc.f90:249-272 (yl0: (2l+1) P_l, l given as l+1) */
static double yl0(int l, double c) {
  switch (l) {
    case 1: return 1.0;
    case 2: return 3.0 * c;
    case 3: return 2.5 * (3 * c * c - 1);
    case 4: return 3.5 * c * (5 * c * c - 3);
    default: return 1.125 * (35 * c * c * c * c - 30 * c * c + 3);
  }
}

/* one-electron Jastrow value at a displaced position of electron iel:
   qmc/vmc/nonloc.f90:432-522 (nonlocj) with psinl/psianl/psibnl
   (qmc/vmc/nonlpsi.f90).  Returns Delta = sum_j fsn(i,j) - fso(i,j)
   (+ diagonal).  gdiff[npj] (optional) receives the change of
   d lnJ/dp (deriv_nonlocj). */
static double jas_1e_value(const orc_sys *o, orc_walker *w, int iel,
                           const double *xn, const double *ren,
                           const double *rvecen, double *gdiff) {
  int N = o->nelec, M = o->ncent, npj = o->npj;
  double val = 0;
  pderiv P = {0, 0, 0, 0, 0}, *PP = NULL;
  double *gnew = NULL, *gold = NULL;
  if (gdiff && npj > 0) {
    gnew = gdiff;
    memset(gnew, 0, sizeof(double) * npj);
    gold = w->gold;
    memset(gold, 0, sizeof(double) * npj);
    P.mode = 1;
    PP = &P;
  }
  for (int jj = 0; jj < N; jj++) {
    if (jj == iel) continue;
    int i, j;
    double rvec[3], rij = 0;
    const double *ri_v, *rj_v, *rvi, *rvj;
    if (jj < iel) { i = iel; j = jj; } else { i = jj; j = iel; }
    for (int k = 0; k < 3; k++) {
      double xi = (i == iel) ? xn[k] : w->x[3 * i + k];
      double xj = (j == iel) ? xn[k] : w->x[3 * j + k];
      rvec[k] = xi - xj;
      rij += rvec[k] * rvec[k];
    }
    rij = sqrt(rij);
    ri_v = (i == iel) ? ren : w->r_en + i * M;
    rj_v = (j == iel) ? ren : w->r_en + j * M;
    rvi = (i == iel) ? rvecen : w->rvec_en + i * M * 3;
    rvj = (j == iel) ? rvecen : w->rvec_en + j * M * 3;
    pair_out po;
    if (PP) P.gv = gnew;
    jas_pair(o, i, j, rij, rvec, ri_v, rj_v, rvi, rvj, &po, PP);
    val += po.f - w->fso[i * N + j];
    if (PP) {
      P.gv = gold;
      jas_pair(o, i, j, w->r_ee[i * N + j], w->rvec_ee + (i * N + j) * 3,
               w->r_en + i * M, w->r_en + j * M, w->rvec_en + i * M * 3,
               w->rvec_en + j * M * 3, &po, PP);
    }
  }
  pair_out po;
  if (PP) P.gv = gnew;
  jas_en(o, ren, rvecen, &po, PP);
  val += po.f - w->fso[iel * N + iel];
  if (PP) {
    P.gv = gold;
    jas_en(o, w->r_en + iel * M, w->rvec_en + iel * M * 3, &po, PP);
    for (int p = 0; p < npj; p++) gdiff[p] = gnew[p] - gold[p];
  }
  return val;
}

/* qmc/vmc/nonloc_pot.f90:1-79 + qmc/vmc/nonloc.f90:1-246 +
   psp_nonloc_pot_bld (qmc/deriv_fast_mod.f90:841-855); with derivs, the
   Jastrow dpe of qmc/fit/deriv_nonloc_pot.f90:59-64 */
static void nonloc_pot_w(const orc_sys *o, orc_walker *w, double *pe,
                         int derivs, int32_t *prng) {
  int N = o->nelec, M = o->ncent, NO = o->norb, nup = o->nup, npj = o->npj;
  double vps[64][QMCW_MAX_LPOT];
  double *orbe = w->orbq, *gd = w->gd, *gtmp = w->scr_basis;
  double ren[256], rvecen[3 * 256], xn[3];
  if (derivs) memset(w->dpe, 0, sizeof(double) * npj);
  memset(w->psp_nonloc_orb, 0, sizeof(double) * N * NO);
  double vpsp = 0;
  /* local part, all electrons */
  for (int i = 0; i < N; i++)
    for (int ic = 0; ic < M; ic++) {
      int ict = o->iwctype[ic];
      double r = fmax(1e-10, w->r_en[i * M + ic]);
      double v = gauss_pot(o, r, o->lpotp1[ict] - 1, ict) - o->znuc[ict] / r;
      *pe += v;
    }
  w->eloc_pot_loc = *pe;
  for (int i = 0; i < N; i++) {
    int up = i < nup;
    int ns = up ? nup : o->ndn, ie = up ? i : i - nup;
    int nocc = up ? o->noccup : o->noccdn;
    const int *occ = up ? o->occup : o->occdn;
    const double *G = up ? w->gup : w->gdn;
    int npt = 0;
    for (int ic = 0; ic < M; ic++) {
      int ict = o->iwctype[ic];
      double r = fmax(1e-10, w->r_en[i * M + ic]);
      int iskip = 1;
      for (int l = 1; l <= o->npotd[ict]; l++) {
        vps[ic][l - 1] = gauss_pot(o, r, l - 1, ict);
        if (l == o->lpotp1[ict]) vps[ic][l - 1] -= o->znuc[ict] / r;
        if (l != o->lpotp1[ict] && fabs(vps[ic][l - 1]) > 1e-4) iskip = 0;
      }
      if (iskip) continue;
      /* DMC: a fresh rotation of the grid for every pair in range
         (qmc/vmc/nonloc.f90:89, l_mode_dmc), drawn from the walker's stream */
      double pxq[QMCW_MAX_NQUAD], pyq[QMCW_MAX_NQUAD], pzq[QMCW_MAX_NQUAD];
      const double *qx = w->xq, *qy = w->yq, *qz = w->zq;
      if (prng) {
        orc_rotqua(prng, o->xq0, o->yq0, o->zq0, o->nquad, pxq, pyq, pzq);
        qx = pxq; qy = pyq; qz = pzq;
      }
      double rei = w->r_en[i * M + ic];
      double ri = 1.0 / rei;
      for (int iq = 0; iq < o->nquad; iq++) {
        const double *rv0 = w->rvec_en + (i * M + ic) * 3;
        double costh = (rv0[0] * qx[iq] + rv0[1] * qy[iq] + rv0[2] * qz[iq]) * ri;
        xn[0] = rei * qx[iq] + o->cent[3 * ic];
        xn[1] = rei * qy[iq] + o->cent[3 * ic + 1];
        xn[2] = rei * qz[iq] + o->cent[3 * ic + 2];
        for (int jc = 0; jc < M; jc++) {
          double s = 0;
          for (int k = 0; k < 3; k++) {
            rvecen[3 * jc + k] = xn[k] - o->cent[3 * jc + k];
            s += rvecen[3 * jc + k] * rvecen[3 * jc + k];
          }
          ren[jc] = (jc == ic) ? rei : sqrt(s);
        }
        orbitals_at(o, rvecen, ren, orbe, NULL, NULL, gtmp);
        double value = jas_1e_value(o, w, i, xn, ren, rvecen, derivs ? gd : NULL);
        double vpot = 0;
        for (int l = 1; l <= o->npotd[ict]; l++)
          if (l != o->lpotp1[ict]) vpot += vps[ic][l - 1] * yl0(l, costh);
        double factor = o->wq[iq] * exp(value);
        for (int q = 0; q < NO; q++) w->psp_nonloc_orb[i * NO + q] += factor * orbe[q] * vpot;
        if (prng) {
          /* DMC mode keeps the point and the Psi ratio for T-moves
             (qmc/vmc/nonloc.f90:99-104, 166-174) */
          double rat = 0;
          for (int j = 0; j < nocc; j++) rat += G[j * ns + ie] * orbe[occ[j]];
          size_t k = (size_t)i * M * QMCW_MAX_NQUAD + npt;
          w->tqr[k] = rat * exp(value);
          w->tqx[3 * k] = qx[iq]; w->tqx[3 * k + 1] = qy[iq]; w->tqx[3 * k + 2] = qz[iq];
          npt++;
        }
        if (derivs) {
          double rat = 0;
          for (int j = 0; j < nocc; j++) rat += G[j * ns + ie] * orbe[occ[j]];
          for (int p = 0; p < npj; p++) w->dpe[p] += factor * vpot * rat * gd[p];
        }
      }
    }
    if (prng) w->tqn[i] = npt;
  }
  for (int i = 0; i < N; i++) {
    int up = i < nup;
    int ns = up ? nup : o->ndn, ie = up ? i : i - nup;
    int nocc = up ? o->noccup : o->noccdn;
    const int *occ = up ? o->occup : o->occdn;
    const double *G = up ? w->gup : w->gdn;
    for (int j = 0; j < nocc; j++) vpsp += G[j * ns + ie] * w->psp_nonloc_orb[i * NO + occ[j]];
  }
  w->vpsp = vpsp;
  *pe += vpsp;
}

static int finite_or_err(double v) { return isfinite(v) ? 0 : QMCW_ERR_NAN; }

/* hpsi: qmc/vmc/hpsi.f90:2-381 (ibasis != 3 branch); prng: DMC mode, the
   pseudopotential grid rotated per pair from this stream */
static int eval_full_w(const orc_sys *o, orc_walker *w, int derivs, int32_t *prng) {
  int N = o->nelec;
  double pe, pe_ee, d2lnd;
  distances(o, w, &pe, &pe_ee);
  jastrow_full_w(o, w, derivs);
  int st = determinant_w(o, w, &d2lnd);
  if (o->nloc > 0) nonloc_pot_w(o, w, &pe, derivs, prng);
  else {
    w->eloc_pot_loc = pe;
    w->vpsp = 0;
    if (derivs) memset(w->dpe, 0, sizeof(double) * o->npj);
  }
  w->pe = pe;
  w->pe_ee = pe_ee;
  w->d2lnpsi = d2lnd + w->d2j;
  double energy = pe - o->hb * w->d2lnpsi;
  for (int i = 0; i < N; i++)
    for (int k = 0; k < 3; k++) {
      double v = w->vj[3 * i + k] + w->vd[3 * i + k];
      energy -= o->hb * v * v;
    }
  w->eloc = energy;
  w->psi2o = 2 * (w->logpsid + w->psij);
  w->pend_ie = -1;
  if (st) return st;
  return finite_or_err(energy);
}

int orc_eval_full(const orc_sys *o, orc_walker *w, const double *x,
                  double *out9, double *grad, double *divv) {
  int N = o->nelec;
  memcpy(w->x, x, sizeof(double) * 3 * N);
  int st = eval_full_w(o, w, o->nparm > 0, NULL);
  if (out9) {
    out9[0] = w->logpsid; out9[1] = w->sgnpsid; out9[2] = w->psij; out9[3] = w->eloc;
    out9[4] = w->pe; out9[5] = w->d2lnpsi; out9[6] = w->vpsp; out9[7] = w->pe_ee;
    out9[8] = w->eloc_pot_loc;
  }
  if (grad)
    for (int q = 0; q < 3 * N; q++) grad[q] = w->vj[q] + w->vd[q];
  if (divv)
    for (int i = 0; i < N; i++) divv[i] = w->divvj[i] + w->divvd[i];
  return st;
}

void orc_jastrow_full(const orc_sys *o, const double *x, double *value,
                      double *v, double *d2, double *divvj, double *gvalue,
                      double *g, double *d2g) {
  orc_walker *w = orc_walker_create(o);
  memcpy(w->x, x, sizeof(double) * 3 * o->nelec);
  double pe, pee;
  distances(o, w, &pe, &pee);
  jastrow_full_w(o, w, gvalue != NULL);
  int N = o->nelec, npj = o->npj;
  if (value) *value = w->psij;
  if (v) memcpy(v, w->vj, sizeof(double) * 3 * N);
  if (d2) *d2 = w->d2j;
  if (divvj) memcpy(divvj, w->divvj, sizeof(double) * N);
  if (gvalue) memcpy(gvalue, w->gvalue, sizeof(double) * npj);
  if (g) memcpy(g, w->g, sizeof(double) * 3 * N * npj);
  if (d2g) memcpy(d2g, w->d2g, sizeof(double) * npj);
  orc_walker_destroy(w);
}

int orc_determinant_full(const orc_sys *o, const double *x, double *logdet,
                         double *sgn, double *vd, double *divvd) {
  orc_walker *w = orc_walker_create(o);
  memcpy(w->x, x, sizeof(double) * 3 * o->nelec);
  double pe, pee, d2;
  distances(o, w, &pe, &pee);
  int st = determinant_w(o, w, &d2);
  if (logdet) *logdet = w->logpsid;
  if (sgn) *sgn = w->sgnpsid;
  if (vd) memcpy(vd, w->vd, sizeof(double) * 3 * o->nelec);
  if (divvd) memcpy(divvd, w->divvd, sizeof(double) * o->nelec);
  orc_walker_destroy(w);
  return st;
}

/* ------------------------------------------------------------------ */
/* one-electron moves                                                  */
/* ------------------------------------------------------------------ */
/* qmc/dete_mod.f90:249-292 (eval_grad) and :294-333 (eval_lapl) for
   electron iel of its spin; d[3*NO] holds dorb of the electron (or ddorb
   with ncomp = 1); tocc is T(:,occ) (ns x nocc); ai is A^-1 col-major */
static void eval_grad_gen(const orc_sys *o, int spin, int ie, const double *d,
                          int ncomp, const double *ai, const double *tocc,
                          const double *y, double *out) {
  int ns = spin ? o->ndn : o->nup, nocc = spin ? o->noccdn : o->noccup;
  const int *ref = spin ? o->iworbddn : o->iworbdup, *occ = spin ? o->occdn : o->occup;
  double v[4096];
  for (int k = 0; k < ncomp; k++) {
    double g = 0;
    for (int q = 0; q < ns; q++) g += ai[q + ns * ie] * d[ref[q] * ncomp + k];
    for (int j = 0; j < nocc; j++) {
      double s = 0;
      for (int q = 0; q < ns; q++) s += d[ref[q] * ncomp + k] * tocc[q * nocc + j];
      v[j] = d[occ[j] * ncomp + k] - s;
    }
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < nocc; j++) g += y[j * ns + i] * ai[i + ns * ie] * v[j];
    out[k] = g;
  }
}

/* T(:,occ) of the current state */
static void tocc_of(const orc_sys *o, const orc_walker *w, int spin, double *tocc) {
  int ns = spin ? o->ndn : o->nup, nocc = spin ? o->noccdn : o->noccup, NO = o->norb;
  const int *occ = spin ? o->occdn : o->occup;
  const double *t = spin ? w->tdn : w->tup;
  for (int q = 0; q < ns; q++)
    for (int j = 0; j < nocc; j++) tocc[q * nocc + j] = t[q * NO + occ[j]];
}

/* qmc/dete_mod.f90:114-214 (dete_update) into the pending buffers */
static int dete_update(const orc_sys *o, orc_walker *w, int iel) {
  int up = iel < o->nup, spin = up ? 0 : 1;
  int ns = up ? o->nup : o->ndn, off = up ? 0 : o->nup, ie = iel - off;
  int nocc = up ? o->noccup : o->noccdn;
  const int *ref = up ? o->iworbdup : o->iworbddn, *occ = up ? o->occup : o->occdn;
  int nd = up ? o->ndetup : o->ndetdn, k2 = o->kmax * o->kmax + 1;
  const int *ord = up ? o->ordup : o->orddn, *rf = up ? o->refup : o->refdn,
            *ex = up ? o->excup : o->excdn;
  memcpy(w->ain, up ? w->aiup : w->aidn, sizeof(double) * ns * ns);
  w->detan = up ? w->deta_up : w->deta_dn;
  tocc_of(o, w, spin, w->tn);
  memcpy(w->invn, up ? w->invup : w->invdn, sizeof(double) * nd * k2);
  memcpy(w->detn, up ? w->detup : w->detdn, sizeof(double) * nd);
  double *ai = w->ain, *t = w->tn;
  double ui[512], vi1[512], vi2[4096];
  double rat = 0;
  for (int q = 0; q < ns; q++) rat += w->orbe[ref[q]] * ai[q + ns * ie];
  for (int q = 0; q < ns; q++) ui[q] = -ai[q + ns * ie] / rat;
  for (int i = 0; i < ns; i++) {
    double s = 0;
    for (int q = 0; q < ns; q++) s += w->orbe[ref[q]] * ai[q + ns * i];
    vi1[i] = s;
  }
  vi1[ie] -= 1.0;
  for (int j = 0; j < nocc; j++) {
    double s = 0;
    for (int q = 0; q < ns; q++) s += w->orbe[ref[q]] * t[q * nocc + j];
    vi2[j] = s - w->orbe[occ[j]];
  }
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) ai[j + ns * i] += ui[j] * vi1[i];
  w->detan *= rat;
  for (int i = 0; i < nocc; i++)
    for (int j = 0; j < ns; j++) t[j * nocc + i] += ui[j] * vi2[i];
  for (int u = 0; u < nd; u++) {
    int k = ord[u];
    for (int ii = 0; ii < k; ii++)
      for (int jj = 0; jj < k; jj++)
        w->invn[u * k2 + jj + k * ii] = t[rf[u * ns + jj] * nocc + ex[u * ns + ii]];
    w->detn[u] = orc_matinv(w->invn + u * k2, k);
  }
  if (up)
    eval_y(o, w->invn, w->invdn, w->detn, w->detdn, &w->chin, w->yupn, w->ydnn);
  else
    eval_y(o, w->invup, w->invn, w->detup, w->detn, &w->chin, w->yupn, w->ydnn);
  return rat == 0.0 ? QMCW_ERR_SINGULAR : 0;
}

/* hpsie (qmc/vmc/hpsie.f90:1-71) = distancese + jastrow4e + determinante;
   the per-electron Jastrow laplacian follows jastrow4.f90's lapjijo terms */
int orc_ratio_1e(const orc_sys *o, orc_walker *w, int32_t iel,
                 const double *xnew, double *out4, double *vnew,
                 double *lap_new) {
  int N = o->nelec, M = o->ncent, NO = o->norb;
  w->pend_ie = iel;
  memcpy(w->xnew, xnew, sizeof(double) * 3);
  for (int ic = 0; ic < M; ic++) {
    double s = 0;
    for (int k = 0; k < 3; k++) {
      w->rvec_en_n[3 * ic + k] = xnew[k] - o->cent[3 * ic + k];
      s += w->rvec_en_n[3 * ic + k] * w->rvec_en_n[3 * ic + k];
    }
    w->r_en_n[ic] = sqrt(s);
  }
  /* jastrow4e */
  memcpy(w->fjn, w->fjo, sizeof(double) * 3 * N);
  double fsumn = w->fsumo;
  for (int jj = 0; jj < N; jj++) {
    if (jj == iel) continue;
    int i, j;
    if (jj < iel) { i = iel; j = jj; } else { i = jj; j = iel; }
    double rvec[3], rij = 0;
    for (int k = 0; k < 3; k++) {
      double xi = (i == iel) ? xnew[k] : w->x[3 * i + k];
      double xj = (j == iel) ? xnew[k] : w->x[3 * j + k];
      rvec[k] = xi - xj;
      rij += rvec[k] * rvec[k];
    }
    rij = sqrt(rij);
    w->r_ee_n[jj] = rij;
    for (int k = 0; k < 3; k++) w->rvec_ee_n[3 * jj + k] = rvec[k];
    pair_out po;
    jas_pair(o, i, j, rij, rvec, (i == iel) ? w->r_en_n : w->r_en + i * M,
             (j == iel) ? w->r_en_n : w->r_en + j * M,
             (i == iel) ? w->rvec_en_n : w->rvec_en + i * M * 3,
             (j == iel) ? w->rvec_en_n : w->rvec_en + j * M * 3, &po, NULL);
    w->fsn[jj] = po.f;
    /* store fijn(:,iel,jj) and fijn(:,jj,iel) */
    const double *g_iel = (i == iel) ? po.gi : po.gj;
    const double *g_jj = (i == iel) ? po.gj : po.gi;
    for (int k = 0; k < 3; k++) {
      w->fijn_i[3 * jj + k] = g_iel[k];
      w->fijn_j[3 * jj + k] = g_jj[k];
    }
    w->lapn_i[jj] = (i == iel) ? po.li : po.lj;
    w->lapn_j[jj] = (i == iel) ? po.lj : po.li;
    fsumn += po.f - w->fso[i * N + j];
    for (int k = 0; k < 3; k++) {
      w->fjn[3 * iel + k] += g_iel[k] - w->fijo[(iel * N + jj) * 3 + k];
      w->fjn[3 * jj + k] += g_jj[k] - w->fijo[(jj * N + iel) * 3 + k];
    }
  }
  {
    pair_out po;
    jas_en(o, w->r_en_n, w->rvec_en_n, &po, NULL);
    w->fsn_diag = po.f;
    for (int k = 0; k < 3; k++) w->fijn_diag[k] = po.gi[k];
    w->lapn_diag = po.li;
    fsumn += po.f - w->fso[iel * N + iel];
    for (int k = 0; k < 3; k++) w->fjn[3 * iel + k] += po.gi[k] - w->fijo[(iel * N + iel) * 3 + k];
  }
  w->fsumn = fsumn;
  /* determinante */
  orbitals_at(o, w->rvec_en_n, w->r_en_n, w->orbe, w->dorbe, w->ddorbe, w->work);
  int st = dete_update(o, w, iel);
  int up = iel < o->nup, spin = up ? 0 : 1, off = up ? 0 : o->nup;
  double psi = w->chin * w->detan * (up ? w->deta_dn : w->deta_up);
  double vd[3], lapd;
  eval_grad_gen(o, spin, iel - off, w->dorbe, 3, w->ain, w->tn, up ? w->yupn : w->ydnn, vd);
  eval_grad_gen(o, spin, iel - off, w->ddorbe, 1, w->ain, w->tn, up ? w->yupn : w->ydnn, &lapd);
  double vj[3] = {w->fjn[3 * iel], w->fjn[3 * iel + 1], w->fjn[3 * iel + 2]};
  double lapj = w->lapn_diag;
  for (int jj = 0; jj < N; jj++) if (jj != iel) lapj += w->lapn_i[jj];
  lapj += vj[0] * vj[0] + vj[1] * vj[1] + vj[2] * vj[2];
  double logn = log(fabs(psi));
  double psio = w->chi * w->deta_up * w->deta_dn;
  if (out4) {
    out4[0] = (psi / psio) * exp(fsumn - w->fsumo);
    out4[1] = logn;
    out4[2] = psi < 0 ? -1.0 : 1.0;
    out4[3] = fsumn;
  }
  if (vnew)
    for (int k = 0; k < 3; k++) vnew[k] = vj[k] + vd[k];
  if (lap_new) *lap_new = lapj + 2 * (vj[0] * vd[0] + vj[1] * vd[1] + vj[2] * vd[2]) + lapd;
  (void)NO;
  return st;
}

/* commit: jassav (qmc/vmc/jassav.f90), detsav (qmc/vmc/detsav.f90),
   dete_save (qmc/dete_mod.f90:70-112), distances of electron ie */
void orc_accept_1e(const orc_sys *o, orc_walker *w, int32_t iel, int accept) {
  if (w->pend_ie != iel) return;
  w->pend_ie = -1;
  if (!accept) return;
  int N = o->nelec, M = o->ncent, NO = o->norb;
  memcpy(w->x + 3 * iel, w->xnew, sizeof(double) * 3);
  for (int ic = 0; ic < M; ic++) {
    w->r_en[iel * M + ic] = w->r_en_n[ic];
    for (int k = 0; k < 3; k++) w->rvec_en[(iel * M + ic) * 3 + k] = w->rvec_en_n[3 * ic + k];
  }
  for (int jj = 0; jj < N; jj++) {
    if (jj == iel) continue;
    int i = jj < iel ? iel : jj, j = jj < iel ? jj : iel;
    w->r_ee[i * N + j] = w->r_ee_n[jj];
    for (int k = 0; k < 3; k++) w->rvec_ee[(i * N + j) * 3 + k] = w->rvec_ee_n[3 * jj + k];
    w->fso[i * N + j] = w->fsn[jj];
    for (int k = 0; k < 3; k++) {
      w->fijo[(iel * N + jj) * 3 + k] = w->fijn_i[3 * jj + k];
      w->fijo[(jj * N + iel) * 3 + k] = w->fijn_j[3 * jj + k];
    }
    w->lapjij[iel * N + jj] = w->lapn_i[jj];
    w->lapjij[jj * N + iel] = w->lapn_j[jj];
  }
  w->fso[iel * N + iel] = w->fsn_diag;
  for (int k = 0; k < 3; k++) w->fijo[(iel * N + iel) * 3 + k] = w->fijn_diag[k];
  w->lapjij[iel * N + iel] = w->lapn_diag;
  w->fsumo = w->fsumn;
  memcpy(w->fjo, w->fjn, sizeof(double) * 3 * N);
  memcpy(w->orb + iel * NO, w->orbe, sizeof(double) * NO);
  memcpy(w->dorb + iel * NO * 3, w->dorbe, sizeof(double) * 3 * NO);
  memcpy(w->ddorb + iel * NO, w->ddorbe, sizeof(double) * NO);
  int up = iel < o->nup;
  int ns = up ? o->nup : o->ndn, nocc = up ? o->noccup : o->noccdn;
  const int *occ = up ? o->occup : o->occdn;
  int nd = up ? o->ndetup : o->ndetdn, k2 = o->kmax * o->kmax + 1;
  memcpy(up ? w->aiup : w->aidn, w->ain, sizeof(double) * ns * ns);
  if (up) w->deta_up = w->detan; else w->deta_dn = w->detan;
  double *t = up ? w->tup : w->tdn;
  for (int q = 0; q < ns; q++)
    for (int j = 0; j < nocc; j++) t[q * NO + occ[j]] = w->tn[q * nocc + j];
  memcpy(up ? w->invup : w->invdn, w->invn, sizeof(double) * nd * k2);
  memcpy(up ? w->detup : w->detdn, w->detn, sizeof(double) * nd);
  w->chi = w->chin;
  memcpy(w->yup, w->yupn, sizeof(double) * o->noccup * o->nup);
  memcpy(w->ydn, w->ydnn, sizeof(double) * o->noccdn * o->ndn);
  w->logpsid = log(fabs(w->chi)) + log(fabs(w->deta_up)) + log(fabs(w->deta_dn));
  w->sgnpsid = (w->chi * w->deta_up * w->deta_dn) < 0 ? -1 : 1;
  w->psij = w->fsumo;
  w->psi2o = 2 * (w->logpsid + w->psij);
}

/* calc_1e_derivatives (qmc/psi_utils_mod.f90:457-483) */
void orc_grad_1e(const orc_sys *o, orc_walker *w, int32_t iel, double *grad,
                 double *lap) {
  int N = o->nelec, NO = o->norb, up = iel < o->nup, spin = up ? 0 : 1;
  int off = up ? 0 : o->nup;
  double tocc[65536];
  tocc_of(o, w, spin, tocc);
  double vd[3], lapd;
  const double *ai = up ? w->aiup : w->aidn, *y = up ? w->yup : w->ydn;
  eval_grad_gen(o, spin, iel - off, w->dorb + iel * NO * 3, 3, ai, tocc, y, vd);
  eval_grad_gen(o, spin, iel - off, w->ddorb + iel * NO, 1, ai, tocc, y, &lapd);
  double vj[3] = {w->fjo[3 * iel], w->fjo[3 * iel + 1], w->fjo[3 * iel + 2]};
  double lapj = 0;
  for (int j = 0; j < N; j++) lapj += w->lapjij[iel * N + j];
  lapj += vj[0] * vj[0] + vj[1] * vj[1] + vj[2] * vj[2];
  if (grad)
    for (int k = 0; k < 3; k++) grad[k] = vd[k] + vj[k];
  if (lap) *lap = lapd + 2 * (vj[0] * vd[0] + vj[1] * vd[1] + vj[2] * vd[2]) + lapj;
}

/* ------------------------------------------------------------------ */
/* parameter derivatives                                               */
/* ------------------------------------------------------------------ */
/* qmc/deriv_fast_mod.f90:174-209 (elocal_orb_bld), :400-425 (dt_bld),
   :468-516 (alpha_deloc_bld), :648-734 (dy_bld), :736-779 (dg_bld),
   :781-823 (csf_derivs_bld); qmc/deriv_orb_mod.f90:1486-1630
   (dpsi_orb_bld, deloc_orb_bld); qmc/vmc/hpsi.f90:366-374 (Jastrow
   denergy); order CSF, Jastrow, orbitals (qmc/deriv_mod.f90:219-239) */
int orc_derivs(const orc_sys *o, orc_walker *w, double *dpsi, double *deloc) {
  int N = o->nelec, NO = o->norb, nup = o->nup, ndn = o->ndn;
  int k2 = o->kmax * o->kmax + 1;
  double hb = o->hb;
  double *B = calloc((size_t)N * NO + 1, sizeof(double));      /* elocal_orb [N][NO] */
  double *dtup = calloc((size_t)nup * NO + 1, sizeof(double));
  double *dtdn = calloc((size_t)ndn * NO + 1, sizeof(double));
  double *delup = calloc(o->ndetup + 1, sizeof(double));
  double *deldn = calloc(o->ndetdn + 1, sizeof(double));
  double *dyup = calloc((size_t)o->noccup * nup + 1, sizeof(double));
  double *dydn = calloc((size_t)o->noccdn * ndn + 1, sizeof(double));
  double *dgup = calloc((size_t)o->noccup * nup + 1, sizeof(double));
  double *dgdn = calloc((size_t)o->noccdn * ndn + 1, sizeof(double));
  /* elocal_orb */
  double vj2 = 0;
  for (int q = 0; q < 3 * N; q++) vj2 += w->vj[q] * w->vj[q];
  double ddjas = (w->d2j + vj2) / N;
  for (int j = 0; j < NO; j++)
    for (int i = 0; i < N; i++) {
      double dv = 0;
      for (int k = 0; k < 3; k++) dv += w->vj[3 * i + k] * w->dorb[(i * NO + j) * 3 + k];
      double kin = -hb * (w->ddorb[i * NO + j] + 2 * dv + ddjas * w->orb[i * NO + j]);
      double pot = w->eloc_pot_loc * w->orb[i * NO + j] / N;
      if (o->nloc > 0) pot += w->psp_nonloc_orb[i * NO + j];
      B[i * NO + j] = kin + pot;
    }
  /* dt_bld, alpha_deloc_bld per spin */
  for (int spin = 0; spin < 2; spin++) {
    int ns = spin ? ndn : nup, off = spin ? nup : 0;
    if (ns == 0) continue;
    const double *ai = spin ? w->aidn : w->aiup, *t = spin ? w->tdn : w->tup;
    const int *ref = spin ? o->iworbddn : o->iworbdup;
    double *dt = spin ? dtdn : dtup;
    double *aib = calloc((size_t)ns * NO + 1, sizeof(double));
    for (int k = 0; k < ns; k++)
      for (int q = 0; q < NO; q++) {
        double s = 0;
        for (int i = 0; i < ns; i++) s += ai[k + ns * i] * B[(off + i) * NO + q];
        aib[k * NO + q] = s;
      }
    for (int k = 0; k < ns; k++)
      for (int q = 0; q < NO; q++) {
        double s = 0;
        for (int r = 0; r < ns; r++) s += aib[k * NO + ref[r]] * t[r * NO + q];
        dt[k * NO + q] = aib[k * NO + q] - s;
      }
    free(aib);
    int nd = spin ? o->ndetdn : o->ndetup;
    const int *ord = spin ? o->orddn : o->ordup, *rf = spin ? o->refdn : o->refup,
              *le = spin ? o->lab_excdn : o->lab_excup;
    const double *inv = spin ? w->invdn : w->invup;
    double *del = spin ? deldn : delup;
    for (int u = 0; u < nd; u++) {
      int k = ord[u];
      double d = 0;
      for (int j = 0; j < k; j++)
        for (int i = 0; i < k; i++)
          d += inv[u * k2 + j + k * i] * dt[rf[u * ns + i] * NO + le[u * ns + j]];
      del[u] = d;
    }
  }
  /* dy_bld */
  double fdt[4096], dti[4096], idti[4096];
  for (int icsf = 0; icsf < o->ncsf; icsf++) {
    double ccsf = o->csf_coef[icsf];
    for (int q = o->csf_det_off[icsf]; q < o->csf_det_off[icsf + 1]; q++) {
      int idet = o->csf_det[q];
      double cdet = o->csf_det_coef[q];
      int iup = o->iwdetup[idet], idn = o->iwdetdn[idet];
      int sgn = o->sgnup[iup] * o->sgndn[idn];
      double coef = sgn * cdet * ccsf * w->detup[iup] * w->detdn[idn];
      double dl = delup[iup] + deldn[idn];
      for (int spin = 0; spin < 2; spin++) {
        int ns = spin ? ndn : nup;
        if (ns == 0) continue;
        int u = spin ? idn : iup;
        int k = spin ? o->orddn[u] : o->ordup[u];
        const int *le = (spin ? o->lab_excdn : o->lab_excup) + u * ns;
        const int *rf = (spin ? o->refdn : o->refup) + u * ns;
        const int *ex = (spin ? o->excdn : o->excup) + u * ns;
        const double *inv = (spin ? w->invdn : w->invup) + u * k2;
        const double *dt = spin ? dtdn : dtup;
        double *dy = spin ? dydn : dyup;
        for (int i = 0; i < k; i++)
          for (int j = 0; j < k; j++) fdt[j + k * i] = dt[rf[j] * NO + le[i]];
        /* fmatmul(prod, mat1, mat2, d): prod(j,i) = sum_k mat1(j,k) mat2(k,i) */
        for (int i = 0; i < k; i++)
          for (int j = 0; j < k; j++) {
            double s = 0;
            for (int kk = 0; kk < k; kk++) s += fdt[j + k * kk] * inv[kk + k * i];
            dti[j + k * i] = s;
          }
        for (int i = 0; i < k; i++)
          for (int j = 0; j < k; j++) {
            double s = 0;
            for (int kk = 0; kk < k; kk++) s += inv[j + k * kk] * dti[kk + k * i];
            idti[j + k * i] = s;
          }
        for (int i = 0; i < k; i++)
          for (int j = 0; j < k; j++)
            dy[ex[j] * ns + rf[i]] += coef * (dl * inv[j + k * i] - idti[j + k * i]);
      }
    }
  }
  double chii = 1.0 / w->chi;
  double esum = 0;
  for (int j = 0; j < o->noccup; j++)
    for (int i = 0; i < nup; i++) esum += w->yup[j * nup + i] * dtup[i * NO + o->occup[j]];
  for (int j = 0; j < o->noccdn; j++)
    for (int i = 0; i < ndn; i++) esum += w->ydn[j * ndn + i] * dtdn[i * NO + o->occdn[j]];
  for (int q = 0; q < o->noccup * nup; q++) dyup[q] = dyup[q] * chii - esum * w->yup[q];
  for (int q = 0; q < o->noccdn * ndn; q++) dydn[q] = dydn[q] * chii - esum * w->ydn[q];
  /* dg_bld */
  for (int spin = 0; spin < 2; spin++) {
    int ns = spin ? ndn : nup, off = spin ? nup : 0;
    if (ns == 0) continue;
    int nocc = spin ? o->noccdn : o->noccup;
    const int *occ = spin ? o->occdn : o->occup, *ref = spin ? o->iworbddn : o->iworbdup,
              *iref = spin ? o->irefdn : o->irefup;
    const double *t = spin ? w->tdn : w->tup, *dt = spin ? dtdn : dtup,
                 *dy = spin ? dydn : dyup, *y = spin ? w->ydn : w->yup,
                 *g = spin ? w->gdn : w->gup, *ai = spin ? w->aidn : w->aiup;
    double *dg = spin ? dgdn : dgup;
    double *aux4 = calloc((size_t)nocc * ns + 1, sizeof(double));
    for (int k = 0; k < ns; k++)
      for (int i = 0; i < ns; i++) {
        double a1 = 0, a2 = 0;
        for (int j = 0; j < nocc; j++) {
          a1 += t[k * NO + occ[j]] * dy[j * ns + i];
          a2 += dt[k * NO + occ[j]] * y[j * ns + i];
        }
        aux4[iref[k] * ns + i] = a1 + a2;
      }
    for (int j = 0; j < nocc; j++)
      for (int i = 0; i < ns; i++) {
        double a3 = 0;
        for (int q = 0; q < ns; q++) a3 += g[j * ns + q] * B[(off + q) * NO + ref[i]];
        aux4[j * ns + i] = dy[j * ns + i] - a3 - aux4[j * ns + i];
      }
    for (int j = 0; j < nocc; j++)
      for (int i = 0; i < ns; i++) {
        double s = 0;
        for (int q = 0; q < ns; q++) s += aux4[j * ns + q] * ai[q + ns * i];
        dg[j * ns + i] = s;
      }
    free(aux4);
  }
  int p = 0;
  /* CSF block */
  for (int ip = 0; ip < o->npcsf; ip++, p++) {
    int icsf = o->pcsf[ip];
    double dd = 0, de = 0;
    for (int q = o->csf_det_off[icsf]; q < o->csf_det_off[icsf + 1]; q++) {
      int idet = o->csf_det[q];
      int iup = o->iwdetup[idet], idn = o->iwdetdn[idet];
      int sgn = o->sgnup[iup] * o->sgndn[idn];
      double coef = sgn * o->csf_det_coef[q] * w->detup[iup] * w->detdn[idn] * chii;
      double dl = delup[iup] + deldn[idn];
      dd += coef;
      de += coef * (dl - esum);
    }
    dpsi[p] = dd;
    deloc[p] = de;
  }
  /* Jastrow block */
  for (int ip = 0; ip < o->npj; ip++, p++) {
    double de = -hb * w->d2g[ip] + (o->nloc > 0 ? w->dpe[ip] : 0.0);
    for (int i = 0; i < N; i++)
      for (int k = 0; k < 3; k++)
        de -= 2 * hb * w->g[(ip * N + i) * 3 + k] * (w->vj[3 * i + k] + w->vd[3 * i + k]);
    dpsi[p] = w->gvalue[ip];
    deloc[p] = de;
  }
  /* orbital block */
  for (int ip = 0; ip < o->nporb; ip++, p++) {
    int oi = o->porb_i[ip], oj = o->porb_j[ip];
    double dp = 0, de = 0;
    int iu = o->ioccup[oi], id = o->ioccdn[oi];
    if (iu >= 0)
      for (int k = 0; k < nup; k++) {
        dp += w->gup[iu * nup + k] * w->orb[k * NO + oj];
        de += w->gup[iu * nup + k] * B[k * NO + oj] + dgup[iu * nup + k] * w->orb[k * NO + oj];
      }
    if (id >= 0)
      for (int k = 0; k < ndn; k++) {
        dp += w->gdn[id * ndn + k] * w->orb[(nup + k) * NO + oj];
        de += w->gdn[id * ndn + k] * B[(nup + k) * NO + oj] +
              dgdn[id * ndn + k] * w->orb[(nup + k) * NO + oj];
      }
    if (o->porb_rev[ip]) {
      int ju = o->ioccup[oj], jd = o->ioccdn[oj];
      if (ju >= 0)
        for (int k = 0; k < nup; k++) {
          dp -= w->gup[ju * nup + k] * w->orb[k * NO + oi];
          de -= w->gup[ju * nup + k] * B[k * NO + oi] + dgup[ju * nup + k] * w->orb[k * NO + oi];
        }
      if (jd >= 0)
        for (int k = 0; k < ndn; k++) {
          dp -= w->gdn[jd * ndn + k] * w->orb[(nup + k) * NO + oi];
          de -= w->gdn[jd * ndn + k] * B[(nup + k) * NO + oi] +
                dgdn[jd * ndn + k] * w->orb[(nup + k) * NO + oi];
        }
    }
    dpsi[p] = dp;
    deloc[p] = de;
  }
  free(B); free(dtup); free(dtdn); free(delup); free(deldn);
  free(dyup); free(dydn); free(dgup); free(dgdn);
  return 0;
}

/* ------------------------------------------------------------------ */
/* The accelerated Metropolis mover (imetro = 6), qmc/vmc/metrop_slat_mov1.f90:
   the electron moves in local (r, theta, phi) coordinates about its nearest
   nucleus: r from sqrt(r)|1 + co r|exp(-zeta r) on [r/deltar, r deltar] by
   rejection, cos(theta) uniform on [1 - deltt, 1] about the electron's
   direction, phi from 1 + term cos(phi) by rejection about the transverse
   velocity.  Forward and reverse densities use the same formulas. */

void orc_set_mover(orc_sys *o, int32_t imetro, double deltar, double deltat) {
  o->imetro = imetro;
  /* the reference resets out-of-range values (qmc/control_mod.f90:374-384) */
  o->deltar = deltar < 1.0 ? 2.0 : deltar;
  o->deltat = (deltat < 0.0 || deltat > 2.0) ? 2.0 : deltat;
}

/* incomplete gamma function gamma(a, x) for a = 3/2, 5/2 with xae = x^a e^-x
   given; returns -gammac (iflag = -1) for x >= a + 4, qmc/vmc/gammai.f90:1-98 */
static double slat_gammai(double a, double x, double xae, int *iflag) {
  if (x < a + 4) {
    *iflag = 1;
    double ap = a, sum = 1 / a, del = sum;
    int iter = (int)(21 + 3 * (x - a));
    for (int n = 1; n <= iter; n++) {
      ap = ap + 1;
      del = del * x / ap;
      sum = sum + del;
    }
    return sum * xae;
  }
  *iflag = -1;
  double a0 = 1, a1 = x, b0 = 0, b1 = 1, fac = 1, g = 0;
  int iter = 20 + (int)lround(a - x);
  if (iter < 7) iter = 7;
  for (int n = 1; n <= iter; n++) {
    double an = (double)n, ana = an - a;
    a0 = (a1 + a0 * ana) * fac;
    b0 = (b1 + b0 * ana) * fac;
    double anf = an * fac;
    a1 = x * a0 + anf * a1;
    b1 = x * b0 + anf * b1;
    fac = 1 / a1;
    g = b1 * fac;
  }
  return -(g * xae);
}

/* zeta and co of the radial function for radial velocity vr at distance r
   (metrop_slat_mov1.f90:125-140, :398-412) */
static void slat_zeta_co(double zcusp, double vr, double r, double *zeta, double *co) {
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
  *zeta = z;
  *co = (z + vr) / (1.0 - (z + vr) * r);
}

/* integral of sqrt(t)|1 + co t|exp(-zeta t) over [rbot, rtop] divided by
   bot (metrop_slat_mov1.f90:198-232) */
static double slat_area(double zeta, double co, double rbot, double rtop, double bot) {
  const double d5b2 = 2.5, d2b3 = .666666666666667, g5b2 = 1.329340388179137;
  int fb, ft, fz;
  double rzero = -1.0 / co;
  double zrbot = zeta * rbot, zrtop = zeta * rtop;
  double zebot = sqrt(zrbot * zrbot * zrbot) * exp(-zrbot);
  double zetop = sqrt(zrtop * zrtop * zrtop) * exp(-zrtop);
  double g52bot = slat_gammai(d5b2, zrbot, zrbot * zebot, &fb);
  double g52top = slat_gammai(d5b2, zrtop, zrtop * zetop, &ft);
  if (rzero < rbot || rzero > rtop) {
    double g52dif = (fb * ft == 1) ? g52top - g52bot : g52top - g52bot + g5b2;
    double g32dif = d2b3 * (g52dif + zetop - zebot);
    return fabs(g32dif + co * g52dif / zeta) / (bot * sqrt(zeta * zeta * zeta));
  }
  double zrzer = zeta * rzero;
  double zezer = sqrt(zrzer * zrzer * zrzer) * exp(-zrzer);
  double g52zer = slat_gammai(d5b2, zrzer, zrzer * zezer, &fz);
  double g52dif1 = (fb * fz == 1) ? g52zer - g52bot : g52zer - g52bot + g5b2;
  double g32dif1 = d2b3 * (g52dif1 + zezer - zebot);
  double g52dif2 = (ft * fz == 1) ? g52top - g52zer : g52top - g52zer + g5b2;
  double g32dif2 = d2b3 * (g52dif2 + zetop - zezer);
  return (fabs(g32dif1 + co * g52dif1 / zeta) + fabs(g32dif2 + co * g52dif2 / zeta)) /
         (bot * sqrt(zeta * zeta * zeta));
}

double orc_slat_area(double zeta, double co, double rbot, double rtop, double bot) {
  return slat_area(zeta, co, rbot, rtop, bot);
}

static double slat_thetamx(const orc_sys *o, double r, double z) {
  return o->deltat + (2.0 - o->deltat) / (1.0 + (z * r) * (z * r));
}

/* nearest nucleus of a point from its e-n vectors rv[M][3] (squared
   distances, the first minimum; metrop_slat_mov1.f90:299-308) */
static int slat_nearest(const orc_sys *o, const double *rv) {
  double best = 99.e9;
  int ic = 0;
  for (int j = 0; j < o->ncent; j++) {
    double d = 0;
    for (int k = 0; k < 3; k++) d += rv[3 * j + k] * rv[3 * j + k];
    if (d < best) { best = d; ic = j; }
  }
  return ic;
}

/* Proposal of electron i from the position whose e-n vectors are rv[M][3],
   velocity v (grad ln Psi): xn and rec = {fxop, areao, rtry}
   (metrop_slat_mov1.f90:78-289). */
static void slat_propose(const orc_sys *o, const double *rv, const double *v, double fbias,
                         int32_t *rng, double *xn, double *rec) {
  const double eps = 1e-10, pi = 4.0 * atan(1.0), deltar = o->deltar, deltri = 1.0 / deltar;
  const int nearo = slat_nearest(o, rv);
  const double *rvm = rv + 3 * nearo;
  const double zn = o->znuc[o->iwctype[nearo]];
  const double zcusp = o->nloc == 0 ? zn : 0.0;
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
  slat_zeta_co(zcusp, vr, rmino, &zeta, &co);
  /* maximum of the radial function on [rbot, rtop] for the rejection */
  const double root = sqrt((1.5 * co - zeta) * (1.5 * co - zeta) + 2 * zeta * co);
  double rmax1 = co != 0 ? (1.5 * co - zeta + root) / (2 * zeta * co) : rtop;
  if (rmax1 < rbot) rmax1 = rbot;
  if (rmax1 > rtop) rmax1 = rtop;
  double gmax = sqrt(rmax1) * fabs(1 + co * rmax1) * exp(-zeta * rmax1);
  if (co < 0) {
    double rmax2 = (1.5 * co - zeta - root) / (2 * zeta * co);
    if (rmax2 < rbot) rmax2 = rbot;
    if (rmax2 > rtop) rmax2 = rtop;
    gmax = fmax(gmax, sqrt(rmax2) * fabs(1 + co * rmax2) * exp(-zeta * rmax2));
  }
  const double bot = sqrt(rmino) * fabs(1 + co * rmino) * exp(-zeta * rmino);
  double rtry = 0, top = 0;
  /* the reference loops without bound; after 2^20 tries of either loop
     (only non-finite input gets there) fxop = NaN and the move is rejected */
  int ok = 0;
  for (int t = 0; t < (1 << 20) && !ok; t++) {
    rtry = ((deltar - deltri) * orc_rannyu(rng) + deltri) * rmino;
    top = sqrt(rtry) * fabs(1 + co * rtry) * exp(-zeta * rtry);
    ok = top / gmax > orc_rannyu(rng);
  }
  double fxop = ok ? top / bot : NAN;
  double areao = slat_area(zeta, co, rbot, rtop, bot);
  const double raver = 0.5 * (rmino + rtry);
  const double deltt = slat_thetamx(o, raver, zn);
  const double costht = 1 - deltt * orc_rannyu(rng);
  const double zprime = rtry * costht;
  const double sintht = sqrt(1 - costht * costht);
  areao *= deltt;
  const double term = fmin(vp * raver * sintht, 1.0);
  const double fm = 1 + term;
  double phitry = 0;
  ok = 0;
  for (int t = 0; t < (1 << 20) && !ok; t++) {
    phitry = pi * orc_rannyu(rng);
    top = 1 + term * cos(phitry);
    ok = top > orc_rannyu(rng) * fm;
  }
  fxop = ok ? fxop * top : NAN;
  const double xprime = rtry * sintht * cos(phitry);
  double yprime = sqrt(fmax(0.0, rtry * rtry - xprime * xprime - zprime * zprime));
  if (orc_rannyu(rng) < 0.5) yprime = -yprime;
  for (int k = 0; k < 3; k++)
    xn[k] = (xa[k] * xprime + ya[k] * yprime + za[k] * zprime) + o->cent[3 * nearo + k];
  rec[0] = fxop;
  rec[1] = areao;
  rec[2] = rtry;
}

/* Density factors of the move from the point with e-n vectors rvf[M][3]
   and velocity vf to the point with e-n vectors rvt[M][3]
   (metrop_slat_mov1.f90:299-336, 376-475): fx and area (the density is
   fx / (area 2 pi r_to^2), r_to the distance of the destination from the
   start's nearest nucleus, returned in *rto).  Returns 0 when the
   destination is outside what the start's sampler reaches (the reference's
   geometrical rejection). */
static int slat_density(const orc_sys *o, const double *rvf, const double *vf, const double *rvt,
                        double fbias, double *fx, double *area, double *rto) {
  const double eps = 1e-10;
  const int nearn = slat_nearest(o, rvf);
  const double *rn = rvf + 3 * nearn, *ro = rvt + 3 * nearn;
  const double zn = o->znuc[o->iwctype[nearn]];
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
  const double delttn = slat_thetamx(o, ravern, zn);
  *rto = rminon;
  if (rminon > rminn * o->deltar || dot < 1 - delttn) return 0;
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
  const double zcusp = o->nloc == 0 ? zn : 0.0;
  vr = fmax(vr, -2 * zn);
  vp *= fbias;
  double zeta, co;
  slat_zeta_co(zcusp, vr, rminn, &zeta, &co);
  const double bot = sqrt(rminn) * fabs(1 + co * rminn) * exp(-zeta * rminn);
  const double top = sqrt(rminon) * fabs(1 + co * rminon) * exp(-zeta * rminon);
  double f = top / bot;
  double a = slat_area(zeta, co, rminn / o->deltar, rminn * o->deltar, bot) * delttn;
  const double term = fmin(vp * ravern * sintht, 1.0);
  double cosphi = 0, rnorm = 0;
  for (int k = 0; k < 3; k++) {
    const double t2 = ro[k] / rminon - costht * rn[k] / rminn;
    rnorm += t2 * t2;
    cosphi += t2 * xa[k];
  }
  cosphi = cosphi / sqrt(rnorm);
  f *= (1 + term * cosphi);
  *fx = f;
  *area = a;
  return 1;
}

/* test access: one proposal of electron i of walker w (its current
   velocity), and the density of a move of electron i from its current
   position to xt (in the same units: fx / (area 2 pi r_to^2)) */
void orc_slat_sample(const orc_sys *o, orc_walker *w, int32_t i, double fbias, int32_t *rng,
                     double *xn, double *rec) {
  double v[3];
  orc_grad_1e(o, w, i, v, NULL);
  slat_propose(o, w->rvec_en + (size_t)i * o->ncent * 3, v, fbias, rng, xn, rec);
}
double orc_slat_density(const orc_sys *o, orc_walker *w, int32_t i, double fbias, const double *xt) {
  double v[3], rvt[3 * 256], fx, area, rto;
  orc_grad_1e(o, w, i, v, NULL);
  for (int j = 0; j < o->ncent; j++)
    for (int k = 0; k < 3; k++) rvt[3 * j + k] = xt[k] - o->cent[3 * j + k];
  if (!slat_density(o, w->rvec_en + (size_t)i * o->ncent * 3, v, rvt, fbias, &fx, &area, &rto)) return 0.0;
  const double pi = 4.0 * atan(1.0);
  return fx / (area * 2 * pi * rto * rto);
}

/* ------------------------------------------------------------------ */
/* VMC driver: qmc/vmc/vmc.f90:250-260 (rotqua before each step),
   proposal/acceptance of qmc/vmc/metrop_mov1.f90:63-99,226-250 with the
   current-state velocity of electron i (eval_grad + Jastrow, as in
   metrop_slat_mov1.f90:205-208) and a complete commit (dete_save). */
static int vmc_walker(const orc_sys *o, orc_walker *w, double *x,
                      int32_t *rng, int nstep, double delta, double fbias,
                      double *etrace, int stride, int64_t *nacc) {
  int N = o->nelec;
  double deltfi = 2 * fbias / delta;
  int st = orc_eval_full(o, w, x, NULL, NULL, NULL);
  if (st) return st;
  int64_t acc = 0;
  for (int is = 0; is < nstep; is++) {
    if (o->nloc > 0) orc_rotqua(rng, o->xq0, o->yq0, o->zq0, o->nquad, w->xq, w->yq, w->zq);
    for (int i = 0; i < N; i++) {
      double vold[3], vnew[3], dx[3], xn[3], fxop = 1, lap;
      orc_grad_1e(o, w, i, vold, NULL);
      if (o->imetro == 6) {
        /* metrop_slat_mov1.f90:78-546 */
        double rec[3], out4[4], fx, area, rto;
        const int M = o->ncent;
        slat_propose(o, w->rvec_en + (size_t)i * M * 3, vold, fbias, rng, xn, rec);
        double rvn[3 * 256];
        for (int j = 0; j < M; j++)
          for (int k = 0; k < 3; k++) rvn[3 * j + k] = xn[k] - o->cent[3 * j + k];
        st = orc_ratio_1e(o, w, i, xn, out4, vnew, &lap);
        double p = 0.0;
        if (slat_density(o, rvn, vnew, w->rvec_en + (size_t)i * M * 3, fbias, &fx, &area, &rto)) {
          const double psi2n = 2 * (out4[1] + out4[3]);
          const double rratio = rec[2] / rto;
          p = rratio * rratio * exp(psi2n - w->psi2o) * fabs((fx * rec[1]) / (rec[0] * area));
          p = p >= 0 ? fmin(1.0, p) : 0.0;      /* NaN (an unsampled proposal): reject */
        }
        int accept = orc_rannyu(rng) < p;
        acc += accept;
        orc_accept_1e(o, w, i, accept);
        continue;
      }
      for (int k = 0; k < 3; k++) {
        double fx = fmin(fabs(vold[k]), deltfi);
        double ftest = fx * delta * 0.5;
        fx = copysign(fx, vold[k]);
        if (ftest <= orc_rannyu(rng)) {
          dx[k] = (orc_rannyu(rng) - 0.5) * delta;
        } else {
          double r1 = orc_rannyu(rng), r2 = orc_rannyu(rng);
          dx[k] = (fmax(r1, r2) - 0.5) * delta;
          if (vold[k] < 0) dx[k] = -dx[k];
        }
        fxop *= (1 + fx * dx[k]);
        xn[k] = w->x[3 * i + k] + dx[k];
      }
      double out4[4];
      st = orc_ratio_1e(o, w, i, xn, out4, vnew, &lap);
      double psi2n = 2 * (out4[1] + out4[3]);
      double fxnp = 1;
      for (int k = 0; k < 3; k++)
        fxnp *= (1 - dx[k] * copysign(fmin(fabs(vnew[k]), deltfi), vnew[k]));
      double p = exp(psi2n - w->psi2o) * fxnp / fxop;
      p = fmin(1.0, p);
      int accept = orc_rannyu(rng) < p;
      acc += accept;
      orc_accept_1e(o, w, i, accept);
    }
    memcpy(x, w->x, sizeof(double) * 3 * N);
    st = eval_full_w(o, w, 0, NULL);
    if (etrace) etrace[(size_t)is * stride] = w->eloc;
    if (st) return st;
  }
  memcpy(x, w->x, sizeof(double) * 3 * N);
  if (nacc) *nacc = acc;
  return 0;
}

int orc_vmc_run(const orc_sys *o, int32_t nw, double *x, int32_t *rng,
                int32_t nstep, double delta, double fbias, double *eloc_trace,
                int64_t *naccept, int32_t nthreads) {
  int N = o->nelec, err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    orc_walker *w = orc_walker_create(o);
#pragma omp for schedule(dynamic, 1)
    for (int iw = 0; iw < nw; iw++) {
      int st = vmc_walker(o, w, x + (size_t)iw * 3 * N, rng + 4 * iw, nstep, delta, fbias,
                          eloc_trace ? eloc_trace + iw : NULL, nw,
                          naccept ? naccept + iw : NULL);
      if (st) {
#pragma omp atomic write
        err = st;
      }
    }
    orc_walker_destroy(w);
  }
  return err;
}

int orc_eval_batch(const orc_sys *o, int32_t nw, const double *x,
                   double *out9, double *grad, double *divv, double *dpsi,
                   double *deloc, int32_t nthreads) {
  int N = o->nelec, P = o->nparm, err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    orc_walker *w = orc_walker_create(o);
#pragma omp for schedule(dynamic, 4)
    for (int iw = 0; iw < nw; iw++) {
      int st = orc_eval_full(o, w, x + (size_t)iw * 3 * N, out9 ? out9 + 9 * iw : NULL,
                             grad ? grad + (size_t)iw * 3 * N : NULL,
                             divv ? divv + (size_t)iw * N : NULL);
      if (!st && dpsi && P > 0)
        st = orc_derivs(o, w, dpsi + (size_t)iw * P, deloc + (size_t)iw * P);
      if (st) {
#pragma omp atomic write
        err = st;
      }
    }
    orc_walker_destroy(w);
  }
  return err;
}

/* copy of a walker (psi_t assignment) */
void orc_walker_copy(const orc_sys *o, orc_walker *d, const orc_walker *s) {
  int N = o->nelec, M = o->ncent, NO = o->norb, nup = o->nup, ndn = o->ndn;
  int k2 = o->kmax * o->kmax + 1;
#define CP(f, n) memcpy(d->f, s->f, sizeof(double) * (size_t)(n))
  CP(x, 3 * N); CP(rvec_en, 3 * N * M); CP(r_en, N * M); CP(rvec_ee, 3 * N * N);
  CP(r_ee, N * N); CP(fso, N * N); CP(fijo, 3 * N * N); CP(d2ijo, N * N);
  CP(lapjij, N * N); CP(fjo, 3 * N); CP(orb, N * NO); CP(dorb, 3 * N * NO);
  CP(ddorb, N * NO); CP(aiup, nup * nup); CP(aidn, ndn * ndn); CP(tup, nup * NO);
  CP(tdn, ndn * NO); CP(invup, o->ndetup * k2); CP(invdn, o->ndetdn * k2);
  CP(detup, o->ndetup); CP(detdn, o->ndetdn); CP(yup, o->noccup * nup);
  CP(ydn, o->noccdn * ndn); CP(gup, o->noccup * nup); CP(gdn, o->noccdn * ndn);
  CP(vj, 3 * N); CP(vd, 3 * N); CP(divvj, N); CP(divvd, N);
  CP(tqx, 3 * N * M * QMCW_MAX_NQUAD); CP(tqr, N * M * QMCW_MAX_NQUAD);
  memcpy(d->tqn, s->tqn, sizeof(int) * (size_t)N);
#undef CP
  d->fsumo = s->fsumo; d->d2o = s->d2o; d->deta_up = s->deta_up; d->deta_dn = s->deta_dn;
  d->chi = s->chi; d->logpsid = s->logpsid; d->sgnpsid = s->sgnpsid; d->psij = s->psij;
  d->eloc = s->eloc; d->pe = s->pe; d->psi2o = s->psi2o;
  memcpy(d->xq, s->xq, sizeof d->xq); memcpy(d->yq, s->yq, sizeof d->yq);
  memcpy(d->zq, s->zq, sizeof d->zq);
  d->pend_ie = -1;
}

void orc_walker_set_quad(orc_walker *w, const double *xq, const double *yq,
                         const double *zq, int nq) {
  memcpy(w->xq, xq, sizeof(double) * nq);
  memcpy(w->yq, yq, sizeof(double) * nq);
  memcpy(w->zq, zq, sizeof(double) * nq);
}

/* ------------------------------------------------------------------ */
/* DMC                                                                 */
/* ------------------------------------------------------------------ */
/* averaged_1e_velocity, qmc/green_functions_mod.f90:598-633, with
   adrift = adrift_default (no modified adrift, no nucleus term) */
static void dmc_vav(double tau, double adrift, int ndrift, const double *g, double *v) {
  double vn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  double vav;
  if (adrift * vn * vn * tau <= 0) vav = vn;
  else vav = (-1 + pow(1 + ndrift * adrift * vn * vn * tau, 1.0 / ndrift)) / (adrift * vn * tau);
  for (int k = 0; k < 3; k++) v[k] = vav * g[k] / vn;
}

/* sample_1e_prob_drift_diffusion, qmc/green_functions_mod.f90:257-265 */
static double dmc_prob(double tau, const double *vav, const double *xo, const double *xn) {
  const double pi = 3.141592653589793;
  double d2 = 0;
  for (int k = 0; k < 3; k++) {
    double d = xn[k] - xo[k] - vav[k] * tau;
    d2 += d * d;
  }
  double nrm = sqrt(2 * pi * tau);
  return exp(-d2 / (2 * tau)) / (nrm * nrm * nrm);
}

/* all-electron mover: geometry of the drift-diffusion move with the nucleus
   cusp from an electron with e-n vectors rv[M][3] (electron minus centre),
   distances r[M] and gradient g.  The local frame and the bent drift point
   follow qmc/green_functions_mod.f90:93-131 (same as :276-302); the velocity
   uses the nucleus adrift of averaged_1e_velocity (:616-622). */
typedef struct {
  double z[3], x[3], xdrift[3], c[3], pgaus, zeta;
} cusp_geo;

static void cusp_geo_at(const orc_sys *o, double tau, int ndrift, const double *g, const double *rv,
                        const double *r, cusp_geo *cg) {
  const double eps = 1e-10, adrift0 = 0.1;
  int M = o->ncent, iw = 0;
  for (int ic = 1; ic < M; ic++)
    if (r[ic] < r[iw]) iw = ic;
  const double *xe = rv + 3 * iw;
  double zn = sqrt(xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2]);
  for (int k = 0; k < 3; k++) cg->z[k] = xe[k] / zn;
  double gz = g[0] * cg->z[0] + g[1] * cg->z[1] + g[2] * cg->z[2];
  for (int k = 0; k < 3; k++) cg->x[k] = g[k] - gz * cg->z[k];
  if (sqrt(cg->x[0] * cg->x[0] + cg->x[1] * cg->x[1] + cg->x[2] * cg->x[2]) < eps) {
    cg->x[0] = eps * (1 - cg->z[0] * cg->z[0]);
    cg->x[1] = eps * (-cg->z[0] * cg->z[1]);
    cg->x[2] = eps * (-cg->z[0] * cg->z[2]);
  }
  double xn = sqrt(cg->x[0] * cg->x[0] + cg->x[1] * cg->x[1] + cg->x[2] * cg->x[2]);
  for (int k = 0; k < 3; k++) cg->x[k] /= xn;
  double Z = o->znuc[o->iwctype[iw]];
  double vn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  double hafzr2 = (Z * zn / 2) * (Z * zn / 2);
  double adrift = (1 + eps + (xe[0] * g[0] + xe[1] * g[1] + xe[2] * g[2]) / (vn * zn)) / 2 +
                  adrift0 * hafzr2 / (1 + hafzr2);
  double v[3];
  dmc_vav(tau, adrift, ndrift, g, v);
  double zi = r[iw];
  double zf = zi + (v[0] * cg->z[0] + v[1] * cg->z[1] + v[2] * cg->z[2]) * tau;
  if (zf >= 0) cg->pgaus = 1 - erfc(zf / sqrt(2 * tau)) / 2;
  else cg->pgaus = erfc(-zf / sqrt(2 * tau)) / 2;
  zf = fmax(0.0, zf);
  double xf = (v[0] * cg->x[0] + v[1] * cg->x[1] + v[2] * cg->x[2]) * tau * fmin(1.0, zf / ((zf + zi) / 2));
  for (int k = 0; k < 3; k++) {
    cg->c[k] = o->cent[3 * iw + k];
    cg->xdrift[k] = cg->c[k] + xf * cg->x[k] + zf * cg->z[k];
  }
  cg->zeta = sqrt(1 / tau + Z * Z);
}

/* sample_1e_prob_drift_diffusion_with_cusp, qmc/green_functions_mod.f90:303-319 */
static double cusp_prob(const cusp_geo *cg, double tau, const double *xn) {
  const double pi = 3.141592653589793;
  double d2 = 0, rn2 = 0;
  for (int k = 0; k < 3; k++) {
    d2 += (xn[k] - cg->xdrift[k]) * (xn[k] - cg->xdrift[k]);
    rn2 += (xn[k] - cg->c[k]) * (xn[k] - cg->c[k]);
  }
  double nrm = sqrt(2 * pi * tau);
  return cg->pgaus * exp(-d2 / (2 * tau)) / (nrm * nrm * nrm) +
         (1 - cg->pgaus) * cg->zeta * cg->zeta * cg->zeta * exp(-2 * cg->zeta * sqrt(rn2)) / pi;
}

/* sample_1e_move_drift_diffusion_with_cusp, qmc/green_functions_mod.f90:132-144:
   gaussian about the drift point with probability pgaus, otherwise the
   exponential zeta^3 exp(-2 zeta r) / pi about the nearest nucleus */
static void cusp_sample(const cusp_geo *cg, double tau, int32_t *rng, double *xn) {
  const double pi = 3.141592653589793;
  if (orc_rannyu(rng) < cg->pgaus) {
    for (int k = 0; k < 3; k++) xn[k] = cg->xdrift[k] + sqrt(tau) * orc_gauss(rng);
  } else {
    double u1 = orc_rannyu(rng), u2 = orc_rannyu(rng), u3 = orc_rannyu(rng);
    double r = (-0.5 / cg->zeta) * log(u1 * u2 * u3);
    double costht = 2 * (orc_rannyu(rng) - 0.5);
    double sintht = sqrt(1 - costht * costht);
    double phi = 2 * pi * orc_rannyu(rng);
    xn[0] = cg->c[0] + r * sintht * cos(phi);
    xn[1] = cg->c[1] + r * sintht * sin(phi);
    xn[2] = cg->c[2] + r * costht;
  }
}

static void en_vectors(const orc_sys *o, const double *xe, double *rv, double *r) {
  for (int ic = 0; ic < o->ncent; ic++) {
    double s = 0;
    for (int k = 0; k < 3; k++) {
      rv[3 * ic + k] = xe[k] - o->cent[3 * ic + k];
      s += rv[3 * ic + k] * rv[3 * ic + k];
    }
    r[ic] = sqrt(s);
  }
}

/* the all-electron mover for an electron at xe with gradient g: one sample
   (xn, or NULL) from the walker stream rng and the proposal density at pt
   (or NULL; return value 0) */
double orc_cusp_move(const orc_sys *o, double tau, int32_t ndrift, const double *g, const double *xe,
                     int32_t *rng, double *xn, const double *pt) {
  double rv[3 * 64], r[64];
  if (o->ncent > 64) return -1;
  en_vectors(o, xe, rv, r);
  cusp_geo cg;
  cusp_geo_at(o, tau, ndrift, g, rv, r, &cg);
  if (xn) cusp_sample(&cg, tau, rng, xn);
  return pt ? cusp_prob(&cg, tau, pt) : 0.0;
}

/* T-moves of one walker, qmc/dmc/dmc_elec/tmove.f90:1-182 with
   discrete_rand (:184-205) and quad_psi (:207-255): for each electron in
   turn, the weights max(0, sum_l w_q (exp(-tau v_l) - 1) P_l(cos) Psi(x_q) /
   Psi(x)) of moving it to the quadrature points of its pairs in range (the
   points and ratios of the last DMC evaluation until a T-move is accepted,
   then fresh rotations and ratios), against 1 for staying; a chosen move is
   accepted with min(1, forward sum / backward sum) (det_balance1; the
   backward points are a fresh rotation about the same centre, as the
   reference, which assumes non-overlapping nonlocal ranges) or always
   (casula1).  Returns 1 if any T-move was accepted. */
static int tmove_w(const orc_sys *o, orc_walker *w, int32_t *rng, double tau, int casula) {
  const int N = o->nelec, M = o->ncent, nq = o->nquad;
  int tacc = 0;
  double *gw = malloc(sizeof(double) * ((size_t)M * nq + 1));
  double *xs = malloc(sizeof(double) * 3 * ((size_t)M * nq + 1));
  double *pr = malloc(sizeof(double) * ((size_t)M * nq + 1));
  int *wc = malloc(sizeof(int) * ((size_t)M * nq + 1));
  double vps[QMCW_MAX_LPOT], gx[QMCW_MAX_NQUAD], gy[QMCW_MAX_NQUAD], gz[QMCW_MAX_NQUAD];
  double out4[4], gn[3], lap;
  for (int i = 0; i < N; i++) {
    gw[0] = 1;
    int n = 0;
    for (int ic = 0; ic < M; ic++) {
      int ict = o->iwctype[ic];
      double r = fmax(1e-10, w->r_en[i * M + ic]);
      int iskip = 1;
      for (int l = 1; l <= o->npotd[ict]; l++) {
        vps[l - 1] = gauss_pot(o, r, l - 1, ict);
        if (l != o->lpotp1[ict] && fabs(vps[l - 1]) > 1e-4) iskip = 0;
      }
      if (iskip) continue;
      if (tacc) orc_rotqua(rng, o->xq0, o->yq0, o->zq0, nq, gx, gy, gz);
      double rei = w->r_en[i * M + ic];
      const double *rv0 = w->rvec_en + (i * M + ic) * 3;
      for (int iq = 0; iq < nq; iq++) {
        double q[3];
        size_t k = (size_t)i * M * QMCW_MAX_NQUAD + n;
        if (tacc) { q[0] = gx[iq]; q[1] = gy[iq]; q[2] = gz[iq]; }
        else { q[0] = w->tqx[3 * k]; q[1] = w->tqx[3 * k + 1]; q[2] = w->tqx[3 * k + 2]; }
        double costh = (rv0[0] * q[0] + rv0[1] * q[1] + rv0[2] * q[2]) / rei;
        double *xt = xs + 3 * n;
        for (int d = 0; d < 3; d++) xt[d] = rei * q[d] + o->cent[3 * ic + d];
        double ratio;
        if (tacc) {
          orc_ratio_1e(o, w, i, xt, out4, gn, &lap);
          ratio = out4[0];
        } else {
          ratio = w->tqr[k];
        }
        double gpsp = 0;
        for (int l = 1; l <= o->npotd[ict]; l++)
          if (l != o->lpotp1[ict]) gpsp += o->wq[iq] * (exp(-tau * vps[l - 1]) - 1) * yl0(l, costh) * ratio;
        pr[n] = ratio;
        gw[1 + n] = fmax(0.0, gpsp);
        wc[n] = ic;
        n++;
      }
    }
    /* discrete_rand over {stay, the n points} */
    double tot = 0;
    for (int k = 0; k <= n; k++) tot += gw[k];
    double rnd = tot * orc_rannyu(rng), psum = 0;
    int which = n;
    for (int k = 0; k <= n; k++) {
      psum += gw[k];
      if (psum > rnd) { which = k; break; }
    }
    if (which == 0) continue;
    int k = which - 1, ic = wc[k], ict = o->iwctype[ic];
    double xnew[3] = {xs[3 * k], xs[3 * k + 1], xs[3 * k + 2]}, xen[3];
    for (int d = 0; d < 3; d++) xen[d] = xnew[d] - o->cent[3 * ic + d];
    double p = 1;
    if (!casula) {
      double bwd = 1;
      double rei = w->r_en[i * M + ic], r = fmax(1e-10, rei);
      for (int l = 1; l <= o->npotd[ict]; l++) vps[l - 1] = gauss_pot(o, r, l - 1, ict);
      orc_rotqua(rng, o->xq0, o->yq0, o->zq0, nq, gx, gy, gz);
      for (int jq = 0; jq < nq; jq++) {
        double costh = (xen[0] * gx[jq] + xen[1] * gy[jq] + xen[2] * gz[jq]) / rei;
        double xt[3] = {rei * gx[jq] + o->cent[3 * ic], rei * gy[jq] + o->cent[3 * ic + 1],
                        rei * gz[jq] + o->cent[3 * ic + 2]};
        orc_ratio_1e(o, w, i, xt, out4, gn, &lap);
        double ratio = out4[0] / pr[k];
        double gpsp = 0;
        for (int l = 1; l <= o->npotd[ict]; l++)
          if (l != o->lpotp1[ict]) gpsp += o->wq[jq] * (exp(-tau * vps[l - 1]) - 1) * yl0(l, costh) * ratio;
        bwd += fmax(0.0, gpsp);
      }
      p = fmin(1.0, tot / bwd);
    }
    if (orc_rannyu(rng) < p || casula) {
      tacc = 1;
      orc_ratio_1e(o, w, i, xnew, out4, gn, &lap);
      orc_accept_1e(o, w, i, 1);
    }
  }
  free(gw); free(xs); free(pr); free(wc);
  return tacc;
}

/* tests: a DMC-mode evaluation of walker w at x (grid rotations from rng),
   then its T-moves; returns 1 if one was accepted, -1 on an evaluation
   error; xout: the positions after the T-moves */
int orc_tmove_run(const orc_sys *o, orc_walker *w, const double *x, int32_t *rng, double tau,
                  int32_t casula, double *xout) {
  memcpy(w->x, x, sizeof(double) * 3 * o->nelec);
  if (eval_full_w(o, w, 0, rng)) return -1;
  int a = tmove_w(o, w, rng, tau, casula);
  if (xout) memcpy(xout, w->x, sizeof(double) * 3 * o->nelec);
  return a;
}

/* tests: the quadrature points (unit vectors) and Psi ratios the last
   DMC-mode evaluation kept for electron i; returns their number */
int32_t orc_walker_tq(const orc_sys *o, const orc_walker *w, int32_t i, double *q, double *r) {
  int n = w->tqn[i];
  size_t k = (size_t)i * o->ncent * QMCW_MAX_NQUAD;
  if (q) memcpy(q, w->tqx + 3 * k, sizeof(double) * 3 * n);
  if (r) memcpy(r, w->tqr + k, sizeof(double) * n);
  return n;
}

/* ene_int_rewt, qmc/green_functions_mod.f90:382-472 (ene_int_v9 or
   no_ene_int) */
static double dmc_eint(const qmcw_dmc_params *p, int N, double tauint, double eo,
                       const double *go, double en, const double *gn) {
  const double pi = 3.141592653589793;
  double ewto, ewtn;
  if (p->ene_int == 1) {
    ewto = eo;
    ewtn = en;
  } else {
    double so = 0, sn = 0;
    for (int q = 0; q < 3 * N; q++) { so += go[q] * go[q]; sn += gn[q] * gn[q]; }
    double facto = fmax(sqrt(so) / pow((double)N, p->pow_rewt) * tauint * p->c_rewt, 1e-9);
    double factn = fmax(sqrt(sn) / pow((double)N, p->pow_rewt) * tauint * p->c_rewt, 1e-9);
    ewto = p->eest + (eo - p->eest) * (sqrt(pi) / 2) * erf(facto) / facto;
    ewtn = p->eest + (en - p->eest) * (sqrt(pi) / 2) * erf(factn) / factn;
  }
  double eint = (ewto + ewtn) / 2 - p->etrial;
  double rtrttau = sqrt(sqrt(p->tau));
  double emin = p->eest - p->etrial - p->limit_rewt * p->esigma / rtrttau;
  return eint < emin ? emin : eint;
}

struct orc_dmc {
  const orc_sys *o;
  int cap, nwalk;
  orc_walker **w;
  int32_t *rng;       /* [cap][4] */
  double *weight;     /* [cap] */
  double *wout;       /* [cap][8] per-walker generation results */
};

orc_dmc *orc_dmc_create(const orc_sys *o, int32_t cap) {
  orc_dmc *d = calloc(1, sizeof(orc_dmc));
  d->o = o;
  d->cap = cap;
  d->w = calloc(cap, sizeof(orc_walker *));
  for (int i = 0; i < cap; i++) d->w[i] = orc_walker_create(o);
  d->rng = calloc((size_t)cap * 4, sizeof(int32_t));
  d->weight = calloc(cap, sizeof(double));
  d->wout = calloc((size_t)cap * 8, sizeof(double));
  return d;
}

void orc_dmc_destroy(orc_dmc *d) {
  if (!d) return;
  for (int i = 0; i < d->cap; i++) orc_walker_destroy(d->w[i]);
  free(d->w); free(d->rng); free(d->weight); free(d->wout);
  free(d);
}

/* population of nwalk walkers at x with their streams and weights (NULL:
   1); each gets a full evaluation (psi_at) in DMC mode: the pseudopotential
   grid is rotated per pair in range, from the walker's stream */
int orc_dmc_set(orc_dmc *d, int32_t nwalk, const double *x, const int32_t *rng,
                const double *weight) {
  int N = d->o->nelec, err = 0;
  if (nwalk > d->cap) return QMCW_ERR_ARG;
  d->nwalk = nwalk;
  for (int i = 0; i < nwalk; i++) {
    memcpy(d->rng + 4 * i, rng + 4 * i, 4 * sizeof(int32_t));
    memcpy(d->w[i]->x, x + (size_t)i * 3 * N, sizeof(double) * 3 * N);
    int st = eval_full_w(d->o, d->w[i], 0, d->rng + 4 * i);
    if (st) err = st;
    d->weight[i] = weight ? weight[i] : 1.0;
  }
  return err;
}

/* one walker's generation: qmc/rundmc_mod.f90:114-303 with the mover of
   sample_1e_move (green_functions_mod.f90:12-38): drift-diffusion
   (:70-80) with pseudopotentials, drift-diffusion with the nucleus cusp
   (:83-145) for all-electron systems; the walker's own rannyu stream for
   the grid rotation, the mover and the Metropolis tests */
static int dmc_walker(const orc_sys *o, orc_walker *w, int32_t *rng, double *weight,
                      const qmcw_dmc_params *p, double *res) {
  int N = o->nelec;
  double go[3 * 128 + 3], eo = w->eloc;
  for (int q = 0; q < 3 * N; q++) go[q] = w->vj[q] + w->vd[q];
  if (o->nloc > 0) orc_rotqua(rng, o->xq0, o->yq0, o->zq0, o->nquad, w->xq, w->yq, w->zq);
  double dd[128], pacc[128], sdd = 0, spd = 0, sad = 0, pa_sum = 0, nacc = 0, ncross = 0;
  int l_accept = 0, st = 0;
  /* rundmc_mod.f90:137: T-moves first; any accepted one makes the walker
     re-evaluated at the end */
  if (p->tmoves && o->nloc > 0) l_accept = tmove_w(o, w, rng, p->tau, p->tmoves == 2);
  double rt = sqrt(p->tau);
  for (int ie = 0; ie < N; ie++) {
    double gold[3], vo[3], vn[3], gnew[3], xo[3], xn[3], out4[4], lap, po, pn, d2 = 0;
    orc_grad_1e(o, w, ie, gold, NULL);
    for (int k = 0; k < 3; k++) xo[k] = w->x[3 * ie + k];
    cusp_geo cg;
    if (o->nloc == 0) {
      cusp_geo_at(o, p->tau, p->n_drift, gold, w->rvec_en + 3 * ie * o->ncent,
                  w->r_en + ie * o->ncent, &cg);
      cusp_sample(&cg, p->tau, rng, xn);
      for (int k = 0; k < 3; k++) d2 += (xn[k] - cg.xdrift[k]) * (xn[k] - cg.xdrift[k]);
    } else {
      dmc_vav(p->tau, p->adrift, p->n_drift, gold, vo);
      for (int k = 0; k < 3; k++) {
        double xd = rt * orc_gauss(rng);
        xn[k] = xo[k] + vo[k] * p->tau + xd;
        d2 += xd * xd;
      }
    }
    dd[ie] = d2;
    int s1 = orc_ratio_1e(o, w, ie, xn, out4, gnew, &lap);
    if (s1) st = s1;
    if (o->nloc == 0) {
      /* green_functions_mod.f90:41-44 for both directions; the backward
         move starts from the proposed point with its e-n vectors */
      po = cusp_prob(&cg, p->tau, xn);
      cusp_geo cgn;
      cusp_geo_at(o, p->tau, p->n_drift, gnew, w->rvec_en_n, w->r_en_n, &cgn);
      pn = cusp_prob(&cgn, p->tau, xo);
    } else {
      dmc_vav(p->tau, p->adrift, p->n_drift, gnew, vn);
      po = dmc_prob(p->tau, vo, xo, xn);
      pn = dmc_prob(p->tau, vn, xn, xo);
    }
    double psirat = out4[0] * out4[0];
    double pa = fmin(1.0, psirat * pn / po);
    if (p->iacc_rej == 0) pa = 1.0;
    if (out4[0] <= 0) {
      if (p->icross <= 0) pa = 0.0;
      ncross += 1;
    }
    pacc[ie] = pa;
    int acc = orc_rannyu(rng) < pa;
    orc_accept_1e(o, w, ie, acc);
    l_accept |= acc;
    nacc += acc;
    pa_sum += pa;
    sdd += d2;
    spd += pa * d2;
    sad += acc * d2;
  }
  if (l_accept) {
    int s2 = eval_full_w(o, w, 0, rng);
    if (s2) st = s2;
  }
  double taunow = p->tau;
  if (p->itau_eff >= 1) taunow = p->tau * (p->itau_eff == 2 ? sad : spd) / sdd;
  double tauint = p->tau;
  if (p->itau_integ >= 1) tauint = (p->itau_eff == 3) ? p->taueff : taunow;
  double wo = *weight;
  if (p->idmc > 0) {
    double gn[3 * 128 + 3];
    for (int q = 0; q < 3 * N; q++) gn[q] = w->vj[q] + w->vd[q];
    double eint = dmc_eint(p, N, tauint, eo, go, w->eloc, gn);
    double enetau = eint * ((p->itau_eff == 3) ? p->taueff : taunow);
    wo = wo * exp(-enetau) * p->pop_ctrl;
  }
  *weight = wo;
  res[0] = taunow; res[1] = pa_sum; res[2] = nacc; res[3] = ncross;
  (void)pacc; (void)dd;
  return st;
}

int orc_dmc_generation(orc_dmc *d, const qmcw_dmc_params *p, qmcw_dmc_stats *stats,
                       int32_t nthreads) {
  int err = 0;
  if (d->o->nelec > 128) return QMCW_ERR_ARG;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < d->nwalk; i++) {
    int st = dmc_walker(d->o, d->w[i], d->rng + 4 * i, d->weight + i, p, d->wout + 8 * i);
    if (st) {
#pragma omp atomic write
      err = st;
    }
  }
  memset(stats, 0, sizeof(*stats));
  for (int i = 0; i < d->nwalk; i++) {
    double w = d->weight[i], e = d->w[i]->eloc;
    stats->wsum += w;
    stats->wesum += w * e;
    stats->we2sum += w * e * e;
    stats->wtausum += w * d->wout[8 * i];
    stats->pacc += d->wout[8 * i + 1];
    stats->nacc += d->wout[8 * i + 2];
    stats->ncross += d->wout[8 * i + 3];
    stats->ntry += d->o->nelec;
  }
  stats->nwalk = d->nwalk;
  return err;
}

/* a split copy continues on a stream drawn from the branching stream:
   four 12-bit limbs, the last one odd as setrn makes it */
static void dmc_new_stream(int32_t *brng, int32_t *dst) {
  for (int k = 0; k < 4; k++) dst[k] = (int32_t)(orc_rannyu(brng) * 4096.0) & 4095;
  dst[3] |= 1;
}

static int cmp_dbl(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* split_join, qmc/dmc/dmc_elec/split_join_mod.f90:3-219: join pairs of
   walkers with weight < 1/2, kill zero weights, split weights above
   max(2, wt_split) into holes or at the end, then close the holes. */
int orc_dmc_branch(orc_dmc *d, int32_t *brng, qmcw_dmc_stats *stats) {
  const orc_sys *o = d->o;
  int n = d->nwalk, MW = d->cap;
  int *iwundr = malloc(sizeof(int) * (MW + 1));
  double *wsav = malloc(sizeof(double) * (MW + 1));
  double *w = d->weight;
  int iunder = 0, ipair = 0, iw2 = 0, nbr = 0;
  for (int iw = 0; iw < n; iw++) {
    if (w[iw] < 0.5) {
      if (w[iw] == 0.0) {
        nbr++;
        iwundr[iunder++] = iw;
      } else if (ipair == 0) {
        ipair = 1;
        iw2 = iw;
      } else {
        nbr++;
        ipair = 0;
        double wt = w[iw] + w[iw2];
        if (orc_rannyu(brng) > w[iw] / wt) {
          w[iw2] = wt;
          iwundr[iunder++] = iw;
        } else {
          w[iw] = wt;
          iwundr[iunder++] = iw2;
        }
      }
    }
  }
  int nafter = n - iunder, nmax_dupl = MW - nafter;
  for (int iw = 0; iw < n; iw++) wsav[iw] = w[iw];
  qsort(wsav, n, sizeof(double), cmp_dbl);
  /* 1-based wt_sav(max(1, n - nmax_dupl)) */
  int i1 = n - nmax_dupl; if (i1 < 1) i1 = 1;
  double wt_split = fmax(2.0, wsav[i1 - 1]);
  int nmore = 0;
  for (int iw = i1; iw <= n; iw++) {
    double r = wsav[iw - 1] / wt_split;
    if (r >= 8.0) nmore += 3;
    else if (r >= 4.0) nmore += 2;
    else if (r >= 2.0) nmore += 1;
  }
  if (n - nmax_dupl + nmore < n) {
    int i2 = n - nmax_dupl + nmore; if (i2 < 1) i2 = 1;
    wt_split = fmax(2.0, wsav[i2 - 1]);
  } else {
    wt_split = 1e99;
  }
  int n2 = n, err = 0;
  for (int iw = 0; iw < n; iw++) {
    if (w[iw] > wt_split) {
      nbr++;
      int dst;
      if (iunder > 0) dst = iwundr[--iunder];
      else {
        if (n2 >= MW) { err = QMCW_ERR_NOMEM; break; }
        dst = n2++;
      }
      w[iw] *= 0.5;
      w[dst] = w[iw];
      orc_walker_copy(o, d->w[dst], d->w[iw]);
      dmc_new_stream(brng, d->rng + 4 * dst);
    }
  }
  for (int j = iunder - 1; j >= 0; j--) {
    int dst = iwundr[j], src = n2 - 1;
    n2--;
    if (dst != src) {
      w[dst] = w[src];
      orc_walker_copy(o, d->w[dst], d->w[src]);
      memcpy(d->rng + 4 * dst, d->rng + 4 * src, 4 * sizeof(int32_t));
    }
  }
  d->nwalk = n2;
  if (stats) { stats->nwalk = n2; stats->nbranch = nbr; }
  free(iwundr);
  free(wsav);
  return err;
}

int32_t orc_dmc_get(const orc_dmc *d, double *x, int32_t *rng, double *weight, double *eloc) {
  int N = d->o->nelec;
  for (int i = 0; i < d->nwalk; i++) {
    if (x) memcpy(x + (size_t)i * 3 * N, d->w[i]->x, sizeof(double) * 3 * N);
    if (rng) memcpy(rng + 4 * i, d->rng + 4 * i, 4 * sizeof(int32_t));
    if (weight) weight[i] = d->weight[i];
    if (eloc) eloc[i] = d->w[i]->eloc;
  }
  return d->nwalk;
}

/* ---------------------------------------------------------------------------
 * Sums over W samples of the products the linear method averages, as CHAMP's
 * per-sample objects define them: dpsi_dpsi(i,j) = dpsi(i) dpsi(j)
 * (qmc/deriv_mod.f90:719), dpsi_dpsi_eloc = dpsi_dpsi * eloc (:924),
 * dpsi_eloc = dpsi * eloc (dpsi_eloc_bld, :960-1000) and dpsi_deloc(i,j) =
 * dpsi(i) deloc(j) (:1406-1410), with the plain sums of dpsi, deloc and eloc.
 * der: [W][2P] rows dpsi[P], deloc[P]; out as the engine's accumulator:
 * [0] W, [1] sum E, [2] sum E^2, sum dpsi[P], sum E dpsi[P], sum deloc[P],
 * then row-major P x P sums of dpsi dpsi, E dpsi dpsi, dpsi deloc.
 */
void orc_accum(int32_t P, int32_t W, const double *der, const double *eloc, double *out) {
  const int64_t PP = (int64_t)P * P;
  double *sd = out + 3, *sde = sd + P, *sl = sde + P;
  double *mdd = sl + P, *mdde = mdd + PP, *mdl = mdde + PP;
  memset(out, 0, (size_t)(3 + 3 * (int64_t)P + 3 * PP) * sizeof(double));
  for (int32_t w = 0; w < W; w++) {
    const double *dp = der + (int64_t)w * 2 * P, *dl = dp + P, e = eloc[w];
    out[0] += 1.0;
    out[1] += e;
    out[2] += e * e;
    for (int32_t i = 0; i < P; i++) {
      sd[i] += dp[i];
      sde[i] += dp[i] * e;
      sl[i] += dl[i];
      for (int32_t j = 0; j < P; j++) {
        const double dd = dp[i] * dp[j];
        mdd[(int64_t)i * P + j] += dd;
        mdde[(int64_t)i * P + j] += dd * e;
        mdl[(int64_t)i * P + j] += dp[i] * dl[j];
      }
    }
  }
}
