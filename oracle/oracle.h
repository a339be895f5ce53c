/*
 * oracle.h -- FP64 CPU restatement of CHAMP's wavefunction hot path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or the
 * CPU baseline).  The product (champ_amd, libchamp_amd.so) never calls it.
 *
 * It follows CHAMP's Fortran routine by routine (file:line cited on each
 * function in oracle.c) on the system description of include/champ_amd.h.
 */
#ifndef QMC_ORACLE_H
#define QMC_ORACLE_H

#include <stdint.h>
#include "../include/champ_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;
typedef struct orc_walker orc_walker;

orc_sys *orc_create(const qmcw_system *s);
void orc_destroy(orc_sys *o);
const char *orc_error(void);

/* derived setup quantities (for tests) */
int32_t orc_nterms(const orc_sys *o);
int32_t orc_nparm(const orc_sys *o);
void orc_asymp(const orc_sys *o, double *asymp_jasa, double *asymp_jasb);
/* cusp constraints: d[neqs][nterms], iwc4[neqs] (1-based) */
int32_t orc_cusp(const orc_sys *o, double *d, int32_t *iwc4);
/* apply cuspexact4 to c[nctype][nterms] in place */
void orc_cuspexact(const orc_sys *o, double *c);
/* error studies: bit 0 rounds the full orbital rows' contraction inputs to
   bfloat16, bit 1 the value-only (quadrature point) rows' */
void orc_set_emulation(orc_sys *o, int32_t mode);
/* VMC mover of orc_vmc_run: imetro 1 (metrop_mov1) or 6 (metrop_slat_mov1,
   with deltar, deltat) */
void orc_set_mover(orc_sys *o, int32_t imetro, double deltar, double deltat);
void orc_set_params(orc_sys *o, const double *a4, const double *b,
                    const double *c, const double *csf_coef,
                    const double *coef);
/* excitation bookkeeping: counts [ndetup, ndetdn, noccup, noccdn, kmax] */
void orc_det_info(const orc_sys *o, int32_t *info);

/* RNG (qmc/vmc/rannyu.f90) on an explicit 4-limb state */
void orc_setrn(int32_t *state, const int32_t *seed);
double orc_rannyu(int32_t *state);
double orc_gauss(int32_t *state);
void orc_quadrature(const orc_sys *o, double *xq, double *yq, double *zq,
                    double *wq);
/* real solid harmonic S_lm (r^l Y_lm, <Y^2> = 4 pi/(2l+1)) at x: out[0]
   the value, out[1..3] the gradient */
void orc_solid_harmonic(int32_t l, int32_t m, const double *x, double *out);
/* tabulated radial function k (value, first and second derivative) at n
   distances: out[3 * i + 0..2] */
void orc_num_radial(const orc_sys *o, int32_t k, int32_t n, const double *r,
                    double *out);
void orc_rotqua(int32_t *state, const double *xq0, const double *yq0,
                const double *zq0, int32_t nq, double *xq, double *yq,
                double *zq);
/* matinv in place on column-major a[n*n]; returns det */
double orc_matinv(double *a, int32_t n);

/* basis values at one point relative to all centres:
   phi[nbasis], dphi[nbasis][3], d2phi[nbasis] */
void orc_basis(const orc_sys *o, const double *xe, double *phi, double *dphi,
               double *d2phi);
/* orbitals at one point: orb[norb], dorb[norb][3], ddorb[norb] */
void orc_orbitals(const orc_sys *o, const double *xe, double *orb,
                  double *dorb, double *ddorb);

/* walker */
orc_walker *orc_walker_create(const orc_sys *o);
void orc_walker_destroy(orc_walker *w);
void orc_walker_copy(const orc_sys *o, orc_walker *dst, const orc_walker *src);

/* hpsi: outputs as in qmcw_eval_full; also fills the walker state.
   out9: [0] ln|psid|, [1] sign, [2] psij, [3] eloc, [4] pe (total pot),
   [5] d2lnpsi, [6] vpsp (nonlocal part), [7] pe_ee, [8] pe local part */
int orc_eval_full(const orc_sys *o, orc_walker *w, const double *x,
                  double *out9, double *grad, double *divv);
/* Jastrow only (jastrow4 / deriv_jastrow4): value, v[N][3], d2, div_vj[N];
   gvalue[PJ], g[PJ][N][3], d2g[PJ] when non-NULL */
void orc_jastrow_full(const orc_sys *o, const double *x, double *value,
                      double *v, double *d2, double *divvj, double *gvalue,
                      double *g, double *d2g);
/* determinant only (determinant.f90): ln|psid|, sign, vd[N][3], divvd[N] */
int orc_determinant_full(const orc_sys *o, const double *x, double *logdet,
                         double *sgn, double *vd, double *divvd);

/* one-electron move (hpsie + determinante + jastrow4e), pending:
   out4: [0] ratio psi_new/psi_old, [1] ln|psid_new|, [2] sign_new,
   [3] psij_new; vnew[3] grad of electron ie; lap_new = (lap Psi)/Psi */
int orc_ratio_1e(const orc_sys *o, orc_walker *w, int32_t ie,
                 const double *xnew, double *out4, double *vnew,
                 double *lap_new);
void orc_accept_1e(const orc_sys *o, orc_walker *w, int32_t ie, int accept);
/* current gradient / (lap Psi)/Psi of electron ie (calc_1e_derivatives) */
void orc_grad_1e(const orc_sys *o, orc_walker *w, int32_t ie, double *grad,
                 double *lap);
/* derivatives at the last full evaluation: dpsi[P], deloc[P] */
int orc_derivs(const orc_sys *o, orc_walker *w, double *dpsi, double *deloc);
/* sums over samples of the linear method's derivative products (layout as qmcw_accum_get) */
void orc_accum(int32_t P, int32_t W, const double *der, const double *eloc, double *out);
/* walker accessors */
void orc_walker_set_quad(orc_walker *w, const double *xq, const double *yq,
                         const double *zq, int nq);
void orc_walker_get(const orc_sys *o, const orc_walker *w, double *x,
                    double *psi2o);

/* Batched VMC (metrop_mov1 proposal, full eval per step) over nw walkers
   with OpenMP across walkers.  x[nw][N][3] in/out, rng[nw][4] in/out,
   eloc_trace[nstep][nw] (may be NULL), naccept[nw] (may be NULL). */
int orc_vmc_run(const orc_sys *o, int32_t nw, double *x, int32_t *rng,
                int32_t nstep, double delta, double fbias, double *eloc_trace,
                int64_t *naccept, int32_t nthreads);

/* Batched full evaluation: out9[nw][9], grad[nw][N][3], divv[nw][N],
   dpsi/deloc[nw][P] (may be NULL) */
int orc_eval_batch(const orc_sys *o, int32_t nw, const double *x,
                   double *out9, double *grad, double *divv, double *dpsi,
                   double *deloc, int32_t nthreads);

/* DMC (qmc/rundmc_mod.f90; drift-diffusion mover with pseudopotentials,
   drift-diffusion with the nucleus cusp for all-electron systems): a
   population of up to cap walkers with per-walker rannyu streams; one
   generation = moves of every electron, full evaluation of walkers that
   moved, reweighting; branching = split_join with its own stream. */
/* The all-electron mover (qmc/green_functions_mod.f90:83-145, 268-323) of an
   electron at xe[3] with gradient g[3]: a sample xn[3] (NULL: none) drawn
   from rng, and the proposal density at pt[3] (NULL: returns 0). */
double orc_cusp_move(const orc_sys *o, double tau, int32_t ndrift, const double *g, const double *xe,
                     int32_t *rng, double *xn, const double *pt);

/* T-moves (qmc/dmc/dmc_elec/tmove.f90), for tests: a DMC-mode evaluation
   of w at x (grid rotations drawn from rng) followed by the T-moves of every
   electron (casula 0: det_balance1, 1: casula1); xout: positions after. */
int orc_tmove_run(const orc_sys *o, orc_walker *w, const double *x, int32_t *rng, double tau,
                  int32_t casula, double *xout);
/* The quadrature points q[n][3] and Psi ratios r[n] kept for electron i by
   the last DMC-mode evaluation; returns n. */
int32_t orc_walker_tq(const orc_sys *o, const orc_walker *w, int32_t i, double *q, double *r);

typedef struct orc_dmc orc_dmc;
orc_dmc *orc_dmc_create(const orc_sys *o, int32_t cap);
void orc_dmc_destroy(orc_dmc *d);
int orc_dmc_set(orc_dmc *d, int32_t nwalk, const double *x, const int32_t *rng,
                const double *weight);
int orc_dmc_generation(orc_dmc *d, const qmcw_dmc_params *p, qmcw_dmc_stats *stats,
                       int32_t nthreads);
int orc_dmc_branch(orc_dmc *d, int32_t *brng, qmcw_dmc_stats *stats);
int32_t orc_dmc_get(const orc_dmc *d, double *x, int32_t *rng, double *weight, double *eloc);

/* accelerated Metropolis mover, test access: a proposal of electron i from
   the walker's current state (rec: fxop, areao, r of the new point), and the
   density of the move of electron i to xt */
void orc_slat_sample(const orc_sys *o, orc_walker *w, int32_t i, double fbias, int32_t *rng,
                     double *xn, double *rec);
double orc_slat_density(const orc_sys *o, orc_walker *w, int32_t i, double fbias, const double *xt);
double orc_slat_area(double zeta, double co, double rbot, double rtop, double bot);

#ifdef __cplusplus
}
#endif
#endif
