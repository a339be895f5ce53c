/*
 * champ_amd.h -- C ABI of the MI355X walker-batched wavefunction engine.
 *
 * The engine replaces CHAMP's per-walker "psi / eloc / deriv" call surface
 * (qmc/psi_utils_mod.f90, qmc/vmc/hpsi.f90, qmc/vmc/hpsie.f90,
 * qmc/deriv_mod.f90) with batched entry points over W walkers that keep
 * their state resident in GPU memory.  Every entry point returns an int
 * status (QMCW_OK = 0).  Host pointers passed in are borrowed and never
 * freed; device state is owned by the handle.  One handle per HIP stream;
 * a handle is not thread-safe.
 *
 * Index conventions: all indices are 0-based.  Electrons 0..nup-1 are spin
 * up, nup..nelec-1 spin down.  Walker-major layouts: x[w][i][k].
 */
#ifndef CHAMP_AMD_H
#define CHAMP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QMCW_OK 0
#define QMCW_ERR_ARG 1       /* invalid argument / unsupported system */
#define QMCW_ERR_HIP 2       /* HIP runtime error */
#define QMCW_ERR_SINGULAR 3  /* a Slater matrix was singular (det = 0) */
#define QMCW_ERR_NAN 4       /* non-finite value produced */
#define QMCW_ERR_NOMEM 5     /* device allocation failed */

#define QMCW_FP64 0          /* orbitals, Jastrow and determinants in FP64 */
#define QMCW_MIXED 1         /* orbitals + Jastrow FP32, determinants/energies FP64 */
#define QMCW_MIXED_BF16 2    /* as MIXED with the pseudopotential projection rounded to BF16 */

#define QMCW_BASIS_SLATER 1
#define QMCW_BASIS_GAUSSIAN 2

#define QMCW_MAX_L 5         /* real solid harmonics up to h */
#define QMCW_MAX_ORD 8       /* maximum Jastrow polynomial order */
#define QMCW_MAX_LPOT 5      /* maximum pseudopotential l components */
#define QMCW_MAX_NQUAD 12

/*
 * System description.  It mirrors the data CHAMP keeps in its common
 * modules (qmc/commons/atom_mod.f90, basis1_mod, coefs_mod, dets_mod,
 * jaspar*_mod, pseudo_mod, gauss_ecp_mod, qua_mod, optim_mod) as flat
 * arrays.  The normalisation of the basis functions is folded into `coef`
 * as CHAMP does (qmc/vmc/basis_norm.f90:84-88).
 */
typedef struct qmcw_system {
  /* electrons */
  int32_t nelec, nup;
  /* nuclei */
  int32_t ncent, nctype;
  const double *cent;       /* [ncent][3] */
  const int32_t *iwctype;   /* [ncent] centre type */
  const double *znuc;       /* [nctype] (effective) nuclear charge */
  /* basis: radial functions r^(n-1) sum_p d_p exp(-zeta_p r) (Slater) or
     r^(n-1) sum_p d_p exp(-zeta_p r^2) (Gaussian), times a real solid
     harmonic (l, m) as in qmc/vmc/basis_fns.f90:200-350 */
  int32_t basis_type;
  int32_t nrad;             /* radial functions over all centre types */
  const int32_t *rad_n;     /* [nrad] */
  const int32_t *rad_prim_off; /* [nrad+1] */
  const double *prim_zeta;  /* [nprim] */
  const double *prim_coef;  /* [nprim] */
  const int32_t *ctype_bas_off; /* [nctype+1] basis functions of one centre of each type */
  const int32_t *bas_rad;   /* [nbas_ctype_total] radial function (global index) */
  const int32_t *bas_l;     /* [nbas_ctype_total] */
  const int32_t *bas_m;     /* [nbas_ctype_total] m in -l..l; cos-like m>=0, sin-like m<0 */
  int32_t nbasis;           /* sum over centres of the per-type counts */
  /* orbitals: orb(i,o) = sum_b coef[o*nbasis + b] phi_b(r_i) */
  int32_t norb;
  const double *coef;       /* [norb][nbasis] */
  /* determinants and CSFs (qmc/csfs_mod.f90 after normalisation) */
  int32_t ndet;
  const int32_t *det_up;    /* [ndet][nup]   orbital labels */
  const int32_t *det_dn;    /* [ndet][nelec-nup] */
  int32_t ncsf;
  const double *csf_coef;   /* [ncsf] */
  const int32_t *csf_det_off; /* [ncsf+1] */
  const int32_t *csf_det;   /* dets of each csf */
  const double *csf_det_coef;
  /* Jastrow, ijas=4 (qmc/vmc/jastrow4.f90) */
  int32_t isc;              /* 2 or 4 */
  int32_t nspin2b;          /* 1: b shared by spin pairs; 2: separate parallel-spin b */
  int32_t nocuspb;          /* 0: parallel-spin e-e cusp is half (sspinn=1/2) */
  int32_t norda, nordb, nordc;
  double scalek;
  const double *a4;         /* [nctype][norda+1] */
  const double *b;          /* [nspin2b][nordb+1] */
  const double *c;          /* [nctype][nterms4(nordc)] */
  /* pseudopotential (nloc = 6: Gaussian, qmc/vmc/readps_gauss.f90); 0 = none */
  int32_t nloc;
  const int32_t *npotd;     /* [nctype] number of l components */
  const int32_t *lpotp1;    /* [nctype] 1-based component that is local */
  const int32_t *ecp_off;   /* [nctype*QMCW_MAX_LPOT + 1] term offsets of (type, l) */
  const double *ecp_coef;   /* [nterm] */
  const int32_t *ecp_power; /* [nterm] */
  const double *ecp_expo;   /* [nterm] */
  int32_t nquad;            /* 6 (octahedron) or 12 (icosahedron) */
  /* parameters whose derivatives are wanted (qmc/optim_mod, optimization menu) */
  int32_t nparma_tot;       /* Jastrow a's varied, over types */
  const int32_t *jas_a_type;/* [nparma_tot] centre type */
  const int32_t *jas_a_idx; /* [nparma_tot] 1-based a4 index (iwjasa) */
  int32_t nparmb_tot;
  const int32_t *jas_b_spin;/* [nparmb_tot] */
  const int32_t *jas_b_idx; /* [nparmb_tot] 1-based b index (iwjasb) */
  int32_t nparmc_tot;
  const int32_t *jas_c_type;/* [nparmc_tot] */
  const int32_t *jas_c_idx; /* [nparmc_tot] 1-based c term index (iwjasc), independent only */
  int32_t nparmcsf;
  const int32_t *csf_parm;  /* [nparmcsf] CSF index varied (iwcsf) */
  int32_t nparmorb;
  const int32_t *orb_parm_i;/* [nparmorb] first orbital label of the rotation */
  const int32_t *orb_parm_j;/* [nparmorb] second orbital label */
  const int32_t *orb_parm_rev; /* [nparmorb] 1: include the reverse excitation */
  double hb;                /* hbar^2/2m = 0.5 */
  /* tabulated radial functions (qmc/vmc/read_bas_num.f90, splfit.f90):
     rad_num[r] = k >= 0 makes radial function r the tabulated R_k(r), and
     its basis functions R_k(r) Y_lm (Y_lm = r^-l S_lm, the solid harmonic
     above; rad_n and the primitives of r are then unused); -1, or rad_num
     NULL, keeps r analytic.  R_k between grid points is the cubic spline
     through num_rwf with second derivatives num_d2; below the first
     interval the polynomial num_ce[k][0..4] in r, beyond the grid
     num_ae[k][0] exp(-num_ae[k][1] r). */
  const int32_t *rad_num;   /* [nrad] or NULL */
  int32_t nnum;             /* tabulated functions */
  const int32_t *num_igrid; /* [nnum] grid: 1 linear, 2 exponential, 3 shifted exponential */
  const int32_t *num_nr;    /* [nnum] points */
  const double *num_h;      /* [nnum] exp_h_bas (the step for igrid 1) */
  const double *num_r0;     /* [nnum] r0_bas as splfit_bas uses it */
  const int32_t *num_off;   /* [nnum+1] offsets of each function's grid values */
  const double *num_rwf;    /* values at the grid points */
  const double *num_d2;     /* spline second derivatives at the grid points */
  const double *num_ce;     /* [nnum][5] */
  const double *num_ae;     /* [nnum][2] */
} qmcw_system;

/* sizeof(qmcw_system) as compiled into the library (ABI check) */
int64_t qmcw_sizeof_system(void);

/* number of e-e-n terms for order nordc (qmc/vmc/jastrow4.f90:495-516) */
int32_t qmcw_nterms4(int32_t nordc);

typedef struct qmcw_handle qmcw_handle;

/* Create a handle for W walkers (replaces the psi_t allocation in
   qmc/psi_type_mod.f90:5-39 and walker_t, qmc/walker_type_mod.f90:9-18).
   precision: QMCW_FP64 / QMCW_MIXED / QMCW_MIXED_BF16. */
int qmcw_create(const qmcw_system *sys, int32_t nwalk, int32_t precision,
                qmcw_handle **out);
int qmcw_destroy(qmcw_handle *h);

/* Replace the wavefunction parameters (a4, b, c, csf_coef, coef) in place;
   the arrays have the shapes given in qmcw_system.  Any pointer may be NULL
   to keep the current values.  (Replaces the parameter update after the
   linear-method step, qmc/optimization_mod.f90.) */
int qmcw_set_params(qmcw_handle *h, const double *a4, const double *b,
                    const double *c, const double *csf_coef,
                    const double *coef);

/* Copy walker positions in/out (host arrays [W][nelec][3]). */
int qmcw_set_walkers(qmcw_handle *h, const double *x);
int qmcw_get_walkers(const qmcw_handle *h, double *x);

/* Full evaluation of every walker at its current position, refreshing the
   resident state (inverses, pair Jastrow arrays, orbitals).  Replaces
   psi_at (qmc/psi_utils_mod.f90:13-42) -> hpsi (qmc/vmc/hpsi.f90:2-381).
   Outputs (host, any may be NULL):
     logpsid[W]  log|Psi_D|, sgn[W] sign of Psi_D, psij[W] ln J,
     grad[W][N][3] grad ln Psi, lap[W][N] laplacian of ln Psi per electron,
     eloc[W], epot[W] (total potential incl. nonlocal). */
int qmcw_eval_full(qmcw_handle *h, double *logpsid, double *sgn,
                   double *psij, double *grad, double *lap, double *eloc,
                   double *epot);

/* Propose moving electron ie of every walker to xnew[W][3] and return the
   ratio Psi(new)/Psi(old) (determinant ratio times exp(Delta ln J)), the new
   grad ln Psi of electron ie and its laplacian.  The proposal is held
   pending until qmcw_accept_1e.  Replaces move_1e
   (qmc/psi_utils_mod.f90:45-90) -> hpsie (qmc/vmc/hpsie.f90:1-71). */
int qmcw_ratio_1e(qmcw_handle *h, int32_t ie, const double *xnew,
                  double *ratio, double *grad_new, double *lap_new);

/* Commit (accept[w] != 0) or drop the pending move of electron ie.
   Replaces copy_psi_1e (qmc/psi_utils_mod.f90:299-387) and the VMC commit
   jassav + detsav + dete_save (qmc/vmc/metrop_slat_mov1.f90:544-546). */
int qmcw_accept_1e(qmcw_handle *h, int32_t ie, const uint8_t *accept);

/* grad ln Psi and laplacian of electron ie in the current state.
   Replaces calc_1e_derivatives (qmc/psi_utils_mod.f90:457-483). */
int qmcw_grad_1e(qmcw_handle *h, int32_t ie, double *grad, double *lap);

/* Number of parameters P (CSF + Jastrow + orbital rotations). */
int32_t qmcw_nparm(const qmcw_handle *h);

/* Full evaluation of every walker at its current position together with the
   parameter derivatives d ln Psi/dp and dE_L/dp, in CHAMP's order CSF,
   Jastrow (a, b, c), orbital rotations (qmc/deriv_mod.f90:219-239).
   Replaces hpsi with ideriv set (qmc/vmc/hpsi.f90:2-381) followed by
   dpsi_bld / deloc_bld (qmc/deriv_mod.f90:248-556).  dpsi, deloc: [W][P]
   host arrays (either may be NULL).  QMCW_ERR_ARG when P = 0. */
int qmcw_derivs(qmcw_handle *h, double *dpsi, double *deloc);

/* Sampling mode: with enable != 0 the full evaluation that ends every step
   of qmcw_vmc_run / qmcw_vmc_run_device also forms the derivatives (CHAMP's
   ideriv during an optimisation run, qmc/vmc/vmc.f90:259-272); they stay on
   the GPU until read with qmcw_get_derivs ([W][P] each, last step).  With
   enable == 2 every step also adds them to the product sums below. */
int qmcw_sample_derivs(qmcw_handle *h, int32_t enable);
int qmcw_get_derivs(qmcw_handle *h, double *dpsi, double *deloc);

/* Sums over walkers (and over calls, until reset) of the products the
   linear method averages: dpsi_dpsi_bld, dpsi_dpsi_eloc_bld, dpsi_eloc_bld
   and dpsi_deloc_bld (qmc/deriv_mod.f90:683-723, 898-926, 960-1000,
   1375-1412), with the sums of dpsi, deloc and E_L behind dpsi_av, deloc_av
   and eloc_av; consumed by ovlp_lin_bld / ham_lin_energy_bld
   (qmc/opt_lin_mod.f90:249-348, 404-620).  qmcw_accum_add adds the last
   derivative sample of every walker; qmcw_sample_derivs(h, 2) makes every
   VMC step do so.  qmcw_accum_get copies QMCW_ACCUM_DOUBLES(P) doubles:
   [0] samples, [1] sum E_L, [2] sum E_L^2, sum dpsi[P], sum E_L dpsi[P],
   sum deloc[P], then row-major P x P: sum dpsi_i dpsi_j,
   sum E_L dpsi_i dpsi_j, sum dpsi_i deloc_j. */
#define QMCW_ACCUM_DOUBLES(P) (3 + 3 * (int64_t)(P) + 3 * (int64_t)(P) * (P))
int qmcw_accum_reset(qmcw_handle *h);
int qmcw_accum_add(qmcw_handle *h);
int qmcw_accum_get(qmcw_handle *h, double *out);

/* qmcw_accum_add with no device allocation.  It has no counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight sampling loops, harnesses that forbid
   hipMalloc).  It adds the last derivative sample of every walker to the
   caller's DEVICE accumulator acc (QMCW_ACCUM_DOUBLES(P) doubles, layout as
   above; the caller clears it), enqueued on the handle's stream with no
   host synchronisation.  Partial sums use a static device array of
   QMCW_ACCUM_STATIC_DOUBLES doubles (32 MiB per GPU, held until the
   process ends): ceil(W/256) * (3 P^2 + 3 P + 3) must fit it (P = 242: up to
   5,888 walkers per handle), else QMCW_ERR_ARG.  Every handle on a GPU
   shares the array, so calls that overlap on one GPU (different streams or
   threads) race; calls in stream order do not.  The result equals
   qmcw_accum_add's bit for bit (same kernels, same chunk order). */
#define QMCW_ACCUM_STATIC_DOUBLES ((int64_t)4 << 20)
int qmcw_accum_add_static(qmcw_handle *h, double *acc);

/* Per-walker random streams (CHAMP's rannyu, qmc/vmc/rannyu.f90:53-74):
   seed[W][4] 12-bit limbs as in setrn (qmc/vmc/rannyu.f90:11-21). */
int qmcw_set_rng(qmcw_handle *h, const int32_t *seed);
int qmcw_get_rng(const qmcw_handle *h, int32_t *seed);

/* Periodic re-inversion: after every `every` one-electron moves (counted
   across sweeps; VMC and DMC moves, not T-moves) rebuild A^-1, T, the alpha
   inverses and Y of every walker from its kept orbital rows, which removes
   the rounding the Sherman-Morrison updates accumulate (dete_update,
   qmc/dete_mod.f90:114-214).  0 (default): only at each step's full
   evaluation, as CHAMP's hpsi does (deriv_fast_mod.f90:367-398).
   qmcw_reinvert does it once, now. */
int qmcw_set_reinversion(qmcw_handle *h, int32_t every);
int qmcw_reinvert(qmcw_handle *h);

/* The mover of qmcw_vmc_run / qmcw_vmc_run_device (CHAMP's imetro,
   qmc/control_mod.f90:78, 374-392): 1, metrop_mov1's box proposal
   (qmc/vmc/metrop_mov1.f90:63-99, the handle's default); 6, CHAMP's
   default accelerated Metropolis mover metrop_slat_mov1
   (qmc/vmc/metrop_slat_mov1.f90:78-546): radial step multiplier deltar and
   cos(theta) step deltat about the nearest nucleus, with the run's fbias.
   Out-of-range steps are reset as CHAMP does (deltar < 1 -> 2; deltat
   outside [0, 2] -> 2).  Other imetro: QMCW_ERR_ARG. */
int qmcw_set_vmc_mover(qmcw_handle *h, int32_t imetro, double deltar, double deltat);

/* Run nstep VMC steps on device.  One step = a Metropolis sweep over all
   electrons with the mover set by qmcw_set_vmc_mover (box step delta
   and force bias fbias for imetro 1; fbias for imetro 6) followed by a full
   evaluation, as in qmc/vmc/metrop_mov1.f90:283-300.  Per walker and step the accumulated
   local energy is written to eloc_trace[step][W] when non-NULL (host).
   naccept (host, may be NULL) receives the accepted move count per walker. */
int qmcw_vmc_run(qmcw_handle *h, int32_t nstep, double delta, double fbias,
                 double *eloc_trace, int64_t *naccept);

/* Device-resident variant of qmcw_vmc_run for timing: no host copies;
   the local energies of the last step stay on the device. */
int qmcw_vmc_run_device(qmcw_handle *h, int32_t nstep, double delta,
                        double fbias);

/* Batches of systems.  No counterpart in the reference, which runs one
   system per process; the batching of the engine's own design: several
   handles (systems of uneven electron and basis counts, each with its own
   walkers, streams and state) advance together, every kernel of the VMC
   step running once over the walkers of all of them (once per lane-group
   width among them: each handle keeps its own widths).  The handles must
   share the precision and the e-e-n order (nordc) and sample no derivatives
   and set no re-inversion interval.  qmcw_batch_vmc_run does what qmcw_vmc_run_device
   does for each handle (a full evaluation, then nstep steps with each
   handle's mover, walkers and random streams; the same walk as the handle
   alone); the handles' later calls are ordered after it.  The batch holds
   pointers to the handles: destroy it before them. */
typedef struct qmcw_batch qmcw_batch;
int qmcw_batch_create(qmcw_handle *const *hs, int32_t n, qmcw_batch **out);
int qmcw_batch_vmc_run(qmcw_batch *b, int32_t nstep, double delta, double fbias);
int qmcw_batch_destroy(qmcw_batch *b);
const char *qmcw_batch_last_error(const qmcw_batch *b);

/* Parameters of one DMC generation (qmc/rundmc_mod.f90:97-330).  The
   defaults in brackets are CHAMP's (qmc/control_mod.f90:11-16) and those of
   its example inputs (idmc 2, itau_eff 1, iacc_rej 1, icross 1). */
typedef struct qmcw_dmc_params {
  double tau;          /* time step */
  double etrial;       /* trial energy */
  double eest;         /* current energy estimate */
  double esigma;       /* energy spread used by the reweighting cut-off */
  double pop_ctrl;     /* population-control factor of this generation */
  double taueff;       /* running effective time step (itau_eff = 3) */
  double c_rewt;       /* [3.5] */
  double pow_rewt;     /* [0.5] */
  double limit_rewt;   /* limit_rewt_dmc [10] */
  double adrift;       /* adrift_default [0.5] */
  int32_t n_drift;     /* n_drift_param [2] */
  int32_t itau_eff;    /* [1] 0: tau; 1: tau sum(pacc dr2)/sum(dr2); 2: accepted only; 3: taueff */
  int32_t itau_integ;  /* [1] */
  int32_t iacc_rej;    /* [1] 0: accept every move */
  int32_t icross;      /* [1] <= 0: reject node-crossing moves */
  int32_t ene_int;     /* [0] 0: ene_int_v9, 1: no_ene_int */
  int32_t idmc;        /* [2] > 0: reweight */
  int32_t tmoves;      /* [0] T-moves before each generation's moves
                          (qmc/dmc/dmc_elec/tmove.f90): 0 off, 1
                          det_balance1, 2 casula1 (pseudopotential systems) */
} qmcw_dmc_params;

/* Sums over the population of one generation (before branching) and the
   population after branching. */
typedef struct qmcw_dmc_stats {
  double wsum;         /* sum of weights */
  double wesum;        /* sum w E_L */
  double we2sum;       /* sum w E_L^2 */
  double wtausum;      /* sum w taunow */
  double pacc;         /* sum of acceptance probabilities */
  double nacc;         /* accepted moves */
  double ntry;         /* proposed moves */
  double ncross;       /* proposed moves that crossed a node */
  int32_t nwalk;       /* population after branching */
  int32_t nbranch;     /* kills + joins + splits */
} qmcw_dmc_stats;

/* DMC population (qmc/rundmc_mod.f90).  qmcw_dmc_begin makes the first
   nwalk walkers (positions set with qmcw_set_walkers, streams with
   qmcw_set_rng) the population, with weights (NULL: 1), and evaluates them;
   the handle's W is the population's capacity (MWALK).  The mover follows
   sample_1e_move (qmc/green_functions_mod.f90:12-38): drift-diffusion for
   pseudopotential systems (nloc > 0), drift-diffusion with the nucleus cusp
   (:83-145, density :268-323) for all-electron ones. */
int qmcw_dmc_begin(qmcw_handle *h, int32_t nwalk, const double *weight);

/* One DMC generation (qmc/rundmc_mod.f90:97-330): with p->tmoves, first the
   T-moves of every electron (qmc/dmc/dmc_elec/tmove.f90); then for every
   walker and electron the drift-diffusion proposal
   (qmc/green_functions_mod.f90:70-145),
   the one-electron ratio, the Metropolis test with the forward/backward
   Green's functions (rundmc_mod.f90:142-196) and the commit; the full
   evaluation of the walkers that moved; the effective time step and the
   reweighting (ene_int_rewt, green_functions_mod.f90:382-472); then
   split_join (qmc/dmc/dmc_elec/split_join_mod.f90:3-219) with the branching
   stream.  Each walker draws from its own rannyu stream; a split copy
   continues on a stream drawn from the branching stream.  stats (may be
   NULL) receives the population sums before branching and the population
   after it.  Positions of the population: qmcw_get_walkers, first nwalk. */
int qmcw_dmc_step(qmcw_handle *h, const qmcw_dmc_params *p, qmcw_dmc_stats *stats);

/* Current population size and weights ([nwalk], may be NULL). */
int qmcw_dmc_get(qmcw_handle *h, int32_t *nwalk, double *weight);

/* rannyu stream of split_join (4 limbs as in setrn). */
int qmcw_set_branch_rng(qmcw_handle *h, const int32_t *seed);
int qmcw_get_branch_rng(const qmcw_handle *h, int32_t *seed);

/* split_join (qmc/dmc/dmc_elec/split_join_mod.f90:3-219) over host arrays,
   the branching qmcw_dmc_step applies between generations; host code only
   (no handle, no GPU).  weight[mwalk]: in, the n walkers' weights; out,
   the *nwalk_out weights after killing, joining and splitting (mwalk plays
   MWALK's part in the split threshold).  origin[mwalk] (out): the index
   among the n old walkers each new walker copies.  rng[mwalk][4]: the
   walkers' streams, moved with them; a split copy gets four 12-bit limbs
   drawn from the branching stream brng[4] (the last made odd), which the
   reference, whose walkers share one stream, does not draw.  brng advances
   by one rannyu per joined pair, as in the reference, plus four per split.
   *nbranch: kills + joins + splits (the reference's nbrnch increment).
   QMCW_ERR_NOMEM when the splits need more than mwalk walkers. */
int qmcw_split_join(int32_t n, int32_t mwalk, double *weight, int32_t *origin, int32_t *rng,
                    int32_t *brng, int32_t *nwalk_out, int32_t *nbranch);

/* The same split_join by the device kernel qmcw_dmc_step runs (one
   workgroup, k_split_join), on host arrays: test access, with the
   arguments and results of qmcw_split_join (it allocates device memory and
   synchronises).  QMCW_HOST_SPLIT_JOIN=1 makes qmcw_dmc_step branch on the
   host instead. */
int qmcw_split_join_gpu(int32_t n, int32_t mwalk, double *weight, int32_t *origin, int32_t *rng,
                        int32_t *brng, int32_t *nwalk_out, int32_t *nbranch);

/* Read back per-walker local energies of the last evaluation. */
int qmcw_get_eloc(const qmcw_handle *h, double *eloc);

/* Kernel timing with HIP events recorded on the handle's stream around every
   launch of the psi kernels (no counterpart in the reference; for
   measurement).  enable != 0 clears and starts accumulation; the events add
   a little host work per launch, so time throughput with profiling off.
   qmcw_profile_get fills ms[QMCW_PROF_NKERNEL] with the summed milliseconds
   and count[QMCW_PROF_NKERNEL] with the number of timed launches, per kernel
   in the order of the QMCW_PROF_* indices. */
#define QMCW_PROF_FULL_JAS 0   /* k_full_jas: distances + full Jastrow */
#define QMCW_PROF_FULL_DET 1   /* k_full_det: basis, orbitals, AMF dets */
#define QMCW_PROF_FULL_NLPP 2  /* k_full_nlpp: pseudopotential quadrature */
#define QMCW_PROF_FULL_FIN 3   /* k_full_fin: local energy assembly */
#define QMCW_PROF_RATIO 4      /* k_ratio_1e: one-electron move ratio */
#define QMCW_PROF_ACCEPT 5     /* k_accept_1e: Metropolis test + update */
#define QMCW_PROF_ROTQUA 6     /* k_rotqua: random rotation of the grid */
#define QMCW_PROF_DERIV_JAS 7  /* k_deriv_jas: Jastrow parameter derivatives */
#define QMCW_PROF_DERIV_NLPP 8 /* k_deriv_nlpp: pseudopotential derivative pieces */
#define QMCW_PROF_DERIVS 9     /* k_derivs: d ln Psi/dp, d E_L/dp assembly */
#define QMCW_PROF_ACCUM 10     /* k_accum_*: sums of the linear method's derivative products */
#define QMCW_PROF_NKERNEL 11
int qmcw_profile(qmcw_handle *h, int32_t enable);
int qmcw_profile_get(qmcw_handle *h, double *ms, int64_t *count);

/* Choose the lane-group width (walkers per wavefront) of each kernel class
   by timing the candidates on the resident walkers with side-effect-free
   launches (no counterpart in the reference).  chosen (may be NULL)
   receives the width of the one-electron moves. */
int qmcw_autotune(qmcw_handle *h, int32_t *chosen);

/* Set the lane-group width (8, 16, 32 or 64 lanes per walker) of every
   kernel explicitly; results are identical for every width (no counterpart
   in the reference). */
int qmcw_set_group_width(qmcw_handle *h, int32_t G);

/* Lane-group widths in use, per kernel class: widths[0] Jastrow,
   [1] orbitals/determinants/local energy, [2] pseudopotential, [3]
   one-electron moves (qmcw_autotune chooses each; no counterpart in the
   reference). */
int qmcw_get_group_widths(const qmcw_handle *h, int32_t *widths);

/* Wait for all work queued on the handle's stream. */
int qmcw_sync(qmcw_handle *h);

/* Last error message of the handle (or of the library when h is NULL). */
const char *qmcw_last_error(const qmcw_handle *h);

#ifdef __cplusplus
}
#endif

#endif /* CHAMP_AMD_H */
