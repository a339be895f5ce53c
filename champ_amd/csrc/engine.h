// engine.h -- device-side data layout of the walker-batched wavefunction
// engine (MI355X / gfx950).  Shared by the kernels (engine_kernels.hip)
// and the host side of the C ABI (engine_api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/champ_amd.h"

namespace qmcw {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kMaxShellM = 11;     // 2l+1 for l <= 5
constexpr int kMaxCtype = 16;
constexpr int kMaxTerms = 64;
constexpr int kMaxKmax = 8;        // excitation order of an AMF alpha block
constexpr int kMaxNs = 512;        // electrons of one spin
// per tabulated radial function: igrid, nr, r0, 1/h (igrid 1) or 1/ln h,
// ce[5] (small-r polynomial), ae[2] (tail a0 exp(-a1 r)), offset of its
// grid points in num_tab
constexpr int kNumPar = 12;

// Flattened system in device memory.  Pointers are device pointers.
struct DevSys {
  int N, nup, ndn, M, nctype, NB, NO;
  int basis_type;
  double pecent, hb;
  const double *cent;        // [M][3]
  const int *iwctype;        // [M]
  const double *znuc;        // [nctype]
  // basis shells: one radial function r^a sum_p d_p exp(-zeta_p r^q) and the
  // real solid harmonics S_lm (m listed in sh_m) that multiply it
  int nshell;                // over all centres
  const int *cent_sh_off;    // [M+1]
  const int *sh_cent, *sh_l, *sh_a, *sh_nm, *sh_prim_off;
  const int *sh_m;           // [nshell][kMaxShellM]
  const int *sh_bas;         // [nshell][kMaxShellM] basis index of each component
  const double *prim_zeta, *prim_coef;
  // tabulated radial functions: a shell with sh_num[s] = k >= 0 has the
  // radial part R_k(r) r^-l (R_k tabulated) in place of r^a g(r)
  int has_num;               // any tabulated shell (sh_num valid)
  const int *sh_num;         // [nshell] table or -1
  const double *num_par;     // [nnum][kNumPar]
  const double *num_tab;     // [sum of nr][4] grid point x, R, R'' (spline), 0
  int maxshell_per_cent;
  const double *coefT;       // [NB][NO] (normalisation folded in)
  const double *coef_occ[2]; // [NB][nocc] coefT of the occupied orbitals of each spin
  // determinants (spin s = 0 up, 1 down)
  int ns[2], nocc[2], ndetu[2], ndet, kmax, noccmax;
  const int *ref[2];         // [ns] reference orbital labels
  const int *occ[2];         // [nocc] occupied orbital labels (sorted union)
  const int *iref[2];        // [ns] position in occ of ref_k
  const int *ivir[2];        // [nocc-ns] occ positions not in ref
  const int *ord[2];         // [ndetu] excitation order k
  const int *dsgn[2];        // [ndetu] permutation sign
  const int *rf[2];          // [ndetu][ns] replaced reference positions
  const int *exl[2];         // [ndetu][ns] excited orbital labels
  const int *exo[2];         // [ndetu][ns] excited orbital positions in occ
  const int *iwdet[2];       // [ndet] unique det of each spin
  const int *udet_off[2];    // [ndetu+1] CSR: dets whose spin-s part is unique det u
  const int *udet_d[2];      // [ndet]
  const int *yent_off[2];    // [nocc*ns+1] CSR: contributions to Y(jo, ir)
  const int *yent_u[2];      // unique det of the contribution
  const int *yent_pos[2];    // its alpha-inverse element (u*kmax2 + jj + k*ii)
  const double *cdet;        // [ndet] cdet_in_wf
  const double *cdd;         // [ndet] cdet * sign(up part) * sign(down part)
  const double *ucf[2];      // per udet_d entry: cdd of the det
  const int *uiu[2], *uid[2];  // per udet_d entry: its up and down unique dets
  int kmax2;                 // kmax*kmax (>= 1)
  int single_det;            // one determinant, no orbital beyond the reference ones: the
                             // one-electron move needs neither T nor Y (Y is zero, T the
                             // fixed permutation ref -> occ) and does not update them
  // Jastrow ijas=4
  int isc, nspin2b, nocuspb, norda, nordb, nordc, nterms;
  double scalek, asymp_r;
  const double *asymp_jasa;  // [nctype]
  double asymp_jasb0, asymp_jasb1;
  const double *a4;          // [nctype][norda+1]
  const double *b;           // [nspin2b][nordb+1]
  const double *c;           // [nctype][nterms]
  const int *term_klm;       // [nterms][3]
  // pseudopotential
  int nloc, nquad;
  int tm_kmax;               // most centres whose nonlocal ranges can hold one point (T-moves)
  const int *npotd, *lpotp1, *ecp_off, *ecp_power;
  const double *ecp_coef, *ecp_expo;
  const double *q0;          // [4][nquad] x0,y0,z0,w
  // derivative parameters
  int npcsf, npa, npb, npc, nporb, npj, nparm;
  const int *pcsf;           // CSF index of each CSF parameter
  const int *csf_det_off, *csf_det;
  const double *csf_det_coef, *csf_coef;
  int ncsf;
  const int *porb_i, *porb_j, *porb_rev;  // orbital rotations (labels)
  const int *iocc_lab[2];    // [NO] position in occ or -1
  // Jastrow derivative slots: one per a-coefficient [nctype][norda+1],
  // b-coefficient [nspin2b][nordb+1] and e-e-n term [nctype][nterms] (the
  // terms in the order of DevSys::c); parameter p of the Jastrow block is
  // sum_q pj_w[q] * slot[pj_slot[q]], q in [pj_off[p], pj_off[p+1]) (the
  // dependent e-e-n coefficients enter through their cusp weights)
  int nslot, slot_b0, slot_c0;
  const int *pj_off, *pj_slot;
  const double *pj_w;
};

// Offsets (in doubles) of one walker's persistent state.
struct WLayout {
  int64_t x, rvec_en, r_en, rvec_ee, r_ee;
  int64_t fso, fijo, lapj, fjo, fsumo;
  int64_t oc;                // [N][noccmax][5] occupied-orbital values
  // determinant blocks in two slots each, [spin][slot]: A^-1 (slot number
  // at ws[par + s]), T, the alpha inverses and their determinants (slot at
  // ws[par + 2 + s]) and Y (both spins, slot at ws[par + 4]).  A one-electron
  // move writes its candidates into the other slots; its commit flips the
  // slot numbers instead of copying them back.
  int64_t ai[2][2], tocc[2][2], inv[2][2], detv[2][2], y[2][2], par, deta[2], chi;
  int64_t psi2o, logpsid, sgn, psij, eloc, epot;
  int64_t quad;              // [3][nquad] rotated quadrature grid
  int64_t stride;
};

// Offsets (in doubles) of one walker's scratch.
struct SLayout {
  int64_t phi;               // [max(NB, points)][5] basis values
  int64_t orb;               // [N][NO][5] all orbitals (derivative evaluation)
  int64_t g[2];              // [nocc][ns] G matrices
  int64_t vd, divv, vj;      // [N][3], [N], [N][3]
  int64_t red;               // reduction scratch (matrix temporaries)
  int64_t phs;               // [64] scalars passed between the phase kernels
  int64_t cup;               // [ndetu0 + ndetu1] det-coefficient sums
  int64_t qd;                // [G][4M] quadrature-point distances
  int64_t vnl;               // [N*M][QMCW_MAX_LPOT+1] nonlocal channel values
  int64_t gproj;             // [N][NB] G(:,i)^T coef of the occupied orbitals
  int64_t bt;                // [M][N] e-e-n variable 1 - k scaled(r_j,ic) (as the kernel's T)
  int64_t uq;                // [64][N] per-lane e-e-n variable of the pairs of a point
  int64_t plist;             // [N*M] ints: (electron, centre) pairs in pseudopotential range
  int64_t prot;              // [N*M][10] DMC: axes of each in-range pair's grid rotation, then
                             // the pair's rank among its electron's pairs
  int64_t fold;              // [N] Jastrow pair sum of each electron at its current position
  int64_t mv;                // pending one-electron move buffers
  int64_t stride;
};

// Derivative scratch (offsets in doubles from the walker's scratch base),
// present when the system has parameters to optimise.
struct DLayout {
  int64_t raw;               // [nslot][N][5] slot pieces: value, grad, lap
  int64_t gvalue, gjas, d2g; // [npj], [npj][N][3], [npj]
  int64_t pbc;               // [N*M][NB] per (electron, centre) weighted basis sums
  int64_t dpe_raw;           // [N*M][nslot]
  int64_t dpe_slot, dpe;     // [nslot], [npj]
  int64_t pnl;               // [N][NO] psp_nonloc_orb
  int64_t btd, uqd;          // [N][M] e-e-n variable B(r_j,ic), [64][N] per-lane U of a point
  int64_t B;                 // [N][NO] orbital-resolved local energy
  int64_t tfull[2], dt[2];   // [ns][NO]
  int64_t aib;               // [ns][NO]
  int64_t del[2], c0[2], c1[2];  // [ndetu]
  int64_t idti[2];           // [ndetu][kmax2]
  int64_t dy[2], dg[2];      // [nocc][ns]
  int64_t aux4;              // [nocc][ns]
  int64_t misc;              // [8]
  int64_t phiall;            // [N][NB][5] every electron's basis values (FP64), for k_orb_mfma
};

// Pending one-electron move buffers inside SLayout::mv
struct MLayout {
  int64_t xn, ren, rven, ree, rvee;       // new distances
  int64_t fsn, fijn_i, fijn_j, lapn_i;    // pair rows [N]...
  int64_t fjn;                            // [N][3]
  int64_t orbe;                           // [noccmax][5]
  int64_t misc;                           // ratio, proposal and Sherman-Morrison vectors
  int64_t slat;                           // accelerated mover: fxop, areao, r of the proposal [4]
  int64_t stride;
};

}  // namespace qmcw
