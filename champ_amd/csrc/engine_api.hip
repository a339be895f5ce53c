// engine_api.hip -- host side of the C ABI (include/champ_amd.h).
//
// qmcw_create flattens the system into device arrays (basis shells, the
// determinant/excitation bookkeeping of CHAMP's AMF formulas, Jastrow
// constants, quadrature grid), lays out the per-walker state and scratch,
// and picks the lane-group width G.  The entry points launch the kernels of
// engine_kernels.h on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine_derivs.h"
#include "engine_accum.h"
#include "engine_dmc.h"
#include "engine_launch.h"

using namespace qmcw;

static thread_local std::string g_lib_err;

struct qmcw_handle {
  int W = 0, Wp = 0, G = 16, precision = QMCW_FP64, chunk = 1;
  // lane-group width of each kernel class (jastrow, determinants, pseudopotential, moves)
  int Gk[4] = {16, 16, 16, 16};
  int Wl = 0;                // walkers the kernels run over (multiple of 8, <= Wp)
  hipStream_t stream = nullptr;
  KArgs A{};
  std::vector<void *> allocs;
  double *io = nullptr;
  size_t io_doubles = 0;
  uint8_t *acc_buf = nullptr;
  long long *nacc = nullptr;
  double *trace = nullptr;
  int trace_steps = 0;
  // host copies of the parameter arrays that set_params can replace
  std::vector<double> a4, b, c, csf_coef, coef, cdet, asymp_jasa;
  double *d_cdd = nullptr, *d_ucf[2] = {nullptr, nullptr};
  std::vector<int> iwdet_h[2], dsgn_h[2], udet_d[2];   // det bookkeeping for re-folding signs
  std::vector<int> occ_h[2];                           // occupied orbital labels per spin
  double *d_coef_occ[2] = {nullptr, nullptr};
  double *d_a4 = nullptr, *d_b = nullptr, *d_c = nullptr, *d_cdet = nullptr,
         *d_coefT = nullptr, *d_asymp = nullptr, *d_csf_coef = nullptr;
  // determinant bookkeeping (host) used to rebuild cdet
  int ncsf = 0, ndet = 0;
  std::vector<int> csf_det_off, csf_det;
  std::vector<double> csf_det_coef;
  int nparm = 0;
  double *der = nullptr;     // [Wp][2 nparm] dpsi, deloc of the last derivative evaluation
  int sample_derivs = 0;     // VMC steps evaluate derivatives (CHAMP's ideriv)
  int der_valid = 0;
  // one VMC step captured as a HIP graph (qmcw_vmc_run_device) and its key
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_ok = 0, graph_G = 0, graph_Wl = 0, graph_derivs = 0;
  double graph_delta = 0.0, graph_fbias = 0.0;
  const double *graph_st = nullptr;
  int use_graph = 1;
  // every walker's occupied-orbital rows in the state match its positions
  // (a full evaluation computed them and only committed moves changed
  // positions since); cleared by new positions or orbital coefficients
  int orb_valid = 0;
  int reinv_every = 0;        // rebuild the Slater inverses after this many one-electron moves (0: only
  int moves_since_inv = 0;    // at the full evaluation) and the moves since the last rebuild
  int no_keep_orb = 0;       // QMCW_NO_KEEP_ORB: recompute them in every full evaluation
  int dmc_unfused = 0;       // QMCW_DMC_UNFUSED: DMC moves as five launches per electron
  int use_lds = 1;           // LDS staging of the moved electron's basis values
  int chunk_override = 0;    // electrons per basis chunk (0: from the L2 budget)
  size_t det_lds_budget = 10240;  // LDS bytes per block for k_full_det's basis values
  int split_move = 0;        // one-electron move as three launches (k_move_a/b/c)
  int split_forced = 0;      // QMCW_SPLIT_MOVE set: the autotuner keeps split_move
  int lds_y = 1;             // LDS staging of eval_y in the move kernels
  int ncu = 256;             // compute units of the device (set at create)
  double *acc = nullptr, *acc_part = nullptr, *acc_partv = nullptr;  // product sums and chunk partials
  int acc_nch = 0;
  // DMC population (qmcw_dmc_begin / qmcw_dmc_step)
  int dmc_on = 0, nact = 0;
  double *dm = nullptr, *dm2 = nullptr, *st2 = nullptr, *dstat = nullptr, *dwt = nullptr;
  // T-moves: per-walker kept quadrature data (double-buffered like the
  // state) and the point columns of one electron's step
  double *tq2 = nullptr;
  double *tmq = nullptr, *tmx = nullptr, *rcol = nullptr, *tmb = nullptr, *tmbq = nullptr,
         *bcol = nullptr, *tmxin = nullptr;
  int *tmint = nullptr;      // [3][Wp]: fresh points to evaluate, backward points, move chosen
  double *tmstale = nullptr; // [Wp]: G, gp and bt of the slot must be rebuilt from the state
  int *dorigin = nullptr;
  double *dsort = nullptr;   // device split_join: sorted weights
  int *dunder = nullptr, *dsj = nullptr;  // its hole stack; results [4] and branching stream [4]
  int npow2 = 1;
  int host_split_join = 0;    // QMCW_HOST_SPLIT_JOIN=1: split_join on the host
  int no_gproj_mfma = 0;      // QMCW_NO_GPROJ_MFMA=1: the projection on VALU in k_full_det
  int no_een_tab = 0;         // QMCW_NO_EEN_TAB=1: the move's e-e-n terms summed per pair
  int32_t brng[4] = {0, 0, 0, 1};
  // event timing of the device driver
  int prof = 0;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_phase;
  size_t ev_used = 0;
  std::string err;
};

// record a start/stop event pair around one phase (profiling only)
static hipEvent_t prof_event(qmcw_handle *h) {
  if (h->ev_used >= h->ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    h->ev.push_back(e);
    h->ev_phase.push_back(-1);
  }
  return h->ev[h->ev_used++];
}
static void prof_begin(qmcw_handle *h, int phase) {
  if (!h->prof) return;
  hipEvent_t e = prof_event(h);
  if (!e) return;
  h->ev_phase[h->ev_used - 1] = phase;
  (void)hipEventRecord(e, h->stream);
}
static void prof_end(qmcw_handle *h) {
  if (!h->prof) return;
  hipEvent_t e = prof_event(h);
  if (!e) return;
  h->ev_phase[h->ev_used - 1] = -1;
  (void)hipEventRecord(e, h->stream);
}

#define HIPCHK(h, call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      if (h) (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);     \
      else g_lib_err = std::string(#call) + ": " + hipGetErrorString(e_);      \
      return QMCW_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

static int fail(qmcw_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  else g_lib_err = msg;
  return code;
}

template <typename V>
static int upload(qmcw_handle *h, const V &vec, const typename V::value_type **out) {
  using T = typename V::value_type;
  size_t n = std::max<size_t>(vec.size(), 1);
  void *p = nullptr;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return fail(h, QMCW_ERR_NOMEM, "hipMalloc failed");
  h->allocs.push_back(p);
  if (!vec.empty() && hipMemcpy(p, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return fail(h, QMCW_ERR_HIP, "hipMemcpy failed");
  *out = (const T *)p;
  return QMCW_OK;
}

extern "C" int64_t qmcw_sizeof_system(void) { return (int64_t)sizeof(qmcw_system); }

extern "C" int32_t qmcw_nterms4(int32_t nord) {
  int i = 0;
  for (int n = 2; n <= nord; n++)
    for (int k = n - 1; k >= 0; k--) {
      int l_hi = (k == 0) ? n - k - 2 : n - k;
      for (int l = l_hi; l >= 0; l--)
        if ((n - k - l) % 2 == 0) i++;
    }
  return i;
}

// -------------------------------------------------------------------------
// host-side preparation
namespace {

struct DetBook {
  int ns[2] = {0, 0}, nocc[2] = {0, 0}, ndetu[2] = {0, 0}, kmax = 0;
  std::vector<int> ref[2], occ[2], iref[2], ivir[2], ord[2], sgn[2], rf[2], exl[2], exo[2],
      iwdet[2], iocc_lab[2];
};

// unique determinants of each spin (qmc/vmc/determinant_up_dn.f90), the
// sorted union of occupied orbitals and the excitation of each unique det
// relative to the first one (qmc/deriv_fast_mod.f90:211-365)
DetBook build_dets(const qmcw_system *s) {
  DetBook D;
  int nsz[2] = {s->nup, s->nelec - s->nup};
  const int *lists[2] = {s->det_up, s->det_dn};
  for (int sp = 0; sp < 2; sp++) {
    int ns = nsz[sp];
    D.ns[sp] = ns;
    std::vector<std::vector<int>> uniq;
    D.iwdet[sp].resize(s->ndet);
    for (int d = 0; d < s->ndet; d++) {
      std::vector<int> v(lists[sp] + (size_t)d * ns, lists[sp] + (size_t)d * ns + ns);
      int f = -1;
      for (size_t u = 0; u < uniq.size(); u++)
        if (uniq[u] == v) { f = (int)u; break; }
      if (f < 0) { uniq.push_back(v); f = (int)uniq.size() - 1; }
      D.iwdet[sp][d] = f;
    }
    if (ns == 0) { uniq.assign(1, std::vector<int>()); std::fill(D.iwdet[sp].begin(), D.iwdet[sp].end(), 0); }
    D.ndetu[sp] = (int)uniq.size();
    std::vector<int> occ;
    for (auto &v : uniq)
      for (int o : v)
        if (std::find(occ.begin(), occ.end(), o) == occ.end()) occ.push_back(o);
    std::sort(occ.begin(), occ.end());
    D.occ[sp] = occ;
    D.nocc[sp] = (int)occ.size();
    D.iocc_lab[sp].assign(s->norb, -1);
    for (size_t j = 0; j < occ.size(); j++) D.iocc_lab[sp][occ[j]] = (int)j;
    const std::vector<int> &ref = uniq[0];
    D.ref[sp] = ref;
    for (int k = 0; k < ns; k++) D.iref[sp].push_back(D.iocc_lab[sp][ref[k]]);
    for (int j = 0; j < D.nocc[sp]; j++)
      if (std::find(D.iref[sp].begin(), D.iref[sp].end(), j) == D.iref[sp].end()) D.ivir[sp].push_back(j);
    int nu = D.ndetu[sp];
    D.ord[sp].assign(nu, 0);
    D.sgn[sp].assign(nu, 1);
    D.rf[sp].assign((size_t)nu * std::max(ns, 1), 0);
    D.exl[sp].assign((size_t)nu * std::max(ns, 1), 0);
    D.exo[sp].assign((size_t)nu * std::max(ns, 1), 0);
    for (int u = 0; u < nu; u++) {
      const std::vector<int> &det = uniq[u];
      std::vector<int> out_ref, in_det;
      for (int o : ref) if (std::find(det.begin(), det.end(), o) == det.end()) out_ref.push_back(o);
      for (int o : det) if (std::find(ref.begin(), ref.end(), o) == ref.end()) in_det.push_back(o);
      int k = (int)out_ref.size();
      D.ord[sp][u] = k;
      D.kmax = std::max(D.kmax, k);
      for (int q = 0; q < k; q++) {
        D.rf[sp][u * ns + q] = (int)(std::find(ref.begin(), ref.end(), out_ref[q]) - ref.begin());
        D.exl[sp][u * ns + q] = in_det[q];
        D.exo[sp][u * ns + q] = D.iocc_lab[sp][in_det[q]];
      }
      // parity of the permutation bringing det into reference order: orbitals
      // shared with the reference are moved to their reference positions
      int n = 1, nsw = 0;
      for (int i = 1; i <= ns; i++) {
        auto it = std::find(ref.begin(), ref.end(), det[i - 1]);
        if (it != ref.end()) {
          int ind = (int)(it - ref.begin()) + 1;
          nsw += (i - n) + (ind - n);
          n++;
        }
      }
      D.sgn[sp][u] = 1 - 2 * (nsw % 2);
    }
  }
  return D;
}

std::vector<double> build_cdet(int ndet, int ncsf, const double *csf_coef, const int *off,
                               const int *dets, const double *cdc) {
  std::vector<double> c(ndet, 0.0);
  for (int k = 0; k < ncsf; k++)
    for (int q = off[k]; q < off[k + 1]; q++) c[dets[q]] += cdc[q] * csf_coef[k];
  return c;
}

}  // namespace

// cdet with the permutation signs of both spin parts folded in (exact: the
// signs are +-1), per determinant and per entry of the dets-of-a-unique-det
// lists, so that the kernels skip the sign lookups
static void fold_det_signs(const qmcw_handle *h, std::vector<double> &cdd,
                           std::vector<double> (&ucf)[2]) {
  cdd.resize(h->cdet.size());
  for (size_t d = 0; d < cdd.size(); d++)
    cdd[d] = h->cdet[d] * h->dsgn_h[0][h->iwdet_h[0][d]] * h->dsgn_h[1][h->iwdet_h[1][d]];
  for (int sp = 0; sp < 2; sp++) {
    ucf[sp].clear();
    for (int d : h->udet_d[sp]) ucf[sp].push_back(cdd[d]);
  }
}

// coefT restricted to the occupied orbitals of each spin, [NB][nocc]
static std::vector<double> coef_occ_table(const std::vector<double> &coefT, int NB, int NO,
                                          const std::vector<int> &occ) {
  std::vector<double> t((size_t)NB * std::max<size_t>(1, occ.size()), 0.0);
  for (int b = 0; b < NB; b++)
    for (size_t j = 0; j < occ.size(); j++) t[(size_t)b * occ.size() + j] = coefT[(size_t)b * NO + occ[j]];
  return t;
}

// e-e-n coefficients from CHAMP's term order (jastrow4.f90:225-244) to the
// kernels' compile-time order (k, then m, then l)
// position of each e-e-n term (CHAMP order) in the order of DevSys::c
static std::vector<int> een_dev_index(int nordc) {
  std::vector<int> ck, cl, cm;
  for (int n = 2; n <= nordc; n++)
    for (int k = n - 1; k >= 0; k--) {
      int l_hi = (k == 0) ? n - k - 2 : n - k;
      for (int l = l_hi; l >= 0; l--) {
        int m = (n - k - l) / 2;
        if (2 * m == n - k - l) { ck.push_back(k); cl.push_back(l); cm.push_back(m); }
      }
    }
  std::vector<int> pos(ck.size(), 0);
  int idx = 0;
  for (int k = 0; k <= nordc; k++)
    for (int m = 0; 2 * m <= nordc - k; m++)
      for (int l = 0; l <= nordc - k - 2 * m; l++) {
        if (!((k >= 1 || m >= 1) && (l + 2 * m >= 1) && (k + l + 2 * m >= 2))) continue;
        for (size_t t = 0; t < ck.size(); t++)
          if (ck[t] == k && cl[t] == l && cm[t] == m) pos[t] = idx;
        idx++;
      }
  return pos;
}

// Jastrow parameter -> derivative slots.  The e-e-n coefficients fixed by
// the cusp conditions move with the varied ones: the cusp matrix d and the
// dependent-term list iwc4 of qmc/fit/cuspinit4.f90:1-60, the weights of
// checkdepend4 (qmc/fit/cuspinit4.f90:62-156).
struct JasMap {
  int nslot = 0, slot_b0 = 0, slot_c0 = 0;
  std::vector<int> off{0}, slot;
  std::vector<double> w;
};

static bool build_jas_map(const qmcw_system *s, int nt, JasMap &jm, std::string &why) {
  const int na = s->norda + 1, nb = s->nordb + 1, nordc = s->nordc;
  jm.slot_b0 = s->nctype * na;
  jm.slot_c0 = jm.slot_b0 + s->nspin2b * nb;
  jm.nslot = jm.slot_c0 + s->nctype * nt;
  for (int p = 0; p < s->nparma_tot; p++) {
    int t = s->jas_a_type[p], iw = s->jas_a_idx[p];
    if (t < 0 || t >= s->nctype || iw < 1 || iw > na) { why = "e-n parameter index out of range"; return false; }
    jm.slot.push_back(t * na + iw - 1); jm.w.push_back(1.0); jm.off.push_back((int)jm.slot.size());
  }
  for (int p = 0; p < s->nparmb_tot; p++) {
    int sp = s->jas_b_spin[p], iw = s->jas_b_idx[p];
    if (sp < 0 || sp >= s->nspin2b || iw < 1 || iw > nb) { why = "e-e parameter index out of range"; return false; }
    jm.slot.push_back(jm.slot_b0 + sp * nb + iw - 1); jm.w.push_back(1.0); jm.off.push_back((int)jm.slot.size());
  }
  if (s->nparmc_tot == 0) return true;
  const int neqs = nordc > 1 ? 2 * (nordc - 1) : 0;
  std::vector<double> d((size_t)std::max(neqs, 1) * nt, 0.0);
  std::vector<int> iwc4(std::max(neqs, 1), 0);
  int i = 0;
  for (int n = 2; n <= nordc; n++)
    for (int k = n - 1; k >= 0; k--) {
      int l_hi = (k == 0) ? n - k - 2 : n - k;
      for (int l = l_hi; l >= 0; l--) {
        int m = (n - k - l) / 2;
        if (2 * m != n - k - l) continue;
        i++;
        if (k == 1 && iwc4[n - 2] == 0) iwc4[n - 2] = i;
        if (k == 0 && iwc4[n + nordc - 3] == 0) iwc4[n + nordc - 3] = i;
        d[(size_t)(l + 2 * m - 1) * nt + i - 1] -= 2 * k;
        if (l + m > 0) d[(size_t)(nordc - 1 + k + m - 1) * nt + i - 1] -= (l + m);
        if (m > 0) d[(size_t)(nordc - 1 + k + l + m - 1) * nt + i - 1] -= m;
      }
    }
  std::vector<int> pos = een_dev_index(nordc);
  auto D = [&](int e, int j) { return d[(size_t)(e - 1) * nt + j - 1]; };   // 1-based
  for (int p = 0; p < s->nparmc_tot; p++) {
    int t = s->jas_c_type[p], idx = s->jas_c_idx[p];
    if (t < 0 || t >= s->nctype || idx < 1 || idx > nt) { why = "e-e-n parameter index out of range"; return false; }
    for (int e = 1; e <= neqs; e++)
      if (iwc4[e - 1] == idx) { why = "a cusp-dependent e-e-n coefficient cannot be varied"; return false; }
    int c0 = jm.slot_c0 + t * nt;
    jm.slot.push_back(c0 + pos[idx - 1]); jm.w.push_back(1.0);
    for (int e = 1; e <= neqs; e++) {
      if (e == nordc) continue;
      if (std::fabs(D(e, idx)) > 1e-10) {
        jm.slot.push_back(c0 + pos[iwc4[e - 1] - 1]);
        jm.w.push_back(-D(e, idx) / D(e, iwc4[e - 1]));
      }
    }
    if (nordc > 1) {
      double v = 0.0;
      for (int e = 1; e <= neqs; e++) {
        if (e == nordc) continue;
        double factor = -D(nordc, iwc4[e - 1]) / D(nordc, iwc4[nordc - 1]);
        v -= factor * (D(e, idx) / D(e, iwc4[e - 1]));
      }
      jm.slot.push_back(c0 + pos[iwc4[nordc - 1] - 1]);
      jm.w.push_back(v);
    }
    jm.off.push_back((int)jm.slot.size());
  }
  return true;
}

static std::vector<double> reorder_een(const std::vector<double> &c, int nctype, int nordc) {
  std::vector<int> champ_k, champ_l, champ_m;
  for (int n = 2; n <= nordc; n++)
    for (int k = n - 1; k >= 0; k--) {
      int l_hi = (k == 0) ? n - k - 2 : n - k;
      for (int l = l_hi; l >= 0; l--) {
        int m = (n - k - l) / 2;
        if (2 * m == n - k - l) { champ_k.push_back(k); champ_l.push_back(l); champ_m.push_back(m); }
      }
    }
  int nt = (int)champ_k.size();
  std::vector<double> out(std::max<size_t>(c.size(), 1), 0.0);
  int idx = 0;
  for (int k = 0; k <= nordc; k++)
    for (int m = 0; 2 * m <= nordc - k; m++)
      for (int l = 0; l <= nordc - k - 2 * m; l++) {
        if (!((k >= 1 || m >= 1) && (l + 2 * m >= 1) && (k + l + 2 * m >= 2))) continue;
        int t = 0;
        while (t < nt && !(champ_k[t] == k && champ_l[t] == l && champ_m[t] == m)) t++;
        for (int it = 0; it < nctype; it++) out[it * nt + idx] = c[it * nt + t];
        idx++;
      }
  return out;
}

static int set_asymp(qmcw_handle *h, const qmcw_system *s, DevSys &S) {
  double k = s->scalek, ar = 1.0 / k;
  S.asymp_r = ar;
  h->asymp_jasa.assign(s->nctype, 0.0);
  for (int it = 0; it < s->nctype; it++) {
    const double *a = h->a4.data() + it * (s->norda + 1);
    double v = a[0] * ar / (1.0 + a[1] * ar);
    for (int iord = 2; iord <= s->norda; iord++) v += a[iord] * std::pow(ar, iord);
    h->asymp_jasa[it] = v;
  }
  for (int i = 0; i < 2; i++) {
    double sspinn = 1.0;
    int isp = 0;
    if (i == 1) {
      sspinn = (s->nspin2b == 1 && s->nocuspb == 0) ? 0.5 : 1.0;
      isp = s->nspin2b - 1;
    }
    const double *b = h->b.data() + isp * (s->nordb + 1);
    double v = sspinn * b[0] * ar / (1.0 + b[1] * ar);
    for (int iord = 2; iord <= s->nordb; iord++) v += b[iord] * std::pow(ar, iord);
    if (i == 0) S.asymp_jasb0 = v; else S.asymp_jasb1 = v;
  }
  return QMCW_OK;
}

// 6-point octahedral rule: +-x, +-y, +-z, weights 1/6 (qmc/vmc/rotqua.f90:396-428)
static void gesqua6(double *q) {
  double *x = q, *y = q + 6, *z = q + 12, *w = q + 18;
  for (int j = 0; j < 6; j++) { x[j] = y[j] = z[j] = 0; w[j] = 1.0 / 6.0; }
  x[0] = 1.0; x[1] = -1.0;
  y[2] = 1.0; y[3] = -1.0;
  z[4] = 1.0; z[5] = -1.0;
}

// 12-point icosahedral rule with CHAMP's value of pi (qmc/vmc/rotqua.f90:300-500)
static void gesqua12(double *q) {
  const double pi = 3.14159265359;
  double *x = q, *y = q + 12, *z = q + 24, *w = q + 36;
  for (int j = 0; j < 12; j++) { x[j] = y[j] = z[j] = 0; w[j] = 1.0 / 12.0; }
  z[0] = 1.0;
  z[1] = -1.0;
  double fi0 = pi / 5.0, s5 = std::sqrt(5.0), c = 1.0 / s5, sn = 2.0 / s5;
  for (int j = 1; j <= 5; j++) {
    double rk2 = 2.0 * (j - 1), rk21 = rk2 + 1.0;
    x[j + 1] = sn * std::cos(rk2 * fi0);
    y[j + 1] = sn * std::sin(rk2 * fi0);
    z[j + 1] = c;
    x[j + 6] = sn * std::cos(rk21 * fi0);
    y[j + 6] = sn * std::sin(rk21 * fi0);
    z[j + 6] = -c;
  }
}

// -------------------------------------------------------------------------
// kernels local to the API
__global__ void k_init_walkers(KArgs A, int W) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.Wp) return;
  double *ws = A.st + (int64_t)w * A.L.stride;
  int nq = A.S.nquad;
  for (int q = 0; q < 3 * nq; q++) ws[A.L.quad + q] = A.S.q0[q];
  for (int q = 0; q < 5; q++) ws[A.L.par + q] = 0.0;   // determinant blocks in slot 0
  if (w >= W) {  // padded walkers copy walker 0
    const double *w0 = A.st + A.L.x;
    for (int q = 0; q < 3 * A.S.N; q++) ws[A.L.x + q] = w0[q];
    for (int q = 0; q < 4; q++) A.rng[4 * w + q] = A.rng[q];
  }
}

__global__ void k_gather_full(KArgs A, double *out, int W) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const int N = A.S.N, stride = 8 + 4 * N;
  const double *ws = A.st + (int64_t)w * A.L.stride;
  const double *ss = A.sc + (int64_t)w * A.SL.stride;
  double *o = out + (int64_t)w * stride;
  o[0] = ws[A.L.logpsid]; o[1] = ws[A.L.sgn]; o[2] = ws[A.L.psij]; o[3] = ws[A.L.eloc];
  o[4] = ws[A.L.epot]; o[5] = ss[A.SL.phs + 60]; o[6] = ss[A.SL.phs + 61]; o[7] = ss[A.SL.phs + 62];
  for (int q = 0; q < 3 * N; q++) o[8 + q] = ss[A.SL.vj + q] + ss[A.SL.vd + q];
  for (int i = 0; i < N; i++) o[8 + 3 * N + i] = ss[A.SL.divv + i];
}

__global__ void k_trace_eloc(KArgs A, double *trace, int step, int W) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  trace[(int64_t)step * W + w] = A.st[(int64_t)w * A.L.stride + A.L.eloc];
}

__global__ void k_copy_strided(double *dst, int64_t dstride, const double *src, int64_t sstride,
                               int n, int rows) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  for (int q = 0; q < n; q++) dst[(int64_t)r * dstride + q] = src[(int64_t)r * sstride + q];
}

// -------------------------------------------------------------------------
enum { KC_JAS = 0, KC_DET = 1, KC_NLPP = 2, KC_MOVE = 3 };

static LaunchCtx make_ctx(const qmcw_handle *h, const KArgs &A, int G) {
  LaunchCtx c;
  c.A = A;
  c.stream = h->stream;
  c.G = G;
  c.Wl = h->Wl;
  c.precision = h->precision;
  c.smem_ratio = 0;
  c.smem_nlpp = 0;
  c.smem_dnlpp = 0;
  return c;
}

template <int G>
static void launch_det(qmcw_handle *h, const KArgs &A, size_t smem) {
  const dim3 grid(h->Wl / (kWave / G)), block(kWave);
  if (h->precision == QMCW_FP64)
    hipLaunchKernelGGL((k_full_det<double, G>), grid, block, smem, h->stream, A);
  else
    hipLaunchKernelGGL((k_full_det<float, G>), grid, block, smem, h->stream, A);
}
template <int G>
static void launch_fin(qmcw_handle *h, const KArgs &A) {
  const dim3 grid(h->Wl / (kWave / G)), block(kWave);
  hipLaunchKernelGGL((k_full_fin<G>), grid, block, 0, h->stream, A);
}
template <int G>
static void launch_derivs_asm(qmcw_handle *h, const KArgs &A) {
  const dim3 grid(h->Wl / (kWave / G)), block(kWave);
  hipLaunchKernelGGL((k_derivs<G>), grid, block, 0, h->stream, A, h->der);
}

#define DISPATCH_GW(gw, fn, ...)                  \
  switch (gw) {                                   \
    case 8: fn<8>(__VA_ARGS__); break;            \
    case 16: fn<16>(__VA_ARGS__); break;          \
    case 32: fn<32>(__VA_ARGS__); break;          \
    default: fn<64>(__VA_ARGS__); break;          \
  }

// Full evaluation of every walker; with derivs also the parameter
// derivatives into h->der (all orbitals, Jastrow slots, pseudopotential
// pieces, assembly).  Each kernel class runs at its own lane-group width.
// k_full_det stages A and the occupied-orbital values of large determinants
// (at least 16 electrons of a spin) in LDS for the inversion and T, after
// the basis values the launch may keep there, when a block's share stays
// within 64 KB.  Small determinants keep the global path, which measured
// faster for the 4 x 4 matrices of the C2-like system.
static void det_lds_inv(const qmcw_handle *h, KArgs &A, size_t &smem) {
  A.lds_inv = 0;
  A.lds_inv_off = 0;
  const int nsm = std::max(A.S.ns[0], A.S.ns[1]);
  const int noccm = std::max(A.S.nocc[0], A.S.nocc[1]);
  if (!h->use_lds || nsm < 16) return;
  const size_t per = (size_t)nsm * nsm + (size_t)nsm * noccm;
  const size_t off = (smem + 7) / 8;
  const size_t tot = (off + (size_t)(kWave / h->Gk[KC_DET]) * per) * 8;
  if (tot > 64 * 1024) return;
  A.lds_inv = (int)per;
  A.lds_inv_off = (int)off;
  smem = tot;
}

// The arguments of the full evaluation's kernels for the handle's current
// state (keep: take the rows the committed moves kept): jas for k_full_jas
// (and k_deriv_jas), det for k_full_det, k_gproj_mfma and k_full_fin (and
// k_derivs), nlpp for k_full_nlpp (and k_deriv_nlpp), with their dynamic LDS
struct EvalArgs {
  KArgs jas, det, nlpp;
  size_t det_smem = 0, nlpp_smem = 0, dnlpp_smem = 0;
};
static void eval_args(const qmcw_handle *h, bool derivs, bool keep, EvalArgs &e) {
  KArgs A = h->A;
  A.full_orb = derivs ? 1 : 0;
  // (the derivatives need all orbitals of every electron, the moves keep
  // only the occupied ones: k_full_det recomputes them under full_orb)
  A.keep_state = keep ? 1 : 0;
  // electrons per basis chunk of k_full_det: a wavefront's basis values
  // (64/G walkers) stay near 16 KB so that they are read back from L2
  const int gd = h->Gk[KC_DET];
  const size_t per_e = (size_t)(kWave / gd) * A.S.NB * 5 *
                       (h->precision == QMCW_FP64 ? sizeof(double) : sizeof(float));
  A.chunk = std::max(1, std::min(h->chunk, (int)((16u << 10) / std::max<size_t>(per_e, 1))));
  if (h->chunk_override > 0) A.chunk = std::min(h->chunk, h->chunk_override);
  // in LDS instead, one electron at a time, when that fits the per-block
  // budget (measured on the C2-like system: one electron in LDS beats both
  // two, which costs occupancy, and L2-sized chunks in global scratch)
  size_t det_smem = 0;
  A.lds_det = 0;
  if (h->use_lds && per_e <= h->det_lds_budget) {
    A.chunk = 1;
    A.lds_det = 1;
    det_smem = per_e;
  }
  det_lds_inv(h, A, det_smem);
  e.jas = A;
  // k_full_nlpp's per-lane point scratch (4M distances, N values of U) in
  // LDS when all of the launch's blocks stay resident with it
  e.nlpp = A;
  {
    const int M = A.S.M, N = A.S.N;
    const size_t tsz = h->precision == QMCW_FP64 ? sizeof(double) : sizeof(float);
    const int per_lane = 4 * M + (int)((N * tsz + 7) / 8);
    const size_t blk = (size_t)kWave * per_lane * sizeof(double);
    const size_t bpc = ((size_t)h->Wl / (kWave / h->Gk[KC_NLPP]) + h->ncu - 1) / h->ncu;
    const size_t cap = std::min<size_t>(48 * 1024, (160 * 1024) / std::max<size_t>(1, bpc));
    if (h->use_lds && blk <= cap) {
      e.nlpp.lds_nlpp = per_lane;
      e.nlpp_smem = blk;
    }
    // k_deriv_nlpp: the same plus N doubles per lane
    const int per_lane_d = per_lane + N;
    const size_t blk_d = (size_t)kWave * per_lane_d * sizeof(double);
    if (derivs && h->use_lds && blk_d <= cap) {
      e.nlpp.lds_dnlpp = per_lane_d;
      e.dnlpp_smem = blk_d;
    }
    // psp_nonloc_orb's contraction on FP64 MFMA after k_deriv_nlpp
    e.nlpp.pnl_mfma = h->no_gproj_mfma ? 0 : 1;
  }
  // the pseudopotential projection on FP64 MFMA, rows = (walker, electron),
  // and under full_orb the contraction of every orbital, rows = (walker,
  // electron, component)
  A.gproj_mfma = (A.S.nloc > 0 && !h->no_gproj_mfma) ? 1 : 0;
  A.orb_mfma = (derivs && A.S.nparm > 0 && !h->no_gproj_mfma) ? 1 : 0;
  e.det = A;
  e.det_smem = det_smem;
}

static void launch_eval(qmcw_handle *h, bool derivs) {
  EvalArgs e;
  eval_args(h, derivs, h->orb_valid && !h->no_keep_orb, e);
  h->orb_valid = 1;
  h->moves_since_inv = 0;
  const int gd = h->Gk[KC_DET];
  const OrdKernels &K = ord_kernels(h->A.S.nordc);
  const LaunchCtx cj = make_ctx(h, e.jas, h->Gk[KC_JAS]);
  LaunchCtx cn = make_ctx(h, e.nlpp, h->Gk[KC_NLPP]);
  cn.smem_nlpp = e.nlpp_smem;
  cn.smem_dnlpp = e.dnlpp_smem;
  const KArgs &A = e.det;
  const size_t det_smem = e.det_smem;
  prof_begin(h, QMCW_PROF_FULL_JAS);
  K.full_jas(cj);
  prof_end(h);
  if (derivs) {
    prof_begin(h, QMCW_PROF_DERIV_JAS);
    K.deriv_jas(cj);
    prof_end(h);
  }
  prof_begin(h, QMCW_PROF_FULL_DET);
  DISPATCH_GW(gd, launch_det, h, A, det_smem);
  if (A.orb_mfma)
    hipLaunchKernelGGL(k_orb_mfma,
                       dim3((unsigned)(((int64_t)h->Wl * A.S.N * 5 + 63) / 64), (A.S.NO + 63) / 64),
                       dim3(64), 0, h->stream, A, h->Wl);
  if (A.gproj_mfma)
    hipLaunchKernelGGL(k_gproj_mfma,
                       dim3((unsigned)(((int64_t)h->Wl * std::max(A.S.ns[0], A.S.ns[1]) + 63) / 64),
                            (A.S.NB + 63) / 64, 2),
                       dim3(64), 0, h->stream, A, h->Wl);
  prof_end(h);
  if (A.S.nloc > 0) {
    prof_begin(h, QMCW_PROF_FULL_NLPP);
    K.full_nlpp(cn);
    prof_end(h);
    if (derivs) {
      prof_begin(h, QMCW_PROF_DERIV_NLPP);
      K.deriv_nlpp(cn);
      if (cn.A.pnl_mfma)
        hipLaunchKernelGGL(k_pnl_mfma,
                           dim3((unsigned)(((int64_t)h->Wl * A.S.N + 63) / 64), (A.S.NO + 63) / 64),
                           dim3(64), 0, h->stream, cn.A, h->Wl);
      prof_end(h);
    }
  }
  prof_begin(h, QMCW_PROF_FULL_FIN);
  DISPATCH_GW(gd, launch_fin, h, A);
  prof_end(h);
  if (derivs) {
    prof_begin(h, QMCW_PROF_DERIVS);
    DISPATCH_GW(gd, launch_derivs_asm, h, A);
    prof_end(h);
    h->der_valid = 1;
  }
}
template <int G>
static void launch_move_bc(qmcw_handle *h, const KArgs &A, size_t smem_b, size_t smem_c, int ie,
                           int mode, double delta, double fbias, double *out, long long *nacc) {
  const dim3 grid(h->Wl / (kWave / G)), block(kWave);
  if (h->precision == QMCW_FP64)
    hipLaunchKernelGGL((k_move_b<double, G>), grid, block, smem_b, h->stream, A, ie);
  else
    hipLaunchKernelGGL((k_move_b<float, G>), grid, block, smem_b, h->stream, A, ie);
  hipLaunchKernelGGL((k_move_c<G>), grid, block, smem_c, h->stream, A, ie, mode, delta, fbias, out,
                     nacc);
}
// The one-electron move's arguments at lane-group width G: LDS for the
// moved electron's basis values (phi_b bytes) and, after them, eval_y's
// inputs (smem: the whole dynamic LDS of the single-kernel move)
static KArgs move_args(const qmcw_handle *h, int G, bool split, size_t &phi_b, size_t &smem) {
  KArgs A = h->A;
  // the moved electron's basis values in LDS when they fit (<= 48 KB per
  // block); in the single-kernel move the same region first holds the
  // e-e-n tables of the Jastrow row (one per centre)
  const size_t tsz = h->precision == QMCW_FP64 ? sizeof(double) : sizeof(float);
  const int ordk = std::max(2, A.S.nordc);
  const int tab = (A.S.nordc > 1 && !split && !h->no_een_tab)
                      ? A.S.M * (3 * ((ordk + 1) * (ordk + 2) / 2) + 8) : 0;
  const int wst = std::max(A.S.NB * 5, tab);
  const size_t lds = (size_t)(kWave / G) * wst * tsz;
  A.lds_phi = (h->use_lds && lds <= 48 * 1024) ? 1 : 0;
  A.een_tab = (A.lds_phi && tab > 0) ? 1 : 0;
  A.lds_wst = A.een_tab ? wst : A.S.NB * 5;
  phi_b = A.lds_phi ? ((size_t)(kWave / G) * A.lds_wst * tsz + 7) / 8 * 8 : 0;
  // eval_y's determinant values and alpha inverses staged in LDS after the
  // basis values, when the two together stay within 48 KB per block and
  // still let every block of the launch be resident at once (160 KB of LDS
  // per compute unit)
  const size_t y_b = (size_t)(kWave / G) *
                     y_lds_doubles(A.S.ndetu[0], A.S.ndetu[1], A.S.kmax2) * sizeof(double);
  const size_t blocks_per_cu = ((size_t)h->Wl / (kWave / G) + h->ncu - 1) / h->ncu;
  const size_t lds_cap = std::min<size_t>(48 * 1024, (160 * 1024) / std::max<size_t>(1, blocks_per_cu));
  const bool ly = h->use_lds && h->lds_y && phi_b + y_b <= lds_cap;
  A.lds_y = ly ? y_lds_doubles(A.S.ndetu[0], A.S.ndetu[1], A.S.kmax2) : 0;
  A.lds_y_off = (int)(phi_b / 8);
  smem = phi_b + (ly ? y_b : 0);
  return A;
}

static void launch_ratio(qmcw_handle *h, int ie, int mode, const double *xn, double delta,
                         double fbias, double *out, long long *nacc = nullptr) {
  const int G = h->Gk[KC_MOVE];
  size_t phi_b = 0, smem = 0;
  const KArgs A = move_args(h, G, h->split_move != 0, phi_b, smem);
  LaunchCtx c = make_ctx(h, A, G);
  c.smem_ratio = smem;
  prof_begin(h, QMCW_PROF_RATIO);
  if (h->split_move) {
    ord_kernels(A.S.nordc).move_a(c, ie, mode, xn, delta, fbias);
    DISPATCH_GW(G, launch_move_bc, h, A, phi_b, c.smem_ratio, ie, mode, delta, fbias, out, nacc);
  } else {
    ord_kernels(A.S.nordc).ratio(c, ie, mode, xn, delta, fbias, out, nacc);
  }
  prof_end(h);
}
template <int G>
static void launch_accept(qmcw_handle *h, int ie, int mode, const uint8_t *acc, double delta,
                          double fbias, long long *nacc) {
  dim3 grid(h->Wl / (kWave / G)), block(kWave);
  prof_begin(h, QMCW_PROF_ACCEPT);
  hipLaunchKernelGGL((k_accept_1e<G>), grid, block, 0, h->stream, h->A, ie, mode, acc, delta,
                     fbias, nacc);
  prof_end(h);
}
template <int G>
static void launch_grad(qmcw_handle *h, int ie, double *out) {
  dim3 grid(h->Wl / (kWave / G)), block(kWave);
  hipLaunchKernelGGL((k_grad_1e<G>), grid, block, 0, h->stream, h->A, ie, out);
}

// one-electron kernels run at the width of the move class
#define DISPATCH_MOVE(h, fn, ...) DISPATCH_GW((h)->Gk[KC_MOVE], fn, h, ##__VA_ARGS__)

// -------------------------------------------------------------------------
extern "C" int qmcw_create(const qmcw_system *s, int32_t nwalk, int32_t precision,
                           qmcw_handle **out) {
  if (!s || !out || nwalk <= 0) return fail(nullptr, QMCW_ERR_ARG, "invalid arguments");
  if (s->basis_type != QMCW_BASIS_SLATER && s->basis_type != QMCW_BASIS_GAUSSIAN)
    return fail(nullptr, QMCW_ERR_ARG, "basis_type must be Slater or Gaussian");
  if (s->isc != 2 && s->isc != 4) return fail(nullptr, QMCW_ERR_ARG, "isc must be 2 or 4");
  if (s->nordc > 7 || s->norda > QMCW_MAX_ORD || s->nordb > QMCW_MAX_ORD)
    return fail(nullptr, QMCW_ERR_ARG, "Jastrow order too large (nordc <= 7)");
  if (s->nloc != 0 && (s->nloc != 6 || (s->nquad != 12 && s->nquad != 6)))
    return fail(nullptr, QMCW_ERR_ARG, "pseudopotential must be nloc=6 with nquad=6 or 12");
  if (s->rad_num && s->nnum > 0) {
    for (int k = 0; k < s->nnum; k++) {
      const int ig = s->num_igrid[k], nr = s->num_nr[k];
      if ((ig != 1 && ig != 2 && ig != 3) || nr < 5 || s->num_off[k + 1] - s->num_off[k] != nr ||
          !(s->num_h[k] > 0) || (ig != 1 && !(s->num_h[k] > 1)) || (ig != 1 && !(s->num_r0[k] > 0)))
        return fail(nullptr, QMCW_ERR_ARG, "invalid tabulated radial function (grid)");
    }
    for (int r = 0; r < s->nrad; r++)
      if (s->rad_num[r] >= s->nnum)
        return fail(nullptr, QMCW_ERR_ARG, "rad_num names a missing tabulated function");
  }
  if (precision < QMCW_FP64 || precision > QMCW_MIXED_BF16)
    return fail(nullptr, QMCW_ERR_ARG, "unknown precision mode");
  int nup = s->nup, ndn = s->nelec - s->nup;
  if (nup > 128 || ndn > 128) return fail(nullptr, QMCW_ERR_ARG, "more than 128 electrons of one spin");
  qmcw_handle *h = new qmcw_handle();
  h->W = nwalk;
  h->precision = precision;
  DevSys &S = h->A.S;
  S.N = s->nelec; S.nup = nup; S.ndn = ndn; S.M = s->ncent; S.nctype = s->nctype;
  S.NO = s->norb; S.basis_type = s->basis_type; S.hb = s->hb > 0 ? s->hb : 0.5;
  S.pecent = 0.0;
  for (int i = 1; i < s->ncent; i++)
    for (int j = 0; j < i; j++) {
      double d2 = 0;
      for (int k = 0; k < 3; k++) {
        double d = s->cent[3 * i + k] - s->cent[3 * j + k];
        d2 += d * d;
      }
      S.pecent += s->znuc[s->iwctype[i]] * s->znuc[s->iwctype[j]] / std::sqrt(d2);
    }
  int st = 0;
#define UP(vec, dst)                                   \
  do {                                                 \
    if ((st = upload(h, vec, &(dst))) != QMCW_OK) {    \
      *out = nullptr;                                  \
      g_lib_err = h->err;                              \
      qmcw_destroy(h);                                 \
      return st;                                       \
    }                                                  \
  } while (0)
  std::vector<double> cent(s->cent, s->cent + 3 * s->ncent);
  std::vector<int> iwctype(s->iwctype, s->iwctype + s->ncent);
  std::vector<double> znuc(s->znuc, s->znuc + s->nctype);
  UP(cent, S.cent); UP(iwctype, S.iwctype); UP(znuc, S.znuc);
  // ---- basis shells: group each type's functions by (radial function
  // content, l); the components of one shell share the radial part
  std::vector<int> sh_cent, sh_l, sh_a, sh_nm, sh_prim_off{0}, sh_m, sh_bas, cent_sh_off{0};
  std::vector<int> sh_num;   // tabulated radial function of the shell or -1
  std::vector<double> pz, pc;
  const bool has_num = s->rad_num && s->nnum > 0;
  int nbas = 0, maxsh = 0;
  for (int ic = 0; ic < s->ncent; ic++) {
    int t = s->iwctype[ic];
    int b0 = s->ctype_bas_off[t], b1 = s->ctype_bas_off[t + 1];
    int first_shell = (int)sh_l.size();
    for (int q = b0; q < b1; q++) {
      int r = s->bas_rad[q], l = s->bas_l[q], m = s->bas_m[q];
      int n = s->rad_n[r];
      const int kn = has_num ? s->rad_num[r] : -1;
      if (l < 0 || l > QMCW_MAX_L || m < -l || m > l) {
        delete h;
        return fail(nullptr, QMCW_ERR_ARG, "basis function with l > 5 or |m| > l");
      }
      if (kn >= 0) n = l + 1;   // tabulated R(r) Y_lm: radial part R r^-l, no r^a factor
      if (n - 1 - l < 0) {
        delete h;
        return fail(nullptr, QMCW_ERR_ARG, "basis function with n < l + 1");
      }
      int p0 = s->rad_prim_off[r], p1 = s->rad_prim_off[r + 1];
      if (kn >= 0) p1 = p0;
      // slot of m in a complete shell: m = 0, 1, -1, 2, -2, ...
      const int slot = m > 0 ? 2 * m - 1 : (m < 0 ? -2 * m : 0);
      int found = -1;
      for (int sh = first_shell; sh < (int)sh_l.size() && found < 0; sh++) {
        if (sh_l[sh] != l || sh_a[sh] != n - 1 - l || sh_num[sh] != kn) continue;
        if (sh_bas[sh * kMaxShellM + slot] >= 0) continue;
        int np = sh_prim_off[sh + 1] - sh_prim_off[sh];
        if (np != p1 - p0) continue;
        bool same = true;
        for (int p = 0; p < np; p++)
          if (pz[sh_prim_off[sh] + p] != s->prim_zeta[p0 + p] ||
              pc[sh_prim_off[sh] + p] != s->prim_coef[p0 + p]) same = false;
        if (same) found = sh;
      }
      if (found < 0) {
        sh_cent.push_back(ic); sh_l.push_back(l); sh_a.push_back(n - 1 - l); sh_nm.push_back(0);
        sh_num.push_back(kn);
        for (int p = p0; p < p1; p++) { pz.push_back(s->prim_zeta[p]); pc.push_back(s->prim_coef[p]); }
        sh_prim_off.push_back((int)pz.size());
        for (int qq = 0; qq < kMaxShellM; qq++) { sh_m.push_back(0); sh_bas.push_back(-1); }
        found = (int)sh_l.size() - 1;
      }
      sh_m[found * kMaxShellM + slot] = m;
      sh_bas[found * kMaxShellM + slot] = nbas + (q - b0);
      sh_nm[found]++;
    }
    nbas += b1 - b0;
    cent_sh_off.push_back((int)sh_l.size());
    maxsh = std::max(maxsh, (int)sh_l.size() - first_shell);
  }
  if (nbas != s->nbasis) {
    delete h;
    return fail(nullptr, QMCW_ERR_ARG, "nbasis does not match the per-type basis lists");
  }
  S.NB = nbas; S.nshell = (int)sh_l.size(); S.maxshell_per_cent = maxsh;
  UP(sh_cent, S.sh_cent); UP(sh_l, S.sh_l); UP(sh_a, S.sh_a); UP(sh_nm, S.sh_nm);
  UP(sh_prim_off, S.sh_prim_off); UP(sh_m, S.sh_m); UP(sh_bas, S.sh_bas);
  UP(cent_sh_off, S.cent_sh_off); UP(pz, S.prim_zeta); UP(pc, S.prim_coef);
  S.has_num = 0;
  for (int v : sh_num) S.has_num |= (v >= 0);
  if (S.has_num) {
    // per table: igrid, nr, r0, 1/h or 1/ln h, ce[5], ae[2], offset; the grid
    // points as qmc/vmc/splfit.f90:13-34 places them
    std::vector<double> par((size_t)s->nnum * kNumPar), tab((size_t)4 * s->num_off[s->nnum]);
    for (int k = 0; k < s->nnum; k++) {
      const int ig = s->num_igrid[k], nr = s->num_nr[k], o = s->num_off[k];
      const double hb = s->num_h[k], r0 = s->num_r0[k];
      double *pk = par.data() + (size_t)k * kNumPar;
      pk[0] = ig; pk[1] = nr; pk[2] = r0;
      pk[3] = ig == 1 ? 1.0 / hb : 1.0 / std::log(hb);
      for (int c = 0; c < 5; c++) pk[4 + c] = s->num_ce[5 * k + c];
      pk[9] = s->num_ae[2 * k]; pk[10] = s->num_ae[2 * k + 1];
      pk[11] = o;
      for (int j = 0; j < nr; j++) {
        double xj = ig == 1 ? r0 + hb * j : (ig == 2 ? r0 * std::pow(hb, j) : r0 * std::pow(hb, j) - r0);
        tab[4 * (size_t)(o + j) + 0] = xj;
        tab[4 * (size_t)(o + j) + 1] = s->num_rwf[o + j];
        tab[4 * (size_t)(o + j) + 2] = s->num_d2[o + j];
        tab[4 * (size_t)(o + j) + 3] = 0.0;
      }
    }
    UP(sh_num, S.sh_num); UP(par, S.num_par); UP(tab, S.num_tab);
  }
  h->coef.assign(s->coef, s->coef + (size_t)s->norb * s->nbasis);
  std::vector<double> coefT((size_t)nbas * s->norb);
  for (int o = 0; o < s->norb; o++)
    for (int b = 0; b < nbas; b++) coefT[(size_t)b * s->norb + o] = h->coef[(size_t)o * nbas + b];
  UP(coefT, S.coefT);
  h->d_coefT = const_cast<double *>(S.coefT);
  // ---- determinants
  DetBook D = build_dets(s);
  if (D.kmax > kMaxKmax) {
    delete h;
    return fail(nullptr, QMCW_ERR_ARG, "excitation order above 8");
  }
  S.ndet = s->ndet; S.kmax = D.kmax; S.kmax2 = std::max(1, D.kmax * D.kmax);
  S.noccmax = std::max(D.nocc[0], D.nocc[1]);
  S.single_det = (s->ndet == 1 && D.kmax == 0 && D.ndetu[0] <= 1 && D.ndetu[1] <= 1 &&
                  D.nocc[0] == D.ns[0] && D.nocc[1] == D.ns[1]) ? 1 : 0;
  if (const char *e = getenv("QMCW_NO_SINGLE_DET")) if (atoi(e)) S.single_det = 0;
  for (int sp = 0; sp < 2; sp++) {
    S.ns[sp] = D.ns[sp]; S.nocc[sp] = D.nocc[sp]; S.ndetu[sp] = D.ndetu[sp];
    UP(D.ref[sp], S.ref[sp]); UP(D.occ[sp], S.occ[sp]); UP(D.iref[sp], S.iref[sp]);
    UP(D.ivir[sp], S.ivir[sp]); UP(D.ord[sp], S.ord[sp]); UP(D.sgn[sp], S.dsgn[sp]);
    UP(D.rf[sp], S.rf[sp]); UP(D.exl[sp], S.exl[sp]); UP(D.exo[sp], S.exo[sp]);
    UP(D.iwdet[sp], S.iwdet[sp]); UP(D.iocc_lab[sp], S.iocc_lab[sp]);
    h->iwdet_h[sp] = D.iwdet[sp]; h->dsgn_h[sp] = D.sgn[sp];
    h->occ_h[sp] = D.occ[sp];
    {
      std::vector<double> co = coef_occ_table(coefT, S.NB, S.NO, D.occ[sp]);
      UP(co, S.coef_occ[sp]);
      h->d_coef_occ[sp] = const_cast<double *>(S.coef_occ[sp]);
    }
    // CSR lists of eval_y: dets of each unique det; contributions to each Y entry
    {
      const int nu = D.ndetu[sp], ns = D.ns[sp], nocc = D.nocc[sp];
      std::vector<int> uoff(nu + 1, 0), ud;
      for (int u = 0; u < nu; u++) {
        for (int d = 0; d < s->ndet; d++)
          if (D.iwdet[sp][d] == u) ud.push_back(d);
        uoff[u + 1] = (int)ud.size();
      }
      std::vector<int> yoff(nocc * ns + 1, 0), yu, yp;
      for (int e = 0; e < nocc * ns; e++) {
        const int jo = e / ns, ir = e % ns;
        for (int u = 1; u < nu; u++) {
          const int k = D.ord[sp][u];
          int jj = -1, ii = -1;
          for (int q = 0; q < k; q++) {
            if (D.exo[sp][u * ns + q] == jo) jj = q;
            if (D.rf[sp][u * ns + q] == ir) ii = q;
          }
          if (jj < 0 || ii < 0) continue;
          yu.push_back(u);
          yp.push_back(u * S.kmax2 + jj + k * ii);
        }
        yoff[e + 1] = (int)yu.size();
      }
      h->udet_d[sp] = ud;
      UP(uoff, S.udet_off[sp]); UP(ud, S.udet_d[sp]);
      UP(yoff, S.yent_off[sp]); UP(yu, S.yent_u[sp]); UP(yp, S.yent_pos[sp]);
    }
  }
  h->ncsf = s->ncsf; h->ndet = s->ndet;
  h->csf_det_off.assign(s->csf_det_off, s->csf_det_off + s->ncsf + 1);
  h->csf_det.assign(s->csf_det, s->csf_det + s->csf_det_off[s->ncsf]);
  h->csf_det_coef.assign(s->csf_det_coef, s->csf_det_coef + s->csf_det_off[s->ncsf]);
  h->csf_coef.assign(s->csf_coef, s->csf_coef + s->ncsf);
  h->cdet = build_cdet(s->ndet, s->ncsf, s->csf_coef, s->csf_det_off, s->csf_det, s->csf_det_coef);
  UP(h->cdet, S.cdet);
  h->d_cdet = const_cast<double *>(S.cdet);
  {
    std::vector<double> cdd, ucf[2];
    fold_det_signs(h, cdd, ucf);
    UP(cdd, S.cdd); h->d_cdd = const_cast<double *>(S.cdd);
    for (int sp = 0; sp < 2; sp++) {
      std::vector<int> iu, id;
      for (int d : h->udet_d[sp]) { iu.push_back(D.iwdet[0][d]); id.push_back(D.iwdet[1][d]); }
      if (iu.empty()) { iu.push_back(0); id.push_back(0); ucf[sp].push_back(0.0); }
      UP(ucf[sp], S.ucf[sp]); h->d_ucf[sp] = const_cast<double *>(S.ucf[sp]);
      UP(iu, S.uiu[sp]); UP(id, S.uid[sp]);
    }
  }
  S.ncsf = s->ncsf;
  UP(h->csf_det_off, S.csf_det_off); UP(h->csf_det, S.csf_det);
  UP(h->csf_det_coef, S.csf_det_coef); UP(h->csf_coef, S.csf_coef);
  h->d_csf_coef = const_cast<double *>(S.csf_coef);
  // ---- Jastrow
  if ((s->isc != 2 && s->isc != 4) || s->norda < 1 || s->nordb < 1 || s->norda > QMCW_MAX_ORD ||
      s->nordb > QMCW_MAX_ORD || s->nordc < 0 || s->nordc > 7 || (s->nspin2b != 1 && s->nspin2b != 2)) {
    delete h;
    return fail(nullptr, QMCW_ERR_ARG,
                "Jastrow: isc must be 2 or 4, 1 <= norda, nordb <= 8, nordc <= 7, nspin2b 1 or 2");
  }
  S.isc = s->isc; S.nspin2b = s->nspin2b; S.nocuspb = s->nocuspb;
  S.norda = s->norda; S.nordb = s->nordb; S.nordc = s->nordc; S.scalek = s->scalek;
  S.nterms = qmcw_nterms4(s->nordc);
  h->a4.assign(s->a4, s->a4 + s->nctype * (s->norda + 1));
  h->b.assign(s->b, s->b + s->nspin2b * (s->nordb + 1));
  h->c.assign(s->c, s->c + std::max(1, s->nctype * S.nterms));
  set_asymp(h, s, S);
  std::vector<double> c_ord = reorder_een(h->c, s->nctype, s->nordc);
  UP(h->a4, S.a4); UP(h->b, S.b); UP(c_ord, S.c); UP(h->asymp_jasa, S.asymp_jasa);
  h->d_a4 = const_cast<double *>(S.a4); h->d_b = const_cast<double *>(S.b);
  h->d_c = const_cast<double *>(S.c); h->d_asymp = const_cast<double *>(S.asymp_jasa);
  std::vector<int> klm;
  for (int n = 2; n <= s->nordc; n++)
    for (int k = n - 1; k >= 0; k--) {
      int l_hi = (k == 0) ? n - k - 2 : n - k;
      for (int l = l_hi; l >= 0; l--) {
        int m = (n - k - l) / 2;
        if (2 * m == n - k - l) { klm.push_back(k); klm.push_back(l); klm.push_back(m); }
      }
    }
  UP(klm, S.term_klm);
  // ---- pseudopotential
  S.nloc = s->nloc;
  S.nquad = (s->nloc > 0 && s->nquad == 6) ? 6 : 12;
  std::vector<double> q0(4 * S.nquad);
  if (S.nquad == 6)
    gesqua6(q0.data());
  else
    gesqua12(q0.data());
  UP(q0, S.q0);
  if (s->nloc > 0) {
    std::vector<int> npotd(s->npotd, s->npotd + s->nctype), lpotp1(s->lpotp1, s->lpotp1 + s->nctype);
    int nt = s->ecp_off[s->nctype * QMCW_MAX_LPOT];
    std::vector<int> off(s->ecp_off, s->ecp_off + s->nctype * QMCW_MAX_LPOT + 1);
    std::vector<int> pw(s->ecp_power, s->ecp_power + nt);
    std::vector<double> cf(s->ecp_coef, s->ecp_coef + nt), ex(s->ecp_expo, s->ecp_expo + nt);
    UP(npotd, S.npotd); UP(lpotp1, S.lpotp1); UP(off, S.ecp_off); UP(pw, S.ecp_power);
    UP(cf, S.ecp_coef); UP(ex, S.ecp_expo);
    // T-moves keep the points of at most tm_kmax pairs per electron: the most
    // centres whose nonlocal ranges (some |v_l| > 1e-4, scanned to 40 bohr
    // in 1e-3 steps) can overlap at one point, with a 0.02 bohr margin
    std::vector<double> rc(s->nctype, 0.0);
    for (int t = 0; t < s->nctype; t++)
      for (int k = 1; k <= 40000; k++) {
        const double r = 1e-3 * k;
        for (int l = 1; l <= npotd[t]; l++) {
          if (l == lpotp1[t]) continue;
          double v = 0.0;
          for (int q = off[t * QMCW_MAX_LPOT + l - 1]; q < off[t * QMCW_MAX_LPOT + l]; q++)
            v += (pw[q] != 2 ? std::pow(r, (double)(pw[q] - 2)) : 1.0) * cf[q] * std::exp(-ex[q] * r * r);
          if (std::fabs(v) > 1e-4) rc[t] = r;
        }
      }
    int km = 1;
    for (int c = 0; c < S.M; c++) {
      const double r1 = rc[s->iwctype[c]];
      if (r1 <= 0.0) continue;
      int n = 0;
      for (int c2 = 0; c2 < S.M; c2++) {
        const double r2 = rc[s->iwctype[c2]];
        double d2 = 0.0;
        for (int k = 0; k < 3; k++) d2 += (s->cent[3 * c + k] - s->cent[3 * c2 + k]) * (s->cent[3 * c + k] - s->cent[3 * c2 + k]);
        if (r2 > 0.0 && std::sqrt(d2) < r1 + r2 + 0.02) n++;
      }
      km = std::max(km, n);
    }
    // QMCW_TM_KMAX lowers the bound (never raises it), to exercise the
    // record-overflow status
    if (const char *e = getenv("QMCW_TM_KMAX")) km = std::max(1, std::min(km, atoi(e)));
    S.tm_kmax = km;
  }
  // ---- parameter counts (derivatives)
  S.npcsf = s->nparmcsf; S.npa = s->nparma_tot; S.npb = s->nparmb_tot; S.npc = s->nparmc_tot;
  S.nporb = s->nparmorb; S.npj = S.npa + S.npb + S.npc;
  S.nparm = S.npcsf + S.npj + S.nporb;
  h->nparm = S.nparm;
  {
    JasMap jm;
    std::string why;
    if (!build_jas_map(s, S.nterms, jm, why)) {
      delete h;
      return fail(nullptr, QMCW_ERR_ARG, why);
    }
    S.nslot = jm.nslot; S.slot_b0 = jm.slot_b0; S.slot_c0 = jm.slot_c0;
    UP(jm.off, S.pj_off); UP(jm.slot, S.pj_slot); UP(jm.w, S.pj_w);
    std::vector<int> pcsf(s->csf_parm, s->csf_parm + s->nparmcsf);
    for (int c : pcsf)
      if (c < 0 || c >= s->ncsf) { delete h; return fail(nullptr, QMCW_ERR_ARG, "CSF parameter index out of range"); }
    std::vector<int> oi(s->orb_parm_i, s->orb_parm_i + s->nparmorb),
        oj(s->orb_parm_j, s->orb_parm_j + s->nparmorb), orv(s->orb_parm_rev, s->orb_parm_rev + s->nparmorb);
    for (int q = 0; q < s->nparmorb; q++)
      if (oi[q] < 0 || oi[q] >= s->norb || oj[q] < 0 || oj[q] >= s->norb) {
        delete h;
        return fail(nullptr, QMCW_ERR_ARG, "orbital rotation index out of range");
      }
    UP(pcsf, S.pcsf); UP(oi, S.porb_i); UP(oj, S.porb_j); UP(orv, S.porb_rev);
  }
  // ---- lane-group width and layouts
  int N = S.N, M = S.M, NO = S.NO, NB = S.NB;
  int G = N <= 4 ? 8 : (N <= 16 ? 16 : (N <= 40 ? 32 : 64));
  if (const char *e = getenv("QMCW_GROUP")) {
    int g = atoi(e);
    if (g == 8 || g == 16 || g == 32 || g == 64) G = g;
  }
  h->G = G;
  for (int k = 0; k < 4; k++) h->Gk[k] = G;
  // padded walker count: a multiple of the walkers per wavefront of every
  // group width (8 for G = 8), so that the autotuner can switch G freely
  h->Wp = ((nwalk + 7) / 8) * 8;
  h->chunk = std::max(1, std::min(N, 32768 / (NB * 5)));
  h->A.chunk = h->chunk;
  h->A.full_orb = 0;
  h->A.derivs = 0;
  h->A.bf16 = precision == QMCW_MIXED_BF16 ? 1 : 0;
  h->A.imetro = 1;
  h->A.deltar = 5.0;
  h->A.deltat = 1.0;
  int nsm = std::max(nup, ndn), noccm = S.noccmax;
  WLayout &L = h->A.L;
  int64_t o = 0;
  auto take = [&o](int64_t n) { int64_t r = o; o += std::max<int64_t>(n, 1); return r; };
  L.x = take(3 * N); L.rvec_en = take(3 * N * M); L.r_en = take(N * M);
  L.rvec_ee = take(3 * N * N); L.r_ee = take(N * N);
  L.fso = take(N * N); L.fijo = take(3 * N * N); L.lapj = take(N * N); L.fjo = take(3 * N);
  L.fsumo = take(1); L.oc = take((int64_t)N * noccm * 5);
  for (int sp = 0; sp < 2; sp++) {
    for (int sl = 0; sl < 2; sl++) {
      L.ai[sp][sl] = take(D.ns[sp] * D.ns[sp]); L.tocc[sp][sl] = take(D.ns[sp] * D.nocc[sp]);
      L.inv[sp][sl] = take(D.ndetu[sp] * S.kmax2); L.detv[sp][sl] = take(D.ndetu[sp]);
      L.y[sp][sl] = take(D.nocc[sp] * D.ns[sp]);
    }
    L.deta[sp] = take(1);
  }
  L.par = take(5);
  L.chi = take(1); L.psi2o = take(1); L.logpsid = take(1); L.sgn = take(1); L.psij = take(1);
  L.eloc = take(1); L.epot = take(1); L.quad = take(36);
  L.stride = (o + 7) / 8 * 8;
  MLayout &ML = h->A.ML;
  o = 0;
  ML.xn = take(3); ML.ren = take(M); ML.rven = take(3 * M); ML.ree = take(N); ML.rvee = take(3 * N);
  ML.fsn = take(N); ML.fijn_i = take(3 * N); ML.fijn_j = take(3 * N); ML.lapn_i = take(2 * N);
  ML.fjn = take(3 * N); ML.orbe = take(noccm * 5);
  ML.misc = take(24 + nsm + noccm); ML.slat = take(4);
  ML.stride = o;
  SLayout &SL = h->A.SL;
  o = 0;
  SL.phi = take((int64_t)h->chunk * NB * 5);
  SL.orb = take(S.nparm > 0 ? (int64_t)N * NO * 5 : 1);
  SL.g[0] = take(D.nocc[0] * D.ns[0]); SL.g[1] = take(D.nocc[1] * D.ns[1]);
  SL.vd = take(3 * N); SL.divv = take(N); SL.vj = take(3 * N);
  SL.red = take(std::max(std::max(64, nsm * nsm), noccm * nsm) + 64);
  SL.cup = take(D.ndetu[0] + D.ndetu[1]);
  SL.phs = take(64);
  SL.qd = take((int64_t)kWave * 4 * M);
  SL.vnl = take((int64_t)N * M * (QMCW_MAX_LPOT + 1));
  SL.gproj = take((int64_t)N * NB);
  SL.bt = take((int64_t)N * M);
  SL.uq = take((int64_t)kWave * N);
  SL.plist = take(((int64_t)N * M + 1) / 2);
  SL.prot = take(S.nloc > 0 ? 10 * (int64_t)N * M : 1);
  SL.fold = take(N);
  SL.mv = take(ML.stride);
  if (S.nparm > 0) {
    DLayout &DL = h->A.DL;
    int64_t NM = (int64_t)N * M;
    DL.raw = take((int64_t)S.nslot * N * 5);
    DL.gvalue = take(S.npj); DL.gjas = take((int64_t)S.npj * N * 3); DL.d2g = take(S.npj);
    DL.pbc = take(NM * NB); DL.dpe_raw = take(NM * S.nslot);
    DL.dpe_slot = take(S.nslot); DL.dpe = take(S.npj); DL.pnl = take((int64_t)N * NO);
    DL.btd = take(NM); DL.uqd = take((int64_t)kWave * N);
    DL.B = take((int64_t)N * NO);
    for (int sp = 0; sp < 2; sp++) {
      DL.tfull[sp] = take((int64_t)D.ns[sp] * NO); DL.dt[sp] = take((int64_t)D.ns[sp] * NO);
      DL.del[sp] = take(D.ndetu[sp]); DL.c0[sp] = take(D.ndetu[sp]); DL.c1[sp] = take(D.ndetu[sp]);
      DL.idti[sp] = take((int64_t)D.ndetu[sp] * S.kmax2);
      DL.dy[sp] = take(D.nocc[sp] * D.ns[sp]); DL.dg[sp] = take(D.nocc[sp] * D.ns[sp]);
    }
    DL.aib = take((int64_t)nsm * NO);
    DL.aux4 = take(noccm * nsm);
    DL.misc = take(8);
    DL.phiall = take((int64_t)N * NB * 5);
  }
  SL.stride = (o + 7) / 8 * 8;
#undef UP
  // ---- device allocations
  if (hipStreamCreate(&h->stream) != hipSuccess) { delete h; return fail(nullptr, QMCW_ERR_HIP, "hipStreamCreate"); }
  size_t stb = (size_t)h->Wp * L.stride * sizeof(double), scb = (size_t)h->Wp * SL.stride * sizeof(double);
  void *p;
  if (hipMalloc(&p, stb) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_NOMEM, "state alloc"); }
  h->allocs.push_back(p); h->A.st = (double *)p;
  if (hipMalloc(&p, scb) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_NOMEM, "scratch alloc"); }
  h->allocs.push_back(p); h->A.sc = (double *)p;
  if (hipMalloc(&p, (size_t)h->Wp * 4 * sizeof(int)) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_NOMEM, "rng alloc"); }
  h->allocs.push_back(p); h->A.rng = (int *)p;
  if (hipMalloc(&p, (size_t)h->Wp * sizeof(int)) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_NOMEM, "flag alloc"); }
  h->allocs.push_back(p); h->A.flags = (int *)p;
  if (hipMalloc(&p, (size_t)h->Wp * sizeof(long long)) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_NOMEM, "counter alloc"); }
  h->allocs.push_back(p); h->nacc = (long long *)p;
  if (hipMalloc(&p, (size_t)h->Wp) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_NOMEM, "flag alloc"); }
  h->allocs.push_back(p); h->acc_buf = (uint8_t *)p;
  if (S.nparm > 0) {
    if (hipMalloc(&p, (size_t)h->Wp * 2 * S.nparm * sizeof(double)) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_NOMEM, "derivative alloc"); }
    h->allocs.push_back(p); h->der = (double *)p;
  }
  h->io_doubles = (size_t)h->Wp * std::max(8 + 4 * N, 8);
  if (hipMalloc(&p, h->io_doubles * sizeof(double)) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_NOMEM, "io alloc"); }
  h->allocs.push_back(p); h->io = (double *)p;
  (void)hipMemsetAsync(h->A.st, 0, stb, h->stream);
  (void)hipMemsetAsync(h->A.sc, 0, scb, h->stream);
  (void)hipMemsetAsync(h->A.rng, 0, (size_t)h->Wp * 4 * sizeof(int), h->stream);
  (void)hipMemsetAsync(h->nacc, 0, (size_t)h->Wp * sizeof(long long), h->stream);
  h->A.Wp = h->Wp;
  h->Wl = h->Wp;
  if (const char *e = getenv("QMCW_NO_GRAPH")) h->use_graph = atoi(e) == 0;
  if (const char *e = getenv("QMCW_NO_LDS")) h->use_lds = atoi(e) == 0;
  if (const char *e = getenv("QMCW_CHUNK")) h->chunk_override = atoi(e);
  if (const char *e = getenv("QMCW_DET_LDS")) h->det_lds_budget = (size_t)atol(e);
  if (const char *e = getenv("QMCW_SPLIT_MOVE")) { h->split_move = atoi(e) != 0; h->split_forced = 1; }
  if (const char *e = getenv("QMCW_NO_LDS_Y")) h->lds_y = atoi(e) == 0;
  if (const char *e = getenv("QMCW_NO_KEEP_ORB")) h->no_keep_orb = atoi(e) != 0;
  if (const char *e = getenv("QMCW_DMC_UNFUSED")) h->dmc_unfused = atoi(e) != 0;
  if (const char *e = getenv("QMCW_HOST_SPLIT_JOIN")) h->host_split_join = atoi(e) != 0;
  if (const char *e = getenv("QMCW_NO_GPROJ_MFMA")) h->no_gproj_mfma = atoi(e) != 0;
  if (const char *e = getenv("QMCW_NO_EEN_TAB")) h->no_een_tab = atoi(e) != 0;
  {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      h->ncu = n;
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) { qmcw_destroy(h); return fail(nullptr, QMCW_ERR_HIP, "init sync"); }
  *out = h;
  return QMCW_OK;
}

static void drop_graph(qmcw_handle *h);

extern "C" int qmcw_destroy(qmcw_handle *h) {
  if (!h) return QMCW_OK;
  drop_graph(h);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->trace) (void)hipFree(h->trace);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return QMCW_OK;
}

extern "C" const char *qmcw_last_error(const qmcw_handle *h) {
  return h ? h->err.c_str() : g_lib_err.c_str();
}

extern "C" int32_t qmcw_nparm(const qmcw_handle *h) { return h ? h->nparm : 0; }

extern "C" int qmcw_sync(qmcw_handle *h) {
  if (!h) return QMCW_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return QMCW_OK;
}

static void drop_graph(qmcw_handle *h);

extern "C" int qmcw_set_params(qmcw_handle *h, const double *a4, const double *b, const double *c,
                               const double *csf_coef, const double *coef) {
  if (!h) return QMCW_ERR_ARG;
  drop_graph(h);   // the captured kernel arguments hold the Jastrow asymptotes
  h->orb_valid = 0;
  DevSys &S = h->A.S;
  if (a4) h->a4.assign(a4, a4 + h->a4.size());
  if (b) h->b.assign(b, b + h->b.size());
  if (c) h->c.assign(c, c + h->c.size());
  if (a4 || b) {
    // asymptotic constants depend on a4 and b (set_scale_dist)
    double ar = S.asymp_r;
    for (int it = 0; it < S.nctype; it++) {
      const double *a = h->a4.data() + it * (S.norda + 1);
      double v = a[0] * ar / (1.0 + a[1] * ar);
      for (int iord = 2; iord <= S.norda; iord++) v += a[iord] * std::pow(ar, iord);
      h->asymp_jasa[it] = v;
    }
    for (int i = 0; i < 2; i++) {
      double sspinn = 1.0;
      int isp = 0;
      if (i == 1) { sspinn = (S.nspin2b == 1 && S.nocuspb == 0) ? 0.5 : 1.0; isp = S.nspin2b - 1; }
      const double *bb = h->b.data() + isp * (S.nordb + 1);
      double v = sspinn * bb[0] * ar / (1.0 + bb[1] * ar);
      for (int iord = 2; iord <= S.nordb; iord++) v += bb[iord] * std::pow(ar, iord);
      if (i == 0) S.asymp_jasb0 = v; else S.asymp_jasb1 = v;
    }
    HIPCHK(h, hipMemcpy(h->d_asymp, h->asymp_jasa.data(), h->asymp_jasa.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_a4, h->a4.data(), h->a4.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_b, h->b.data(), h->b.size() * 8, hipMemcpyHostToDevice));
  }
  if (c) {
    std::vector<double> c_ord = reorder_een(h->c, S.nctype, S.nordc);
    HIPCHK(h, hipMemcpy(h->d_c, c_ord.data(), c_ord.size() * 8, hipMemcpyHostToDevice));
  }
  if (csf_coef) {
    h->csf_coef.assign(csf_coef, csf_coef + h->ncsf);
    h->cdet = build_cdet(h->ndet, h->ncsf, h->csf_coef.data(), h->csf_det_off.data(),
                         h->csf_det.data(), h->csf_det_coef.data());
    HIPCHK(h, hipMemcpy(h->d_cdet, h->cdet.data(), h->cdet.size() * 8, hipMemcpyHostToDevice));
    std::vector<double> cdd, ucf[2];
    fold_det_signs(h, cdd, ucf);
    HIPCHK(h, hipMemcpy(h->d_cdd, cdd.data(), cdd.size() * 8, hipMemcpyHostToDevice));
    for (int sp = 0; sp < 2; sp++)
      if (!ucf[sp].empty())
        HIPCHK(h, hipMemcpy(h->d_ucf[sp], ucf[sp].data(), ucf[sp].size() * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_csf_coef, h->csf_coef.data(), h->csf_coef.size() * 8, hipMemcpyHostToDevice));
  }
  if (coef) {
    int NB = S.NB, NO = S.NO;
    h->coef.assign(coef, coef + (size_t)NO * NB);
    std::vector<double> coefT((size_t)NB * NO);
    for (int o = 0; o < NO; o++)
      for (int bb = 0; bb < NB; bb++) coefT[(size_t)bb * NO + o] = h->coef[(size_t)o * NB + bb];
    HIPCHK(h, hipMemcpy(h->d_coefT, coefT.data(), coefT.size() * 8, hipMemcpyHostToDevice));
    for (int sp = 0; sp < 2; sp++) {
      std::vector<double> co = coef_occ_table(coefT, NB, NO, h->occ_h[sp]);
      HIPCHK(h, hipMemcpy(h->d_coef_occ[sp], co.data(), co.size() * 8, hipMemcpyHostToDevice));
    }
  }
  return QMCW_OK;
}

extern "C" int qmcw_set_walkers(qmcw_handle *h, const double *x) {
  if (!h || !x) return QMCW_ERR_ARG;
  int N = h->A.S.N;
  h->dmc_on = 0;            // a new configuration ends a DMC population
  h->orb_valid = 0;
  h->Wl = h->Wp;
  h->A.Wp = h->Wp;
  HIPCHK(h, hipMemcpyAsync(h->io, x, (size_t)h->W * 3 * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_copy_strided, dim3((h->W + 63) / 64), dim3(64), 0, h->stream,
                     h->A.st + h->A.L.x, (int64_t)h->A.L.stride, (const double *)h->io,
                     (int64_t)(3 * N), 3 * N, h->W);
  hipLaunchKernelGGL(k_init_walkers, dim3((h->Wp + 63) / 64), dim3(64), 0, h->stream, h->A, h->W);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return QMCW_OK;
}

extern "C" int qmcw_get_walkers(const qmcw_handle *hc, double *x) {
  qmcw_handle *h = const_cast<qmcw_handle *>(hc);
  if (!h || !x) return QMCW_ERR_ARG;
  int N = h->A.S.N;
  hipLaunchKernelGGL(k_copy_strided, dim3((h->W + 63) / 64), dim3(64), 0, h->stream, h->io,
                     (int64_t)(3 * N), (const double *)(h->A.st + h->A.L.x), (int64_t)h->A.L.stride,
                     3 * N, h->W);
  HIPCHK(h, hipMemcpyAsync(x, h->io, (size_t)h->W * 3 * N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return QMCW_OK;
}

extern "C" int qmcw_set_rng(qmcw_handle *h, const int32_t *seed) {
  if (!h || !seed) return QMCW_ERR_ARG;
  std::vector<int> s4((size_t)h->Wp * 4);
  for (int w = 0; w < h->Wp; w++)
    for (int q = 0; q < 4; q++) s4[4 * w + q] = seed[4 * std::min(w, h->W - 1) + q];
  HIPCHK(h, hipMemcpy(h->A.rng, s4.data(), s4.size() * sizeof(int), hipMemcpyHostToDevice));
  return QMCW_OK;
}

extern "C" int qmcw_get_rng(const qmcw_handle *hc, int32_t *seed) {
  qmcw_handle *h = const_cast<qmcw_handle *>(hc);
  if (!h || !seed) return QMCW_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(seed, h->A.rng, (size_t)h->W * 4 * sizeof(int), hipMemcpyDeviceToHost));
  return QMCW_OK;
}

static int check_flags(qmcw_handle *h) {
  const int n = h->dmc_on ? h->nact : h->W;
  std::vector<int> fl(n);
  HIPCHK(h, hipMemcpyAsync(fl.data(), h->A.flags, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  int any = 0;
  for (int f : fl) any |= f;
  if (any & 1) return fail(h, QMCW_ERR_SINGULAR, "a Slater matrix is singular");
  if (any & 2) return fail(h, QMCW_ERR_NAN, "non-finite local energy");
  if (any & 4) return fail(h, QMCW_ERR_ARG, "an electron is in range of more pseudopotential centres than the T-move record holds");
  return QMCW_OK;
}

extern "C" int qmcw_eval_full(qmcw_handle *h, double *logpsid, double *sgn, double *psij,
                              double *grad, double *lap, double *eloc, double *epot) {
  if (!h) return QMCW_ERR_ARG;
  launch_eval(h, false);
  HIPCHK(h, hipGetLastError());
  int N = h->A.S.N, stride = 8 + 4 * N;
  hipLaunchKernelGGL(k_gather_full, dim3((h->W + 63) / 64), dim3(64), 0, h->stream, h->A, h->io, h->W);
  std::vector<double> o((size_t)h->W * stride);
  HIPCHK(h, hipMemcpyAsync(o.data(), h->io, o.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int w = 0; w < h->W; w++) {
    const double *r = o.data() + (size_t)w * stride;
    if (logpsid) logpsid[w] = r[0];
    if (sgn) sgn[w] = r[1];
    if (psij) psij[w] = r[2];
    if (eloc) eloc[w] = r[3];
    if (epot) epot[w] = r[4];
    if (grad) memcpy(grad + (size_t)w * 3 * N, r + 8, 3 * N * sizeof(double));
    if (lap) memcpy(lap + (size_t)w * N, r + 8 + 3 * N, N * sizeof(double));
  }
  return check_flags(h);
}

extern "C" int qmcw_ratio_1e(qmcw_handle *h, int32_t ie, const double *xnew, double *ratio,
                             double *grad_new, double *lap_new) {
  if (!h || !xnew || ie < 0 || ie >= h->A.S.N) return QMCW_ERR_ARG;
  // xnew staged after the 8*Wp output block
  double *xin = h->io + (size_t)h->Wp * 8;
  std::vector<double> xp((size_t)h->Wp * 3);
  for (int w = 0; w < h->Wp; w++)
    for (int q = 0; q < 3; q++) xp[3 * w + q] = xnew[3 * std::min(w, h->W - 1) + q];
  HIPCHK(h, hipMemcpyAsync(xin, xp.data(), xp.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  launch_ratio(h, ie, 0, (const double *)xin, 1.0, 1.0, h->io);
  HIPCHK(h, hipGetLastError());
  std::vector<double> o((size_t)h->W * 8);
  HIPCHK(h, hipMemcpyAsync(o.data(), h->io, o.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int w = 0; w < h->W; w++) {
    if (ratio) ratio[w] = o[8 * w];
    if (grad_new) for (int q = 0; q < 3; q++) grad_new[3 * w + q] = o[8 * w + 4 + q];
    if (lap_new) lap_new[w] = o[8 * w + 7];
  }
  return QMCW_OK;
}

extern "C" int qmcw_accept_1e(qmcw_handle *h, int32_t ie, const uint8_t *accept) {
  if (!h || !accept || ie < 0 || ie >= h->A.S.N) return QMCW_ERR_ARG;
  std::vector<uint8_t> a(h->Wp, 0);
  for (int w = 0; w < h->W; w++) a[w] = accept[w];
  HIPCHK(h, hipMemcpyAsync(h->acc_buf, a.data(), h->Wp, hipMemcpyHostToDevice, h->stream));
  DISPATCH_MOVE(h, launch_accept, ie, 0, (const uint8_t *)h->acc_buf, 1.0, 1.0, (long long *)nullptr);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return QMCW_OK;
}

extern "C" int qmcw_grad_1e(qmcw_handle *h, int32_t ie, double *grad, double *lap) {
  if (!h || ie < 0 || ie >= h->A.S.N) return QMCW_ERR_ARG;
  DISPATCH_MOVE(h, launch_grad, ie, h->io);
  HIPCHK(h, hipGetLastError());
  std::vector<double> o((size_t)h->W * 4);
  HIPCHK(h, hipMemcpyAsync(o.data(), h->io, o.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int w = 0; w < h->W; w++) {
    if (grad) for (int q = 0; q < 3; q++) grad[3 * w + q] = o[4 * w + q];
    if (lap) lap[w] = o[4 * w + 3];
  }
  return QMCW_OK;
}

// one VMC step: grid rotation, sweep over the electrons, full evaluation
// Product sums of the derivative samples (engine_accum.h); buffers sized by
// W and P at first use, outside any graph capture.
static int accum_alloc(qmcw_handle *h) {
  if (h->acc) return QMCW_OK;
  const int P = h->nparm;
  if (P <= 0) return fail(h, QMCW_ERR_ARG, "the system has no parameters to differentiate");
  h->acc_nch = (h->W + kAccChunk - 1) / kAccChunk;
  void *p;
  if (hipMalloc(&p, accum_doubles(P) * sizeof(double)) != hipSuccess) return fail(h, QMCW_ERR_NOMEM, "accumulator alloc");
  h->allocs.push_back(p); h->acc = (double *)p;
  if (hipMalloc(&p, (size_t)h->acc_nch * 3 * P * P * sizeof(double)) != hipSuccess) return fail(h, QMCW_ERR_NOMEM, "accumulator alloc");
  h->allocs.push_back(p); h->acc_part = (double *)p;
  if (hipMalloc(&p, (size_t)h->acc_nch * (3 + 3 * P) * sizeof(double)) != hipSuccess) return fail(h, QMCW_ERR_NOMEM, "accumulator alloc");
  h->allocs.push_back(p); h->acc_partv = (double *)p;
  const int64_t n = accum_doubles(P);
  hipLaunchKernelGGL(k_accum_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->acc, n);
  return QMCW_OK;
}

static void launch_accum(qmcw_handle *h) {
  const int P = h->nparm, nb = (P + kAccBlk - 1) / kAccBlk, nch = h->acc_nch;
  const int64_t n = accum_doubles(P);
  prof_begin(h, QMCW_PROF_ACCUM);
  hipLaunchKernelGGL(k_accum_gemm, dim3(nb * nb, nch), dim3(64), 0, h->stream, (const double *)h->der,
                     2 * P, P, (const double *)h->A.st, (int64_t)h->A.L.stride, (int64_t)h->A.L.eloc,
                     h->W, h->acc_part);
  hipLaunchKernelGGL(k_accum_vec, dim3((P + 1 + 63) / 64, nch), dim3(64), 0, h->stream,
                     (const double *)h->der, 2 * P, P, (const double *)h->A.st,
                     (int64_t)h->A.L.stride, (int64_t)h->A.L.eloc, h->W, h->acc_partv);
  hipLaunchKernelGGL(k_accum_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                     (const double *)h->acc_part, (const double *)h->acc_partv, nch, P, h->acc);
  prof_end(h);
}

// The Slater inverses, T, the alpha inverses, Y and G rebuilt from the
// orbital rows the committed moves kept (the determinant part of the full
// evaluation, without its orbitals), which removes the Sherman-Morrison
// updates' accumulated rounding.
static void launch_reinvert(qmcw_handle *h) {
  if (!h->orb_valid) return;
  KArgs A = h->A;
  A.full_orb = 0;
  A.keep_state = 1;
  A.lds_det = 0;
  A.chunk = 1;
  A.gproj_mfma = 1;     // no projection: the full evaluation that follows forms it
  size_t smem = 0;
  det_lds_inv(h, A, smem);
  prof_begin(h, QMCW_PROF_FULL_DET);
  DISPATCH_GW(h->Gk[KC_DET], launch_det, h, A, smem);
  prof_end(h);
  h->moves_since_inv = 0;
}

// one committed-or-rejected one-electron move of every walker: re-invert
// when the interval set by qmcw_set_reinversion is reached (not after a
// sweep's last move, which the full evaluation follows)
static void count_move(qmcw_handle *h, bool last) {
  if (h->reinv_every > 0 && ++h->moves_since_inv >= h->reinv_every && !last) launch_reinvert(h);
}

static void vmc_one_step(qmcw_handle *h, double delta, double fbias) {
  const int N = h->A.S.N;
  if (h->A.S.nloc > 0) {
    prof_begin(h, QMCW_PROF_ROTQUA);
    hipLaunchKernelGGL(k_rotqua, dim3((h->Wl + 63) / 64), dim3(64), 0, h->stream, h->A);
    prof_end(h);
  }
  for (int ie = 0; ie < N; ie++) {  // proposal, ratio, Metropolis test and commit in one launch
    launch_ratio(h, ie, 2, (const double *)nullptr, delta, fbias, (double *)nullptr,
               h->nacc);
    count_move(h, ie + 1 == N);
  }
  launch_eval(h, h->sample_derivs != 0);
  if (h->sample_derivs == 2) launch_accum(h);
}

static void drop_graph(qmcw_handle *h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  h->graph_ok = 0;
}

// The step's launches (kernels only; the arguments, including the walker
// state pointer, are captured by value) as a graph, rebuilt when a captured
// value changes.
static int step_graph(qmcw_handle *h, double delta, double fbias) {
  const int derivs = h->sample_derivs;
  const int gkey = h->Gk[0] | (h->Gk[1] << 8) | (h->Gk[2] << 16) | (h->Gk[3] << 24);
  if (h->graph_ok && h->graph_G == gkey && h->graph_Wl == h->Wl && h->graph_derivs == derivs &&
      h->graph_delta == delta && h->graph_fbias == fbias && h->graph_st == h->A.st)
    return QMCW_OK;
  drop_graph(h);
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  vmc_one_step(h, delta, fbias);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (e != hipSuccess) return fail(h, QMCW_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  h->graph = g;
  HIPCHK(h, hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  h->graph_ok = 1;
  h->graph_G = gkey; h->graph_Wl = h->Wl; h->graph_derivs = derivs;
  h->graph_delta = delta; h->graph_fbias = fbias; h->graph_st = h->A.st;
  return QMCW_OK;
}

static int vmc_steps(qmcw_handle *h, int nstep, double delta, double fbias, bool trace) {
  launch_eval(h, false);
  if (!trace && !h->prof && h->use_graph && nstep > 0) {
    int st = step_graph(h, delta, fbias);
    if (st) return st;
    for (int is = 0; is < nstep; is++) HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
    HIPCHK(h, hipGetLastError());
    return QMCW_OK;
  }
  for (int is = 0; is < nstep; is++) {
    vmc_one_step(h, delta, fbias);
    if (trace)
      hipLaunchKernelGGL(k_trace_eloc, dim3((h->W + 63) / 64), dim3(64), 0, h->stream, h->A,
                         h->trace, is, h->W);
  }
  HIPCHK(h, hipGetLastError());
  return QMCW_OK;
}

extern "C" int qmcw_vmc_run(qmcw_handle *h, int32_t nstep, double delta, double fbias,
                            double *eloc_trace, int64_t *naccept) {
  if (!h || nstep < 0 || delta <= 0) return QMCW_ERR_ARG;
  if (eloc_trace && nstep > h->trace_steps) {
    if (h->trace) (void)hipFree(h->trace);
    HIPCHK(h, hipMalloc(&h->trace, (size_t)nstep * h->W * sizeof(double)));
    h->trace_steps = nstep;
  }
  HIPCHK(h, hipMemsetAsync(h->nacc, 0, (size_t)h->Wp * sizeof(long long), h->stream));
  int st = vmc_steps(h, nstep, delta, fbias, eloc_trace != nullptr);
  if (st) return st;
  if (eloc_trace)
    HIPCHK(h, hipMemcpyAsync(eloc_trace, h->trace, (size_t)nstep * h->W * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
  if (naccept)
    HIPCHK(h, hipMemcpyAsync(naccept, h->nacc, (size_t)h->W * sizeof(long long), hipMemcpyDeviceToHost,
                             h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return check_flags(h);
}

extern "C" int qmcw_set_reinversion(qmcw_handle *h, int32_t every) {
  if (!h || every < 0) return QMCW_ERR_ARG;
  h->reinv_every = every;
  drop_graph(h);             // the captured sweep holds the re-inversion launches
  return QMCW_OK;
}

extern "C" int qmcw_reinvert(qmcw_handle *h) {
  if (!h) return QMCW_ERR_ARG;
  if (!h->orb_valid) return fail(h, QMCW_ERR_ARG, "no full evaluation since the walkers were set");
  launch_reinvert(h);
  HIPCHK(h, hipGetLastError());
  return QMCW_OK;
}

extern "C" int qmcw_set_vmc_mover(qmcw_handle *h, int32_t imetro, double deltar, double deltat) {
  if (!h) return QMCW_ERR_ARG;
  if (imetro != 1 && imetro != 6) return fail(h, QMCW_ERR_ARG, "imetro must be 1 or 6");
  h->A.imetro = imetro;
  h->A.deltar = deltar < 1.0 ? 2.0 : deltar;
  h->A.deltat = (deltat < 0.0 || deltat > 2.0) ? 2.0 : deltat;
  drop_graph(h);             // the captured sweep holds the kernel arguments
  return QMCW_OK;
}

extern "C" int qmcw_vmc_run_device(qmcw_handle *h, int32_t nstep, double delta, double fbias) {
  if (!h || nstep < 0 || delta <= 0) return QMCW_ERR_ARG;
  return vmc_steps(h, nstep, delta, fbias, false);
}

// ---------------------------------------------------------------------------
// Batches: the VMC step of several handles (systems of uneven electron and
// basis counts) in one launch per kernel and lane-group width
// (engine_batch.h).  Argument sets: 0 the first full evaluation (each
// handle's kept rows as it left them), 1 the steps.  Kernels per set: move,
// jas, det (also gproj, fin), nlpp, rotqua; block offsets per kernel (move,
// jas, det, nlpp, rotqua, gproj) and width (8, 16, 32, 64: the systems of
// each width in their own launch; rotqua and gproj use the first).
enum { KB_MOVE, KB_JAS, KB_DET, KB_NLPP, KB_ROT, KB_NARGS, KB_GPROJ = KB_NARGS, KB_NOFF };
constexpr int kBatchWidths[4] = {8, 16, 32, 64};
static int width_index(int G) { return G == 8 ? 0 : (G == 16 ? 1 : (G == 32 ? 2 : 3)); }

struct qmcw_batch {
  std::vector<qmcw_handle *> hs;
  hipStream_t stream = nullptr;              // the first handle's stream
  hipEvent_t done = nullptr;                 // recorded after the batch's work
  int n = 0, maxN = 0, precision = 0, nordc = 0;
  bool any_nloc = false, any_gproj = false;
  bool sweep = false;                        // QMCW_BATCH_SWEEP=1: each system's sweep in one launch
  KArgs *dA[2] = {nullptr, nullptr};         // [KB_NARGS][n] per set
  int *dboff = nullptr;                      // [KB_NOFF][4][n + 1]
  int *dnw = nullptr;                        // [n]
  long long **dnacc = nullptr;               // [n]
  int nblk[KB_NOFF][4] = {};
  size_t smem_move = 0, smem_det = 0, smem_nlpp = 0;
  std::string err;
};

static int batch_fail(qmcw_batch *b, int code, const std::string &msg) {
  if (b) b->err = msg;
  g_lib_err = msg;
  return code;
}

#define BHIPCHK(b, call)                                                             \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return batch_fail(b, QMCW_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// the argument set `set` of every handle (keep: each handle's kept rows,
// else the rows the moves keep during the steps) and, for set 1, the block
// offsets, walker counts and counters; synchronous copies
static int batch_upload(qmcw_batch *b, int set, bool first) {
  const int n = b->n;
  std::vector<KArgs> A((size_t)KB_NARGS * n);
  std::vector<int> boff((size_t)KB_NOFF * 4 * (n + 1), 0), nw(n);
  std::vector<long long *> nacc(n);
  for (int s = 0; s < n; s++) {
    qmcw_handle *h = b->hs[s];
    size_t phi_b = 0, sm = 0;
    A[(size_t)KB_MOVE * n + s] = move_args(h, h->Gk[KC_MOVE], false, phi_b, sm);
    b->smem_move = std::max(b->smem_move, sm);
    EvalArgs e;
    const bool keep = first ? (h->orb_valid != 0) : true;
    eval_args(h, false, keep && !h->no_keep_orb, e);
    A[(size_t)KB_JAS * n + s] = e.jas;
    A[(size_t)KB_DET * n + s] = e.det;
    A[(size_t)KB_NLPP * n + s] = e.nlpp;
    A[(size_t)KB_ROT * n + s] = h->A;
    b->smem_det = std::max(b->smem_det, e.det_smem);
    b->smem_nlpp = std::max(b->smem_nlpp, e.nlpp_smem);
    const int Wl = h->Wl;
    const int gx = (int)(((int64_t)Wl * std::max(h->A.S.ns[0], h->A.S.ns[1]) + 63) / 64);
    const int gy = (h->A.S.NB + 63) / 64;
    const int cnt[KB_NOFF] = {Wl / (kWave / h->Gk[KC_MOVE]), Wl / (kWave / h->Gk[KC_JAS]),
                              Wl / (kWave / h->Gk[KC_DET]), Wl / (kWave / h->Gk[KC_NLPP]),
                              (Wl + kWave - 1) / kWave, e.det.gproj_mfma ? gx * gy * 2 : 0};
    const int wi[KB_NOFF] = {width_index(h->Gk[KC_MOVE]), width_index(h->Gk[KC_JAS]),
                             width_index(h->Gk[KC_DET]), width_index(h->Gk[KC_NLPP]), 0, 0};
    for (int k = 0; k < KB_NOFF; k++)
      for (int g = 0; g < 4; g++) {
        int *o = boff.data() + ((size_t)k * 4 + g) * (n + 1);
        o[s + 1] = o[s] + (g == wi[k] ? cnt[k] : 0);
      }
    nw[s] = Wl;
    nacc[s] = h->nacc;
  }
  BHIPCHK(b, hipMemcpy(b->dA[set], A.data(), A.size() * sizeof(KArgs), hipMemcpyHostToDevice));
  if (set == 1) {
    for (int k = 0; k < KB_NOFF; k++)
      for (int g = 0; g < 4; g++) b->nblk[k][g] = boff[((size_t)k * 4 + g) * (n + 1) + n];
    BHIPCHK(b, hipMemcpy(b->dboff, boff.data(), boff.size() * sizeof(int), hipMemcpyHostToDevice));
    BHIPCHK(b, hipMemcpy(b->dnw, nw.data(), nw.size() * sizeof(int), hipMemcpyHostToDevice));
    BHIPCHK(b, hipMemcpy(b->dnacc, nacc.data(), nacc.size() * sizeof(long long *), hipMemcpyHostToDevice));
  }
  return QMCW_OK;
}

static BatchArgs batch_args(const qmcw_batch *b, int set, int kargs, int koff, int g = 0) {
  BatchArgs B;
  B.A = b->dA[set] + (size_t)kargs * b->n;
  B.boff = b->dboff + ((size_t)koff * 4 + g) * (b->n + 1);
  B.nw = b->dnw;
  B.nacc = b->dnacc;
  B.n = b->n;
  return B;
}

template <int G>
static void batch_det_fin(const qmcw_batch *b, const BatchArgs &B, bool fin) {
  const dim3 grid(b->nblk[KB_DET][width_index(G)]), block(kWave);
  if (fin)
    hipLaunchKernelGGL((k_full_fin_batch<G>), grid, block, 0, b->stream, B);
  else if (b->precision == QMCW_FP64)
    hipLaunchKernelGGL((k_full_det_batch<double, G>), grid, block, b->smem_det, b->stream, B);
  else
    hipLaunchKernelGGL((k_full_det_batch<float, G>), grid, block, b->smem_det, b->stream, B);
}

// the full evaluation of every handle with argument set `set`
static void batch_eval(const qmcw_batch *b, int set) {
  const OrdKernels &K = ord_kernels(b->nordc);
  BatchCtx c;
  c.stream = b->stream;
  c.precision = b->precision;
  c.smem = 0;
  for (int g = 0; g < 4; g++)
    if (b->nblk[KB_JAS][g] > 0) {
      c.B = batch_args(b, set, KB_JAS, KB_JAS, g);
      c.G = kBatchWidths[g];
      c.nblocks = b->nblk[KB_JAS][g];
      K.full_jas_batch(c);
    }
  for (int g = 0; g < 4; g++)
    if (b->nblk[KB_DET][g] > 0) {
      const BatchArgs Bd = batch_args(b, set, KB_DET, KB_DET, g);
      DISPATCH_GW(kBatchWidths[g], batch_det_fin, b, Bd, false);
    }
  if (b->any_gproj && b->nblk[KB_GPROJ][0] > 0)
    hipLaunchKernelGGL(k_gproj_mfma_batch, dim3(b->nblk[KB_GPROJ][0]), dim3(kWave), 0, b->stream,
                       batch_args(b, set, KB_DET, KB_GPROJ));
  if (b->any_nloc) {
    c.smem = b->smem_nlpp;
    for (int g = 0; g < 4; g++)
      if (b->nblk[KB_NLPP][g] > 0) {
        c.B = batch_args(b, set, KB_NLPP, KB_NLPP, g);
        c.G = kBatchWidths[g];
        c.nblocks = b->nblk[KB_NLPP][g];
        K.full_nlpp_batch(c);
      }
  }
  for (int g = 0; g < 4; g++)
    if (b->nblk[KB_DET][g] > 0) {
      const BatchArgs Bd = batch_args(b, set, KB_DET, KB_DET, g);
      DISPATCH_GW(kBatchWidths[g], batch_det_fin, b, Bd, true);
    }
}

static void batch_step(const qmcw_batch *b, double delta, double fbias) {
  const OrdKernels &K = ord_kernels(b->nordc);
  if (b->any_nloc)
    hipLaunchKernelGGL(k_rotqua_batch, dim3(b->nblk[KB_ROT][0]), dim3(kWave), 0, b->stream,
                       batch_args(b, 1, KB_ROT, KB_ROT));
  BatchCtx c;
  c.stream = b->stream;
  c.precision = b->precision;
  c.smem = b->smem_move;
  if (b->sweep) {
    for (int g = 0; g < 4; g++)
      if (b->nblk[KB_MOVE][g] > 0) {
        c.B = batch_args(b, 1, KB_MOVE, KB_MOVE, g);
        c.G = kBatchWidths[g];
        c.nblocks = b->nblk[KB_MOVE][g];
        K.sweep_batch(c, delta, fbias);
      }
  } else {
    for (int ie = 0; ie < b->maxN; ie++)
      for (int g = 0; g < 4; g++)
        if (b->nblk[KB_MOVE][g] > 0) {
          c.B = batch_args(b, 1, KB_MOVE, KB_MOVE, g);
          c.G = kBatchWidths[g];
          c.nblocks = b->nblk[KB_MOVE][g];
          K.ratio_batch(c, ie, delta, fbias);
        }
  }
  batch_eval(b, 1);
}

extern "C" int qmcw_batch_destroy(qmcw_batch *b) {
  if (!b) return QMCW_ERR_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (int set = 0; set < 2; set++)
    if (b->dA[set]) (void)hipFree(b->dA[set]);
  if (b->dboff) (void)hipFree(b->dboff);
  if (b->dnw) (void)hipFree(b->dnw);
  if (b->dnacc) (void)hipFree(b->dnacc);
  if (b->done) (void)hipEventDestroy(b->done);
  delete b;
  return QMCW_OK;
}

extern "C" int qmcw_batch_create(qmcw_handle *const *hs, int32_t n, qmcw_batch **out) {
  if (!hs || n <= 0 || !out) return fail(nullptr, QMCW_ERR_ARG, "invalid arguments");
  *out = nullptr;
  for (int s = 0; s < n; s++) {
    const qmcw_handle *h = hs[s];
    if (!h) return fail(nullptr, QMCW_ERR_ARG, "null handle in the batch");
    if (h->precision != hs[0]->precision)
      return fail(nullptr, QMCW_ERR_ARG, "the handles of a batch must share the precision");
    if (&ord_kernels(h->A.S.nordc) != &ord_kernels(hs[0]->A.S.nordc))
      return fail(nullptr, QMCW_ERR_ARG, "the handles of a batch must share the e-e-n order (nordc)");
    if (h->sample_derivs || h->reinv_every > 0)
      return fail(nullptr, QMCW_ERR_ARG,
                  "a batch runs plain VMC steps: no derivative sampling or re-inversion interval");
    for (int t = 0; t < s; t++)
      if (hs[t] == h) return fail(nullptr, QMCW_ERR_ARG, "a handle appears twice in the batch");
  }
  qmcw_batch *b = new qmcw_batch;
  b->n = n;
  b->hs.assign(hs, hs + n);
  b->stream = hs[0]->stream;
  b->precision = hs[0]->precision;
  b->nordc = hs[0]->A.S.nordc;
  if (const char *e = getenv("QMCW_BATCH_SWEEP")) b->sweep = atoi(e) != 0;
  // each handle keeps its lane-group widths: the systems of each width run
  // in a launch of their own
  for (qmcw_handle *h : b->hs) {
    b->maxN = std::max(b->maxN, h->A.S.N);
    b->any_nloc = b->any_nloc || h->A.S.nloc > 0;
    b->any_gproj = b->any_gproj || (h->A.S.nloc > 0 && !h->no_gproj_mfma);
  }
  const size_t na = (size_t)KB_NARGS * n * sizeof(KArgs);
  if (hipMalloc(&b->dA[0], na) != hipSuccess || hipMalloc(&b->dA[1], na) != hipSuccess ||
      hipMalloc(&b->dboff, (size_t)KB_NOFF * 4 * (n + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc(&b->dnw, (size_t)n * sizeof(int)) != hipSuccess ||
      hipMalloc(&b->dnacc, (size_t)n * sizeof(long long *)) != hipSuccess ||
      hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
    qmcw_batch_destroy(b);
    return fail(nullptr, QMCW_ERR_NOMEM, "batch allocation");
  }
  *out = b;
  return QMCW_OK;
}

extern "C" int qmcw_batch_vmc_run(qmcw_batch *b, int32_t nstep, double delta, double fbias) {
  if (!b || nstep < 0 || delta <= 0) return QMCW_ERR_ARG;
  // the arguments are copied synchronously: wait for the handles' own work
  for (qmcw_handle *h : b->hs) BHIPCHK(b, hipStreamSynchronize(h->stream));
  int st = batch_upload(b, 0, true);
  if (st) return st;
  st = batch_upload(b, 1, false);
  if (st) return st;
  // the first full evaluation (as qmcw_vmc_run starts), then the steps
  batch_eval(b, 0);
  for (qmcw_handle *h : b->hs) {
    h->orb_valid = 1;
    h->moves_since_inv = 0;
  }
  for (int is = 0; is < nstep; is++) batch_step(b, delta, fbias);
  BHIPCHK(b, hipGetLastError());
  // the handles' later work follows the batch's
  BHIPCHK(b, hipEventRecord(b->done, b->stream));
  for (qmcw_handle *h : b->hs)
    if (h->stream != b->stream) BHIPCHK(b, hipStreamWaitEvent(h->stream, b->done, 0));
  return QMCW_OK;
}

extern "C" const char *qmcw_batch_last_error(const qmcw_batch *b) {
  return b ? b->err.c_str() : g_lib_err.c_str();
}

extern "C" int qmcw_profile(qmcw_handle *h, int32_t enable) {
  if (!h) return QMCW_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->prof = enable != 0;
  h->ev_used = 0;
  return QMCW_OK;
}

extern "C" int qmcw_profile_get(qmcw_handle *h, double *ms, int64_t *count) {
  if (!h || !ms || !count) return QMCW_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int p = 0; p < QMCW_PROF_NKERNEL; p++) { ms[p] = 0.0; count[p] = 0; }
  for (size_t i = 0; i + 1 < h->ev_used; i += 2) {
    int ph = h->ev_phase[i];
    if (ph < 0 || ph >= QMCW_PROF_NKERNEL) continue;
    float t = 0.0f;
    HIPCHK(h, hipEventElapsedTime(&t, h->ev[i], h->ev[i + 1]));
    ms[ph] += t;
    count[ph] += 1;
  }
  return QMCW_OK;
}

// Group-width autotuner: for each admissible width, times every kernel of the
// full evaluation and one ratio launch per electron (mode 0 at the current
// positions: neither changes the walkers or their random streams) with HIP
// events, and keeps the fastest width of each kernel class.
extern "C" int qmcw_autotune(qmcw_handle *h, int32_t *chosen) {
  if (!h) return QMCW_ERR_ARG;
  const int N = h->A.S.N;
  const int cand[4] = {8, 16, 32, 64};
  // kernel classes of the profiled kernels
  const int cls[QMCW_PROF_NKERNEL] = {KC_JAS, KC_DET, KC_NLPP, KC_DET, KC_MOVE, KC_MOVE, -1,
                                      KC_JAS, KC_NLPP, KC_DET, -1};
  double best[4] = {1e30, 1e30, 1e30, 1e30};
  int bestG[4] = {h->Gk[0], h->Gk[1], h->Gk[2], h->Gk[3]};
  int best_split = h->split_move;
  double *xin = h->io + (size_t)h->Wp * 8;
  const int prof_save = h->prof;
  drop_graph(h);
  // the one-electron move is also timed as three launches (k_move_a/b/c)
  // unless QMCW_SPLIT_MOVE fixed the choice; the other classes once per width
  const int sm0 = h->split_forced ? h->split_move : 0, sm1 = h->split_forced ? h->split_move : 1;
  for (int sm = sm0; sm <= sm1; sm++)
  for (int c = 0; c < 4; c++) {
    const int G = cand[c];
    h->split_move = sm;
    if (G > 16 && N <= 8) continue;      // lanes would mostly idle
    if (G < 16 && N > 48) continue;
    for (int k = 0; k < 4; k++) h->Gk[k] = G;
    double t[4] = {0, 0, 0, 0};
    for (int rep = 0; rep < 2; rep++) {
      HIPCHK(h, hipStreamSynchronize(h->stream));
      h->prof = rep == 1;          // the first repetition warms up
      h->ev_used = 0;
      launch_eval(h, false);
      for (int ie = 0; ie < N; ie++) {
        hipLaunchKernelGGL(k_copy_strided, dim3((h->Wl + 63) / 64), dim3(64), 0, h->stream, xin,
                           (int64_t)3, (const double *)(h->A.st + h->A.L.x + 3 * ie),
                           (int64_t)h->A.L.stride, 3, h->Wl);
        launch_ratio(h, ie, 0, (const double *)xin, 1.0, 1.0, (double *)nullptr);
      }
      HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    for (size_t i = 0; i + 1 < h->ev_used; i += 2) {
      const int ph = h->ev_phase[i];
      if (ph < 0 || ph >= QMCW_PROF_NKERNEL || cls[ph] < 0) continue;
      float ms = 0.f;
      HIPCHK(h, hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]));
      t[cls[ph]] += ms;
    }
    for (int k = 0; k < 4; k++) {
      if (k != KC_MOVE && sm != sm0) continue;
      if (t[k] > 0 && t[k] < best[k]) {
        best[k] = t[k];
        bestG[k] = G;
        if (k == KC_MOVE) best_split = sm;
      }
    }
  }
  h->prof = prof_save;
  h->ev_used = 0;
  h->split_move = best_split;
  for (int k = 0; k < 4; k++) h->Gk[k] = bestG[k];
  h->G = bestG[KC_MOVE];
  // restore a consistent resident state with the chosen widths
  launch_eval(h, false);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (chosen) *chosen = h->G;
  return QMCW_OK;
}

extern "C" int qmcw_get_group_widths(const qmcw_handle *h, int32_t *widths) {
  if (!h || !widths) return QMCW_ERR_ARG;
  for (int k = 0; k < 4; k++) widths[k] = h->Gk[k];
  return QMCW_OK;
}

extern "C" int qmcw_set_group_width(qmcw_handle *h, int32_t G) {
  if (!h) return QMCW_ERR_ARG;
  if (G != 8 && G != 16 && G != 32 && G != 64)
    return fail(h, QMCW_ERR_ARG, "group width must be 8, 16, 32 or 64");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->G = G;
  for (int k = 0; k < 4; k++) h->Gk[k] = G;
  return QMCW_OK;
}

extern "C" int qmcw_get_eloc(const qmcw_handle *hc, double *eloc) {
  qmcw_handle *h = const_cast<qmcw_handle *>(hc);
  if (!h || !eloc) return QMCW_ERR_ARG;
  hipLaunchKernelGGL(k_copy_strided, dim3((h->W + 63) / 64), dim3(64), 0, h->stream, h->io, (int64_t)1,
                     (const double *)(h->A.st + h->A.L.eloc), (int64_t)h->A.L.stride, 1, h->W);
  HIPCHK(h, hipMemcpyAsync(eloc, h->io, (size_t)h->W * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return QMCW_OK;
}

static int copy_derivs(qmcw_handle *h, double *dpsi, double *deloc) {
  const int P = h->nparm;
  std::vector<double> o((size_t)h->W * 2 * P);
  HIPCHK(h, hipMemcpyAsync(o.data(), h->der, o.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int w = 0; w < h->W; w++) {
    const double *r = o.data() + (size_t)w * 2 * P;
    if (dpsi) memcpy(dpsi + (size_t)w * P, r, P * sizeof(double));
    if (deloc) memcpy(deloc + (size_t)w * P, r + P, P * sizeof(double));
  }
  return QMCW_OK;
}

extern "C" int qmcw_derivs(qmcw_handle *h, double *dpsi, double *deloc) {
  if (!h) return QMCW_ERR_ARG;
  if (h->nparm == 0) return fail(h, QMCW_ERR_ARG, "the system has no parameters to differentiate");
  launch_eval(h, true);
  HIPCHK(h, hipGetLastError());
  int st = copy_derivs(h, dpsi, deloc);
  if (st) return st;
  return check_flags(h);
}

extern "C" int qmcw_sample_derivs(qmcw_handle *h, int32_t enable) {
  if (!h) return QMCW_ERR_ARG;
  if (enable && h->nparm == 0) return fail(h, QMCW_ERR_ARG, "the system has no parameters to differentiate");
  if (enable < 0 || enable > 2) return fail(h, QMCW_ERR_ARG, "sample_derivs takes 0, 1 or 2");
  if (enable == 2) {
    int st = accum_alloc(h);
    if (st) return st;
  }
  h->sample_derivs = enable;
  return QMCW_OK;
}

extern "C" int qmcw_get_derivs(qmcw_handle *h, double *dpsi, double *deloc) {
  if (!h) return QMCW_ERR_ARG;
  if (!h->der_valid) return fail(h, QMCW_ERR_ARG, "no derivative evaluation has run on this handle");
  return copy_derivs(h, dpsi, deloc);
}

extern "C" int qmcw_accum_reset(qmcw_handle *h) {
  if (!h) return QMCW_ERR_ARG;
  if (!h->acc) return accum_alloc(h);
  const int64_t n = accum_doubles(h->nparm);
  hipLaunchKernelGGL(k_accum_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->acc, n);
  return QMCW_OK;
}

extern "C" int qmcw_accum_add(qmcw_handle *h) {
  if (!h) return QMCW_ERR_ARG;
  if (!h->der_valid) return fail(h, QMCW_ERR_ARG, "no derivative evaluation has run on this handle");
  int st = accum_alloc(h);
  if (st) return st;
  launch_accum(h);
  return QMCW_OK;
}

extern "C" int qmcw_accum_get(qmcw_handle *h, double *out) {
  if (!h || !out) return QMCW_ERR_ARG;
  int st = accum_alloc(h);
  if (st) return st;
  HIPCHK(h, hipMemcpyAsync(out, h->acc, accum_doubles(h->nparm) * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return QMCW_OK;
}

// ---------------------------------------------------------------------------
// Product sums with no allocation: the chunk partial sums live in a
// file-scope device array (one copy per GPU, held until the process ends).
// Its host-side name is not a device address: kernels get the address
// hipGetSymbolAddress returns for the current device, looked up once per
// device.  The partial sums are overwritten in full by every call, so
// nothing needs clearing; calls that overlap on one GPU share the array and
// race.

static __device__ __align__(256) double scratch[QMCW_ACCUM_STATIC_DOUBLES];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64];

static int scratch_for_device(qmcw_handle *h, double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return fail(h, QMCW_ERR_HIP, "hipGetDevice failed");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(h, QMCW_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return QMCW_OK;
}

extern "C" int qmcw_accum_add_static(qmcw_handle *h, double *acc) {
  if (!h || !acc) return QMCW_ERR_ARG;
  const int P = h->nparm;
  if (P <= 0) return fail(h, QMCW_ERR_ARG, "the system has no parameters to differentiate");
  const int nch = (h->W + kAccChunk - 1) / kAccChunk;
  const int64_t need = (int64_t)nch * 3 * P * P + (int64_t)nch * (3 + 3 * (int64_t)P);
  if (need > (int64_t)QMCW_ACCUM_STATIC_DOUBLES)
    return fail(h, QMCW_ERR_ARG, "qmcw_accum_add_static: " + std::to_string(h->W) + " walkers x " +
                std::to_string(P) + " parameters need " + std::to_string(need) +
                " doubles of partial sums, more than the static array's " +
                std::to_string((long long)QMCW_ACCUM_STATIC_DOUBLES));
  if (!h->der_valid) return fail(h, QMCW_ERR_ARG, "no derivative evaluation has run on this handle");
  double *part = nullptr;
  int st = scratch_for_device(h, &part);
  if (st) return st;
  double *partv = part + (int64_t)nch * 3 * P * P;
  const int nb = (P + kAccBlk - 1) / kAccBlk;
  const int64_t n = accum_doubles(P);
  hipLaunchKernelGGL(k_accum_gemm, dim3(nb * nb, nch), dim3(64), 0, h->stream, (const double *)h->der,
                     2 * P, P, (const double *)h->A.st, (int64_t)h->A.L.stride, (int64_t)h->A.L.eloc,
                     h->W, part);
  hipLaunchKernelGGL(k_accum_vec, dim3((P + 1 + 63) / 64, nch), dim3(64), 0, h->stream,
                     (const double *)h->der, 2 * P, P, (const double *)h->A.st,
                     (int64_t)h->A.L.stride, (int64_t)h->A.L.eloc, h->W, partv);
  hipLaunchKernelGGL(k_accum_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                     (const double *)part, (const double *)partv, nch, P, acc);
  HIPCHK(h, hipGetLastError());
  return QMCW_OK;
}

// ---------------------------------------------------------------------------
// DMC

static int dmc_alloc(qmcw_handle *h) {
  if (h->dm) return QMCW_OK;
  void *p;
  size_t Wp = h->Wp;
  HIPCHK(h, hipMalloc(&p, Wp * kDmStride * sizeof(double))); h->allocs.push_back(p); h->dm = (double *)p;
  HIPCHK(h, hipMalloc(&p, Wp * kDmStride * sizeof(double))); h->allocs.push_back(p); h->dm2 = (double *)p;
  HIPCHK(h, hipMalloc(&p, Wp * h->A.L.stride * sizeof(double))); h->allocs.push_back(p); h->st2 = (double *)p;
  HIPCHK(h, hipMalloc(&p, 8 * sizeof(double))); h->allocs.push_back(p); h->dstat = (double *)p;
  HIPCHK(h, hipMalloc(&p, Wp * sizeof(double))); h->allocs.push_back(p); h->dwt = (double *)p;
  HIPCHK(h, hipMalloc(&p, Wp * sizeof(int))); h->allocs.push_back(p); h->dorigin = (int *)p;
  // device split_join: sort buffer (next power of two), hole stack, results, branching stream
  h->npow2 = 1;
  while (h->npow2 < h->W) h->npow2 <<= 1;
  HIPCHK(h, hipMalloc(&p, (size_t)h->npow2 * sizeof(double))); h->allocs.push_back(p); h->dsort = (double *)p;
  HIPCHK(h, hipMalloc(&p, Wp * sizeof(int))); h->allocs.push_back(p); h->dunder = (int *)p;
  HIPCHK(h, hipMalloc(&p, 8 * sizeof(int))); h->allocs.push_back(p); h->dsj = (int *)p;
  if (h->A.S.nloc > 0) {
    const int64_t ts = (int64_t)4 * h->A.S.N * h->A.S.tm_kmax * h->A.S.nquad;
    HIPCHK(h, hipMalloc(&p, Wp * ts * sizeof(double))); h->allocs.push_back(p); h->A.tq = (double *)p;
    HIPCHK(h, hipMalloc(&p, Wp * ts * sizeof(double))); h->allocs.push_back(p); h->tq2 = (double *)p;
    h->A.tq_stride = ts;
  }
  return QMCW_OK;
}

// point columns of the T-move steps (allocated at the first generation with T-moves)
static int tm_alloc(qmcw_handle *h) {
  if (h->tmq) return QMCW_OK;
  void *p;
  const size_t Wp = h->Wp, ncol = (size_t)h->A.S.tm_kmax * h->A.S.nquad, nb = h->A.S.nquad;
  auto get = [&](size_t n, double *&dst) -> int {
    HIPCHK(h, hipMalloc(&p, n * sizeof(double)));
    h->allocs.push_back(p);
    dst = (double *)p;
    return QMCW_OK;
  };
  int st = get(ncol * Wp * 3, h->tmq);
  if (!st) st = get(ncol * Wp * 3, h->tmx);
  if (!st) st = get(ncol * Wp * 8, h->rcol);
  if (!st) st = get(nb * Wp * 3, h->tmb);
  if (!st) st = get(nb * Wp * 3, h->tmbq);
  if (!st) st = get(nb * Wp * 8, h->bcol);
  if (!st) st = get(Wp * 3, h->tmxin);
  if (!st) st = get(Wp, h->tmstale);
  if (!st) {
    HIPCHK(h, hipMalloc(&p, 3 * Wp * sizeof(int)));
    h->allocs.push_back(p);
    h->tmint = (int *)p;
  }
  return st;
}

extern "C" int qmcw_set_branch_rng(qmcw_handle *h, const int32_t *seed) {
  if (!h || !seed) return QMCW_ERR_ARG;
  for (int k = 0; k < 4; k++) h->brng[k] = seed[k] & 4095;
  h->brng[3] |= 1;     // setrn makes the last limb odd (qmc/vmc/rannyu.f90:11-21)
  return QMCW_OK;
}

extern "C" int qmcw_get_branch_rng(const qmcw_handle *h, int32_t *seed) {
  if (!h || !seed) return QMCW_ERR_ARG;
  for (int k = 0; k < 4; k++) seed[k] = h->brng[k];
  return QMCW_OK;
}

extern "C" int qmcw_dmc_begin(qmcw_handle *h, int32_t nwalk, const double *weight) {
  if (!h || nwalk <= 0 || nwalk > h->W) return fail(h, QMCW_ERR_ARG, "nwalk must be in 1..W");
  int st = dmc_alloc(h);
  if (st) return st;
  h->nact = nwalk;
  h->Wl = (nwalk + 7) / 8 * 8;
  h->A.Wp = h->Wl;
  std::vector<double> wt(h->Wp, 1.0);
  if (weight) std::copy(weight, weight + nwalk, wt.begin());
  HIPCHK(h, hipMemcpyAsync(h->dwt, wt.data(), h->Wp * sizeof(double), hipMemcpyHostToDevice, h->stream));
  // DMC-mode evaluation (per-pair grid rotations) of every walker
  h->A.dmc_rot = h->A.S.nloc > 0 ? 1 : 0;
  h->A.dmove = nullptr;
  launch_eval(h, false);
  h->A.dmc_rot = 0;
  hipLaunchKernelGGL(k_dmc_init, dim3((h->Wl + 63) / 64), dim3(64), 0, h->stream, h->A, h->dm,
                     (const double *)h->dwt, nwalk);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->dmc_on = 1;
  st = check_flags(h);
  return st;
}

// split_join (qmc/dmc/dmc_elec/split_join_mod.f90:3-219) on the host: the
// decisions are sequential in the walker index and draw from the branching
// stream.  Produces the new population as copies of old records: origin[d],
// the new weights and the streams (a split copy continues on four 12-bit
// limbs drawn from the branching stream, the last made odd).
// The loops follow the reference's order (joins in walker order with one
// rannyu per joined pair, the split threshold from the sorted weights,
// splits into the holes last-first, then compaction); a split copy gets a
// fresh stream of four 12-bit limbs drawn from the branching stream.
extern "C" int qmcw_split_join(int32_t n, int32_t mwalk, double *w, int32_t *origin, int32_t *rng,
                               int32_t *brng, int32_t *nwalk_out, int32_t *nbranch) {
  if (n < 0 || mwalk < n || !w || !origin || !rng || !brng || !nwalk_out || !nbranch)
    return fail(nullptr, QMCW_ERR_ARG, "qmcw_split_join: bad arguments");
  const int MW = mwalk;
  for (int i = 0; i < MW; i++) origin[i] = i;
  std::vector<int> under;
  int ipair = 0, iw2 = 0, nbr = 0;
  for (int iw = 0; iw < n; iw++) {
    if (!(w[iw] < 0.5)) continue;
    if (w[iw] == 0.0) {
      nbr++;
      under.push_back(iw);
    } else if (ipair == 0) {
      ipair = 1;
      iw2 = iw;
    } else {
      nbr++;
      ipair = 0;
      const double wt = w[iw] + w[iw2];
      if (rannyu(brng) > w[iw] / wt) {
        w[iw2] = wt;
        under.push_back(iw);
      } else {
        w[iw] = wt;
        under.push_back(iw2);
      }
    }
  }
  const int nafter = n - (int)under.size(), nmax_dupl = MW - nafter;
  std::vector<double> srt(w, w + n);
  std::sort(srt.begin(), srt.end());
  auto at1 = [&](int i) { return n > 0 ? srt[std::max(1, i) - 1] : 0.0; };   // 1-based, clamped below
  double wt_split = std::max(2.0, at1(n - nmax_dupl));
  int nmore = 0;
  for (int iw = std::max(1, n - nmax_dupl); iw <= n; iw++) {
    const double r = srt[iw - 1] / wt_split;
    if (r >= 8.0) nmore += 3;
    else if (r >= 4.0) nmore += 2;
    else if (r >= 2.0) nmore += 1;
  }
  if (n - nmax_dupl + nmore < n) wt_split = std::max(2.0, at1(n - nmax_dupl + nmore));
  else wt_split = 1e99;
  int n2 = n;
  for (int iw = 0; iw < n; iw++) {
    if (!(w[iw] > wt_split)) continue;
    nbr++;
    int dst;
    if (!under.empty()) {
      dst = under.back();
      under.pop_back();
    } else {
      if (n2 >= MW) return fail(nullptr, QMCW_ERR_NOMEM, "DMC population exceeds the walker capacity");
      dst = n2++;
    }
    w[iw] *= 0.5;
    w[dst] = w[iw];
    origin[dst] = origin[iw];
    for (int k = 0; k < 4; k++) rng[4 * dst + k] = (int32_t)(rannyu(brng) * 4096.0) & 4095;
    rng[4 * dst + 3] |= 1;
  }
  while (!under.empty()) {
    const int dst = under.back(), src = n2 - 1;
    under.pop_back();
    n2--;
    if (dst != src) {
      w[dst] = w[src];
      origin[dst] = origin[src];
      for (int k = 0; k < 4; k++) rng[4 * dst + k] = rng[4 * src + k];
    }
  }
  *nwalk_out = n2;
  *nbranch = nbr;
  return QMCW_OK;
}

// the device kernel on host arrays (test access; allocates and synchronises)
extern "C" int qmcw_split_join_gpu(int32_t n, int32_t mwalk, double *w, int32_t *origin, int32_t *rng,
                                   int32_t *brng, int32_t *nwalk_out, int32_t *nbranch) {
  if (n < 0 || mwalk < n || !w || !origin || !rng || !brng || !nwalk_out || !nbranch)
    return fail(nullptr, QMCW_ERR_ARG, "qmcw_split_join_gpu: bad arguments");
  const int mp = (mwalk + 7) / 8 * 8;
  int npow2 = 1;
  while (npow2 < mwalk) npow2 <<= 1;
  double *dw = nullptr, *ds = nullptr;
  int *dr = nullptr, *dor = nullptr, *du = nullptr, *dsj = nullptr;
  int st = QMCW_OK;
  std::vector<double> wh(mp, 0.0);
  std::vector<int32_t> rh((size_t)mp * 4, 0);
  std::copy(w, w + n, wh.begin());
  std::copy(rng, rng + (size_t)4 * n, rh.begin());
  int res[8];
  if (hipMalloc((void **)&dw, mp * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&ds, npow2 * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&dr, (size_t)mp * 4 * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&dor, mp * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&du, mp * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&dsj, 8 * sizeof(int)) != hipSuccess) {
    st = fail(nullptr, QMCW_ERR_NOMEM, "qmcw_split_join_gpu: hipMalloc");
  } else if (hipMemcpy(dw, wh.data(), mp * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(dr, rh.data(), rh.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(dsj + 4, brng, 4 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    st = fail(nullptr, QMCW_ERR_HIP, "qmcw_split_join_gpu: hipMemcpy");
  } else {
    hipLaunchKernelGGL(k_split_join, dim3(1), dim3(1024), 0, 0, dw, n, mwalk, dsj + 4, dr, dor, ds, npow2,
                       du, dsj);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(res, dsj, 8 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(wh.data(), dw, mp * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(rh.data(), dr, rh.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      st = fail(nullptr, QMCW_ERR_HIP, "qmcw_split_join_gpu: kernel or copy");
    } else if (res[2]) {
      st = fail(nullptr, QMCW_ERR_NOMEM, "DMC population exceeds the walker capacity");
    } else {
      std::vector<int> oh(mp);
      (void)hipMemcpy(oh.data(), dor, mp * sizeof(int), hipMemcpyDeviceToHost);
      std::copy(wh.begin(), wh.begin() + mwalk, w);
      std::copy(oh.begin(), oh.begin() + mwalk, origin);
      std::copy(rh.begin(), rh.begin() + (size_t)4 * mwalk, rng);
      for (int k = 0; k < 4; k++) brng[k] = res[4 + k];
      *nwalk_out = res[0];
      *nbranch = res[1];
    }
  }
  (void)hipFree(dw); (void)hipFree(ds); (void)hipFree(dr); (void)hipFree(dor); (void)hipFree(du);
  (void)hipFree(dsj);
  return st;
}

static int dmc_split_join(qmcw_handle *h, std::vector<double> &w, std::vector<int32_t> &rng,
                          std::vector<int> &origin, int &n_out, int &nbranch) {
  origin.resize(h->W);
  int32_t no = 0, nb = 0;
  const int st = qmcw_split_join(h->nact, h->W, w.data(), origin.data(), rng.data(), h->brng, &no, &nb);
  if (st) return fail(h, st, qmcw_last_error(nullptr));
  n_out = no;
  nbranch = nb;
  return QMCW_OK;
}

extern "C" int qmcw_dmc_step(qmcw_handle *h, const qmcw_dmc_params *p, qmcw_dmc_stats *stats) {
  if (!h || !p) return QMCW_ERR_ARG;
  if (!h->dmc_on) return fail(h, QMCW_ERR_ARG, "qmcw_dmc_begin has not been called");
  if (!(p->tau > 0.0) || p->n_drift < 1) return fail(h, QMCW_ERR_ARG, "tau > 0 and n_drift >= 1 required");
  if (p->tmoves < 0 || p->tmoves > 2) return fail(h, QMCW_ERR_ARG, "tmoves must be 0, 1 or 2");
  const int N = h->A.S.N;
  const dim3 g1((h->Wl + 63) / 64), b1(64);
  double *gbuf = h->io, *xin = h->io + (size_t)h->Wp * 8;
  if (h->A.S.nloc > 0) hipLaunchKernelGGL(k_rotqua, g1, b1, 0, h->stream, h->A);
  hipLaunchKernelGGL(k_dmc_gen_begin, g1, b1, 0, h->stream, h->A, h->dm);
  if (p->tmoves && h->A.S.nloc > 0) {
    // T-moves of every electron in turn (rundmc_mod.f90:137, tmove.f90):
    // points, their ratios when fresh (none is for the first electron), the
    // choice, the backward ratios (det_balance1), the decision and the
    // commit (k_ratio_1e at the chosen point, k_accept_1e)
    int st = tm_alloc(h);
    if (st) return st;
    // Ratios at fresh and backward points come from one k_tm_ratios launch
    // each (a walker with nothing to evaluate leaves at once); a host read
    // of the choice flags skips the backward, decision and commit launches
    // of an electron for which no walker chose a move.
    const int cw = h->Wp;
    const OrdKernels &K = ord_kernels(h->A.S.nordc);
    const LaunchCtx ct = make_ctx(h, h->A, h->Gk[KC_NLPP]);
    int *nfwd = h->tmint, *nbwd = h->tmint + cw, *sel = h->tmint + 2 * cw;
    hipLaunchKernelGGL(k_fill_d, dim3((h->Wl + 255) / 256), dim3(256), 0, h->stream, h->tmstale, 1.0, h->Wl);
    std::vector<int> fl(h->Wl);
    for (int ie = 0; ie < N; ie++) {
      hipLaunchKernelGGL(k_tm_points, g1, b1, 0, h->stream, h->A, ie, cw, h->tmq, h->tmx, h->dm, nfwd);
      if (ie > 0)
        K.tm_ratios(ct, ie, cw, (const double *)h->tmx, (const int *)nfwd, h->rcol, h->tmstale);
      hipLaunchKernelGGL(k_tm_select, g1, b1, 0, h->stream, h->A, ie, cw, (const double *)h->tmq,
                         (const double *)h->tmx, (const double *)h->rcol, h->tmb, h->tmbq, h->dm,
                         nbwd, sel, *p);
      HIPCHK(h, hipMemcpyAsync(fl.data(), sel, h->Wl * sizeof(int), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      int anysel = 0;
      for (int v : fl) anysel |= v;
      if (!anysel) continue;
      if (p->tmoves == 1)
        K.tm_ratios(ct, ie, cw, (const double *)h->tmb, (const int *)nbwd, h->bcol, h->tmstale);
      hipLaunchKernelGGL(k_tm_decide, g1, b1, 0, h->stream, h->A, ie, cw, (const double *)h->tmbq,
                         (const double *)h->bcol, h->tmxin, h->acc_buf, h->dm, h->tmstale, *p);
      if (!h->split_move && !h->dmc_unfused) {
        // only the walkers that accepted evaluate and commit (mode 4)
        h->A.acc_in = h->acc_buf;
        launch_ratio(h, ie, 4, (const double *)h->tmxin, 1.0, 1.0, (double *)nullptr);
        h->A.acc_in = nullptr;
      } else {
        launch_ratio(h, ie, 0, (const double *)h->tmxin, 1.0, 1.0, (double *)nullptr);
        DISPATCH_MOVE(h, launch_accept, ie, 0, (const uint8_t *)h->acc_buf, 1.0, 1.0, (long long *)nullptr);
      }
    }
  }
  if (!h->split_move && !h->dmc_unfused) {
    // gradient, proposal, ratio, test and commit of each electron in one
    // k_ratio_1e launch (mode 3)
    h->A.dm = h->dm;
    h->A.dmp = *p;
    for (int ie = 0; ie < N; ie++) {
      launch_ratio(h, ie, 3, nullptr, 1.0, 1.0, (double *)nullptr);
      count_move(h, ie + 1 == N);
    }
    h->A.dm = nullptr;
  } else {
    for (int ie = 0; ie < N; ie++) {
      DISPATCH_MOVE(h, launch_grad, ie, gbuf);
      hipLaunchKernelGGL(k_dmc_propose, g1, b1, 0, h->stream, h->A, ie, (const double *)gbuf, xin,
                         h->dm, *p);
      launch_ratio(h, ie, 0, (const double *)xin, 1.0, 1.0, (double *)nullptr);
      hipLaunchKernelGGL(k_dmc_decide, g1, b1, 0, h->stream, h->A, ie, h->dm, h->acc_buf, *p);
      DISPATCH_MOVE(h, launch_accept, ie, 0, (const uint8_t *)h->acc_buf, 1.0, 1.0, (long long *)nullptr);
      count_move(h, ie + 1 == N);
    }
  }
  // DMC-mode evaluation: walkers that moved draw a grid rotation per pair
  h->A.dmc_rot = h->A.S.nloc > 0 ? 1 : 0;
  h->A.dmove = h->dm + kDmAcc;
  h->A.dmove_stride = kDmStride;
  launch_eval(h, false);
  h->A.dmc_rot = 0;
  h->A.dmove = nullptr;
  hipLaunchKernelGGL(k_dmc_finish, g1, b1, 0, h->stream, h->A, h->dm, *p);
  hipLaunchKernelGGL(k_dmc_sums, dim3(1), dim3(256), 0, h->stream, h->A, (const double *)h->dm,
                     h->nact, h->dstat);
  hipLaunchKernelGGL(k_dmc_weights, g1, b1, 0, h->stream, (const double *)h->dm, h->nact, h->dwt);
  HIPCHK(h, hipGetLastError());
  double sums[8];
  if (!h->host_split_join) {
    // split_join and the gather on the device; one synchronisation
    int res[8];
    std::vector<int> fl(h->nact);
    HIPCHK(h, hipMemcpyAsync(h->dsj + 4, h->brng, 4 * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_split_join, dim3(1), dim3(1024), 0, h->stream, h->dwt, h->nact, h->W, h->dsj + 4,
                       h->A.rng, h->dorigin, h->dsort, h->npow2, h->dunder, h->dsj);
    hipLaunchKernelGGL(k_dmc_gather, dim3(h->Wp), dim3(256), 0, h->stream, (const double *)h->A.st,
                       h->st2, (int64_t)h->A.L.stride, (const double *)h->dm, h->dm2,
                       (const int *)h->dorigin, (const double *)h->dwt, 0,
                       (const double *)h->A.tq, h->tq2, h->A.tq_stride, (const int *)h->dsj);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(sums, h->dstat, sizeof(sums), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(fl.data(), h->A.flags, h->nact * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(res, h->dsj, 8 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int any = 0;
    for (int f : fl) any |= f;
    if (any & 1) return fail(h, QMCW_ERR_SINGULAR, "a Slater matrix is singular");
    if (any & 2) return fail(h, QMCW_ERR_NAN, "non-finite local energy");
    if (any & 4) return fail(h, QMCW_ERR_ARG, "an electron is in range of more pseudopotential centres than the T-move record holds");
    if (res[2]) return fail(h, QMCW_ERR_NOMEM, "DMC population exceeds the handle's walkers");
    for (int k = 0; k < 4; k++) h->brng[k] = res[4 + k];
    std::swap(h->A.st, h->st2);
    std::swap(h->dm, h->dm2);
    std::swap(h->A.tq, h->tq2);
    h->nact = res[0];
    h->Wl = res[3];
    h->A.Wp = res[3];
    if (stats) {
      stats->wsum = sums[0]; stats->wesum = sums[1]; stats->we2sum = sums[2];
      stats->wtausum = sums[3]; stats->pacc = sums[4]; stats->nacc = sums[5];
      stats->ncross = sums[6]; stats->ntry = sums[7];
      stats->nwalk = res[0];
      stats->nbranch = res[1];
    }
    return QMCW_OK;
  }
  std::vector<double> w(h->Wp, 0.0);
  std::vector<int32_t> rng((size_t)h->Wp * 4, 0);   // the padding records reach Wp
  HIPCHK(h, hipMemcpyAsync(sums, h->dstat, sizeof(sums), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(w.data(), h->dwt, h->nact * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(rng.data(), h->A.rng, (size_t)h->nact * 4 * sizeof(int32_t),
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  int st = check_flags(h);
  if (st) return st;
  std::vector<int> origin;
  int n2 = 0, nbr = 0;
  st = dmc_split_join(h, w, rng, origin, n2, nbr);
  if (st) return st;
  const int Wl2 = (n2 + 7) / 8 * 8;
  std::vector<int> org(Wl2);
  std::vector<double> wn(Wl2, 0.0);
  for (int d = 0; d < Wl2; d++) {
    org[d] = d < n2 ? origin[d] : origin[0];
    wn[d] = d < n2 ? w[d] : 0.0;
    if (d >= n2)
      for (int k = 0; k < 4; k++) rng[4 * d + k] = rng[k];
  }
  HIPCHK(h, hipMemcpyAsync(h->dorigin, org.data(), Wl2 * sizeof(int), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->dwt, wn.data(), Wl2 * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->A.rng, rng.data(), (size_t)Wl2 * 4 * sizeof(int32_t),
                           hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_dmc_gather, dim3(Wl2), dim3(256), 0, h->stream, (const double *)h->A.st,
                     h->st2, (int64_t)h->A.L.stride, (const double *)h->dm, h->dm2,
                     (const int *)h->dorigin, (const double *)h->dwt, Wl2,
                     (const double *)h->A.tq, h->tq2, h->A.tq_stride, (const int *)nullptr);
  HIPCHK(h, hipGetLastError());
  std::swap(h->A.st, h->st2);
  std::swap(h->dm, h->dm2);
  std::swap(h->A.tq, h->tq2);
  h->nact = n2;
  h->Wl = Wl2;
  h->A.Wp = Wl2;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (stats) {
    stats->wsum = sums[0]; stats->wesum = sums[1]; stats->we2sum = sums[2];
    stats->wtausum = sums[3]; stats->pacc = sums[4]; stats->nacc = sums[5];
    stats->ncross = sums[6]; stats->ntry = sums[7];
    stats->nwalk = n2;
    stats->nbranch = nbr;
  }
  return QMCW_OK;
}

extern "C" int qmcw_dmc_get(qmcw_handle *h, int32_t *nwalk, double *weight) {
  if (!h) return QMCW_ERR_ARG;
  if (!h->dmc_on) return fail(h, QMCW_ERR_ARG, "qmcw_dmc_begin has not been called");
  if (nwalk) *nwalk = h->nact;
  if (weight) {
    hipLaunchKernelGGL(k_dmc_weights, dim3((h->nact + 63) / 64), dim3(64), 0, h->stream,
                       (const double *)h->dm, h->nact, h->dwt);
    HIPCHK(h, hipMemcpyAsync(weight, h->dwt, h->nact * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return QMCW_OK;
}
