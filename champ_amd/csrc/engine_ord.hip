// engine_ord.hip -- the launcher table of one e-e-n order; compiled with
// -DQMCW_ORD=<2..7>.  See engine_launch.h.
#include <hip/hip_runtime.h>

#include "engine_derivs.h"
#include "engine_launch.h"

#ifndef QMCW_ORD
#error "compile with -DQMCW_ORD=<order of the e-e-n polynomial>"
#endif

#define QMCW_CAT2(a, b) a##b
#define QMCW_CAT(a, b) QMCW_CAT2(a, b)

namespace qmcw {
namespace {

constexpr int ORD = QMCW_ORD;

template <int G>
void full_jas_g(const LaunchCtx &c) {
  const dim3 grid(c.Wl / (kWave / G)), block(kWave);
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_full_jas<double, G, ORD>), grid, block, 0, c.stream, c.A);
  else
    hipLaunchKernelGGL((k_full_jas<float, G, ORD>), grid, block, 0, c.stream, c.A);
}
// systems of this many centres or more take the instance whose point
// Jastrow goes two centres at a time
constexpr int kPairCentres = 8;
template <int G>
void full_nlpp_g(const LaunchCtx &c) {
  const dim3 grid(c.Wl / (kWave / G)), block(kWave);
  const bool pair = c.A.S.M >= kPairCentres;
  if (c.precision == QMCW_FP64) {
    if (pair)
      hipLaunchKernelGGL((k_full_nlpp<double, G, ORD, true>), grid, block, c.smem_nlpp, c.stream, c.A);
    else
      hipLaunchKernelGGL((k_full_nlpp<double, G, ORD, false>), grid, block, c.smem_nlpp, c.stream, c.A);
  } else {
    if (pair)
      hipLaunchKernelGGL((k_full_nlpp<float, G, ORD, true>), grid, block, c.smem_nlpp, c.stream, c.A);
    else
      hipLaunchKernelGGL((k_full_nlpp<float, G, ORD, false>), grid, block, c.smem_nlpp, c.stream, c.A);
  }
}
template <int G>
void ratio_g(const LaunchCtx &c, int ie, int mode, const double *xn, double delta, double fbias,
             double *out, long long *nacc) {
  const dim3 grid(c.Wl / (kWave / G)), block(kWave);
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_ratio_1e<double, G, ORD>), grid, block, c.smem_ratio, c.stream, c.A, ie,
                       mode, xn, delta, fbias, out, nacc);
  else
    hipLaunchKernelGGL((k_ratio_1e<float, G, ORD>), grid, block, c.smem_ratio, c.stream, c.A, ie,
                       mode, xn, delta, fbias, out, nacc);
}
template <int G>
void deriv_jas_g(const LaunchCtx &c) {
  const dim3 grid(c.Wl / (kWave / G)), block(kWave);
  hipLaunchKernelGGL((k_deriv_jas<G, ORD>), grid, block, 0, c.stream, c.A);
}
template <int G>
void deriv_nlpp_g(const LaunchCtx &c) {
  const dim3 grid(c.Wl / (kWave / G)), block(kWave);
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_deriv_nlpp<double, G, ORD>), grid, block, c.smem_dnlpp, c.stream, c.A);
  else
    hipLaunchKernelGGL((k_deriv_nlpp<float, G, ORD>), grid, block, c.smem_dnlpp, c.stream, c.A);
}

#define QMCW_BY_G(fn, c, ...)                  \
  switch ((c).G) {                             \
    case 8: fn<8>(c, ##__VA_ARGS__); break;    \
    case 16: fn<16>(c, ##__VA_ARGS__); break;  \
    case 32: fn<32>(c, ##__VA_ARGS__); break;  \
    default: fn<64>(c, ##__VA_ARGS__); break;  \
  }

void full_jas(const LaunchCtx &c) { QMCW_BY_G(full_jas_g, c); }
void full_nlpp(const LaunchCtx &c) { QMCW_BY_G(full_nlpp_g, c); }
void ratio(const LaunchCtx &c, int ie, int mode, const double *xn, double delta, double fbias,
           double *out, long long *nacc) {
  QMCW_BY_G(ratio_g, c, ie, mode, xn, delta, fbias, out, nacc);
}
template <int G>
void move_a_g(const LaunchCtx &c, int ie, int mode, const double *xn, double delta, double fbias) {
  const dim3 grid(c.Wl / (kWave / G)), block(kWave);
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_move_a<double, G, ORD>), grid, block, 0, c.stream, c.A, ie, mode, xn, delta, fbias);
  else
    hipLaunchKernelGGL((k_move_a<float, G, ORD>), grid, block, 0, c.stream, c.A, ie, mode, xn, delta, fbias);
}
void move_a(const LaunchCtx &c, int ie, int mode, const double *xn, double delta, double fbias) {
  QMCW_BY_G(move_a_g, c, ie, mode, xn, delta, fbias);
}
template <int G>
void tm_ratios_g(const LaunchCtx &c, int ie, int cw, const double *pts, const int *cnt, double *out,
                 double *stale) {
  const dim3 grid(c.Wl / (kWave / G)), block(kWave);
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_tm_ratios<double, G, ORD>), grid, block, 0, c.stream, c.A, ie, cw, pts, cnt, out, stale);
  else
    hipLaunchKernelGGL((k_tm_ratios<float, G, ORD>), grid, block, 0, c.stream, c.A, ie, cw, pts, cnt, out, stale);
}
void tm_ratios(const LaunchCtx &c, int ie, int cw, const double *pts, const int *cnt, double *out,
               double *stale) {
  QMCW_BY_G(tm_ratios_g, c, ie, cw, pts, cnt, out, stale);
}
void deriv_jas(const LaunchCtx &c) { QMCW_BY_G(deriv_jas_g, c); }
void deriv_nlpp(const LaunchCtx &c) { QMCW_BY_G(deriv_nlpp_g, c); }

#define QMCW_BY_GB(fn, c, ...)                 \
  switch ((c).G) {                             \
    case 8: fn<8>(c, ##__VA_ARGS__); break;    \
    case 16: fn<16>(c, ##__VA_ARGS__); break;  \
    case 32: fn<32>(c, ##__VA_ARGS__); break;  \
    default: fn<64>(c, ##__VA_ARGS__); break;  \
  }
template <int G>
void ratio_batch_g(const BatchCtx &c, int ie, double delta, double fbias) {
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_ratio_1e_batch<double, G, ORD>), dim3(c.nblocks), dim3(kWave), c.smem,
                       c.stream, c.B, ie, delta, fbias);
  else
    hipLaunchKernelGGL((k_ratio_1e_batch<float, G, ORD>), dim3(c.nblocks), dim3(kWave), c.smem,
                       c.stream, c.B, ie, delta, fbias);
}
template <int G>
void sweep_batch_g(const BatchCtx &c, double delta, double fbias) {
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_sweep_1e_batch<double, G, ORD>), dim3(c.nblocks), dim3(kWave), c.smem,
                       c.stream, c.B, delta, fbias);
  else
    hipLaunchKernelGGL((k_sweep_1e_batch<float, G, ORD>), dim3(c.nblocks), dim3(kWave), c.smem,
                       c.stream, c.B, delta, fbias);
}
template <int G>
void full_jas_batch_g(const BatchCtx &c) {
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_full_jas_batch<double, G, ORD>), dim3(c.nblocks), dim3(kWave), 0, c.stream, c.B);
  else
    hipLaunchKernelGGL((k_full_jas_batch<float, G, ORD>), dim3(c.nblocks), dim3(kWave), 0, c.stream, c.B);
}
template <int G>
void full_nlpp_batch_g(const BatchCtx &c) {
  if (c.precision == QMCW_FP64)
    hipLaunchKernelGGL((k_full_nlpp_batch<double, G, ORD>), dim3(c.nblocks), dim3(kWave), c.smem,
                       c.stream, c.B);
  else
    hipLaunchKernelGGL((k_full_nlpp_batch<float, G, ORD>), dim3(c.nblocks), dim3(kWave), c.smem,
                       c.stream, c.B);
}
void ratio_batch(const BatchCtx &c, int ie, double delta, double fbias) {
  QMCW_BY_GB(ratio_batch_g, c, ie, delta, fbias);
}
void sweep_batch(const BatchCtx &c, double delta, double fbias) {
  QMCW_BY_GB(sweep_batch_g, c, delta, fbias);
}
void full_jas_batch(const BatchCtx &c) { QMCW_BY_GB(full_jas_batch_g, c); }
void full_nlpp_batch(const BatchCtx &c) { QMCW_BY_GB(full_nlpp_batch_g, c); }

}  // namespace

#ifndef __HIP_DEVICE_COMPILE__   // a host table of host launchers
extern const OrdKernels QMCW_CAT(kOrdKernels, QMCW_ORD) = {full_jas, full_nlpp, ratio, deriv_jas,
                                                            deriv_nlpp, move_a, tm_ratios,
                                                            ratio_batch, sweep_batch,
                                                            full_jas_batch, full_nlpp_batch};
#endif

}  // namespace qmcw
