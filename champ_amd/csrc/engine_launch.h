// engine_launch.h -- launchers of the kernels whose code depends on the
// order of the e-e-n polynomial.  engine_ord.hip is compiled once per order
// (QMCW_ORD = 2..7, one translation unit each, built in parallel) and each
// exports a table of these launchers; the host picks the table of the
// system's nordc, so every kernel carries only the registers of its order.
#pragma once
#include "engine_batch.h"

namespace qmcw {

// What a launcher needs from the handle
struct LaunchCtx {
  KArgs A;              // kernel arguments (by value)
  hipStream_t stream;
  int G;                // lane-group width
  int Wl;               // walker records launched over (multiple of 8)
  int precision;        // QMCW_FP64 / QMCW_MIXED / QMCW_MIXED_BF16
  size_t smem_ratio;    // dynamic LDS of k_ratio_1e (0 when A.lds_phi == 0)
  size_t smem_nlpp;     // dynamic LDS of k_full_nlpp (0 when A.lds_nlpp == 0)
  size_t smem_dnlpp;    // dynamic LDS of k_deriv_nlpp (0 when A.lds_dnlpp == 0)
};

// What a batch launcher needs (engine_batch.h): the device arrays of the
// systems' arguments and block offsets, and the grid
struct BatchCtx {
  BatchArgs B;
  hipStream_t stream;
  int G;                // lane-group width of the kernel class
  int nblocks;          // blocks of all systems
  int precision;
  size_t smem;          // dynamic LDS: the largest any system's blocks need
};

struct OrdKernels {
  void (*full_jas)(const LaunchCtx &);
  void (*full_nlpp)(const LaunchCtx &);
  void (*ratio)(const LaunchCtx &, int ie, int mode, const double *xn, double delta,
                double fbias, double *out, long long *nacc);
  void (*deriv_jas)(const LaunchCtx &);
  void (*deriv_nlpp)(const LaunchCtx &);
  void (*move_a)(const LaunchCtx &, int ie, int mode, const double *xn, double delta, double fbias);
  void (*tm_ratios)(const LaunchCtx &, int ie, int cw, const double *pts, const int *cnt,
                    double *out, double *stale);
  void (*ratio_batch)(const BatchCtx &, int ie, double delta, double fbias);
  void (*sweep_batch)(const BatchCtx &, double delta, double fbias);
  void (*full_jas_batch)(const BatchCtx &);
  void (*full_nlpp_batch)(const BatchCtx &);
};

extern const OrdKernels kOrdKernels2, kOrdKernels3, kOrdKernels4, kOrdKernels5, kOrdKernels6,
    kOrdKernels7;

inline const OrdKernels &ord_kernels(int nordc) {
  switch (nordc <= 2 ? 2 : nordc) {
    case 2: return kOrdKernels2;
    case 3: return kOrdKernels3;
    case 4: return kOrdKernels4;
    case 5: return kOrdKernels5;
    case 6: return kOrdKernels6;
    default: return kOrdKernels7;
  }
}

}  // namespace qmcw
