// Instantiations of the persistent transport kernels, one part per
// compilation (LMC_TK_PART 0..4), so the build compiles them in parallel:
//   0: FP32 Philox forward       1: FP32 xoshiro forward    2: FP32 adjoint
//   3: FP64 Philox fwd + adjoint 4: FP64 xoshiro fwd + adjoint
// Every kernel instantiation lives in exactly one part.

#define LMC_TRANSPORT_PARTS  // the plain kernels live in lmc_transport.hip
#include "lmc_launch.h"

#ifndef LMC_TK_PART
#error "compile with -DLMC_TK_PART=<0..4>"
#endif

namespace lmc {
namespace {

// The transport kernel of a launch: the estimator, 1024-thread or
// 512/256-thread body; a batch or a single scene; with or without reverse-path
// tracking; tables in LDS or read through the caches.  kAux as in the
// kernels: 2 (Absorption grid), 0 for FP32 Philox runs only;
// nullptr where a combination does not exist.
template <class R, bool kXo, bool kAdj, int kAux>
TransportFn<R> pick_transport(bool est, bool wide, bool lds_tables, bool batch, bool track) {
  if constexpr (kAux == 0 && (sizeof(R) != 4 || kXo)) {
    return nullptr;
  } else {
    if (est) {
      if constexpr (kAdj || kAux == 0) return nullptr;
      else return lds_tables ? k_transport_est<R, kXo, true, kAux> : k_transport_est<R, kXo, false, kAux>;
    }
    if (batch) {
      if (wide) {
        if constexpr (sizeof(R) == 4 && !kAdj)
          return track ? k_transport_wide<R, kXo, true, kAdj, true, kAux>
                       : k_transport_wide<R, kXo, false, kAdj, true, kAux>;
        else return nullptr;
      }
      if (track)
        return lds_tables ? k_transport<R, kXo, true, true, kAdj, true, kAux>
                          : k_transport<R, kXo, true, false, kAdj, true, kAux>;
      return lds_tables ? k_transport<R, kXo, false, true, kAdj, true, kAux>
                        : k_transport<R, kXo, false, false, kAdj, true, kAux>;
    }
    if (wide) {
      if constexpr (sizeof(R) == 4 && !kAdj)
        return track ? k_transport_wide<R, kXo, true, kAdj, false, kAux> : k_transport_wide<R, kXo, false, kAdj, false, kAux>;
      else return nullptr;
    }
    if (track)
      return lds_tables ? k_transport<R, kXo, true, true, kAdj, false, kAux>
                        : k_transport<R, kXo, true, false, kAdj, false, kAux>;
    return lds_tables ? k_transport<R, kXo, false, true, kAdj, false, kAux>
                      : k_transport<R, kXo, false, false, kAdj, false, kAux>;
  }
}

}  // namespace

template <class R, bool kXo, bool kAdj>
TransportFn<R> transport_kernel(int aux_mode, bool est, bool wide, bool lds_tables, bool batch, bool track) {
  if (aux_mode == 2) return pick_transport<R, kXo, kAdj, 2>(est, wide, lds_tables, batch, track);
  if (aux_mode == 1) return pick_transport<R, kXo, kAdj, 1>(est, wide, lds_tables, batch, track);
  return pick_transport<R, kXo, kAdj, 0>(est, wide, lds_tables, batch, track);
}

#define LMC_TK(R, XO, ADJ) \
  template TransportFn<R> transport_kernel<R, XO, ADJ>(int, bool, bool, bool, bool, bool);
#if LMC_TK_PART == 0
LMC_TK(float, false, false)
#elif LMC_TK_PART == 1
LMC_TK(float, true, false)
#elif LMC_TK_PART == 2
LMC_TK(float, false, true)
LMC_TK(float, true, true)
#elif LMC_TK_PART == 3
LMC_TK(double, false, false)
LMC_TK(double, false, true)
#elif LMC_TK_PART == 4
LMC_TK(double, true, false)
LMC_TK(double, true, true)
#else
#error "LMC_TK_PART out of range"
#endif

}  // namespace lmc
