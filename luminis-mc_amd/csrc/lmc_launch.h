// Transport-kernel selection shared by the host code (lmc_transport.hip) and
// the translation units that instantiate the transport kernels
// (lmc_transport_kernels.hip, compiled once per precision / RNG / gradient
// part so the kernels build in parallel).
#pragma once

#include "lmc_kernels.h"

namespace lmc {

template <class R> using TransportFn = void (*)(KParams<R>);

// The transport kernel of a launch (nullptr where the combination does not
// exist).  aux_mode is the kernels' kAux: 0 no per-photon extras (FP32
// Philox forward runs), 1 penetration depth, 2 depth and the Absorption grid.
// Instantiated for R in {float, double} x kXo x kAdj by
// lmc_transport_kernels.hip.
template <class R, bool kXo, bool kAdj>
TransportFn<R> transport_kernel(int aux_mode, bool est, bool wide, bool lds_tables, bool batch, bool track);

}  // namespace lmc
