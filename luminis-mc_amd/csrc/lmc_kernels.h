// Device code of the gfx950 transport path: kernel parameters, per-lane
// photon state, the per-event body, sensor deposits and the kernels.
//
// Register budget.  A photon's live state is what the reference keeps in
// Photon (photon.hpp:65-243) minus what is constant or dead on this path:
// the launch frame P0 and Jones vector E0 are scene constants (scalar
// registers), the exit frame P_n is the current frame, Pn2 and the
// normalised T product never reach a detector, and the frames after the
// first scatter and before the last one keep only their m and s rows
// (n = s x m for the right-handed frames).  Event counters are wave-uniform
// (scalar registers); crossing counters go to LDS.
#pragma once

#include "lmc_device.h"

namespace lmc {

// Laser constants of one scene (Laser, laser.hpp:45-90), with the vacuum
// wavenumber and the absolute index of the layer holding the launch point.
template <class R> struct LaunchConst {
  double lpos_x, lpos_y, sigma, pulse, rep, t_off;
  double k0d;  // vacuum wavenumber in FP64: the planar field's phase k0 * optical path
  R lpos_z;
  R ldir[3], lm[3], ln[3];
  R pol[4];
  R k0;
  R hand;  // +1 when the launch frame is right-handed (m x n = s), else -1
  int32_t source, temporal, launch_layer;
};

// Field order matters: the backend merges neighbouring kernel-argument loads
// into wide scalar loads whose registers live (or spill) together, so the
// pointers the event loop reads stay apart from those only the prologue
// copies (layers, species, sensors, planes, laser constants), and the
// per-event scalars apart from the per-photon refill keys (a spilled group is
// reloaded whole: 8 lane reads where one value is needed).
template <class R> struct KParams {
  // the phase and amplitude tables in one blob: theta CDFs (R), S1/S2 nodes
  // (STab<R>), species-selection CDFs (R) and the CDFs' guide tables (u16),
  // each section 32-byte aligned; species and layers hold offsets from the
  // blob's start in units of the element type, so one base address serves all
  // four.  The first tab_bytes (all but the guides) are what a launch with
  // LDS-resident tables copies to LDS.
  const unsigned char* tab;
  double* out;
  unsigned long long* hits;
  unsigned long long* counter;
  unsigned long long* stats;  // photons, events, trials, crossings, cbs
  double* grad;
  R* qf;
  double* qd;                   // FP64 optical path at the crossing
  int32_t* qi;
  unsigned long long* q_count;  // [0] slots handed out, [1] overflowed
  int32_t max_events, q_chunk;  // read every event / every crossing: loaded apart from the refill keys
  int64_t q_cap;
  int64_t n_out;
  uint64_t n_photons, offset;
  uint64_t local_offset;        // batched runs: first photon index within every scene
  int32_t n_layers, n_species, n_sensors, n_planes, n_scenes;
  int32_t tab_bytes;
  int32_t lds_grid, claim;
  int32_t lds_grad;             // adjoint runs: privatised grids keep their gradient rows in LDS too
  uint32_t key0, key1, idx_hi;
  // staged into LDS by the prologue (read in the loop only by batched runs)
  const LayerDev<R>* layers;
  const SpeciesDev<R>* species;
  const SensorDev* sensors;
  const PlaneDev* planes;
  const LaunchConst<R>* lcs;     // one per scene
  const uint64_t* prefix;        // batched runs: first global photon index of each scene, n_scenes + 1
  const R* aux;                  // estimator bin geometry and statistics histogram parameters
  const AbsDev<R>* abs_scenes;   // batched runs: each scene's Absorption grid
  R abs_radius, abs_depth, abs_dr, abs_dz, abs_dt;  // Absorption grid of a single-scene run
  int32_t abs_on, abs_nr, abs_nz, abs_nt;
  int64_t abs_off;
  int32_t est_lds_n;             // doubles of the estimator grids staged in LDS
  int32_t direct;                // 1: the transport kernel deposits the direct sensors (deposits_direct)
  R pz_lo, pz_hi;                // lowest and highest sensor plane (as R) of all scenes of the launch
  // PhotonRecordSensor records of one sensor-kernel launch (LMC_RECORD_FIELDS
  // doubles each), [0] records written, [1] records that found no room
  double* rec;
  unsigned long long* rec_count;
  int64_t rec_cap;
  // Binned deposits (forward runs): the sensor kernel runs twice over a range
  // of records, the first time counting per workgroup the deposits into each
  // tile of 2^bin_shift accumulators of the grids that live in HBM, the second
  // time writing them as (accumulator index, value) entries grouped by tile;
  // k_bin_accumulate then sums each tile in LDS and adds it into `out` once
  // (see k_bin_accumulate for why: FP64 atomics into HBM are the sensor
  // kernel's bound).
  uint32_t* bin_keys;    // entry: accumulator index
  double* bin_vals;      //        and value
  uint32_t* bin_pos;     // [tile][workgroup]: deposit counts, then the first entry of each
  uint32_t* bin_tile;    // first entry of each tile (n_tiles + 1), from the tile totals
  uint4* bin_items;      // accumulation work items: tile, first entry, entries, 1 if the tile has one item
  uint32_t* bin_n_items;
  uint32_t* bin_flag;    // set when the two passes disagree on a workgroup's deposits (never expected)
  int64_t bin_rec_lo, bin_rec_per_wg;  // records [lo + wg * per_wg, + per_wg) of this pass, per workgroup
  int64_t bin_cap;       // entries the buffers hold
  int32_t bin_tiles, bin_shift, bin_wgs, bin_chunk;
};

// One queued plane crossing accepted by at least one sensor of its plane:
// everything the sensors' process_hit needs (detector.cpp:178-270 builds the
// same InteractionInfo), so the transport kernel carries none of the sensor
// code.  Array of structures: each record is kRecQ quads of R written with
// 16-byte vector stores, so a record fills whole cache lines at once.
//   Q0 hit x, hit y, arrival path, weight      Q1 E_m, E_n (re, im)
//   Q2-Q4 exit frame rows m, n, s (+ first-frame m)
//   Q4-Q9 CBS history: frame after the first scatter (m, s rows), frame
//         before the last one (m, s), raw interior product T (4 complex),
//         r_n - r_1                             (photons with >= 2 events at
//                                               a plane holding a CBS sensor)
//   Q10-Q11 PhotonRecordSensor extras: r_1, launch time, r_n, path at the
//           crossing                            (planes holding one)
//   Q12-Q14 adjoint scores (adjoint runs only)
// The plane, sensor mask, event count and species pair go to int arrays, the
// FP64 optical path to a double array (both indexed by slot).
constexpr int kRecQ = 15;
constexpr int kRecQLight = 5, kRecQCbs = 10, kRecQScore = 12;

template <class R> __device__ __forceinline__ void put4(R* f, int q, R a, R b, R c, R d) {
  if constexpr (sizeof(R) == 4) {
    reinterpret_cast<float4*>(f)[q] = make_float4(a, b, c, d);
  } else {
    reinterpret_cast<double2*>(f)[2 * q] = make_double2(a, b);
    reinterpret_cast<double2*>(f)[2 * q + 1] = make_double2(c, d);
  }
}
template <class R> struct Quad { R a, b, c, d; };
template <class R> __device__ __forceinline__ Quad<R> get4(const R* f, int q) {
  if constexpr (sizeof(R) == 4) {
    const float4 v = reinterpret_cast<const float4*>(f)[q];
    return {v.x, v.y, v.z, v.w};
  } else {
    const double2 u = reinterpret_cast<const double2*>(f)[2 * q];
    const double2 v = reinterpret_cast<const double2*>(f)[2 * q + 1];
    return {u.x, u.y, v.x, v.y};
  }
}
enum RecI { I_PLANE = 0, I_MASK = 1, I_EVENTS = 2, I_SP = 3, I_DEPTH = 4, kRecI = 5 };  // I_DEPTH: float bits (kernels with kAux >= 1)

template <class R> struct SceneView {
  const LaunchConst<R>* lc;
  const LayerDev<R>* L;
  const SpeciesDev<R>* S;
  const R* cdf;
  const uint16_t* guide;
  const STab<R>* stab;
  const R* sel;
  const SensorDev* sens;
  const PlaneDev* planes;
  double* lgrid;     // LDS accumulators
  unsigned* lhits;   // LDS hit counters [kMaxSensors], then crossings, cbs
  double* lrefl;     // single scenes: LDS reflectance sums of the direct sensors [kMaxSensors]
  unsigned long long* wq_base;  // per wave: first slot of its current queue chunk
  unsigned* wq_used;            // per wave: slots used in that chunk
};

// xoshiro state exists only in the sequential-stream build.
template <bool kXo> struct XoState {
  __device__ __forceinline__ void init(uint64_t, uint64_t) {}
};
template <> struct XoState<true> {
  XoshiroStream s;
  __device__ __forceinline__ void init(uint64_t seed, uint64_t photon) { s.init(seed, 0, photon); }
};

template <class R, bool kXo> struct Photon {
  V3<R> pos;
  Frame<R> P;
  cx<R> Em, En;
  R w, gpl, t0;
  R depth;  // maximum z reached (penetration depth, simulation.cpp:396-400)
  double opl;
  int32_t events, layer;
  int32_t cur_sp;    // species of the pending vertex between the two parts of an estimator event
  Xoshiro128ss rng;  // this photon's event stream (seeded from Philox at launch)
  bool alive;
  // FP32 kernels: 1 when the photon's last flight segment left the sample
  // with its plane crossings not yet recorded (the segment's start is kept in
  // w_th, w_ph, w_az), 2 when that segment also started on an interface
  int32_t pend;
  // CBS history
  V3<R> m1, s1;          // frame after the first scatter
  R qc, qs;              // cos and sin of the last scattering angle: the frame before the last scatter is A^T P
  V3<R> r1, rn;
  cx<R> T[4];            // raw interior product J_{n-1}...J_2 (Frobenius-normalised)
  cx<R> b2, b1;          // last event's J_n = diag(b2, b1) R(bc, bs), pending
  R bc, bs;
  int32_t sp;            // (first species << 16) | last species
  // words of the current event kept past the flight (theta, the two azimuth
  // draws, the shared 24-bit auxiliary draw)
  uint32_t w_th, w_ph, w_az;
  XoState<kXo> xo;
  R aL[kMaxGradLayers], aN[kMaxGradLayers], aG[kMaxGradLayers];
  // the current layer's path length, scatter count and HG score not yet added
  // into aL, aN, aG[layer] (added when the photon changes layer): one add per
  // event instead of a masked add per layer slot
  R cL, cN, cG;
};

template <class R> __device__ __forceinline__ V3<R> cross3(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ uint4 event_block(const uint32_t key0, const uint32_t key1, uint32_t idx, uint32_t idx_hi,
                                             uint32_t ev, uint32_t sub) {
  // The key is uniform: a scalar copy the compiler cannot see through keeps
  // the round-key schedule on the scalar unit (SALU adds) without hoisting
  // all 20 round keys into scalar registers for the whole event loop.
  uint32_t k0v, k1v;
  asm volatile("s_mov_b32 %0, %1" : "=s"(k0v) : "s"(key0));
  asm volatile("s_mov_b32 %0, %1" : "=s"(k1v) : "s"(key1));
  return philox4x32_10(make_uint4(idx, idx_hi, ev, sub), k0v, k1v);
}

// Random numbers of the production path.  Photon i of a run owns a stream:
// Philox4x32-10 of the counter (i, high word of i, 0xFFFFFFFE, 0) under the
// seed is the initial state of a xoshiro128** generator (a uniformly random
// point of its 2^128 - 1 cycle).  The emission draws come first (four
// outputs for an extended source, two for a pulse profile); then each event
// takes the next four outputs as words x, y, z, w: x -> free path (high 24
// bits), y -> theta (high 24), z -> azimuth component and radius (high 24),
// w -> azimuth angle (high 23) and half period (bit 8); the low bytes of x,
// y, z form one auxiliary 24-bit uniform (the species choice in a mixture
// layer, otherwise the roulette draw).  Every draw is a function of the seed
// and the photon index alone, whatever the launch shape.
__device__ __forceinline__ Xoshiro128ss seed_stream(uint32_t key0, uint32_t key1, uint32_t idx, uint32_t idx_hi) {
  const uint4 sd = event_block(key0, key1, idx, idx_hi, 0xFFFFFFFEu, 0u);
  return {sd.x | ((sd.x | sd.y | sd.z | sd.w) == 0u ? 1u : 0u), sd.y, sd.z, sd.w};  // never the zero state
}

template <class R, bool kXo> __device__ __forceinline__ R aux_uniform(const Photon<R, kXo>& ph) {
  const uint32_t x = ((ph.w_th & 0xFFu) << 16) | ((ph.w_az & 0xFFu) << 8) | (ph.w_ph & 0xFFu);
  return static_cast<R>(static_cast<float>(x) + 0.5f) * static_cast<R>(5.9604644775390625e-8);
}

// a / b: hardware reciprocal in FP32 (~1 ulp), exact division in FP64.
template <class R> __device__ __forceinline__ R qdiv(R a, R b) {
  if constexpr (sizeof(R) == 4)
    return a * __builtin_amdgcn_rcpf(b);
  else
    return a / b;
}

// Azimuth with density F(phi) = a cos^2 phi + b sin^2 phi + c sin phi cos phi,
// the density the reference samples by rejection (phase.cpp:27-49), here
// exactly and without a loop.  F is a positive semi-definite quadratic form
// in (cos phi, sin phi) with eigenvalues l1 = M + Rm >= l2 = M - Rm (M =
// (a+b)/2, Rm = |((a-b)/2, c/2)|), so F(phi0 + chi) = l2 + (l1 - l2) cos^2 chi
// about the principal direction phi0: a mixture of the uniform density
// (weight 1 - e) and cos^2 chi / pi (weight e = Rm / M).  The angle chi of a
// point uniform in the unit disk centred at (1, 0) has density proportional to
// cos^2 chi on (-pi/2, pi/2) (the disk is r < 2 cos chi in polar form), and a
// half-period flip extends it to the whole circle.  Draws: v (bits 8-31 of the
// first word) picks the component, v M < Rm, and is then uniform again as
// v M / Rm, the disk point's squared radius; bits 9-31 of the second word are
// the disk point's angle (the whole direction for the uniform component), bit
// 8 the half period.  Returns cos phi, sin phi and F(phi).
template <class R>
__device__ __forceinline__ void azimuth_sample(R a, R b, R c, uint32_t wv, uint32_t wa, R& cph, R& sph, R& F) {
  const R Mm = R(0.5) * (a + b), A = R(0.5) * (a - b), B = R(0.5) * c;
  const R Rm = M<R>::sqrt(A * A + B * B);
  // principal eigenvector: (A + Rm, B) for A >= 0, else (B, Rm - A); its
  // squared norm is 2 Rm (Rm + |A|)
  const R t = Rm + M<R>::fabs(A);
  // (1, 0) for an isotropic form (Rm = 0); selects, not a branch
  const bool ani = Rm > R(0);
  const R nrm0 = M<R>::rsqrt(ani ? R(2) * Rm * t : R(1));
  const R c0 = ani ? (A >= R(0) ? t : B) * nrm0 : R(1);
  const R s0 = ani ? (A >= R(0) ? B : t) * nrm0 : R(0);
  const R v = u01<R>(wv);
  const bool lobe = v * Mm < Rm;
  const R r = lobe ? M<R>::sqrt(qdiv(v * Mm, Rm)) : R(1);
  R sa, ca;
  M<R>::sincos_rev(static_cast<R>(static_cast<float>(wa >> 9) + 0.5f) * static_cast<R>(1.1920928955078125e-7), &sa, &ca);
  const R px = (lobe ? R(1) : R(0)) + r * ca, py = r * sa;
  R nrm = M<R>::rsqrt(px * px + py * py);
  if (wa & 0x100u) nrm = -nrm;
  const R cx = px * nrm, sx = py * nrm;
  cph = c0 * cx - s0 * sx;
  sph = s0 * cx + c0 * sx;
  F = a * cph * cph + b * sph * sph + c * cph * sph;
}

// Inverse-CDF lookup: SamplingTable::Sample (src/table.cpp:50-66).  The
// lower_bound is narrowed by a guide table g[j] = lower_bound(j/G) (stored
// shifted right by guide_shift for tables beyond 65536 nodes, the bracket
// widened to match), so the result is exactly the reference's index on the
// same CDF values.
template <class R>
__device__ __forceinline__ R table_sample(const SceneView<R>& v, const SpeciesDev<R>& s, R u) {
  const R* c = v.cdf + s.cdf_off;
  const uint16_t* g = v.guide + s.guide_off;
  int j = static_cast<int>(u * static_cast<R>(s.guide_n));
  j = j < 0 ? 0 : (j >= s.guide_n ? s.guide_n - 1 : j);
  const int sh = s.guide_shift;
  int lo = static_cast<int>(g[j]) << sh, hi = (static_cast<int>(g[j + 1]) << sh) + ((1 << sh) - 1);
  hi = hi < s.cdf_n - 1 ? hi : s.cdf_n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] < u)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo == 0) return s.cdf_min;
  const R ca = c[lo - 1], cb = c[lo];
  const R xa = s.cdf_min + static_cast<R>(lo - 1) * s.cdf_step;
  const R t = qdiv(u - ca, cb - ca);
  return xa + t * s.cdf_step;
}

// DataTable::Sample on the uniform theta grid (src/table.cpp:100-118).
template <class R>
__device__ __forceinline__ void s_lookup(const STab<R>* stab, const SpeciesDev<R>& s, R th, cx<R>& s1, cx<R>& s2) {
  const STab<R>* tb = stab + s.s_off;
  const R pos = th * s.inv_ds;
  const int last = s.s_n - 1;
  if constexpr (sizeof(R) == 4) {
    // FP32: one clamped interpolation (s_n >= 2) instead of three branches;
    // pos <= 0 and NaN give node 0, pos >= last gives node last (t = 1)
    const R pc = fminf(pos > R(0) ? pos : R(0), static_cast<R>(last));
    const int i = min(static_cast<int>(pc), last - 1);
    const R t = pc - static_cast<R>(i);
    cx<R> a1, a2, b1, b2;
    stab_get(tb[i], a1, a2);
    stab_get(tb[i + 1], b1, b2);
    s1 = {a1.r + t * (b1.r - a1.r), a1.i + t * (b1.i - a1.i)};
    s2 = {a2.r + t * (b2.r - a2.r), a2.i + t * (b2.i - a2.i)};
    return;
  }
  if (!(pos > R(0))) {  // also maps a NaN angle to node 0: indices stay in bounds
    stab_get(tb[0], s1, s2);
    return;
  }
  if (pos >= static_cast<R>(last)) {
    stab_get(tb[last], s1, s2);
    return;
  }
  const int i = static_cast<int>(pos);
  const R t = pos - static_cast<R>(i);
  cx<R> a1, a2, b1, b2;
  stab_get(tb[i], a1, a2);
  stab_get(tb[i + 1], b1, b2);
  s1 = {a1.r + t * (b1.r - a1.r), a1.i + t * (b1.i - a1.i)};
  s2 = {a2.r + t * (b2.r - a2.r), a2.i + t * (b2.i - a2.i)};
}

template <class R> __device__ __forceinline__ R clamp1(R x) { return x > R(1) ? R(1) : (x < R(-1) ? R(-1) : x); }

// Exit polar angle from the -z axis.  FP64: acos(-s_z) as the reference
// (detector.cpp:1274); FP32: atan2(|s_perp|, -s_z), which keeps ~1e-7 rad
// resolution near exact backscatter where acos in FP32 would quantise to 3e-4.
template <class R> __device__ __forceinline__ R exit_theta(V3<R> s) {
  if constexpr (sizeof(R) == 8) {
    return M<R>::acos(-s.z);
  } else {
    return M<R>::atan2(M<R>::sqrt(s.x * s.x + s.y * s.y), -s.z);
  }
}
template <class R> __device__ __forceinline__ R exit_phi(V3<R> s) {
  R phi = M<R>::atan2(s.y, s.x);
  if (phi < R(0)) phi += R(2.0 * 3.14159265358979323846);
  return phi;
}

// Reverse-path Jones vector: Xu (2008) three-stage construction,
// reverse_field_prefix/suffix (detector.cpp:1080-1161).  P0/E0 are the
// launch frame and Jones vector; Pq is the frame before the last scatter
// (P_{n-1}), P1 the frame after the first scatter, Pn the exit frame and T the
// raw interior product J_{n-1}...J_2.  The prefix (stages A and B, scatter at
// r_n and the interior by reciprocity) does not depend on the exit frame.
template <class R>
__device__ __forceinline__ void reverse_prefix(const LaunchConst<R>& lc, const STab<R>* stab, const SpeciesDev<R>& spn,
                                               const Frame<R>& Pq, const cx<R> T[4], cx<R>& xm, cx<R>& xn) {
  const V3<R> s0{lc.ldir[0], lc.ldir[1], lc.ldir[2]};
  const V3<R> m0{lc.lm[0], lc.lm[1], lc.lm[2]};
  const V3<R> n0{lc.ln[0], lc.ln[1], lc.ln[2]};
  // Stage A: scatter at r_n, s0 -> -s_{n-1}
  const V3<R> s_out_a{-Pq.s.x, -Pq.s.y, -Pq.s.z};
  const R th_a = M<R>::acos(clamp1(dot(s0, s_out_a)));
  cx<R> a1, a2;
  s_lookup(stab, spn, th_a, a1, a2);
  const V3<R> np = unit_cross_or(s0, s_out_a, n0);
  const V3<R> mp_in = unit_cross_or(np, s0, m0);
  const V3<R> mp_out = unit_cross_or(np, s_out_a, Pq.m);
  const cx<R> E0m{lc.pol[0], lc.pol[1]}, E0n{lc.pol[2], lc.pol[3]};
  // E = S_a * rot2(mp_in, np, m0, n0) * E0
  R r00 = dot(mp_in, m0), r01 = dot(mp_in, n0), r10 = dot(np, m0), r11 = dot(np, n0);
  xm = cadd(cscale(E0m, r00), cscale(E0n, r01));
  xn = cadd(cscale(E0m, r10), cscale(E0n, r11));
  xm = cmul(a2, xm);
  xn = cmul(a1, xn);
  // rot2(m_{n-1}, -n_{n-1}, mp_out, np)
  const V3<R> nq{-Pq.n.x, -Pq.n.y, -Pq.n.z};
  r00 = dot(Pq.m, mp_out);
  r01 = dot(Pq.m, np);
  r10 = dot(nq, mp_out);
  r11 = dot(nq, np);
  cx<R> ym = cadd(cscale(xm, r00), cscale(xn, r01));
  cx<R> yn = cadd(cscale(xm, r10), cscale(xn, r11));
  // Stage B: Q T^T Q with Q = diag(1,-1): [[T00, -T10], [-T01, T11]]
  xm = cadd(cmul(T[0], ym), cscale(cmul(T[2], yn), R(-1)));
  xn = cadd(cscale(cmul(T[1], ym), R(-1)), cmul(T[3], yn));
  if constexpr (sizeof(R) == 4) {
    // the field's scale cancels in the suffix's normalisation; bringing it to
    // one here keeps the suffix's squared norm inside FP32 range whatever the
    // scale of T (renormalised every other event in FP32 runs) and of S
    const R n2 = cnorm(xm) + cnorm(xn);
    if (n2 > R(0)) {
      const R inv = M<R>::rsqrt(n2);
      xm = cscale(xm, inv);
      xn = cscale(xn, inv);
    }
  }
}

// Stage C of the reverse field and its normalisation (reverse_field_suffix,
// detector.cpp:1123-1161), from the prefix (xm, xn).
template <class R>
__device__ __forceinline__ void reverse_suffix(const STab<R>* stab, const SpeciesDev<R>& sp1, const Frame<R>& P1,
                                               const Frame<R>& Pn, R hand, cx<R> xm, cx<R> xn, cx<R>& Erm,
                                               cx<R>& Ern) {
  // Stage C: scatter at r_1, -s1 -> s_n
  const V3<R> s_in_c{-P1.s.x, -P1.s.y, -P1.s.z};
  const R th_c = M<R>::acos(clamp1(dot(s_in_c, Pn.s)));
  cx<R> c1, c2;
  s_lookup(stab, sp1, th_c, c1, c2);
  const V3<R> n1neg{-P1.n.x, -P1.n.y, -P1.n.z};
  const V3<R> npp = unit_cross_or(s_in_c, Pn.s, n1neg);
  R r00, r01, r10, r11;
  if constexpr (sizeof(R) == 4) {
    // FP32: mpp_in = unit(npp x s_in) is not formed.  npp is a unit vector
    // normal to s1, and P1 = (m1, s1 x m1, s1) is right-handed, so
    // (npp x s_in) . m1 = npp . (s_in x m1) = npp . (-n1) and
    // (npp x s_in) . (-n1) = -npp . m1
    r10 = dot(npp, P1.m);
    r11 = dot(npp, n1neg);
    r00 = r11;
    r01 = -r10;
  } else {
    const V3<R> mpp_in = unit_cross_or(npp, s_in_c, P1.m);
    r00 = dot(mpp_in, P1.m);
    r01 = dot(mpp_in, n1neg);
    r10 = dot(npp, P1.m);
    r11 = dot(npp, n1neg);
  }
  cx<R> ym = cadd(cscale(xm, r00), cscale(xn, r01));
  cx<R> yn = cadd(cscale(xm, r10), cscale(xn, r11));
  ym = cmul(c2, ym);
  yn = cmul(c1, yn);
  if constexpr (sizeof(R) == 4) {
    // mpp_out = unit(npp x s_n) likewise: the photon frame has the launch
    // frame's handedness h (m x n = h s), so m_n . (npp x s_n) = h n_n . npp
    // and n_n . (npp x s_n) = -h m_n . npp
    r01 = dot(Pn.m, npp);
    r11 = dot(Pn.n, npp);
    r00 = hand * r11;
    r10 = -hand * r01;
  } else {
    const V3<R> mpp_out = unit_cross_or(npp, Pn.s, Pn.m);
    r00 = dot(Pn.m, mpp_out);
    r01 = dot(Pn.m, npp);
    r10 = dot(Pn.n, mpp_out);
    r11 = dot(Pn.n, npp);
  }
  xm = cadd(cscale(ym, r00), cscale(yn, r01));
  xn = cadd(cscale(ym, r10), cscale(yn, r11));
  const R nrm2 = cnorm(xm) + cnorm(xn);
  if (nrm2 > R(0) && M<R>::sqrt(nrm2) > static_cast<R>(sizeof(R) == 8 ? 1e-300 : 1e-30)) {
    const R inv = M<R>::rsqrt(nrm2);
    Erm = cscale(xm, inv);
    Ern = cscale(xn, inv);
  } else {
    Erm = {R(0), R(0)};
    Ern = {R(0), R(0)};
  }
}

// reverse_field (detector.cpp:1165-1173) of a detected photon.
template <class R>
__device__ __forceinline__ void reverse_field(const LaunchConst<R>& lc, const STab<R>* stab, const SpeciesDev<R>& spn,
                                              const SpeciesDev<R>& sp1, const Frame<R>& P1, const Frame<R>& Pq,
                                              const Frame<R>& Pn, const cx<R> T[4], cx<R>& Erm, cx<R>& Ern) {
  cx<R> xm, xn;
  reverse_prefix(lc, stab, spn, Pq, T, xm, xn);
  reverse_suffix(stab, sp1, P1, Pn, lc.hand, xm, xn, Erm, Ern);
}

// Accumulate one value: LDS copy (ds_add_f64) for privatised grids, HBM
// (global_atomic_add_f64) otherwise.  The two pointers keep their address
// spaces so neither path degrades to a flat atomic.
typedef __attribute__((address_space(3))) double lds_double;
__device__ __forceinline__ void acc_add(double* lds_base, double* hbm_base, bool lds, int64_t off, double val) {
  if (val != 0.0) {
    if (lds)
      __hip_atomic_fetch_add((lds_double*)(lds_base + off), val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      unsafeAtomicAdd(hbm_base + off, val);
  }
}

// Adjoint scores of one detected photon for every layer parameter:
// d ln(path density x weight)/d mu_s = N_l/mu_s - L_l, d/d mu_a = -L_l,
// d/dg = sum of d ln p_HG/dg over the layer's scatters.
template <class R, bool kXo>
__device__ __forceinline__ void adj_scores(const KParams<R>& p, const SceneView<R>& v, const Photon<R, kXo>& ph,
                                           R corr, R score[3 * kMaxGradLayers]) {
#pragma unroll
  for (int l = 0; l < kMaxGradLayers; ++l) {
    const R mus = l < p.n_layers ? v.L[l].mu_s : R(1);
    const bool cur = l == ph.layer;
    const R Ll = ph.aL[l] + (cur ? ph.cL - corr : R(0));
    const R Nl = ph.aN[l] + (cur ? ph.cN : R(0));
    score[3 * l + 0] = (mus > R(0) ? Nl / mus : R(0)) - Ll;
    score[3 * l + 1] = -Ll;
    score[3 * l + 2] = ph.aG[l] + (cur ? ph.cG : R(0));
  }
}

// kBin: 0 deposits (LDS or an FP64 atomic in HBM); 1 counts the deposits into
// grids marked for binning (sd.bin) per tile in the workgroup's LDS counters
// `lbin` (no side effect otherwise); 2 writes each of those as an entry at its
// tile's next slot (`lbin`: cursors) and makes the other deposits as with 0.
// Binned deposits keep zero values, so both passes see the same ones.
template <class R, bool kAdj, int kBin = 0>
__device__ __forceinline__ void deposit(const KParams<R>& p, const SensorDev& sd, double* lds_grid, int64_t off,
                                        double val, const R* score, unsigned* lbin = nullptr) {
  if constexpr (kBin != 0) {
    static_assert(!kAdj, "binned deposits serve forward runs");
    if (!(p.direct ? sd.bin_d : sd.bin)) {  // LDS copy or FP64 atomic, in the writing pass only
      if constexpr (kBin == 2) acc_add(lds_grid + sd.lds_off, p.out + sd.out_off, sd.lds != 0, off, val);
      return;
    }
    const uint32_t key = static_cast<uint32_t>(sd.out_off + off);
    if constexpr (kBin == 1) {
      atomicAdd(&lbin[key >> p.bin_shift], 1u);
    } else {
      const unsigned e = atomicAdd(&lbin[key >> p.bin_shift], 1u);
      if (e < p.bin_cap) {
        p.bin_keys[e] = key;
        p.bin_vals[e] = val;
      }
    }
    return;
  }
  acc_add(lds_grid + sd.lds_off, p.out + sd.out_off, sd.lds != 0, off, val);
  if constexpr (kAdj) {
    // gradient row q of this value: LDS copy after the value grids when the
    // grid is privatised (a one-cell grid would otherwise take every
    // photon's atomics on one address), HBM otherwise
    const int np = 3 * (p.n_layers < kMaxGradLayers ? p.n_layers : kMaxGradLayers);
    const bool lg = sd.lds != 0 && p.lds_grad != 0;
    for (int q = 0; q < np; ++q)
      if (score[q] != R(0))
        acc_add(lds_grid + static_cast<int64_t>(p.lds_grid) * (1 + q) + sd.lds_off, p.grad + q * p.n_out + sd.out_off,
                lg, off, val * static_cast<double>(score[q]));
  }
}

// Stokes S0..S3 of a lab-frame field (Ex, Ey): |Ex|^2 + |Ey|^2, |Ex|^2 - |Ey|^2,
// 2 Re(Ex Ey*), 2 Im(Ex Ey*) (detector.cpp:823-827).
template <class R> __device__ __forceinline__ R stokes(int q, cx<R> Ex, cx<R> Ey) {
  if (q == 0) return cnorm(Ex) + cnorm(Ey);
  if (q == 1) return cnorm(Ex) - cnorm(Ey);
  if (q == 2) return R(2) * (Ex.r * Ey.r + Ex.i * Ey.i);
  return R(2) * (Ex.i * Ey.r - Ex.r * Ey.i);
}

// StatisticsSensor::process_hit (detector.cpp:1703-1800): one count in each
// set histogram (events, exit theta and phi, penetration depth, each also in
// its time window; arrival time; weight).  Parameters at p.aux + sd.aux in the
// order of the scene descriptor's block; an index the reference could compute
// one past the end (a value just below the upper edge) is clamped.
template <class R, bool kAdj>
__device__ __forceinline__ void stats_deposit(const KParams<R>& p, const SensorDev& sd, double* lds_grid, int events,
                                              V3<R> dir, R arr, R w, float depth, const R* score) {
  const R* a = p.aux + sd.aux;
  const int Nt = static_cast<int>(a[24]);
  const R dt = a[25], tmax = a[26];
  int t_idx = -1;
  if (Nt > 1 && dt > R(0) && arr >= R(0) && arr < tmax) {
    t_idx = static_cast<int>(arr / dt) + 1;
    if (t_idx >= Nt) t_idx = -1;
  }
  int64_t base = 0;
  auto temporal = [&](int n, int idx) {
    idx = idx < n - 1 ? idx : n - 1;
    deposit<R, kAdj>(p, sd, lds_grid, base + idx, 1.0, score);
    if (t_idx >= 1)
      deposit<R, kAdj>(p, sd, lds_grid, base + static_cast<int64_t>(t_idx) * n + idx, 1.0, score);
  };
  if (a[0] != R(0)) {
    const int me = static_cast<int>(a[6]);
    if (events >= 0 && events < me) temporal(me, events);
    base += static_cast<int64_t>(Nt) * me;
  }
  if (a[1] != R(0)) {
    const int n = static_cast<int>(a[9]);
    const R theta = exit_theta(dir);
    if (theta >= a[7] && theta < a[8]) temporal(n, static_cast<int>((theta - a[7]) / a[10]));
    base += static_cast<int64_t>(Nt) * n;
  }
  if (a[2] != R(0)) {
    const int n = static_cast<int>(a[13]);
    const R phi = exit_phi(dir);
    if (phi >= a[11] && phi < a[12]) temporal(n, static_cast<int>((phi - a[11]) / a[14]));
    base += static_cast<int64_t>(Nt) * n;
  }
  if (a[3] != R(0)) {
    const int n = static_cast<int>(a[16]);
    const R d = static_cast<R>(depth);
    if (d >= R(0) && d < a[15]) temporal(n, static_cast<int>(d / a[17]));
    base += static_cast<int64_t>(Nt) * n;
  }
  if (a[4] != R(0)) {
    const int n = static_cast<int>(a[19]);
    const int idx = static_cast<int>(arr / a[20]);
    if (idx >= 1 && idx < n) deposit<R, kAdj>(p, sd, lds_grid, base + idx, 1.0, score);
    base += n;
  }
  if (a[5] != R(0)) {
    const int n = static_cast<int>(a[22]);
    if (w >= R(0) && w < a[21]) {
      const int idx = static_cast<int>(w / a[23]);
      deposit<R, kAdj>(p, sd, lds_grid, base + (idx < n - 1 ? idx : n - 1), 1.0, score);
    }
  }
}

// Deposit the 8 CBS grids for one detection: coherent |E_f + E_r|^2 terms and
// incoherent |E_f|^2 + |E_r|^2 terms (accumulate_stokes, detector.cpp:1194-1226).
// A time-resolved hit goes to its time bin only; the reference's second copy
// in the time-integrated bin 0 is the sum of bins 1.. and is formed when the
// accumulators are read (fold_time_bins), which halves these atomics.
template <class R, bool kAdj, int kBin = 0>
__device__ __forceinline__ void cbs_deposit(const KParams<R>& p, const SensorDev& sd, double* lds_grid, int64_t bin,
                                            int t_idx, cx<R> Efx, cx<R> Efy, cx<R> Erx, cx<R> Ery, const R* score,
                                            unsigned* lbin = nullptr) {
  const int64_t tplane = static_cast<int64_t>(sd.n0) * sd.n1;
  const int64_t plane = tplane * sd.n2;
  const cx<R> Ex = cadd(Efx, Erx), Ey = cadd(Efy, Ery);
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    const R val = q < 4 ? stokes(q, Ex, Ey) : stokes(q - 4, Efx, Efy) + stokes(q - 4, Erx, Ery);
    deposit<R, kAdj, kBin>(p, sd, lds_grid, q * plane + t_idx * tplane + bin, static_cast<double>(val), score, lbin);
  }
}

// Unit right-handed frame rows m, n = s x m, s from stored m and s rows.
template <class R> __device__ __forceinline__ Frame<R> frame_of(V3<R> m, V3<R> s) {
  Frame<R> f;
  f.m = m;
  f.s = s;
  f.n = cross3(s, m);
  return f;
}

// Sensor::check_conditions (detector.cpp:374-417).
template <class R>
__device__ __forceinline__ bool accepts(const SensorDev& sd, V3<R> hp, V3<R> dir, int cdir, int events) {
  if (sd.f_dir_en && sd.f_dir != LMC_DIR_BOTH && sd.f_dir != cdir) return false;
  if (sd.f_ev_en && (events < sd.f_ev_min || events > sd.f_ev_max)) return false;
  if (sd.f_th_en) {
    const R ct = -dir.z;
    if (ct < static_cast<R>(sd.cos_th_min) || ct > static_cast<R>(sd.cos_th_max)) return false;
  }
  if (sd.f_ph_en) {
    // the direction is invariant over a flight; without the opaque copy the
    // compiler evaluates this atan2 ahead of every flight, hit or not
    V3<R> d = dir;
    asm volatile("" : "+v"(d.x), "+v"(d.y));
    const R phi = exit_phi(d);
    if (phi < static_cast<R>(sd.ph_min) || phi > static_cast<R>(sd.ph_max)) return false;
  }
  if (sd.f_pos_en) {
    if (hp.x < static_cast<R>(sd.x_min) || hp.x > static_cast<R>(sd.x_max)) return false;
    if (hp.y < static_cast<R>(sd.y_min) || hp.y > static_cast<R>(sd.y_max)) return false;
  }
  return true;
}

// Grid cell of one accepted crossing, with the binning of the sensors'
// process_hit: time bin (detector.cpp:787-797, :1182-1190), exit angles
// (:1274-1281), pixel (:546-556, :800-806).  Returns false when the crossing
// falls outside the grid, where the reference deposits nothing and counts no
// hit.  t_idx: time bin >= 1 of a time-resolved grid, else 0 (the grid row a
// deposit goes to; row 0 of a time-resolved grid is completed at read-out).
// Without FMA contraction: the two binned passes of the sensor kernel must bin
// every crossing identically, and contraction decided by each inlining
// context could round a coordinate differently in each.
template <class R>
__device__ __forceinline__ bool sensor_bin(const SensorDev& sd, R hx, R hy, R arr, V3<R> dir, int64_t& bin,
                                           int& t_idx) {
#pragma clang fp contract(off)
  const int kind = sd.kind;
  t_idx = 0;  // time bin; 0 = steady state (the only row when dt == 0)
  bin = 0;
  if ((kind == LMC_SENSOR_FARFIELD_CBS || kind == LMC_SENSOR_PLANAR_FLUENCE) && sd.g[5] > 0) {
    const double a = static_cast<double>(arr);
    if (a < 0 || a >= sd.g[2]) return false;
    t_idx = static_cast<int>(a / sd.g[5]) + 1;
    if (t_idx >= sd.n2) return false;
  }
  if (kind == LMC_SENSOR_FARFIELD_CBS || kind == LMC_SENSOR_FARFIELD_FLUENCE) {
    const bool cbs = kind == LMC_SENSOR_FARFIELD_CBS;
    const int it = static_cast<int>(exit_theta(dir) / static_cast<R>(cbs ? sd.g[3] : sd.g[2]));
    const int jp = static_cast<int>(exit_phi(dir) / static_cast<R>(cbs ? sd.g[4] : sd.g[3]));
    if (it < 0 || it >= sd.n0 || jp < 0 || jp >= sd.n1) return false;
    bin = static_cast<int64_t>(it) * sd.n1 + jp;
    return true;
  }
  if (kind == LMC_SENSOR_PLANAR_FLUENCE || kind == LMC_SENSOR_PLANAR_FIELD) {
    const bool flu = kind == LMC_SENSOR_PLANAR_FLUENCE;
    const int xi = static_cast<int>((static_cast<double>(hx) + 0.5 * sd.g[0]) / (flu ? sd.g[3] : sd.g[2]));
    const int yi = static_cast<int>((static_cast<double>(hy) + 0.5 * sd.g[1]) / (flu ? sd.g[4] : sd.g[3]));
    if (xi < 0 || xi >= sd.n0 || yi < 0 || yi >= sd.n1) return false;
    bin = static_cast<int64_t>(xi) * sd.n1 + yi;
    return true;
  }
  return kind == LMC_SENSOR_REFLECTANCE;
}

// Transport kernels that deposit the light sensors' crossings themselves
// (direct_deposit): forward single-scene runs, where the deposit code sits
// outside the flight loop (kAux 0: exits recorded at the wave's refill), and
// forward runs where registers are not the bound (FP64 and xoshiro check
// runs, which thus take the same code path as the production kernels).
// Gradient runs queue every crossing, their deposits carry scores; the other
// FP32 Philox kernels queue them too (the deposit code would spill registers
// the event loop needs: the flight-loop plane code of kAux >= 1, the batched
// kernel's per-lane scene state).  The host reads the same function
// (KParams::direct).
constexpr bool deposits_direct(bool fp64, bool xo, bool adj, int aux, bool batch) {
  return !adj && (fp64 || xo || (aux == 0 && !batch));
}

// The light sensors' process_hit at the crossing, inside the transport kernel
// of forward runs: PlanarFluenceSensor (detector.cpp:780-841, lab x negated
// at :812), FarFieldFluenceSensor (the forward field's Stokes vector in the
// exit direction's bin) and ReflectanceSensor (the exit weight).  Same binning
// and arithmetic as the sensor kernel's; the grid deposits are FP64 atomics
// into HBM that the memory system completes while the kernel goes on
// scattering, so these crossings take no queue record and no sensor-kernel
// pass.  The reflectance weight and the hit counts, a few addresses every
// photon adds to, are summed in the workgroup's LDS (single scenes; added to
// HBM when the kernel ends) instead of by atomics on one HBM address each.
template <class R, bool kXo, bool kBatch>
__device__ __forceinline__ void direct_deposit(const KParams<R>& p, const SceneView<R>& v, const Photon<R, kXo>& ph,
                                               int si, V3<R> hp, R arr) {
  const SensorDev& sd = v.sens[si];
  int64_t bin;
  int t_idx;
  if (!sensor_bin(sd, hp.x, hp.y, arr, ph.P.s, bin, t_idx)) return;
  double* g = p.out + sd.out_off;
  double refl = 0.0;
  if (sd.kind == LMC_SENSOR_REFLECTANCE) {
    refl = static_cast<double>(ph.w);
  } else {
    const R ws = M<R>::sqrt(ph.w);
    const cx<R> fm = cscale(ph.Em, ws), fn = cscale(ph.En, ws);
    const cx<R> Efx = cadd(cscale(fm, ph.P.m.x), cscale(fn, ph.P.n.x));
    const cx<R> Efy = cadd(cscale(fm, ph.P.m.y), cscale(fn, ph.P.n.y));
    const bool flu = sd.kind == LMC_SENSOR_PLANAR_FLUENCE;
    const int64_t tplane = static_cast<int64_t>(sd.n0) * sd.n1;
    const int64_t plane = flu ? tplane * sd.n2 : tplane;
    const cx<R> Ex = flu ? cscale(Efx, R(-1)) : Efx;
    const int64_t cell = static_cast<int64_t>(t_idx) * tplane + bin;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double val = static_cast<double>(stokes(q, Ex, Efy));
      if (val != 0.0) unsafeAtomicAdd(g + q * plane + cell, val);
    }
  }
  if constexpr (kBatch) {  // a batch's many sensors (FP64 and xoshiro check runs only)
    if (refl != 0.0) unsafeAtomicAdd(g, refl);
    atomicAdd(p.hits + si, 1ull);
  } else {
    if (refl != 0.0) __hip_atomic_fetch_add((lds_double*)(v.lrefl + si), refl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    atomicAdd(&v.lhits[si], 1u);
  }
}

// A crossing of plane `pl` at hit point hp, corr (the overshoot from hp to
// the segment's end) back along the flight: the overshoot-corrected paths
// (detector.cpp:194-235); the sensors' filters decide acceptance and
// absorption here, the light sensors of forward runs deposit here
// (direct_deposit), the others' deposits happen in k_crossings from the
// queued record.  Returns true when an absorbing sensor accepted the photon.
template <class R, bool kXo, bool kAdj, int kAux, bool kBatch>
__device__ __forceinline__ bool plane_record(const KParams<R>& p, const SceneView<R>& v, Photon<R, kXo>& ph, int pl,
                                             V3<R> hp, R corr, int cdir) {
  {
    const uint64_t act = __ballot(1);
    if (__lane_id() == __ffsll(static_cast<unsigned long long>(act)) - 1)
      atomicAdd(&v.lhits[kMaxSensors], static_cast<unsigned>(__popcll(act)));
  }
  const R gpl_hit = corr > R(0) ? ph.gpl - corr : ph.gpl;
  const R arr = ph.t0 + gpl_hit;
  bool killed = false;
  int mask = 0;
  const int first = v.planes[pl].first, cnt = v.planes[pl].count;
  for (int k = 0; k < cnt; ++k) {
    const SensorDev& sd = v.sens[first + k];
    if (!accepts(sd, hp, ph.P.s, cdir, ph.events)) continue;
    if (sd.absorb) killed = true;
    if constexpr (deposits_direct(sizeof(R) == 8, kXo, kAdj, kAux, kBatch)) {
      if (sd.direct) {
        direct_deposit<R, kXo, kBatch>(p, v, ph, first + k, hp, arr);
        continue;
      }
    }
    // only crossings that can land in a sensor's grid are queued
    if (-ph.P.s.z >= sd.q_cmin && hp.x >= sd.q_x0 && hp.x <= sd.q_x1 && hp.y >= sd.q_y0 && hp.y <= sd.q_y1 &&
        arr >= sd.q_t0 && arr <= sd.q_t1)
      mask |= 1 << k;
  }
  if (mask == 0) return killed;
  // Queue slots come from a per-wave chunk of q_chunk consecutive slots, so
  // the global counter sees one atomic per chunk instead of one per crossing
  // event; the recording lanes get consecutive slots, so their records fill
  // neighbouring cache lines.  A wave's last chunk may stay partly unused: those
  // slots keep the plane sentinel -1 and the sensor kernel skips them.
  const uint64_t act = __ballot(1);
  const int leader = __ffsll(static_cast<unsigned long long>(act)) - 1;
  const int lane = __lane_id();
  const unsigned k = static_cast<unsigned>(__popcll(act));
  const unsigned rank = static_cast<unsigned>(__popcll(act & ((1ull << lane) - 1ull)));
  const int wid = static_cast<int>(threadIdx.x >> 6);
  unsigned long long b_old = 0, b_new = 0;
  unsigned room = 0;
  if (lane == leader) {
    const unsigned used = v.wq_used[wid];
    room = static_cast<unsigned>(p.q_chunk) - used;
    b_old = v.wq_base[wid] + used;
    if (k <= room) {
      v.wq_used[wid] = used + k;
    } else {
      // as many whole chunks as the lanes beyond the room need, in one
      // contiguous claim: more recording lanes than a chunk holds (16-slot
      // chunks of the no-alloc queue) stay within this wave's slots
      const unsigned need = k - room, q = static_cast<unsigned>(p.q_chunk);
      const unsigned nch = (need + q - 1) / q;
      b_new = atomicAdd(&p.q_count[0], static_cast<unsigned long long>(nch) * q);
      v.wq_base[wid] = b_new + static_cast<unsigned long long>(nch - 1) * q;
      v.wq_used[wid] = need - (nch - 1) * q;
    }
  }
  b_old = __shfl(b_old, leader);
  b_new = __shfl(b_new, leader);
  room = __shfl(room, leader);
  const unsigned long long slot = rank < room ? b_old + rank : b_new + (rank - room);
  if (static_cast<int64_t>(slot) >= p.q_cap) {
    atomicAdd(&p.q_count[1], 1ull);
    return killed;
  }
  const int64_t c = p.q_cap;
  R* f = p.qf + static_cast<int64_t>(slot) * (4 * kRecQ);
  double opl_hit = ph.opl;
  if (corr > R(0)) opl_hit -= static_cast<double>(corr) * layer_n64(v.L[ph.layer]);
  put4(f, 0, hp.x, hp.y, arr, ph.w);
  put4(f, 1, ph.Em.r, ph.Em.i, ph.En.r, ph.En.i);
  put4(f, 2, ph.P.m.x, ph.P.m.y, ph.P.m.z, ph.P.n.x);
  put4(f, 3, ph.P.n.y, ph.P.n.z, ph.P.s.x, ph.P.s.y);
  put4(f, 4, ph.P.s.z, ph.m1.x, ph.m1.y, ph.m1.z);
  p.qd[slot] = opl_hit;
  if (ph.events >= 2 && (v.planes[pl].flags & kPlaneCbs)) {
    // the frame before the last scatter, Po = A(theta, phi)^T P with the
    // last event's angles (A is orthogonal)
    const Frame<R>& P = ph.P;
    const R c = ph.qc, s = ph.qs, cp = ph.bc, sn = ph.bs;
    const R cc = c * cp, sc = s * cp;
    const V3<R> mq{cc * P.m.x - sn * P.n.x + sc * P.s.x, cc * P.m.y - sn * P.n.y + sc * P.s.y,
                   cc * P.m.z - sn * P.n.z + sc * P.s.z};
    const V3<R> sq{c * P.s.x - s * P.m.x, c * P.s.y - s * P.m.y, c * P.s.z - s * P.m.z};
    put4(f, 5, ph.s1.x, ph.s1.y, ph.s1.z, mq.x);
    put4(f, 6, mq.y, mq.z, sq.x, sq.y);
    put4(f, 7, sq.z, ph.T[0].r, ph.T[0].i, ph.T[1].r);
    put4(f, 8, ph.T[1].i, ph.T[2].r, ph.T[2].i, ph.T[3].r);
    put4(f, 9, ph.T[3].i, ph.rn.x - ph.r1.x, ph.rn.y - ph.r1.y, ph.rn.z - ph.r1.z);
  }
  if constexpr (kAux >= 1) {  // what a PhotonRecordSensor reads beyond the sensors' common record
    if (v.planes[pl].flags & kPlaneRec) {
      put4(f, 10, ph.r1.x, ph.r1.y, ph.r1.z, ph.t0);
      put4(f, 11, ph.rn.x, ph.rn.y, ph.rn.z, gpl_hit);
    }
  }
  if constexpr (kAdj) {
    R score[3 * kMaxGradLayers];
    adj_scores(p, v, ph, corr, score);
#pragma unroll
    for (int q = 0; q < 3; ++q)
      put4(f, kRecQScore + q, score[4 * q], score[4 * q + 1], score[4 * q + 2], score[4 * q + 3]);
  }
  int32_t* iq = p.qi + slot;
  iq[I_PLANE * c] = pl;
  iq[I_MASK * c] = mask;
  iq[I_EVENTS * c] = ph.events;
  iq[I_SP * c] = ph.sp;
  if constexpr (kAux >= 1) iq[I_DEPTH * c] = __float_as_int(static_cast<float>(ph.depth));
  return killed;
}

// A crossing of plane `pl` by the segment prev -> pos: hit point by linear
// interpolation in z (detector.cpp:194-212), then plane_record.
template <class R, bool kXo, bool kAdj, int kAux, bool kBatch>
__device__ __forceinline__ bool plane_hit(const KParams<R>& p, const SceneView<R>& v, Photon<R, kXo>& ph, int pl,
                                          V3<R> prev, int cdir) {
  const R zp = plane_z<R>(v.planes[pl]);
  const V3<R> d{ph.pos.x - prev.x, ph.pos.y - prev.y, ph.pos.z - prev.z};
  const R t = qdiv(zp - prev.z, d.z);
  const V3<R> hp{prev.x + d.x * t, prev.y + d.y * t, prev.z + d.z * t};
  const V3<R> ov{hp.x - ph.pos.x, hp.y - ph.pos.y, hp.z - ph.pos.z};
  return plane_record<R, kXo, kAdj, kAux, kBatch>(p, v, ph, pl, hp, M<R>::sqrt(dot(ov, ov)), cdir);
}

// SensorsGroup::record_hit (detector.cpp:178-270) for the segment prev->pos:
// planes within [min z, max z] in increasing z.  from_iface: the segment
// starts on the layer interface the photon has just crossed (FP32 runs place
// it exactly there; the reference, and the FP64 build, nudge it 1e-12 past),
// so a plane at that interface, already recorded by the sub-step that ended
// there, is not recorded again.  A single scene's planes are the same for
// every lane: the loop runs over them with a wave-uniform index (kernel
// argument bound, scalar loop), so the common no-crossing case costs two
// compares per plane.  A batch's lanes can sit in different scenes and take
// their plane range from their layer.
template <class R, bool kXo, bool kAdj, int kAux, bool kBatch>
__device__ __forceinline__ bool record_hit(const KParams<R>& p, const SceneView<R>& v, Photon<R, kXo>& ph, V3<R> prev,
                                           bool from_iface) {
  const R z1 = prev.z, z2 = ph.pos.z;
  if (M<R>::fabs(z2 - z1) < static_cast<R>(1e-12)) return false;
  const R zmin = z1 < z2 ? z1 : z2, zmax = z1 < z2 ? z2 : z1;
  // no plane of any scene of the launch within the segment, for every lane
  // of the wave: the common case, one compare pair and a uniform branch
  if (!__ballot(zmin <= p.pz_hi && zmax >= p.pz_lo)) return false;
  bool killed = false;
  int pl0, pl1;
  if constexpr (kBatch) {
    pl0 = v.L[ph.layer].pl_first;
    pl1 = pl0 + v.L[ph.layer].pl_count;
  } else {
    pl0 = 0;
    pl1 = p.n_planes;
  }
  for (int pl = pl0; pl < pl1; ++pl) {
    const R zp = plane_z<R>(v.planes[pl]);
    if (zp >= zmin && zp <= zmax && !(from_iface && zp == z1))
      killed |= plane_hit<R, kXo, kAdj, kAux, kBatch>(p, v, ph, pl, prev, (z2 > z1) ? LMC_DIR_FORWARD : LMC_DIR_BACKWARD);
  }
  return killed;
}

// optical_depth_to_detector (detector.cpp:53-101) for a ray leaving the
// vertex toward the detector plane above it: sum of mu_t over the layers it
// crosses, from the vertex's layer upward.
template <class R>
__device__ __forceinline__ R est_optical_depth(const SceneView<R>& v, int n_layers, int layer, V3<R> p, V3<R> dir,
                                               R L_total) {
  if (n_layers == 1) return v.L[layer].mu_t * L_total;
  R tau = R(0), traveled = R(0);
  const R dz = dir.z;
  const R eps = static_cast<R>(1e-15);
  for (int guard = 0; guard < 2 * kMaxLayers && traveled < L_total - eps; ++guard) {
    const R remaining = L_total - traveled;
    R seg = remaining;
    if (M<R>::fabs(dz) > eps) {
      // next interface between p.z and p.z + dz * remaining: the layer's top
      // face going up, its bottom face going down
      const R zt = p.z + dz * remaining;
      const bool up = dz < R(0);
      const bool has = up ? (layer > 0 && v.L[layer].z_min >= zt) : (v.L[layer].nbr & 1) && v.L[layer].z_max <= zt;
      if (has) {
        const R sd = qdiv((up ? v.L[layer].z_min : v.L[layer].z_max) - p.z, dz);
        if (sd > eps && sd < remaining) seg = sd;
      }
    }
    tau += v.L[layer].mu_t * seg;
    p.x += dir.x * seg;
    p.y += dir.y * seg;
    p.z += dir.z * seg;
    traveled += seg;
    if (seg >= remaining - eps) break;
    const int next = dz > R(0) ? layer + 1 : layer - 1;
    if (next < 0 || next >= n_layers) break;
    layer = next;
  }
  return tau;
}

// Lane l's value in every lane (a wave-uniform value the compiler keeps in
// scalar registers).
__device__ __forceinline__ float bcast(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double bcast(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ int bcast(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
template <class R> __device__ __forceinline__ V3<R> bcast(V3<R> a, int l) {
  return {bcast(a.x, l), bcast(a.y, l), bcast(a.z, l)};
}
template <class R> __device__ __forceinline__ cx<R> bcast(cx<R> a, int l) { return {bcast(a.r, l), bcast(a.i, l)}; }

// FarFieldCBSSensor::process_estimation (detector.cpp:1315-1566) for the
// vertices of a wave: every lane whose photon stands at a scattering vertex
// (species chosen, not yet scattered) gets its next scatter forced into every
// (theta, phi) bin of the sensor, with the weight w * albedo * exp(-tau) *
// F / (pi I_norm) * dOmega that lands there, the forward field of that
// scatter and, from the second order on, its reverse-path partner.  Each lane
// computes its vertex's reverse prefix (stages A, B) once; then the wave takes
// the vertices one at a time, broadcast to scalar registers, with the bins
// spread over the lanes (distinct bins per lane: no LDS add conflicts within a
// wave, and a long photon no longer runs its bins alone).  The common phase
// exp(i k L) of the forward and reverse fields cancels in every Stokes term
// and is not formed.
template <class R, bool kXo>
__device__ void estimate_wave(const KParams<R>& p, const SceneView<R>& v, const Photon<R, kXo>& ph, int si) {
  const SensorDev& sd = v.sens[si];
  const R tiny = static_cast<R>(sizeof(R) == 8 ? 1e-300 : 0.0);
  bool mine = ph.alive;
  if (mine && sd.f_ev_en && (ph.events < sd.f_ev_min || ph.events > sd.f_ev_max)) mine = false;
  const SpeciesDev<R>& S = v.S[mine ? ph.cur_sp : 0];
  const bool single = ph.events == 0;
  const R w_sc = ph.w * (single ? v.L[mine ? ph.layer : 0].albedo : S.alb);
  if (mine && (!(S.pi_inorm > R(0)) || w_sc < tiny)) mine = false;
  const LaunchConst<R>& lc = *v.lc;
  cx<R> pm{R(0), R(0)}, pn{R(0), R(0)};
  if (mine && !single) {
    cx<R> Tm[4] = {ph.T[0], ph.T[1], ph.T[2], ph.T[3]};
    if (ph.events >= 2) {  // interior product of the estimated path: J_k ... J_2
      const cx<R> r0 = cadd(cscale(ph.T[0], ph.bc), cscale(ph.T[2], ph.bs));
      const cx<R> r1 = cadd(cscale(ph.T[1], ph.bc), cscale(ph.T[3], ph.bs));
      const cx<R> r2 = cadd(cscale(ph.T[0], -ph.bs), cscale(ph.T[2], ph.bc));
      const cx<R> r3 = cadd(cscale(ph.T[1], -ph.bs), cscale(ph.T[3], ph.bc));
      Tm[0] = cmul(ph.b2, r0);
      Tm[1] = cmul(ph.b2, r1);
      Tm[2] = cmul(ph.b1, r2);
      Tm[3] = cmul(ph.b1, r3);
    }
    reverse_prefix(lc, v.stab, S, ph.P, Tm, pm, pn);
  }
  uint64_t todo = __ballot(mine);
  if (todo == 0) return;
  const int lane = __lane_id();
  if (lane == __ffsll(static_cast<unsigned long long>(todo)) - 1)
    atomicAdd(&v.lhits[si], static_cast<unsigned>(__popcll(todo)));  // hits += 1 per estimated vertex
  const V3<R> e1{lc.lm[0], lc.lm[1], lc.lm[2]}, e2{lc.ln[0], lc.ln[1], lc.ln[2]};
  const V3<R> s_in{lc.ldir[0], lc.ldir[1], lc.ldir[2]};
  const int Nth = sd.n0, Nph = sd.n1;
  // cos th, sin th, dOmega (Nth each), cos ph, sin ph (Nph each); then, from
  // the next multiple of 4, per bin (so.x, so.y, so.z, dOmega dphi) in FP32
  const R* gc = p.aux + sd.aux;
  const R dphi = static_cast<R>(sd.g[4]);
  const R tmax = static_cast<R>(sd.g[2]), dt = static_cast<R>(sd.g[5]);
  const R zp = static_cast<R>(v.planes[sd.plane].z);
  const int64_t tplane = static_cast<int64_t>(Nth) * Nph, plane = tplane * sd.n2;
  const bool in_lds = sd.est_lds >= 0;
  lds_double* lgs = in_lds ? (lds_double*)(v.lgrid + sd.est_lds) : nullptr;
  double* hg = p.out + sd.out_off;
  const int nb = sd.est_imax * Nph;
  const int so_off = (sd.aux + 3 * Nth + 2 * Nph + 3) & ~3;
  const int it_l = lane / Nph, jp_l = lane - it_l * Nph;  // this lane's first bin
  const int d_it = 64 / Nph, d_jp = 64 - d_it * Nph;
  while (todo) {
    const int l = __ffsll(static_cast<unsigned long long>(todo)) - 1;
    todo &= todo - 1;
    // lane l's vertex, wave-uniform
    const V3<R> pos = bcast(ph.pos, l), r1 = bcast(ph.r1, l);
    Frame<R> Pc;
    Pc.m = bcast(ph.P.m, l);
    Pc.n = bcast(ph.P.n, l);
    Pc.s = bcast(ph.P.s, l);
    const cx<R> Em = bcast(ph.Em, l), En = bcast(ph.En, l);
    const cx<R> xm = bcast(pm, l), xn = bcast(pn, l);
    const R w = bcast(w_sc, l), gpl = bcast(ph.gpl, l), t0 = bcast(ph.t0, l);
    const int vsp = bcast(ph.cur_sp, l), vlayer = bcast(ph.layer, l), vev = bcast(ph.events, l);
    const int vsp1 = bcast(ph.sp, l) >> 16;
    const bool vsingle = vev == 0;
    Frame<R> P1;
    if (!vsingle) P1 = frame_of(bcast(ph.m1, l), bcast(ph.s1, l));
    const SpeciesDev<R>& Sv = v.S[vsp];
    const SpeciesDev<R>& S1 = v.S[vsp1];
    const R Emm = cnorm(Em), Enn = cnorm(En), E12 = R(2) * (Em.r * En.r + Em.i * En.i);
    int it_n = it_l, jp_n = jp_l;
    for (int b = lane; b < nb; b += 64) {
      // bin b = it * Nph + jp; this lane's next bin is 64 further on
      const int it = it_n, jp = jp_n;
      it_n += d_it;
      jp_n += d_jp;
      if (jp_n >= Nph) {
        jp_n -= Nph;
        ++it_n;
      }
      V3<R> so;
      R dwt;  // dOmega dphi
      if constexpr (sizeof(R) == 4) {  // the bin's precomputed record
        const float4 g4 = reinterpret_cast<const float4*>(p.aux + so_off)[b];
        so = {g4.x, g4.y, g4.z};
        dwt = g4.w;
      } else {
        const R c_th = gc[it], s_th = gc[Nth + it], dom = gc[2 * Nth + it];
        const R c_ph = gc[3 * Nth + jp], s_ph = gc[3 * Nth + Nph + jp];
        const R a1 = s_th * c_ph, a2 = s_th * s_ph;
        so = {e1.x * a1 + e2.x * a2 - s_in.x * c_th, e1.y * a1 + e2.y * a2 - s_in.y * c_th,
              e1.z * a1 + e2.z * a2 - s_in.z * c_th};
        dwt = dom * dphi;
      }
      // intersect_plane (detector.cpp:36-47) with the sensor plane, normal +z
      if (M<R>::fabs(so.z) < static_cast<R>(1e-14)) continue;
      const R L = qdiv(zp - pos.z, so.z);
      if (!(L > R(0))) continue;
      const R Tr = M<R>::exp(-est_optical_depth(v, p.n_layers, vlayer, pos, so, L));
      if (Tr < static_cast<R>(1e-20)) continue;
      R sin_phi, cos_phi, th, ct_v = R(0), st_v = R(0);
      if constexpr (sizeof(R) == 4) {
        // FP32: both angles from so's components in the vertex frame, x = m.so,
        // y = n.so: |s x so| = |(x, y)|, and with m x n = h s the reference's
        // -m.unit(s x so) and n.unit(s x so) are h y / |(x, y)|, h x / |(x, y)|
        const R x = dot(Pc.m, so), y = dot(Pc.n, so);
        const R p2 = x * x + y * y;
        if (!(p2 > R(0))) continue;
        const R ip = M<R>::rsqrt(p2);
        cos_phi = lc.hand * x * ip;
        sin_phi = lc.hand * y * ip;
        ct_v = clamp1(dot(Pc.s, so));
        st_v = p2 * ip;
        th = M<R>::acos(ct_v);
      } else {
        const V3<R> pv = cross_unit(Pc.s, so);
        const R pnrm = M<R>::sqrt(dot(pv, pv));
        if (pnrm < static_cast<R>(1e-12)) continue;
        const V3<R> ph_hat = vscale(pv, qdiv(R(1), pnrm));
        sin_phi = -dot(Pc.m, ph_hat);
        cos_phi = dot(Pc.n, ph_hat);
        th = M<R>::acos(clamp1(dot(Pc.s, so)));
      }
      cx<R> s1, s2;
      s_lookup(v.stab, Sv, th, s1, s2);
      // phase_F (detector.cpp:1012-1023)
      const R s22 = cnorm(s2), s11 = cnorm(s1);
      const R cc = cos_phi * cos_phi, ss = sin_phi * sin_phi;
      const R F = Emm * (s22 * cc + s11 * ss) + Enn * (s22 * ss + s11 * cc) + E12 * (s22 - s11) * sin_phi * cos_phi;
      if (F < tiny) continue;
      const R w_bin = w * Tr * (qdiv(F, Sv.pi_inorm) * dwt);
      if (w_bin < tiny) continue;
      int t_idx = 0;  // time_bin (detector.cpp:1182-1190); photon speed 1 in path units
      if (dt > R(0)) {
        const R arr = t0 + (gpl + L);
        if (arr < R(0) || arr >= tmax) continue;
        t_idx = static_cast<int>(qdiv(arr, dt)) + 1;
        if (t_idx >= sd.n2) continue;
      }
      const R amp = M<R>::sqrt(w_bin);
      // exit frame (scatter_frame, detector.cpp:1027-1044) and the normalised
      // forward field (apply_scatter_normalized, :121-161)
      R st = st_v, ct = ct_v;
      if constexpr (sizeof(R) == 8) M<R>::sincos(th, &st, &ct);
      Frame<R> Pe;
      {
        const R ctc = ct * cos_phi, cts = ct * sin_phi, stc = st * cos_phi, sts = st * sin_phi;
        Pe.m = {ctc * Pc.m.x + cts * Pc.n.x - st * Pc.s.x, ctc * Pc.m.y + cts * Pc.n.y - st * Pc.s.y,
                ctc * Pc.m.z + cts * Pc.n.z - st * Pc.s.z};
        Pe.n = {-sin_phi * Pc.m.x + cos_phi * Pc.n.x, -sin_phi * Pc.m.y + cos_phi * Pc.n.y,
                -sin_phi * Pc.m.z + cos_phi * Pc.n.z};
        // the rotated direction is the forced one, so (FP32) so itself
        if constexpr (sizeof(R) == 4)
          Pe.s = so;
        else
          Pe.s = {stc * Pc.m.x + sts * Pc.n.x + ct * Pc.s.x, stc * Pc.m.y + sts * Pc.n.y + ct * Pc.s.y,
                  stc * Pc.m.z + sts * Pc.n.z + ct * Pc.s.z};
      }
      const R fi = qdiv(amp, M<R>::sqrt(F));
      const cx<R> um = cadd(cscale(Em, cos_phi), cscale(En, sin_phi));
      const cx<R> un = cadd(cscale(Em, -sin_phi), cscale(En, cos_phi));
      const cx<R> fm = cscale(cmul(s2, um), fi), fn = cscale(cmul(s1, un), fi);
      const cx<R> Efx = cadd(cscale(fm, Pe.m.x), cscale(fn, Pe.n.x));
      const cx<R> Efy = cadd(cscale(fm, Pe.m.y), cscale(fn, Pe.n.y));
      cx<R> Erx{R(0), R(0)}, Ery{R(0), R(0)};
      if (!vsingle) {
        cx<R> rm, rn;
        reverse_suffix(v.stab, S1, P1, Pe, lc.hand, xm, xn, rm, rn);
        // CBS path phase exp(i k (s_out + s_in) . (r_vertex - r_1))
        const R arg = lc.k0 * ((so.x + s_in.x) * (pos.x - r1.x) + (so.y + s_in.y) * (pos.y - r1.y) +
                               (so.z + s_in.z) * (pos.z - r1.z));
        R sa, ca;
        M<R>::sincos_phase(arg, &sa, &ca);
        const cx<R> pp{ca * amp, sa * amp};
        rm = cmul(rm, pp);
        rn = cmul(rn, pp);
        Erx = cadd(cscale(rm, Pe.m.x), cscale(rn, Pe.n.x));
        Ery = cadd(cscale(rm, Pe.m.y), cscale(rn, Pe.n.y));
      }
      const cx<R> Ex = cadd(Efx, Erx), Ey = cadd(Efy, Ery);
      R val[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        val[q] = q < 4 ? stokes(q, Ex, Ey) : stokes(q - 4, Efx, Efy) + stokes(q - 4, Erx, Ery);
      if (in_lds) {  // wave-uniform; LDS offsets in 32 bits
        lds_double* a = lgs + (static_cast<int>(t_idx) * static_cast<int>(tplane) + b);
#pragma unroll
        for (int q = 0; q < 8; ++q)  // no zero test: LDS adds are cheap and the values are almost never zero
          __hip_atomic_fetch_add(a + q * static_cast<int>(plane), static_cast<double>(val[q]), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
      } else {
        double* a = hg + (t_idx * tplane + b);
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (val[q] != R(0)) unsafeAtomicAdd(a + q * plane, static_cast<double>(val[q]));
      }
    }
  }
}

// Laser::emit_photon (laser.cpp:77-99) and sample_emission_time (:42-75).
// The production path draws the emission uniforms from the photon's stream.
template <class R, bool kXo>
__device__ __forceinline__ void launch(const KParams<R>& p, const LaunchConst<R>& L, Photon<R, kXo>& ph, uint32_t idx) {
  if constexpr (!kXo) ph.rng = seed_stream(p.key0, p.key1, idx, p.idx_hi);
  R x = static_cast<R>(L.lpos_x), y = static_cast<R>(L.lpos_y);
  R t0 = static_cast<R>(L.t_off * 1e-3);
  const R pulse = static_cast<R>(L.pulse * 1e-3);
  const R fwhm_to_sigma = static_cast<R>(1.0 / (2.0 * 1.1774100225154747));  // 1/(2 sqrt(2 ln 2))
  if constexpr (kXo) {
    ph.xo.init((static_cast<uint64_t>(p.key1) << 32) | p.key0, (static_cast<uint64_t>(p.idx_hi) << 32) | idx);
    double xd = L.lpos_x, yd = L.lpos_y, td = L.t_off * 1e-3;
    const double kTwoPi = 2.0 * 3.14159265358979323846;
    if (L.source == LMC_SOURCE_UNIFORM) {
      const double th = ph.xo.s.uniform() * kTwoPi;
      const double r = ::sqrt(ph.xo.s.uniform()) * L.sigma;
      xd += r * ::cos(th);
      yd += r * ::sin(th);
    } else if (L.source == LMC_SOURCE_GAUSSIAN) {
      xd = ph.xo.s.normal(xd, L.sigma);
      yd = ph.xo.s.normal(yd, L.sigma);
    }
    switch (L.temporal) {
      case LMC_TIME_GAUSSIAN: td = ph.xo.s.normal(td, L.pulse * 1e-3 / (2.0 * 1.1774100225154747)); break;
      case LMC_TIME_TOPHAT: td = td + (ph.xo.s.uniform() - 0.5) * L.pulse * 1e-3; break;
      case LMC_TIME_EXPONENTIAL: td = td - ::log(ph.xo.s.uniform()) * L.pulse * 1e-3; break;
      case LMC_TIME_PULSE_TRAIN: if (L.rep > 0.0) td = td + ph.xo.s.uniform() * (1e9 / L.rep); break;
      case LMC_TIME_CW: if (L.pulse > 0.0) td = td + ph.xo.s.uniform() * L.pulse * 1e-3; break;
      default: break;
    }
    x = static_cast<R>(xd);
    y = static_cast<R>(yd);
    t0 = static_cast<R>(td);
  } else {
    if (L.source != LMC_SOURCE_POINT) {
      uint4 b;
      b.x = ph.rng.next();
      b.y = ph.rng.next();
      b.z = ph.rng.next();
      b.w = ph.rng.next();
      const R sg = static_cast<R>(L.sigma);
      if (L.source == LMC_SOURCE_UNIFORM) {
        R sn, cs;
        M<R>::sincos_rev(u01<R>(b.x), &sn, &cs);
        const R r = M<R>::sqrt(u01<R>(b.y)) * sg;
        x += r * cs;
        y += r * sn;
      } else {  // Box-Muller, cosine branch for each axis (rng.cpp:18-28)
        R sn, cs;
        M<R>::sincos_rev(u01<R>(b.y), &sn, &cs);
        x += M<R>::sqrt(R(-2) * M<R>::log(u01<R>(b.x))) * cs * sg;
        M<R>::sincos_rev(u01<R>(b.w), &sn, &cs);
        y += M<R>::sqrt(R(-2) * M<R>::log(u01<R>(b.z))) * cs * sg;
      }
    }
    if (L.temporal != LMC_TIME_DELTA) {
      uint2 b;
      b.x = ph.rng.next();
      b.y = ph.rng.next();
      const R ua = u01<R>(b.x);
      switch (L.temporal) {
        case LMC_TIME_GAUSSIAN: {
          R sn, cs;
          M<R>::sincos_rev(u01<R>(b.y), &sn, &cs);
          t0 += M<R>::sqrt(R(-2) * M<R>::log(ua)) * cs * pulse * fwhm_to_sigma;
          break;
        }
        case LMC_TIME_TOPHAT: t0 += (ua - R(0.5)) * pulse; break;
        case LMC_TIME_EXPONENTIAL: t0 -= M<R>::log(ua) * pulse; break;
        case LMC_TIME_PULSE_TRAIN: if (L.rep > 0.0) t0 += ua * static_cast<R>(1e9 / L.rep); break;
        case LMC_TIME_CW: if (L.pulse > 0.0) t0 += ua * pulse; break;
        default: break;
      }
    }
  }
  ph.pos = {x, y, L.lpos_z};
  ph.P.m = {L.lm[0], L.lm[1], L.lm[2]};
  ph.P.n = {L.ln[0], L.ln[1], L.ln[2]};
  ph.P.s = {L.ldir[0], L.ldir[1], L.ldir[2]};
  ph.Em = {L.pol[0], L.pol[1]};
  ph.En = {L.pol[2], L.pol[3]};
  ph.w = R(1);
  ph.depth = R(0);
  ph.gpl = R(0);
  ph.t0 = t0;
  ph.opl = 0.0;
  ph.events = 0;
  ph.layer = L.launch_layer;
  ph.alive = true;
  ph.pend = 0;
  ph.m1 = ph.P.m;
  ph.s1 = ph.P.s;
  ph.qc = R(1);
  ph.qs = R(0);
  ph.r1 = ph.pos;
  ph.rn = ph.pos;
  ph.T[0] = {R(1), R(0)};
  ph.T[1] = {R(0), R(0)};
  ph.T[2] = {R(0), R(0)};
  ph.T[3] = {R(1), R(0)};
  ph.b2 = {R(1), R(0)};
  ph.b1 = {R(1), R(0)};
  ph.bc = R(1);
  ph.bs = R(0);
  ph.sp = 0;
#pragma unroll
  for (int l = 0; l < kMaxGradLayers; ++l) {
    ph.aL[l] = R(0);
    ph.aN[l] = R(0);
    ph.aG[l] = R(0);
  }
  ph.cL = R(0);
  ph.cN = R(0);
  ph.cG = R(0);
}

template <class R, bool kXo, bool kAdj> __device__ __forceinline__ void add_path(Photon<R, kXo>& ph, R d) {
  if constexpr (kAdj) ph.cL += d;
}

// the current layer's score shares into its slots, before the photon leaves it
template <class R, bool kXo, bool kAdj> __device__ __forceinline__ void flush_scores(Photon<R, kXo>& ph) {
  if constexpr (kAdj) {
#pragma unroll
    for (int l = 0; l < kMaxGradLayers; ++l) {
      if (l == ph.layer) {
        ph.aL[l] += ph.cL;
        ph.aN[l] += ph.cN;
        ph.aG[l] += ph.cG;
      }
    }
    ph.cL = R(0);
    ph.cN = R(0);
    ph.cG = R(0);
  }
}

// One iteration of run_photon's loop (simulation.cpp:297-614).  Returns the
// number of azimuth trials used.
template <class R, bool kXo, bool kTrack, bool kAdj, int kPart, int kAux = 1, bool kBatch = false>
__device__ __forceinline__ int transport_event(const KParams<R>& p, const SceneView<R>& v, Photon<R, kXo>& ph,
                                               bool renorm = true) {
  // kPart 0: the whole event; 1: the flight up to the species of the next
  // vertex (left in ph.cur_sp); 2: the scatter at that vertex.  Estimator
  // runs split the event so the whole wave can work on each vertex in between.
  int sp;
  if constexpr (kPart == 2) {
    sp = ph.cur_sp;
  } else {
    R u_fp;
    if constexpr (kXo) {
      u_fp = static_cast<R>(ph.xo.s.uniform());
    } else {
      uint4 b;
      b.x = ph.rng.next();
      b.y = ph.rng.next();
      b.z = ph.rng.next();
      b.w = ph.rng.next();
      u_fp = u01<R>(b.x);
      ph.w_th = (b.y & 0xFFFFFF00u) | (b.x & 0xFFu);  // theta: high 24 bits; low byte: aux bits 16-23
      ph.w_ph = b.z;                                   // azimuth component: bits 8-31; low byte: aux bits 0-7
      ph.w_az = (b.w & 0xFFFFFF00u) | (b.y & 0xFFu);  // azimuth angle: bits 8-31; low byte: aux bits 8-15
    }
    R step = -v.L[ph.layer].mfp * M<R>::log(u_fp);
    const V3<R> dir = ph.P.s;
    bool on_iface = false;  // FP32: the photon stands exactly on the interface it has just crossed
    // The flight in segments (simulation.cpp:317-407): a segment ends at the
    // current layer's face when the remaining step crosses it (the remaining
    // optical depth then carries into the next layer, :336,373) and otherwise
    // at the step's end; each segment's sensor-plane crossings are recorded
    // (one record_hit per segment, a single copy of the plane code).  The layer
    // index is tracked explicitly (the reference nudges z by 1e-12 past the
    // interface and re-derives it; the FP64 build nudges too).
    for (;;) {
      const LayerDev<R>& Lc = v.L[ph.layer];
      const R new_z = ph.pos.z + dir.z * step;
      // the face the step heads for (z_max going deeper, z_min going back)
      // and whether it is crossed, as selects: no branch per lane
      const bool fwd = new_z > ph.pos.z, bwd = new_z < ph.pos.z;
      const R zi = fwd ? Lc.z_max : Lc.z_min;
      const int nb = Lc.nbr;
      const R dzt = dir.z * step;
      const R t = qdiv(zi - ph.pos.z, dzt);
      const bool has = ((fwd && (nb & 1) && zi <= new_z) || (bwd && (nb & 2) && zi >= new_z)) &&
                       !(M<R>::fabs(dzt) < static_cast<R>(1e-15)) && !(t < R(0) || t > R(1));
      const R seg = has ? step * t : step;
      ph.gpl += seg;
      ph.opl += layer_n64(Lc) * static_cast<double>(seg);
      add_path<R, kXo, kAdj>(ph, seg);
      const V3<R> prev = ph.pos;
      ph.pos.x += dir.x * seg;
      ph.pos.y += dir.y * seg;
      const bool up = dir.z > R(0);
      R z_iface = zi;
      if constexpr (sizeof(R) == 8) z_iface = zi + (up ? 1e-12 : -1e-12);
      ph.pos.z = has ? z_iface : ph.pos.z + dir.z * seg;
      if constexpr (kAux >= 1) {  // else 0 in every record: no StatisticsSensor reads it
        ph.depth = (!has && ph.pos.z > ph.depth) ? ph.pos.z : ph.depth;
      }
      if constexpr (sizeof(R) == 4) {
        // Leaving the sample (Sample::is_inside, sample.cpp:268-288): the
        // photon dies here whatever its crossings do, so they are recorded
        // later, together with the other lanes' (a wave runs the plane code
        // once for all its exits instead of once per exit).
        if (!has && (!(ph.pos.z >= R(0)) || ph.pos.z >= Lc.z_top || !isfinite(ph.pos.x) || !isfinite(ph.pos.y) ||
                     !isfinite(ph.pos.z))) {
          ph.w_th = __float_as_uint(prev.x);
          ph.w_ph = __float_as_uint(prev.y);
          ph.w_az = __float_as_uint(prev.z);
          ph.pend = on_iface ? 2 : 1;
          ph.alive = false;
          return 0;
        }
      }
      if constexpr (kAux != 0) {  // kAux == 0: no plane a flight inside the sample can cross
        if (record_hit<R, kXo, kAdj, kAux, kBatch>(p, v, ph, prev, on_iface)) {
          ph.alive = false;
          return 0;
        }
      }
      if (!has) break;
      const R rem = (step - seg) * Lc.mu_t;
      on_iface = sizeof(R) == 4;
      flush_scores<R, kXo, kAdj>(ph);
      ph.layer += up ? 1 : -1;
      const R mt = v.L[ph.layer].mu_t;
      step = mt > R(0) ? qdiv(rem, mt) : rem;
    }
    // Sample::is_inside (sample.cpp:268-288)
    if (!(ph.pos.z >= R(0)) || ph.pos.z >= v.L[ph.layer].z_top || !isfinite(ph.pos.x) ||
        !isfinite(ph.pos.y) || !isfinite(ph.pos.z)) {
      ph.alive = false;
      return 0;
    }
    // species of this vertex (MixtureLayer::select_scatter_medium, sample.cpp:96-104)
    const LayerDev<R>& Lv = v.L[ph.layer];
    sp = Lv.sp_first;
    if (Lv.sp_count > 1) {
      R u;
      if constexpr (kXo)
        u = static_cast<R>(ph.xo.s.uniform());
      else
        u = aux_uniform<R>(ph);
      // lower_bound of u in the selection CDF (std::lower_bound in the
      // reference), narrowed by the layer's guide g[j] = lower_bound(j / G)
      // stored after the CDF (g[G] = K): the bracket [g[j], g[j+1]] holds the
      // answer, usually one or two nodes, so a 32-species mixture takes one
      // search round instead of five dependent loads
      const R* c = v.sel + Lv.sel_off;
      const int K = Lv.sp_count;
      const uint16_t* gd = reinterpret_cast<const uint16_t*>(c + K);
      int j = static_cast<int>(u * static_cast<R>(kSelGuide));
      j = j < 0 ? 0 : (j >= kSelGuide ? kSelGuide - 1 : j);
      int lo = gd[j], hi = gd[j + 1];
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (c[mid] < u)
          lo = mid + 1;
        else
          hi = mid;
      }
      sp += lo < K ? lo : K - 1;
    }
    if constexpr (kPart == 1) {
      ph.cur_sp = sp;
      return 0;
    }
  }
  const LayerDev<R>& Ld = v.L[ph.layer];
  const SpeciesDev<R>& S = v.S[sp];
  if constexpr (kTrack) {
    // the pending J_{n-1} of the last event is committed before this event's
    // sampling (its factors are no longer needed once the new ones exist)
    const int evt = ph.events + 1;
    if (evt >= 3) {
      // commit the pending J_{n-1} = diag(b2, b1) [[bc, bs], [-bs, bc]]:  T <- J T, Frobenius-normalised.
      // Only T's direction is used (the reverse field is normalised), so the
      // FP32 Philox path normalises every other wave iteration (renorm, a
      // wave-uniform flag): between normalisations |T| stays within one
      // factor max|S1|,|S2| of one.  The check modes normalise every event.
      const cx<R> r0 = cadd(cscale(ph.T[0], ph.bc), cscale(ph.T[2], ph.bs));
      const cx<R> r1 = cadd(cscale(ph.T[1], ph.bc), cscale(ph.T[3], ph.bs));
      const cx<R> r2 = cadd(cscale(ph.T[0], -ph.bs), cscale(ph.T[2], ph.bc));
      const cx<R> r3 = cadd(cscale(ph.T[1], -ph.bs), cscale(ph.T[3], ph.bc));
      const cx<R> t0 = cmul(ph.b2, r0), t1 = cmul(ph.b2, r1), t2 = cmul(ph.b1, r2), t3 = cmul(ph.b1, r3);
      if (kXo || sizeof(R) == 8 || renorm) {
        const R fr2 = cnorm(t0) + cnorm(t1) + cnorm(t2) + cnorm(t3);
        const R inv = fr2 > R(0) ? M<R>::rsqrt(fr2) : R(1);
        ph.T[0] = cscale(t0, inv);
        ph.T[1] = cscale(t1, inv);
        ph.T[2] = cscale(t2, inv);
        ph.T[3] = cscale(t3, inv);
      } else {
        ph.T[0] = t0;
        ph.T[1] = t1;
        ph.T[2] = t2;
        ph.T[3] = t3;
      }
    }
  }
  // theta (medium.cpp:51-58 -> phase.cpp)
  R uth;
  if constexpr (kXo)
    uth = static_cast<R>(ph.xo.s.uniform());
  else
    uth = u01<R>(ph.w_th);
  R theta, st, ct;
  const int kind = S.kind;
  if constexpr (sizeof(R) == 4) {
    // One divergent branch only (the table search of the table kinds); the
    // closed forms are selects over the whole wave, so a wave whose lanes sit
    // in layers of different phase kinds pays no extra branch per kind.
    // Cosine of the uniform (phase.cpp:78-83) and HG (:108-115) kinds:
    R c = R(2) * uth - R(1);
    if (__ballot(kind == LMC_PHASE_HG)) {  // wave-uniform
      const bool hg = kind == LMC_PHASE_HG;
      const R g = hg ? S.hg_g : R(0.5), gg = g * g;
      const R frac = qdiv(R(1) - gg, R(1) - g + R(2) * g * uth);
      const R ch = qdiv(R(1) + gg - frac * frac, R(2) * g);
      c = hg ? ch : c;
      if constexpr (kAdj) {
        // d ln p_HG / dg at the sampled cosine (hardware reciprocals in FP32)
        const R cc = clamp1(ch);
        const R sg = -R(2) * qdiv(g, R(1) - gg) - R(3) * qdiv(g - cc, R(1) + gg - R(2) * g * cc);
        if (hg) ph.cG += sg;
      }
    }
    R t = R(0);
    if (kind == LMC_PHASE_TABLE_THETA || kind == LMC_PHASE_TABLE_COS) t = table_sample(v, S, uth);
    const bool direct = kind == LMC_PHASE_TABLE_THETA;
    c = clamp1(kind == LMC_PHASE_TABLE_COS ? t : c);
    theta = t;
    if (__ballot(!direct)) {  // wave-uniform
      const R tc = M<R>::acos(c);
      theta = direct ? t : tc;
    }
    M<R>::sincos(theta, &st, &ct);
  } else {
    if (kind == LMC_PHASE_HG) {  // phase.cpp:108-115
      const R g = S.hg_g, gg = g * g;
      const R frac = qdiv(R(1) - gg, R(1) - g + R(2) * g * uth);
      R c = qdiv(R(1) + gg - frac * frac, R(2) * g);
      theta = M<R>::acos(c);
      if constexpr (kAdj) ph.cG += -R(2) * g / (R(1) - gg) - R(3) * (g - c) / (R(1) + gg - R(2) * g * c);
    } else if (kind == LMC_PHASE_TABLE_THETA) {
      theta = table_sample(v, S, uth);
    } else if (kind == LMC_PHASE_TABLE_COS) {
      theta = M<R>::acos(table_sample(v, S, uth));
    } else {
      theta = M<R>::acos(R(2) * uth - R(1));
    }
    M<R>::sincos(theta, &st, &ct);
  }
  cx<R> s1, s2;
  s_lookup(v.stab, S, theta, s1, s2);
  // azimuth by rejection against F(phi) = a cos^2 + b sin^2 + c sin cos (phase.cpp:27-49)
  const R s2sq = cnorm(s2), s1sq = cnorm(s1);
  const R e1sq = cnorm(ph.Em), e2sq = cnorm(ph.En);
  const R e12 = ph.Em.r * ph.En.r + ph.Em.i * ph.En.i;
  const R a = s2sq * e1sq + s1sq * e2sq;
  const R b = s1sq * e1sq + s2sq * e2sq;
  const R c = R(2) * (s2sq - s1sq) * e12;
  R cph = R(1), sph = R(0), F = R(0);
  int trials = 0;
  if constexpr (kXo) {
    // the reference's rejection sampler in its draw order
    const R diff = a - b;
    const R Fmax = R(0.5) * (a + b) + R(0.5) * M<R>::sqrt(diff * diff + c * c);
    while (trials < 256) {
      const R pa = static_cast<R>(ph.xo.s.uniform());
      const R aa = static_cast<R>(ph.xo.s.uniform());
      R sp_, cp_;
      M<R>::sincos_rev(pa, &sp_, &cp_);
      const R Ft = a * cp_ * cp_ + b * sp_ * sp_ + c * cp_ * sp_;
      ++trials;
      if (aa * Fmax <= (Ft > R(0) ? Ft : R(0))) {
        cph = cp_;
        sph = sp_;
        F = Ft;
        break;
      }
    }
  } else {
    azimuth_sample(a, b, c, ph.w_ph, ph.w_az, cph, sph, F);
    trials = 1;
  }
  // frame update P <- A(theta, phi) P (simulation.cpp:438-453)
  const Frame<R> Po = ph.P;
  if constexpr (sizeof(R) == 4) {
    // FP32: through u = cos phi m + sin phi n, the azimuthal rotation's new
    // m before the polar tilt (24 operations instead of 28)
    const V3<R> u{cph * Po.m.x + sph * Po.n.x, cph * Po.m.y + sph * Po.n.y, cph * Po.m.z + sph * Po.n.z};
    ph.P.m = {ct * u.x - st * Po.s.x, ct * u.y - st * Po.s.y, ct * u.z - st * Po.s.z};
    ph.P.n = {-sph * Po.m.x + cph * Po.n.x, -sph * Po.m.y + cph * Po.n.y, -sph * Po.m.z + cph * Po.n.z};
    ph.P.s = {st * u.x + ct * Po.s.x, st * u.y + ct * Po.s.y, st * u.z + ct * Po.s.z};
  } else {
    const R ctc = ct * cph, cts = ct * sph, stc = st * cph, sts = st * sph;
    ph.P.m = {ctc * Po.m.x + cts * Po.n.x - st * Po.s.x, ctc * Po.m.y + cts * Po.n.y - st * Po.s.y,
              ctc * Po.m.z + cts * Po.n.z - st * Po.s.z};
    ph.P.n = {-sph * Po.m.x + cph * Po.n.x, -sph * Po.m.y + cph * Po.n.y, -sph * Po.m.z + cph * Po.n.z};
    ph.P.s = {stc * Po.m.x + sts * Po.n.x + ct * Po.s.x, stc * Po.m.y + sts * Po.n.y + ct * Po.s.y,
              stc * Po.m.z + sts * Po.n.z + ct * Po.s.z};
  }
  // Jones update E <- S R(phi) E / sqrt(F) (simulation.cpp:456-492); F is the
  // accepted trial's value of the same quadratic form.
  {
    const R inv = M<R>::rsqrt(F);
    const cx<R> um = cadd(cscale(ph.Em, cph), cscale(ph.En, sph));
    const cx<R> un = cadd(cscale(ph.Em, -sph), cscale(ph.En, cph));
    ph.Em = cscale(cmul(s2, um), inv);
    ph.En = cscale(cmul(s1, un), inv);
  }
  if constexpr (kTrack) {  // CBS bookkeeping (simulation.cpp:495-567)
    const int evt = ph.events + 1;
    // a first scatter happens in few waves' iterations (after refills): a
    // wave-uniform branch instead of ten selects in every iteration
    if (__ballot(evt == 1)) {
      if (evt == 1) {
        // an empty volatile asm keeps this block a branch: speculated, it
        // becomes ten selects in every event
        asm volatile("");
        ph.r1 = ph.pos;
        ph.m1 = ph.P.m;
        ph.s1 = ph.P.s;
        ph.sp = (sp << 16) | sp;
      }
    }
    ph.rn = ph.pos;
    ph.sp = (ph.sp & 0xFFFF0000) | sp;
    ph.qc = ct;
    ph.qs = st;
    ph.b2 = s2;
    ph.b1 = s1;
    ph.bc = cph;
    ph.bs = sph;
  }
  if constexpr (kAdj) ph.cN += R(1);
  // weight, roulette, cap (simulation.cpp:578-613)
  ph.events++;
  const R w_in = ph.w;
  if (Ld.mu_t > R(0)) ph.w *= Ld.albedo;
  if constexpr (kAux == 2) {  // Absorption::record_absorption (absortion.cpp:47-77) of w * mu_a / mu_t
    // a single scene's grid from the kernel arguments; a batch's per scene
    AbsDev<R> A;
    if constexpr (kBatch) {
      A = p.abs_scenes[Ld.scene];
    } else {
      A = {p.abs_radius, p.abs_depth, p.abs_dr, p.abs_dz, p.abs_dt, p.abs_nr, p.abs_nz, p.abs_nt, 1, p.abs_off};
    }
    const R d_w = Ld.mu_t > R(0) ? w_in * (Ld.mu_a / Ld.mu_t) : R(1);
    const R r = M<R>::sqrt(ph.pos.x * ph.pos.x + ph.pos.y * ph.pos.y), z = ph.pos.z;
    if (A.on && !(r > A.radius) && z >= R(0) && !(z > A.depth)) {
      const int ir = static_cast<int>(r / A.dr), iz = static_cast<int>(z / A.dz);
      if (ir < A.nr && iz < A.nz) {
        double* g = p.out + A.off;
        const int64_t cell = static_cast<int64_t>(ir) * A.nz + iz;
        unsafeAtomicAdd(g + cell, static_cast<double>(d_w));
        if (A.dt > R(0)) {
          const int k = static_cast<int>(M<R>::floor((ph.t0 + ph.gpl) / A.dt)) + 1;
          if (k >= 1 && k < A.nt)
            unsafeAtomicAdd(g + static_cast<int64_t>(k) * A.nr * A.nz + cell, static_cast<double>(d_w));
        }
      }
    }
  }
  if (ph.w < static_cast<R>(1e-4)) {
    R ur;
    if constexpr (kXo) {
      ur = static_cast<R>(ph.xo.s.uniform());
    } else if (Ld.sp_count > 1) {
      // the auxiliary draw chose the species; where it fell inside that
      // species' CDF interval is uniform again (24 bits times its probability)
      const R* c = v.sel + Ld.sel_off;
      const int k = sp - Ld.sp_first;
      const R lo = k > 0 ? c[k - 1] : R(0);
      ur = qdiv(aux_uniform<R>(ph) - lo, c[k] - lo);
    } else {
      ur = aux_uniform<R>(ph);
    }
    if (ur < static_cast<R>(0.1))
      ph.w = sizeof(R) == 4 ? ph.w * R(10) : ph.w / static_cast<R>(0.1);  // FP32: w * 10 is w / 0.1 to 1e-17
    else
      ph.alive = false;
  }
  if (ph.events > p.max_events) ph.alive = false;
  return trials;
}

__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// The transport kernel.  kBatch: the photons of one launch belong to several
// scenes, concatenated in photon-index order (scene i owns global indices
// [prefix[i], prefix[i+1])); layers, species, sensors and tables of all
// scenes then stay in global memory (cache resident: the index order keeps
// only a few scenes in flight at a time) and each photon launches with its
// own scene's laser.  Otherwise the scene constants are staged in LDS.
// Fixed LDS layout of the transport kernel's scene header: compile-time
// offsets from the LDS base, so the event loop addresses the staged layers,
// species, sensors and planes with immediate offsets instead of keeping their
// addresses in scalar registers.  Staged tables (when they fit) follow `end`.
constexpr size_t al16(size_t x) { return (x + 15) & ~static_cast<size_t>(15); }
// Dead lanes a wave gathers before it refills them: refilling on every death
// runs the launch and exit-recording code in ~1 of 4 iterations on 240-event
// paths; waiting for six idles ~4 % of the lanes and runs it a sixth as
// often (on the 3-media multilayer run 1.6 % faster than three, 1 % faster
// than eight).
constexpr int kRefillMin = 6;
template <class R> struct LdsHdr {
  static constexpr size_t hits = 0;
  static constexpr size_t wq_base = al16(hits + sizeof(unsigned) * (kMaxSensors + 2));
  static constexpr size_t wq_used = al16(wq_base + sizeof(unsigned long long) * 16);  // up to 16 waves
  static constexpr size_t common_end = al16(wq_used + sizeof(unsigned) * 16);  // batched runs stop here
  static constexpr size_t layers = common_end;
  static constexpr size_t species = al16(layers + sizeof(LayerDev<R>) * kMaxLayers);
  static constexpr size_t sensors = al16(species + sizeof(SpeciesDev<R>) * kMaxSpecies);
  static constexpr size_t planes = al16(sensors + sizeof(SensorDev) * kMaxSensors);
  static constexpr size_t lc = al16(planes + sizeof(PlaneDev) * kMaxSensors);
  static constexpr size_t refl = al16(lc + sizeof(LaunchConst<R>));
  static constexpr size_t end = al16(refl + sizeof(double) * kMaxSensors);
};

// The transport kernel.  kBatch: the photons of one launch belong to several
// scenes, concatenated in photon-index order (scene i owns global indices
// [prefix[i], prefix[i+1])); layers, species, sensors and tables of all
// scenes then stay in global memory (cache resident: the index order keeps
// only a few scenes in flight at a time) and each photon launches with its
// own scene's laser.  Otherwise the scene constants are staged in LDS.
// The body is shared by two entry points: 512-thread workgroups (two per CU
// when their LDS allows) and 1024-thread workgroups (one per CU, the whole
// LDS for tables that two workgroups could not hold).
template <class R, bool kXo, bool kTrack, bool kLdsTables, bool kAdj, bool kBatch, bool kEst = false, int kAux = 1>
__device__ __forceinline__ void transport_body(const KParams<R>& p) {
  extern __shared__ __align__(16) unsigned char smem[];
  using H = LdsHdr<R>;
  const int tid = threadIdx.x, nt = blockDim.x;
  auto copy = [&](void* dst, const void* src, size_t bytes) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    for (int i = tid; i < static_cast<int>(bytes / 4); i += nt) d[i] = s[i];
  };
  unsigned* sHits = reinterpret_cast<unsigned*>(smem + H::hits);
  for (int i = tid; i < kMaxSensors + 2; i += nt) sHits[i] = 0u;
  const int n_waves = (nt + 63) >> 6;
  unsigned long long* sWB = reinterpret_cast<unsigned long long*>(smem + H::wq_base);
  unsigned* sWU = reinterpret_cast<unsigned*>(smem + H::wq_used);
  for (int i = tid; i < n_waves; i += nt) {
    sWB[i] = 0ull;
    sWU[i] = static_cast<unsigned>(p.q_chunk);  // full: the first crossing takes a chunk
  }
  SceneView<R> v;
  v.lgrid = nullptr;
  v.lrefl = nullptr;
  v.lhits = sHits;
  v.wq_base = sWB;
  v.wq_used = sWU;
  uint64_t* sPre = nullptr;
  unsigned char* q;
  auto carve = [&](size_t bytes) {
    unsigned char* r = q;
    q += al16(bytes);
    return r;
  };
  if constexpr (kBatch) {
    q = smem + H::common_end;
    sPre = reinterpret_cast<uint64_t*>(carve(sizeof(uint64_t) * (p.n_scenes + 1)));
    for (int i = tid; i <= p.n_scenes; i += nt) sPre[i] = p.prefix[i];
    v.lc = p.lcs;
    v.sens = p.sensors;
    if constexpr (kLdsTables) {
      // a batch stages the per-event scene arrays of all its scenes (the
      // phase and amplitude tables stay in HBM, read through the caches)
      LayerDev<R>* sL = reinterpret_cast<LayerDev<R>*>(carve(sizeof(LayerDev<R>) * p.n_layers));
      SpeciesDev<R>* sS = reinterpret_cast<SpeciesDev<R>*>(carve(sizeof(SpeciesDev<R>) * p.n_species));
      PlaneDev* sPl = reinterpret_cast<PlaneDev*>(carve(sizeof(PlaneDev) * p.n_planes));
      copy(sL, p.layers, sizeof(LayerDev<R>) * p.n_layers);
      copy(sS, p.species, sizeof(SpeciesDev<R>) * p.n_species);
      copy(sPl, p.planes, sizeof(PlaneDev) * p.n_planes);
      v.L = sL;
      v.S = sS;
      v.planes = sPl;
    } else {
      v.L = p.layers;
      v.S = p.species;
      v.planes = p.planes;
    }
  } else {
    LayerDev<R>* sL = reinterpret_cast<LayerDev<R>*>(smem + H::layers);
    SpeciesDev<R>* sS = reinterpret_cast<SpeciesDev<R>*>(smem + H::species);
    SensorDev* sSen = reinterpret_cast<SensorDev*>(smem + H::sensors);
    PlaneDev* sPl = reinterpret_cast<PlaneDev*>(smem + H::planes);
    LaunchConst<R>* sLC = reinterpret_cast<LaunchConst<R>*>(smem + H::lc);
    copy(sL, p.layers, sizeof(LayerDev<R>) * p.n_layers);
    copy(sS, p.species, sizeof(SpeciesDev<R>) * p.n_species);
    copy(sSen, p.sensors, sizeof(SensorDev) * p.n_sensors);
    copy(sPl, p.planes, sizeof(PlaneDev) * p.n_planes);
    copy(sLC, p.lcs, sizeof(LaunchConst<R>));
    v.lc = sLC;
    v.L = sL;
    v.S = sS;
    v.sens = sSen;
    v.planes = sPl;
    double* sRefl = reinterpret_cast<double*>(smem + H::refl);
    for (int i = tid; i < kMaxSensors; i += nt) sRefl[i] = 0.0;
    v.lrefl = sRefl;
    q = smem + H::end;
  }
  if constexpr (kLdsTables && !kBatch) {
    unsigned char* tb = carve(p.tab_bytes);
    copy(tb, p.tab, p.tab_bytes);
    v.cdf = reinterpret_cast<const R*>(tb);
    // the guide tables stay in global memory (cache resident): the LDS copy
    // ends where they begin
    v.guide = reinterpret_cast<const uint16_t*>(p.tab);
    v.stab = reinterpret_cast<const STab<R>*>(tb);
    v.sel = reinterpret_cast<const R*>(tb);
  } else {
    v.cdf = reinterpret_cast<const R*>(p.tab);
    v.guide = reinterpret_cast<const uint16_t*>(p.tab);
    v.stab = reinterpret_cast<const STab<R>*>(p.tab);
    v.sel = reinterpret_cast<const R*>(p.tab);
  }
  if constexpr (kEst) {  // the workgroup's copy of the estimator grids that fit in LDS
    double* eg = reinterpret_cast<double*>(carve(sizeof(double) * p.est_lds_n));
    for (int i = tid; i < p.est_lds_n; i += nt) eg[i] = 0.0;
    v.lgrid = eg;
  }
  __syncthreads();

  // persistent photon loop with wave-level refill from a claimed index pool
  Photon<R, kXo> ph;
  ph.alive = false;
  ph.pend = 0;
  uint64_t w_ph = 0, w_ev = 0, w_tr = 0;  // wave-uniform counters
  uint64_t pool_next = 0, pool_end = 0;
  bool exhausted = false;
  const int lane = __lane_id();
  const uint64_t below_mask = (1ull << lane) - 1ull;
  uint32_t iter = 0;
  for (;;) {
    const bool renorm = (++iter & 1u) == 0u;  // wave-uniform: T normalised every other iteration
    // Refill once a few lanes are dead (the launch code runs for the whole
    // wave whenever any lane takes a photon): kRefillMin dead lanes, or all.
    const uint64_t dead = __ballot(!ph.alive);
    const bool refill = !exhausted && (__popcll(dead) >= kRefillMin || dead == ~0ull);
    if constexpr (sizeof(R) == 4) {
      // record the crossings of the photons that left the sample (their
      // state is frozen) before their lanes are reused, or while draining
      if ((refill || exhausted) && __ballot(ph.pend != 0)) {
        if (ph.pend != 0) {
          const V3<R> prev{__uint_as_float(ph.w_th), __uint_as_float(ph.w_ph), __uint_as_float(ph.w_az)};
          record_hit<R, kXo, kAdj, kAux, kBatch>(p, v, ph, prev, ph.pend == 2);
          ph.pend = 0;
        }
      }
    }
    if (refill) {
      if (pool_next >= pool_end) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(p.counter, static_cast<unsigned long long>(p.claim));
        base = uniform_u64(__shfl(base, 0));
        if (base >= p.n_photons) {
          exhausted = true;
          pool_next = pool_end = p.n_photons;
        } else {
          pool_next = base;
          pool_end = base + p.claim < p.n_photons ? base + p.claim : p.n_photons;
        }
      }
      const uint64_t my = pool_next + static_cast<uint64_t>(__popcll(dead & below_mask));
      const bool take = !ph.alive && my < pool_end;
      if (take) {
        if constexpr (kBatch) {
          // scene of global index g: prefix[sc] <= g < prefix[sc + 1]
          const uint64_t g = p.offset + my;
          int lo = 0, hi = p.n_scenes;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sPre[mid] <= g)
              lo = mid;
            else
              hi = mid;
          }
          launch<R, kXo>(p, p.lcs[lo], ph, static_cast<uint32_t>(p.local_offset + (g - sPre[lo])));
        } else {
          launch<R, kXo>(p, *v.lc, ph, static_cast<uint32_t>(p.offset + my));
        }
      }
      const uint64_t nd = static_cast<uint64_t>(__popcll(dead));
      const uint64_t avail = pool_end - pool_next;
      const uint64_t used = nd < avail ? nd : avail;
      pool_next += used;
      w_ph += used;
    }
    const uint64_t alive = __ballot(ph.alive);
    if (alive == 0) {
      if (exhausted) break;
      continue;
    }
    int32_t ev0 = ph.events;
    if constexpr (kEst) {
      if (ph.alive) transport_event<R, kXo, kTrack, kAdj, 1, kAux, kBatch>(p, v, ph, renorm);
      // SensorsGroup::run_estimators (detector.cpp:272-281) with the whole wave
      for (int si = 0; si < p.n_sensors; ++si)
        if (v.sens[si].est && v.sens[si].kind == LMC_SENSOR_FARFIELD_CBS) estimate_wave<R, kXo>(p, v, ph, si);
      if (ph.alive) w_tr += transport_event<R, kXo, kTrack, kAdj, 2, kAux, kBatch>(p, v, ph, renorm);
    } else if (ph.alive) {
      const int tr = transport_event<R, kXo, kTrack, kAdj, 0, kAux, kBatch>(p, v, ph, renorm);
      w_tr += tr;  // per-lane trial counts summed below
    }
    w_ev += static_cast<uint64_t>(__popcll(__ballot(ph.events != ev0)));
  }
  // reduce the per-lane trial sums across the wave
  for (int o = 32; o > 0; o >>= 1) w_tr += __shfl_xor(w_tr, o);
  if (lane == 0) {
    atomicAdd(&p.stats[0], static_cast<unsigned long long>(w_ph));
    atomicAdd(&p.stats[2], static_cast<unsigned long long>(w_tr));
  }
  __syncthreads();
  if (tid == 0) atomicAdd(&p.stats[3], static_cast<unsigned long long>(sHits[kMaxSensors]));
  if constexpr (!kBatch && deposits_direct(sizeof(R) == 8, kXo, kAdj, kAux, false)) {
    // the workgroup's hit counts and reflectance sums of the direct sensors
    for (int si = tid; si < p.n_sensors; si += nt) {
      const SensorDev& sd = v.sens[si];
      if (!sd.direct) continue;
      if (sHits[si]) atomicAdd(&p.hits[si], static_cast<unsigned long long>(sHits[si]));
      if (v.lrefl[si] != 0.0) unsafeAtomicAdd(p.out + sd.out_off, v.lrefl[si]);
    }
  }
  if (lane == 0) atomicAdd(&p.stats[1], static_cast<unsigned long long>(w_ev));
  if constexpr (kEst) {  // estimator hits and the LDS grid copies go to HBM
    for (int si = 0; si < p.n_sensors; ++si) {
      const SensorDev& sd = v.sens[si];
      if (!sd.est) continue;
      if (tid == 0 && sHits[si]) atomicAdd(&p.hits[si], static_cast<unsigned long long>(sHits[si]));
      if (sd.kind != LMC_SENSOR_FARFIELD_CBS || sd.est_lds < 0) continue;
      const int64_t len = 8ll * sd.n0 * sd.n1 * sd.n2;
      for (int64_t i = tid; i < len; i += nt) {
        const double val = v.lgrid[sd.est_lds + i];
        if (val != 0.0) unsafeAtomicAdd(p.out + sd.out_off + i, val);
      }
    }
  }
}

// FP32 runs: at most 128 VGPRs, so 16 waves of a CU stay resident (4 per
// SIMD); the adjoint's per-layer scores then spill 6-10 dwords to scratch,
// which costs less than a fourth of the waves (3 per SIMD ran 20 % slower).
// kAux: per-photon outputs beyond the sensors' crossings.  1 tracks the
// penetration depth (maximum z) that StatisticsSensor histograms; 2 also
// records the Absorption grid at every vertex; 0 (FP32 Philox forward runs
// without either, whose sensor planes all lie on or outside the stack's faces
// and off the launch point) carries no trace of them and records plane
// crossings only when a photon leaves the sample (no plane code in the
// flight loop).
template <class R, bool kXo, bool kTrack, bool kLdsTables, bool kAdj, bool kBatch, int kAux = 1>
__global__ void __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 2)))
k_transport(KParams<R> p) {
  transport_body<R, kXo, kTrack, kLdsTables, kAdj, kBatch, false, kAux>(p);
}

template <class R, bool kXo, bool kTrack, bool kAdj, bool kBatch, int kAux = 1>
__global__ void __launch_bounds__(1024) k_transport_wide(KParams<R> p) {
  transport_body<R, kXo, kTrack, true, kAdj, kBatch, false, kAux>(p);
}

// Estimator runs (a FARFIELD_CBS sensor in estimator mode): the forced-scatter
// loop over the sensor's bins dominates, so 256-thread workgroups and (FP32)
// 128 VGPRs, 4 waves per SIMD: a few dwords spill, 8 % faster than 168 VGPRs
// at 3 waves; tables in LDS when they fit beside the grid.
template <class R, bool kXo, bool kLdsTables, int kAux = 1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 2)))
k_transport_est(KParams<R> p) {
  transport_body<R, kXo, true, kLdsTables, false, false, true, kAux>(p);
}

// Sensor kernel: one thread per queued crossing.  For every sensor that
// accepted it, the process_hit of its kind: FarFieldCBSSensor with the
// time-reversed partner's Jones vector and the CBS path phase
// exp(i k0 (s_out + s_in).(r_n - r_1)) (detector.cpp:1231-1310),
// PlanarFluenceSensor (:780-841), PlanarFieldSensor with the FP64
// propagation phase exp(i k0 optical_path) (:546-568), and the far-field
// fluence and reflectance sensors.  Small grids accumulate in a per-workgroup
// LDS copy (ds_add_f64), the rest straight to HBM (global_atomic_add_f64).
// Time-resolved CBS and planar-fluence hits are deposited in their time bin
// only; bin 0 (time-integrated) is completed at read-out.
// kBin 0: one pass, deposits by FP64 atomics (grids in HBM) or into the
// workgroup's LDS copy.  kBin 1 and 2: the binned passes over this
// workgroup's contiguous share of the record range [bin_rec_lo, + wgs *
// per_wg) (the same records both times): 1 only counts the HBM deposits per
// tile, 2 does everything else and writes those deposits as entries.
template <class R, bool kAdj, int kBin = 0>
__global__ void __launch_bounds__(256) k_crossings(KParams<R> p) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* sgrid = reinterpret_cast<double*>(smem);
  const int n_gp = 3 * (p.n_layers < kMaxGradLayers ? p.n_layers : kMaxGradLayers);
  const int64_t lds_vals = static_cast<int64_t>(p.lds_grid) * (1 + ((kAdj && p.lds_grad) ? n_gp : 0));
  unsigned* shits = reinterpret_cast<unsigned*>(smem + ((sizeof(double) * lds_vals + 15) & ~static_cast<size_t>(15)));
  unsigned* lbin = shits + ((p.n_sensors + 1 + 3) & ~3);  // binned passes: per-tile counters / cursors
  const int tid = threadIdx.x, nt = blockDim.x;
  if constexpr (kBin != 1)
    for (int64_t i = tid; i < lds_vals; i += nt) sgrid[i] = 0.0;
  for (int i = tid; i < p.n_sensors + 1; i += nt) shits[i] = 0u;
  if constexpr (kBin == 1)
    for (int t = tid; t < p.bin_tiles; t += nt) lbin[t] = 0u;
  if constexpr (kBin == 2)  // first entry of this workgroup's deposits in each tile
    for (int t = tid; t < p.bin_tiles; t += nt)
      lbin[t] = p.bin_tile[t] + p.bin_pos[static_cast<int64_t>(t) * p.bin_wgs + blockIdx.x];
  __syncthreads();
  unsigned* s_cbs = shits + p.n_sensors;
  const int64_t n_rec = static_cast<int64_t>(p.q_count[0]) < p.q_cap ? static_cast<int64_t>(p.q_count[0]) : p.q_cap;
  const int64_t c = p.q_cap;
  int64_t r_first, r_end, r_step;
  if constexpr (kBin == 0) {
    r_first = static_cast<int64_t>(blockIdx.x) * nt + tid;
    r_end = n_rec;
    r_step = static_cast<int64_t>(gridDim.x) * nt;
  } else {
    const int64_t b0 = p.bin_rec_lo + static_cast<int64_t>(blockIdx.x) * p.bin_rec_per_wg;
    r_first = b0 + tid;
    r_end = b0 + p.bin_rec_per_wg < n_rec ? b0 + p.bin_rec_per_wg : n_rec;
    r_step = nt;
  }
  for (int64_t i = r_first; i < r_end; i += r_step) {
    const R* f = p.qf + i * (4 * kRecQ);
    const int32_t* iq = p.qi + i;
    const int pl = iq[I_PLANE * c];
    if (pl < 0) continue;  // slot of a partly used chunk
    const int mask = iq[I_MASK * c];
    const int events = iq[I_EVENTS * c];
    const Quad<R> q0 = get4(f, 0), q1 = get4(f, 1), q2 = get4(f, 2), q3 = get4(f, 3), q4 = get4(f, 4);
    const R hx = q0.a, hy = q0.b, arr = q0.c, w = q0.d;
    const cx<R> Em{q1.a, q1.b}, En{q1.c, q1.d};
    Frame<R> P;
    P.m = {q2.a, q2.b, q2.c};
    P.n = {q2.d, q3.a, q3.b};
    P.s = {q3.c, q3.d, q4.a};
    R score[3 * kMaxGradLayers];
    if constexpr (kAdj) {
#pragma unroll
      for (int qq = 0; qq < 3; ++qq) {
        const Quad<R> sq = get4(f, kRecQScore + qq);
        score[4 * qq] = sq.a;
        score[4 * qq + 1] = sq.b;
        score[4 * qq + 2] = sq.c;
        score[4 * qq + 3] = sq.d;
      }
    }
    const LaunchConst<R>& lc = p.lcs[p.planes[pl].scene];
    const R ws = M<R>::sqrt(w);
    const cx<R> fm = cscale(Em, ws), fn = cscale(En, ws);
    // forward field on the lab axes (project_to_lab, detector.cpp:1048-1053)
    const cx<R> Efx = cadd(cscale(fm, P.m.x), cscale(fn, P.n.x));
    const cx<R> Efy = cadd(cscale(fm, P.m.y), cscale(fn, P.n.y));
    const int first = p.planes[pl].first, cnt = p.planes[pl].count;
    // the photon's polarization_reverse as a PhotonRecordSensor sees it: set
    // by a CBS sensor earlier in the plane's list (detector.cpp:1281-1288)
    cx<R> pr_m{R(0), R(0)}, pr_n{R(0), R(0)};
    for (int k = 0; k < cnt; ++k) {
      if (!((mask >> k) & 1)) continue;
      const int si = first + k;
      const SensorDev& sd = p.sensors[si];
      if (sd.kind == LMC_SENSOR_PHOTON_RECORD) {  // PhotonRecordSensor::process_hit (detector.cpp:463-485)
        if constexpr (kBin == 1) continue;
        const unsigned long long slot = atomicAdd(&p.rec_count[0], 1ull);
        if (static_cast<int64_t>(slot) < p.rec_cap) {
          const Quad<R> q10 = get4(f, 10), q11 = get4(f, 11);
          double* o = p.rec + static_cast<int64_t>(slot) * LMC_RECORD_FIELDS;
          // field by field (layout of LMC_RECORD_FIELDS), so no 34-double
          // array raises the kernel's register peak
          auto put = [&](int j, double x) { o[j] = x; };
          auto put3 = [&](int j, R x, R y, R z) {
            o[j] = static_cast<double>(x);
            o[j + 1] = static_cast<double>(y);
            o[j + 2] = static_cast<double>(z);
          };
          put(0, static_cast<double>(sd.id));
          put(1, static_cast<double>(events));
          put(2, static_cast<double>(__int_as_float(iq[I_DEPTH * c])));
          put(3, static_cast<double>(q10.d));
          put(4, static_cast<double>(arr));
          put(5, static_cast<double>(q11.d));
          put(6, static_cast<double>(w));
          put(7, static_cast<double>(si));  // the host replaces it with the scene's k
          put3(8, q10.a, q10.b, q10.c);
          put3(11, q11.a, q11.b, q11.c);
          put(14, static_cast<double>(hx));
          put(15, static_cast<double>(hy));
          put(16, p.planes[pl].z);
          put3(17, P.s.x, P.s.y, P.s.z);
          put3(20, P.m.x, P.m.y, P.m.z);
          put3(23, P.n.x, P.n.y, P.n.z);
          put3(26, Em.r, Em.i, En.r);
          put3(29, En.i, pr_m.r, pr_m.i);
          put(32, static_cast<double>(pr_n.r));
          put(33, static_cast<double>(pr_n.i));
        } else {
          atomicAdd(&p.rec_count[1], 1ull);
        }
        atomicAdd(&shits[si], 1u);
        continue;
      }
      if (sd.kind == LMC_SENSOR_STATISTICS) {  // not binned: deposits in the second pass (LDS or atomics)
        if constexpr (kBin != 1) {
          stats_deposit<R, kAdj>(p, sd, sgrid, events, P.s, arr, w, __int_as_float(iq[I_DEPTH * c]), score);
          atomicAdd(&shits[si], 1u);
        }
        continue;
      }
      int64_t bin;
      int t_idx;
      if (!sensor_bin(sd, hx, hy, arr, P.s, bin, t_idx)) continue;
      if (sd.kind == LMC_SENSOR_FARFIELD_CBS) {
        cx<R> Erx{R(0), R(0)}, Ery{R(0), R(0)};
        if (events >= 2) {
          const int spk = iq[I_SP * c];
          const Quad<R> q5 = get4(f, 5), q6 = get4(f, 6), q7 = get4(f, 7), q8 = get4(f, 8), q9 = get4(f, 9);
          const Frame<R> P1 = frame_of(V3<R>{q4.b, q4.c, q4.d}, V3<R>{q5.a, q5.b, q5.c});
          const Frame<R> Pq = frame_of(V3<R>{q5.d, q6.a, q6.b}, V3<R>{q6.c, q6.d, q7.a});
          const cx<R> T[4] = {{q7.b, q7.c}, {q7.d, q8.a}, {q8.b, q8.c}, {q8.d, q9.a}};
          const V3<R> d{q9.b, q9.c, q9.d};
          cx<R> rm, rn;
          reverse_field(lc, reinterpret_cast<const STab<R>*>(p.tab), p.species[spk & 0xFFFF], p.species[spk >> 16], P1, Pq, P, T, rm, rn);
          pr_m = rm;
          pr_n = rn;
          const R arg =
              lc.k0 * ((P.s.x + lc.ldir[0]) * d.x + (P.s.y + lc.ldir[1]) * d.y + (P.s.z + lc.ldir[2]) * d.z);
          R sa, ca;
          M<R>::sincos_phase(arg, &sa, &ca);
          const cx<R> pp{ca * ws, sa * ws};
          rm = cmul(rm, pp);
          rn = cmul(rn, pp);
          Erx = cadd(cscale(rm, P.m.x), cscale(rn, P.n.x));
          Ery = cadd(cscale(rm, P.m.y), cscale(rn, P.n.y));
          if constexpr (kBin != 1) atomicAdd(s_cbs, 1u);
        }
        cbs_deposit<R, kAdj, kBin>(p, sd, sgrid, bin, t_idx, Efx, Efy, Erx, Ery, score, lbin);
      } else if (sd.kind == LMC_SENSOR_PLANAR_FLUENCE) {
        const int64_t tplane = static_cast<int64_t>(sd.n0) * sd.n1;
        const int64_t plane = tplane * sd.n2;
        const cx<R> Ex = cscale(Efx, R(-1));  // lab x negated (detector.cpp:812)
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
          deposit<R, kAdj, kBin>(p, sd, sgrid, q * plane + t_idx * tplane + bin,
                                 static_cast<double>(stokes(q, Ex, Efy)), score, lbin);
        }
      } else if (sd.kind == LMC_SENSOR_PLANAR_FIELD) {
        double sp, cp;
        ::sincos(lc.k0d * p.qd[i], &sp, &cp);
        const double wsd = ::sqrt(static_cast<double>(w));
        const double exr = static_cast<double>(Em.r * P.m.x + En.r * P.n.x);
        const double exi = static_cast<double>(Em.i * P.m.x + En.i * P.n.x);
        const double eyr = static_cast<double>(Em.r * P.m.y + En.r * P.n.y);
        const double eyi = static_cast<double>(Em.i * P.m.y + En.i * P.n.y);
        const int64_t plane = static_cast<int64_t>(sd.n0) * sd.n1;
        deposit<R, kAdj, kBin>(p, sd, sgrid, 2 * bin, (exr * cp - exi * sp) * wsd, score, lbin);
        deposit<R, kAdj, kBin>(p, sd, sgrid, 2 * bin + 1, (exr * sp + exi * cp) * wsd, score, lbin);
        deposit<R, kAdj, kBin>(p, sd, sgrid, 2 * (plane + bin), (eyr * cp - eyi * sp) * wsd, score, lbin);
        deposit<R, kAdj, kBin>(p, sd, sgrid, 2 * (plane + bin) + 1, (eyr * sp + eyi * cp) * wsd, score, lbin);
      } else if (sd.kind == LMC_SENSOR_FARFIELD_FLUENCE) {
        const int64_t plane = static_cast<int64_t>(sd.n0) * sd.n1;
#pragma unroll 1
        for (int q = 0; q < 4; ++q)
          deposit<R, kAdj, kBin>(p, sd, sgrid, q * plane + bin, static_cast<double>(stokes(q, Efx, Efy)), score,
                                 lbin);
      } else {  // LMC_SENSOR_REFLECTANCE
        deposit<R, kAdj, kBin>(p, sd, sgrid, 0, static_cast<double>(w), score, lbin);
      }
      if constexpr (kBin != 1) atomicAdd(&shits[si], 1u);
    }
  }
  __syncthreads();
  if constexpr (kBin == 1) {  // this workgroup's counts, [tile][workgroup]
    for (int t = tid; t < p.bin_tiles; t += nt) p.bin_pos[static_cast<int64_t>(t) * p.bin_wgs + blockIdx.x] = lbin[t];
    return;
  }
  if constexpr (kBin == 2) {  // each cursor ends where the next workgroup's (or tile's) entries begin
    for (int t = tid; t < p.bin_tiles; t += nt) {
      const uint32_t end = blockIdx.x + 1 < static_cast<unsigned>(p.bin_wgs)
                               ? p.bin_tile[t] + p.bin_pos[static_cast<int64_t>(t) * p.bin_wgs + blockIdx.x + 1]
                               : p.bin_tile[t + 1];
      if (lbin[t] != end) atomicOr(p.bin_flag, 1u);
    }
  }
  for (int i = tid; i < p.n_sensors; i += nt)
    if (shits[i]) atomicAdd(&p.hits[i], static_cast<unsigned long long>(shits[i]));
  if (tid == 0 && *s_cbs) atomicAdd(&p.stats[4], static_cast<unsigned long long>(*s_cbs));
  for (int si = 0; si < p.n_sensors; ++si) {
    const SensorDev& sd = p.sensors[si];
    if (!sd.lds) continue;
    const int64_t len = (sd.kind == LMC_SENSOR_FARFIELD_CBS)
                            ? 8ll * sd.n0 * sd.n1 * sd.n2
                            : (sd.kind == LMC_SENSOR_REFLECTANCE
                                   ? 1ll
                                   : (sd.kind == LMC_SENSOR_STATISTICS ? static_cast<int64_t>(sd.n0)
                                                                       : 4ll * sd.n0 * sd.n1));
    for (int64_t i = tid; i < len; i += nt) {
      const double val = sgrid[sd.lds_off + i];
      if (val != 0.0) unsafeAtomicAdd(p.out + sd.out_off + i, val);
    }
    if (kAdj && p.lds_grad) {
      for (int q = 0; q < n_gp; ++q)
        for (int64_t i = tid; i < len; i += nt) {
          const double val = sgrid[static_cast<int64_t>(p.lds_grid) * (1 + q) + sd.lds_off + i];
          if (val != 0.0) unsafeAtomicAdd(p.grad + q * p.n_out + sd.out_off + i, val);
        }
    }
  }
}

// Plain kernels: defined once, in the host translation unit.
#ifndef LMC_TRANSPORT_PARTS
// Binned deposits, step 2: for each tile (one workgroup per tile), the
// exclusive scan of the per-workgroup counts of pass 1 in place, and the
// tile's total into tile_tot.
__global__ void __launch_bounds__(256) k_bin_scan_wgs(uint32_t* pos, uint32_t* tile_tot, int n_wg) {
  __shared__ uint32_t part[256];
  uint32_t* row = pos + static_cast<int64_t>(blockIdx.x) * n_wg;
  const int per = (n_wg + 255) / 256;
  const int a = threadIdx.x * per, b = a + per < n_wg ? a + per : n_wg;
  uint32_t sum = 0;
  for (int i = a; i < b; ++i) sum += row[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int k = 0; k < 256; ++k) {
      const uint32_t v = part[k];
      part[k] = run;
      run += v;
    }
    tile_tot[blockIdx.x] = run;
  }
  __syncthreads();
  uint32_t run = part[threadIdx.x];
  for (int i = a; i < b; ++i) {
    const uint32_t v = row[i];
    row[i] = run;
    run += v;
  }
}

// Step 3 (one workgroup): tile totals -> first entry of each tile (exclusive
// scan; tile[n_tiles] = all entries), and the accumulation work items: a tile
// of c entries becomes ceil(c / chunk) items of at most `chunk` entries (item:
// tile, first entry, entries, 1 when it is the tile's only item).
__global__ void __launch_bounds__(1024) k_bin_scan_tiles(uint32_t* tile, uint4* items, uint32_t* n_items, int n_tiles,
                                                         int chunk, int max_items) {
  __shared__ uint32_t pe[1024], pi[1024];
  const int per = (n_tiles + 1023) / 1024;
  const int a = threadIdx.x * per, b = a + per < n_tiles ? a + per : n_tiles;
  uint32_t se = 0, si = 0;
  for (int k = a; k < b; ++k) {
    se += tile[k];
    si += (tile[k] + chunk - 1) / chunk;
  }
  pe[threadIdx.x] = se;
  pi[threadIdx.x] = si;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t re = 0, ri = 0;
    for (int k = 0; k < 1024; ++k) {
      const uint32_t ve = pe[k], vi = pi[k];
      pe[k] = re;
      pi[k] = ri;
      re += ve;
      ri += vi;
    }
    tile[n_tiles] = re;
    *n_items = ri < static_cast<uint32_t>(max_items) ? ri : static_cast<uint32_t>(max_items);
  }
  __syncthreads();
  uint32_t re = pe[threadIdx.x], ri = pi[threadIdx.x];
  for (int k = a; k < b; ++k) {
    const uint32_t cnt = tile[k];
    tile[k] = re;
    const uint32_t nch = (cnt + chunk - 1) / chunk;
    for (uint32_t j = 0; j < nch; ++j, ++ri) {
      const uint32_t first = re + j * static_cast<uint32_t>(chunk);
      const uint32_t len = cnt - j * static_cast<uint32_t>(chunk) < static_cast<uint32_t>(chunk)
                               ? cnt - j * static_cast<uint32_t>(chunk) : static_cast<uint32_t>(chunk);
      if (ri < static_cast<uint32_t>(max_items)) items[ri] = make_uint4(static_cast<uint32_t>(k), first, len, nch == 1);
    }
    re += cnt;
  }
}

// Step 5: sum one work item's entries in LDS (2^shift accumulators of its
// tile) and add the sums into `out`: a plain read-add-write per cell when the
// item is its tile's only one (no other workgroup touches the tile), FP64
// atomics for the cells of a tile split into several items.  Measured on the
// multilayer run, scattered device FP64 atomics ran at ~1.1e10/s into the
// 136 MB fluence grid and ~2.4e10/s into an L2-sized grid: the sensor kernel's
// whole cost.  Entries are read once, coalesced; each tile is written once.
__global__ void __launch_bounds__(256) k_bin_accumulate(double* out, int64_t n_out, const uint32_t* keys,
                                                        const double* vals, const uint4* items,
                                                        const uint32_t* n_items, int shift) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* acc = reinterpret_cast<double*>(smem);
  if (blockIdx.x >= *n_items) return;
  const uint4 w = items[blockIdx.x];
  const int T = 1 << shift;
  const int64_t base = static_cast<int64_t>(w.x) << shift;
  for (int k = threadIdx.x; k < T; k += blockDim.x) acc[k] = 0.0;
  __syncthreads();
  for (uint32_t e = w.y + threadIdx.x; e < w.y + w.z; e += blockDim.x) {
    const uint32_t k = keys[e] - static_cast<uint32_t>(base);  // < T by construction; checked, not trusted
    if (k < static_cast<uint32_t>(T))
      __hip_atomic_fetch_add((lds_double*)(acc + k), vals[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < T && base + k < n_out; k += blockDim.x) {
    const double v = acc[k];
    if (v == 0.0) continue;
    if (w.w)
      out[base + k] += v;
    else
      unsafeAtomicAdd(out + base + k, v);
  }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t val) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = val;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

template <class R>
__global__ void k_azimuth_probe(int n, const double* abc, const uint32_t* words, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    R cph, sph, F;
    azimuth_sample<R>(static_cast<R>(abc[3 * i]), static_cast<R>(abc[3 * i + 1]), static_cast<R>(abc[3 * i + 2]),
                      words[2 * i], words[2 * i + 1], cph, sph, F);
    out[3 * i] = static_cast<double>(cph);
    out[3 * i + 1] = static_cast<double>(sph);
    out[3 * i + 2] = static_cast<double>(F);
  }
}

// Stage C of the reverse field on given frames: in[20 i ..] = m1, s1 (P1 =
// frame_of(m1, s1)), m_n, n_n, s_n, (xm, xn) as 4 reals, the launch frame's
// handedness; constant amplitudes S1 = 0.6 + 0.2i, S2 = 0.3 - 0.5i on a
// two-node table; out[4 i ..] = Erm, Ern.
template <class R>
__global__ void k_reverse_probe(int n, const double* in, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double* q = in + 20 * static_cast<int64_t>(i);
    auto v3 = [&](int o) { return V3<R>{static_cast<R>(q[o]), static_cast<R>(q[o + 1]), static_cast<R>(q[o + 2])}; };
    const Frame<R> P1 = frame_of(v3(0), v3(3));
    Frame<R> Pn;
    Pn.m = v3(6);
    Pn.n = v3(9);
    Pn.s = v3(12);
    const cx<R> xm{static_cast<R>(q[15]), static_cast<R>(q[16])}, xn{static_cast<R>(q[17]), static_cast<R>(q[18])};
    STab<R> tb[2];
    for (int k = 0; k < 2; ++k) {
      if constexpr (sizeof(R) == 4) {
        tb[k].v = make_float4(0.6f, 0.2f, 0.3f, -0.5f);
      } else {
        tb[k].a = make_double2(0.6, 0.2);
        tb[k].b = make_double2(0.3, -0.5);
      }
    }
    SpeciesDev<R> sp{};
    sp.s_n = 2;
    sp.s_off = 0;
    sp.inv_ds = static_cast<R>(1.0 / 3.14159265358979323846);
    cx<R> rm, rn;
    reverse_suffix(tb, sp, P1, Pn, static_cast<R>(q[19]), xm, xn, rm, rn);
    out[4 * i] = rm.r;
    out[4 * i + 1] = rm.i;
    out[4 * i + 2] = rn.r;
    out[4 * i + 3] = rn.i;
  }
}

// sin and cos of FP32 path phases as the CBS reverse-path factor forms them.
__global__ void k_phase_probe(int n, const float* x, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float s, c;
    M<float>::sincos_phase(x[i], &s, &c);
    out[2 * i] = s;
    out[2 * i + 1] = c;
  }
}

// First n outputs of photon `photon`'s stream under key (k0, k1), one lane.
__global__ void k_stream_probe(uint32_t k0, uint32_t k1, uint64_t photon, int n, uint32_t* out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Xoshiro128ss g = seed_stream(k0, k1, static_cast<uint32_t>(photon), static_cast<uint32_t>(photon >> 32));
    for (int i = 0; i < n; ++i) out[i] = g.next();
  }
}

__global__ void k_philox_probe(int n, const uint4* ctr, uint32_t k0, uint32_t k1, uint4* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = philox4x32_10(ctr[i], k0, k1);
}

#endif  // LMC_TRANSPORT_PARTS

}  // namespace lmc
