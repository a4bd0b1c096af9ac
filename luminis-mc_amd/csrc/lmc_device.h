// Device-side building blocks of the gfx950 transport kernels: counter-based
// RNG, small complex/vector algebra in registers, scene views and the
// per-event transport body.  Everything here is templated on the transport
// real type R (float: production; double: FP64 check build).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/luminis_hip.h"

namespace lmc {

constexpr int kMaxLayers = 8;
constexpr int kMaxSensors = 8;
constexpr int kMaxSpecies = 64;
constexpr int kMaxGradLayers = 4;
// guide entries of a mixture layer's species-selection CDF (u16, kSelGuide + 1
// after the CDF's kMaxSpecies-at-most values)
constexpr int kSelGuide = 256;

// ------------------------------------------------------------ math ------
template <class R> struct M;
template <> struct M<float> {
  // hardware log2 (v_log_f32, ~1 ulp) scaled to ln
  static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
  // hardware square root (v_sqrt_f32, ~1 ulp): the correctly rounded form
  // costs five more instructions for denormal scaling the path never needs
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float rsqrt(float x) { return __frsqrt_rn(x); }
  // acos by Abramowitz & Stegun 4.4.46 (Hastings): acos|x| = sqrt(1 - |x|) P7(|x|),
  // |error| <= 2e-8 in exact arithmetic (4e-7 abs, 2e-7 rel near 0 in FP32);
  // acos x = pi - acos|x| for x < 0.
  static __device__ __forceinline__ float acos(float x) {
    const float ax = fabsf(x);
    float p = -0.0012624911f;
    p = __builtin_fmaf(p, ax, 0.0066700901f);
    p = __builtin_fmaf(p, ax, -0.0170881256f);
    p = __builtin_fmaf(p, ax, 0.0308918810f);
    p = __builtin_fmaf(p, ax, -0.0501743046f);
    p = __builtin_fmaf(p, ax, 0.0889789874f);
    p = __builtin_fmaf(p, ax, -0.2145988016f);
    p = __builtin_fmaf(p, ax, 1.5707963050f);
    const float r = __builtin_amdgcn_sqrtf(1.0f - ax) * p;
    return x < 0.0f ? 3.14159265358979f - r : r;
  }
  static __device__ __forceinline__ float atan2(float y, float x) { return atan2f(y, x); }
  // sin/cos of 2*pi*u for u in [0,1): the hardware evaluates in revolutions.
  static __device__ __forceinline__ void sincos_rev(float u, float* s, float* c) {
    *s = __builtin_amdgcn_sinf(u);
    *c = __builtin_amdgcn_cosf(u);
  }
  // angles within a few revolutions (scattering angles)
  static __device__ __forceinline__ void sincos(float x, float* s, float* c) { __sincosf(x, s, c); }
  // path phases of any size: reduced to [0, 1) revolutions first (v_fract is
  // exact), since v_sin / v_cos take revolutions on a limited domain (+-256)
  static __device__ __forceinline__ void sincos_phase(float x, float* s, float* c) {
    const float r = __builtin_amdgcn_fractf(x * 0.159154943091895336f);
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
  }
  static __device__ __forceinline__ float fabs(float x) { return fabsf(x); }
  static __device__ __forceinline__ float exp(float x) { return __expf(x); }
  static __device__ __forceinline__ float floor(float x) { return floorf(x); }
};
template <> struct M<double> {
  static __device__ __forceinline__ double log(double x) { return ::log(x); }
  static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
  static __device__ __forceinline__ double rsqrt(double x) { return 1.0 / ::sqrt(x); }
  static __device__ __forceinline__ double acos(double x) { return ::acos(x); }
  static __device__ __forceinline__ double atan2(double y, double x) { return ::atan2(y, x); }
  static __device__ __forceinline__ void sincos_rev(double u, double* s, double* c) {
    const double p = 2.0 * 3.14159265358979323846 * u;
    ::sincos(p, s, c);
  }
  static __device__ __forceinline__ void sincos(double x, double* s, double* c) { ::sincos(x, s, c); }
  static __device__ __forceinline__ void sincos_phase(double x, double* s, double* c) { ::sincos(x, s, c); }
  static __device__ __forceinline__ double fabs(double x) { return ::fabs(x); }
  static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
  static __device__ __forceinline__ double floor(double x) { return ::floor(x); }
};

template <class R> struct cx {
  R r, i;
};
template <class R> __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <class R> __device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return {a.r + b.r, a.i + b.i}; }
template <class R> __device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return {a.r * s, a.i * s}; }
template <class R> __device__ __forceinline__ R cnorm(cx<R> a) { return a.r * a.r + a.i * a.i; }

template <class R> struct V3 {
  R x, y, z;
};
template <class R> __device__ __forceinline__ R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> __device__ __forceinline__ V3<R> vscale(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
// cross product normalised when non-zero (reference src/vec.cpp:15-28)
template <class R> __device__ __forceinline__ V3<R> cross_unit(V3<R> a, V3<R> b) {
  V3<R> r{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  const R n2 = dot(r, r);
  if (n2 > R(0)) r = vscale(r, M<R>::rsqrt(n2));
  return r;
}
// safe_unit (src/utils.cpp:39-51); inputs here are unit or zero vectors.
template <class R> __device__ __forceinline__ V3<R> safe_unit(V3<R> v, V3<R> fb) {
  const R n2 = dot(v, v);
  if (n2 < R(1e-20)) {
    const R f2 = dot(fb, fb);
    if (f2 < R(1e-20)) return {R(1), R(0), R(0)};
    return vscale(fb, M<R>::rsqrt(f2));
  }
  return vscale(v, M<R>::rsqrt(n2));
}

// safe_unit(cross_unit(a, b), fb): the unit cross product, or fb normalised
// when the cross product is exactly zero.  FP32: one normalisation (the second
// one of the composed form rescales a unit vector by rsqrt(1 +- ulps));
// FP64: the composed form, operation for operation.
template <class R> __device__ __forceinline__ V3<R> unit_cross_or(V3<R> a, V3<R> b, V3<R> fb) {
  if constexpr (sizeof(R) == 8) {
    return safe_unit(cross_unit(a, b), fb);
  } else {
    const V3<R> r{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    const R n2 = dot(r, r);
    if (n2 > R(0)) return vscale(r, M<R>::rsqrt(n2));
    const R f2 = dot(fb, fb);
    if (f2 < R(1e-20)) return {R(1), R(0), R(0)};
    return vscale(fb, M<R>::rsqrt(f2));
  }
}

// Local frame: rows m, n, s (photon.hpp P_local).
template <class R> struct Frame {
  V3<R> m, n, s;
};

// ------------------------------------------------------------ RNG -------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c.x;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c.z;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
// xoshiro128** (Blackman & Vigna, "Scrambled linear pseudorandom number
// generators", ACM TOMS 47, 2021): 128-bit state, 32-bit outputs, one cycle
// of length 2^128 - 1 through every non-zero state.  Shifts, rotates and
// xors only (the multiplications by 5 and 9 are shift-adds), so a 32-bit
// output costs about nine full-rate vector instructions, where a Philox
// block spends twenty 32x32->64 multiplies on 128 bits.
struct Xoshiro128ss {
  uint32_t s0, s1, s2, s3;
  static __device__ __forceinline__ uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
  __device__ __forceinline__ uint32_t next() {
    const uint32_t x5 = (s1 << 2) + s1;
    const uint32_t r7 = rotl(x5, 7);
    const uint32_t out = (r7 << 3) + r7;
    const uint32_t t = s1 << 9;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = rotl(s3, 11);
    return out;
  }
};
// 24-bit uniform on the open interval (0,1): never 0 (log) and never 1.
template <class R> __device__ __forceinline__ R u01(uint32_t x) {
  return static_cast<R>((x >> 8) + 0.5f) * static_cast<R>(5.9604644775390625e-8);
}

// Draw streams.  The transport body asks for uniforms in the reference's
// order (free path, [species], theta, phi trials, [roulette]); the two
// generators map those requests to their own counters.
//
// Philox: one 4-word block per event (free path, species, theta, roulette)
// at counter (photon_lo, photon_hi, event, 0), and phi-trial pairs from
// blocks (photon_lo, photon_hi, event, 1 + j): the stream of every event is
// a pure function of (seed, photon, event), which keeps common random numbers
// aligned across parameter perturbations.
template <class R> struct PhiloxStream {
  static constexpr bool kSequential = false;
  uint32_t k0, k1, c0, c1, ev, trial;
  R u0, u1, u2, u3;
  R t2, t3;  // second trial of the current pair
  __device__ __forceinline__ void init(uint64_t seed_key0, uint32_t seed_key1, uint64_t photon) {
    k0 = static_cast<uint32_t>(seed_key0);
    k1 = seed_key1;
    c0 = static_cast<uint32_t>(photon);
    c1 = static_cast<uint32_t>(photon >> 32);
  }
  __device__ __forceinline__ void emit_block(uint32_t j, R& a, R& b, R& c, R& d) {
    const uint4 o = philox4x32_10(make_uint4(c0, c1, 0xFFFFFFFFu, j), k0, k1);
    a = u01<R>(o.x);
    b = u01<R>(o.y);
    c = u01<R>(o.z);
    d = u01<R>(o.w);
  }
  __device__ __forceinline__ void begin_event(uint32_t e) {
    ev = e;
    trial = 0;
    const uint4 o = philox4x32_10(make_uint4(c0, c1, e, 0u), k0, k1);
    u0 = u01<R>(o.x);
    u1 = u01<R>(o.y);
    u2 = u01<R>(o.z);
    u3 = u01<R>(o.w);
  }
  __device__ __forceinline__ R free_path() { return u0; }
  __device__ __forceinline__ R species() { return u1; }
  __device__ __forceinline__ R theta() { return u2; }
  __device__ __forceinline__ R roulette() { return u3; }
  __device__ __forceinline__ void phi_trial(R& uphi, R& uacc) {
    if ((trial & 1u) == 0u) {
      const uint4 o = philox4x32_10(make_uint4(c0, c1, ev, 1u + (trial >> 1)), k0, k1);
      uphi = u01<R>(o.x);
      uacc = u01<R>(o.y);
      t2 = u01<R>(o.z);
      t3 = u01<R>(o.w);
    } else {
      uphi = t2;
      uacc = t3;
    }
    ++trial;
  }
  __device__ __forceinline__ R normal01(R ua, R ub) {
    return M<R>::sqrt(R(-2) * M<R>::log(ua)) * cos_rev(ub);
  }
  static __device__ __forceinline__ R cos_rev(R u) {
    R s, c;
    M<R>::sincos_rev(u, &s, &c);
    return c;
  }
};

// xoshiro256++ re-seeded per photon with mix_seed (reference src/rng.cpp:5-16,
// include/luminis/math/rng.hpp:59-85); strictly sequential draws.
struct XoshiroStream {
  static constexpr bool kSequential = true;
  uint64_t s0, s1, s2, s3;
  __device__ __forceinline__ static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  __device__ __forceinline__ void init(uint64_t seed, uint32_t, uint64_t photon) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (photon + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    uint64_t st[4];
    for (int i = 0; i < 4; ++i) {
      z += 0x9E3779B97F4A7C15ULL;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
      st[i] = x ^ (x >> 31);
    }
    s0 = st[0];
    s1 = st[1];
    s2 = st[2];
    s3 = st[3];
  }
  __device__ __forceinline__ uint64_t next() {
    const uint64_t r = rotl(s0 + s3, 23) + s0;
    const uint64_t t = s1 << 17;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = rotl(s3, 45);
    return r;
  }
  __device__ __forceinline__ double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
  __device__ __forceinline__ void begin_event(uint32_t) {}
  __device__ __forceinline__ double free_path() { return uniform(); }
  __device__ __forceinline__ double species() { return uniform(); }
  __device__ __forceinline__ double theta() { return uniform(); }
  __device__ __forceinline__ double roulette() { return uniform(); }
  __device__ __forceinline__ void phi_trial(double& uphi, double& uacc) {
    uphi = uniform();
    uacc = uniform();
  }
  // Rng::normal (src/rng.cpp:18-28)
  __device__ __forceinline__ double normal(double mean, double sd) {
    double ua = uniform();
    while (ua <= 0.0) ua = uniform();
    const double ub = uniform();
    return ::sqrt(-2.0 * ::log(ua)) * ::cos(2.0 * 3.14159265358979323846 * ub) * sd + mean;
  }
};

// --------------------------------------------------------- scene views --
// Per-species constants (one lmc_species_desc, converted).
template <class R> struct SpeciesDev {
  int32_t kind;
  int32_t cdf_n;     // nodes
  int32_t cdf_off;   // offset in the R cdf pool
  int32_t guide_off; // offset in the guide pool (guide_n + 1 entries)
  int32_t guide_n;
  int32_t s_n;
  int32_t s_off;     // offset in the S pool (one STab entry per node)
  int32_t guide_shift;  // guide entries hold lower_bound >> guide_shift (tables beyond 65536 nodes)
  R cdf_min, cdf_step, hg_g, inv_ds;
  R pi_inorm;  // pi * I_norm, the estimator's angular normalisation (detector.cpp:105-118, 1436)
  R alb;       // mu_s / (mu_a + mu_s) of the species (estimator weight, detector.cpp:1466)
  R pad_[2];   // 64 bytes in FP32: the event loop's species address is a shift
};
static_assert(sizeof(SpeciesDev<float>) == 64, "FP32 species record: 64 bytes");

// S1/S2 at one node: (s1.re, s1.im, s2.re, s2.im).
template <class R> struct STab;
template <> struct STab<float> {
  float4 v;
};
template <> struct STab<double> {
  double2 a, b;
};
__device__ __forceinline__ void stab_get(const STab<float>& t, cx<float>& s1, cx<float>& s2) {
  s1 = {t.v.x, t.v.y};
  s2 = {t.v.z, t.v.w};
}
__device__ __forceinline__ void stab_get(const STab<double>& t, cx<double>& s1, cx<double>& s2) {
  s1 = {t.a.x, t.a.y};
  s2 = {t.b.x, t.b.y};
}

// One layer with what the event loop needs of its scene, so that layers of
// several scenes can share one array: neighbour flags (bit 0: a layer above in
// +z, bit 1: a layer below), the scene's top face (+inf when semi-infinite)
// and the scene's sensor planes [pl_first, pl_first + pl_count).
// 64 bytes in FP32 (a power of two: the event loop's layer addresses are a
// shift, not a multiply).  The refractive index is kept in FP64 only.
template <class R> struct LayerDev {
  R z_min, z_max;  // z_max = +inf for a semi-infinite last layer
  R mfp, mu_t, albedo, z_top;
  R mu_s, mu_a;
  int32_t sp_first, sp_count, sel_off, pl_first;
  int16_t nbr, pl_count;
  int32_t scene;  // index of the layer's scene in its set (a batch's per-scene Absorption grid)
  double n64;  // the refractive index in FP64, for the FP64 optical path
};
static_assert(sizeof(LayerDev<float>) == 64, "FP32 layer record: 64 bytes");

// One scene's Absorption grid (Absorption::record_absorption,
// absortion.cpp:47-77) for batched runs: r-z(-t) cells from `off` in the
// accumulators; on == 0 when the scene records none.
template <class R> struct AbsDev {
  R radius, depth, dr, dz, dt;
  int32_t nr, nz, nt, on;
  int64_t off;
};
// Refractive index of a layer in FP64 (stored so in both builds): the optical
// path, whose k0 multiple is the planar field's phase, accumulates in FP64
// without the ~3e-8 relative bias of an FP32 index.
template <class R> __device__ __forceinline__ double layer_n64(const LayerDev<R>& L) { return L.n64; }

struct SensorDev {
  int32_t kind, absorb, plane, id;
  int32_t f_dir_en, f_dir, f_ev_en, f_ev_min, f_ev_max, f_th_en, f_ph_en, f_pos_en;
  int32_t n0, n1, n2, lds;  // grid sizes; lds != 0: accumulate in the LDS copy
  float cos_th_min, cos_th_max, ph_min, ph_max;
  float x_min, x_max, y_min, y_max;
  double g[8];
  int64_t out_off;
  int32_t lds_off;
  // forward runs: deposits binned by tile and summed in LDS (grids far beyond
  // the caches); bin_d the same plan for runs whose transport kernel deposits
  // the direct sensors itself (those sensors then take no queued deposits)
  int32_t bin, bin_d;
  // estimator mode: est != 0 takes no crossings; a FARFIELD_CBS sensor then
  // deposits every vertex's forced scatter into its first est_imax theta rows
  // (bin geometry at aux in the auxiliary pool), in the transport kernel's
  // LDS copy at est_lds (doubles) or, when est_lds < 0, in HBM.  A STATISTICS
  // sensor keeps its 27 histogram parameters at aux.
  int32_t est, est_imax, aux, est_lds;
  // conservative grid bounds for the transport kernel's queueing test: exit
  // cosine -s_z >= q_cmin, hit point in [q_x0, q_x1] x [q_y0, q_y1], arrival
  // in [q_t0, q_t1] (slightly wider than the grid; the sensor kernel bins exactly)
  float q_cmin, q_x0, q_x1, q_y0, q_y1, q_t0, q_t1;
  // a light sensor (planar fluence, far-field fluence, reflectance): the
  // transport kernels of forward runs that deposit_direct() names deposit
  // its crossings themselves, which are then never queued
  int32_t direct, pad4;
};

// flags: kPlaneCbs a FarFieldCBSSensor (analog) reads the reverse-path
// history of its crossings; kPlaneRec a PhotonRecordSensor reads r_1, r_n and
// the launch time.
enum : int32_t { kPlaneCbs = 1, kPlaneRec = 2 };
struct PlaneDev {
  double z;
  int32_t first, count, flags, scene;  // scene: index of its laser constants
  float zf;                            // z rounded to FP32: the FP32 kernels' crossing tests
  int32_t pad;
};
template <class R> __device__ __forceinline__ R plane_z(const PlaneDev& P);
template <> __device__ __forceinline__ float plane_z<float>(const PlaneDev& P) { return P.zf; }
template <> __device__ __forceinline__ double plane_z<double>(const PlaneDev& P) { return P.z; }

}  // namespace lmc
