// gfx950 photon-transport kernels and the C ABI of libluminis_hip.so.
//
// One persistent kernel carries a photon per lane.  Each lane runs the
// reference's per-event loop (src/simulation.cpp:297-614); when its photon
// dies it is refilled from a wave-private pool of photon indices that the
// wave claims from one global counter, so heavy-tailed paths do not idle the
// wave.  Scene constants (layers, species, sensors) and, when they fit, the
// theta-CDF, its guide table and the S1/S2 tables are staged in LDS.  Small
// sensor grids are privatised per workgroup in LDS (FP64 ds_add) and flushed
// with global FP64 atomics; large grids take global FP64 atomics directly.
// Transport arithmetic runs in R = float (FP64 for the optical path that
// feeds the field phase, and for every accumulation) or R = double.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lmc_device.h"

using namespace lmc;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define LMC_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(LMC_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

namespace lmc {

// ------------------------------------------------------- kernel params --
template <class R> struct KParams {
  const LayerDev<R>* layers;
  const SpeciesDev<R>* species;
  const R* cdf;
  const uint16_t* guide;
  const STab<R>* stab;
  const R* sel;
  const SensorDev* sensors;
  const PlaneDev* planes;
  int32_t n_layers, n_species, n_sensors, n_planes;
  int32_t cdf_total, guide_total, stab_total, sel_total;
  int32_t lds_grid;    // doubles of LDS-privatised accumulators
  int32_t launch_layer;
  int32_t source, temporal;
  double lpos[3];
  R ldir[3], lm[3], ln[3];
  R pol[4];
  R k0;
  double sigma, pulse, rep, t_off;
  uint64_t n_photons, offset, seed;
  uint32_t key1;
  int32_t max_events, claim;
  R z_top;
  int32_t top_finite;
  double* out;
  unsigned long long* hits;
  unsigned long long* counter;
  unsigned long long* stats;  // photons, events, trials, crossings, cbs
  double* grad;
  int64_t n_out;
  // CBS reverse-path record queue (structure of arrays, kRecF R-fields and
  // kRecI int-fields per record, field-major with stride q_cap)
  R* qf;
  int32_t* qi;
  unsigned long long* q_count;  // [0] records written, [1] overflowed
  int64_t q_cap;
};

// Fields of one queued CBS detection (photons with >= 2 scattering events
// that landed in a valid bin): everything the reverse-path reconstruction
// needs, so the transport kernel carries none of its code.
enum RecF {
  F_W = 0, F_EM = 1, F_P = 5, F_P1 = 14, F_PN1 = 23, F_T = 32, F_D = 40, F_SCORE = 43, kRecF = 55
};
enum RecI { I_SENSOR = 0, I_BIN = 1, I_TIDX = 2, I_SP = 3, kRecI = 4 };

// LDS-resident views of the scene.
template <class R> struct SceneView {
  const LayerDev<R>* L;
  const SpeciesDev<R>* S;
  const R* cdf;
  const uint16_t* guide;
  const STab<R>* stab;
  const R* sel;
  const SensorDev* sens;
  const PlaneDev* planes;
  double* lgrid;       // LDS accumulators
  unsigned* lhits;     // LDS hit counters
};

template <class R, class RNG> struct Photon {
  V3<R> pos;
  Frame<R> P;
  cx<R> Em, En;
  R w, gpl, t0;
  double opl;
  int32_t events, layer;
  bool alive;
  Frame<R> P1, Pn1;
  V3<R> r1, rn;
  cx<R> T[4], Tb[4];
  int32_t sp1, spn;
  RNG rng;
  R aL[kMaxGradLayers], aN[kMaxGradLayers], aG[kMaxGradLayers];
};

// Inverse-CDF lookup: SamplingTable::Sample (src/table.cpp:50-66).  The
// lower_bound is narrowed by a guide table g[j] = lower_bound(j/G), so the
// result is exactly the reference's index on the same CDF values.
template <class R>
__device__ __forceinline__ R table_sample(const SceneView<R>& v, const SpeciesDev<R>& s, R u) {
  const R* c = v.cdf + s.cdf_off;
  const uint16_t* g = v.guide + s.guide_off;
  int j = static_cast<int>(u * static_cast<R>(s.guide_n));
  j = j < 0 ? 0 : (j >= s.guide_n ? s.guide_n - 1 : j);
  int lo = g[j], hi = g[j + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] < u)
      lo = mid + 1;
    else
      hi = mid;
  }
  const int idx = lo;
  if (idx == 0) return s.cdf_min;
  const R ca = c[idx - 1], cb = c[idx];
  const R xa = s.cdf_min + static_cast<R>(idx - 1) * s.cdf_step;
  const R xb = s.cdf_min + static_cast<R>(idx) * s.cdf_step;
  const R t = (u - ca) / (cb - ca);
  return xa + t * (xb - xa);
}

// DataTable::Sample on the uniform theta grid (src/table.cpp:100-118).
template <class R>
__device__ __forceinline__ void s_lookup(const SceneView<R>& v, const SpeciesDev<R>& s, R th, cx<R>& s1,
                                         cx<R>& s2) {
  const STab<R>* tb = v.stab + s.s_off;
  const R pos = th * s.inv_ds;
  const int last = s.s_n - 1;
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
// raw interior product J_{n-1}...J_2.  Returns the normalised reverse field in
// the exit basis (m_n, n_n).
template <class R>
__device__ __forceinline__ void reverse_field(const KParams<R>& p, const STab<R>* stab, const SpeciesDev<R>& spn,
                                              const SpeciesDev<R>& sp1, const Frame<R>& P1, const Frame<R>& Pq,
                                              const Frame<R>& Pn, const cx<R> T[4], cx<R>& Erm, cx<R>& Ern) {
  SceneView<R> v;
  v.stab = stab;
  const V3<R> s0{p.ldir[0], p.ldir[1], p.ldir[2]};
  const V3<R> m0{p.lm[0], p.lm[1], p.lm[2]};
  const V3<R> n0{p.ln[0], p.ln[1], p.ln[2]};
  // Stage A: scatter at r_n, s0 -> -s_{n-1}
  const V3<R> s_out_a{-Pq.s.x, -Pq.s.y, -Pq.s.z};
  const R th_a = M<R>::acos(clamp1(dot(s0, s_out_a)));
  cx<R> a1, a2;
  s_lookup(v, spn, th_a, a1, a2);
  const V3<R> np = safe_unit(cross_unit(s0, s_out_a), n0);
  const V3<R> mp_in = safe_unit(cross_unit(np, s0), m0);
  const V3<R> mp_out = safe_unit(cross_unit(np, s_out_a), Pq.m);
  const cx<R> E0m{p.pol[0], p.pol[1]}, E0n{p.pol[2], p.pol[3]};
  // E = S_a * rot2(mp_in, np, m0, n0) * E0
  R r00 = dot(mp_in, m0), r01 = dot(mp_in, n0), r10 = dot(np, m0), r11 = dot(np, n0);
  cx<R> xm = cadd(cscale(E0m, r00), cscale(E0n, r01));
  cx<R> xn = cadd(cscale(E0m, r10), cscale(E0n, r11));
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
  // Stage C: scatter at r_1, -s1 -> s_n
  const V3<R> s_in_c{-P1.s.x, -P1.s.y, -P1.s.z};
  const R th_c = M<R>::acos(clamp1(dot(s_in_c, Pn.s)));
  cx<R> c1, c2;
  s_lookup(v, sp1, th_c, c1, c2);
  const V3<R> n1neg{-P1.n.x, -P1.n.y, -P1.n.z};
  const V3<R> npp = safe_unit(cross_unit(s_in_c, Pn.s), n1neg);
  const V3<R> mpp_in = safe_unit(cross_unit(npp, s_in_c), P1.m);
  const V3<R> mpp_out = safe_unit(cross_unit(npp, Pn.s), Pn.m);
  r00 = dot(mpp_in, P1.m);
  r01 = dot(mpp_in, n1neg);
  r10 = dot(npp, P1.m);
  r11 = dot(npp, n1neg);
  ym = cadd(cscale(xm, r00), cscale(xn, r01));
  yn = cadd(cscale(xm, r10), cscale(xn, r11));
  ym = cmul(c2, ym);
  yn = cmul(c1, yn);
  r00 = dot(Pn.m, mpp_out);
  r01 = dot(Pn.m, npp);
  r10 = dot(Pn.n, mpp_out);
  r11 = dot(Pn.n, npp);
  xm = cadd(cscale(ym, r00), cscale(yn, r01));
  xn = cadd(cscale(ym, r10), cscale(yn, r11));
  const R nrm2 = cnorm(xm) + cnorm(xn);
  if (nrm2 > R(0) && M<R>::sqrt(nrm2) > static_cast<R>(sizeof(R) == 8 ? 1e-300 : 1e-30)) {
    const R inv = R(1) / M<R>::sqrt(nrm2);
    Erm = cscale(xm, inv);
    Ern = cscale(xn, inv);
  } else {
    Erm = {R(0), R(0)};
    Ern = {R(0), R(0)};
  }
}

// Accumulate one value: LDS copy (ds_add_f64) for privatised grids, HBM
// (global_atomic_add_f64) otherwise.  The two pointers keep their address
// spaces so neither path degrades to a flat atomic.
typedef __attribute__((address_space(3))) double lds_double;
__device__ __forceinline__ void acc_add(double* lds_base, double* hbm_base, bool lds, int64_t off, double val) {
  if (val != 0.0) {
    if (lds)
      __hip_atomic_fetch_add((lds_double*)(lds_base + off), val, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      unsafeAtomicAdd(hbm_base + off, val);
  }
}

// Adjoint scores of one detected photon for every layer parameter:
// d ln(path density x weight)/d mu_s = N_l/mu_s - L_l, d/d mu_a = -L_l,
// d/dg = sum of d ln p_HG/dg over the layer's scatters.
template <class R, class RNG>
__device__ __forceinline__ void adj_scores(const KParams<R>& p, const SceneView<R>& v, const Photon<R, RNG>& ph,
                                           R score[3 * kMaxGradLayers]) {
#pragma unroll
  for (int l = 0; l < kMaxGradLayers; ++l) {
    const R mus = l < p.n_layers ? v.L[l].mu_s : R(1);
    score[3 * l + 0] = (mus > R(0) ? ph.aN[l] / mus : R(0)) - ph.aL[l];
    score[3 * l + 1] = -ph.aL[l];
    score[3 * l + 2] = ph.aG[l];
  }
}

template <class R, class RNG, bool kAdj>
__device__ __forceinline__ void deposit(const KParams<R>& p, const SensorDev& sd, double* lds_grid, int64_t off,
                                        double val, const R* score) {
  acc_add(lds_grid + sd.lds_off, p.out + sd.out_off, sd.lds != 0, off, val);
  if constexpr (kAdj) {
    const int np = 3 * (p.n_layers < kMaxGradLayers ? p.n_layers : kMaxGradLayers);
    for (int q = 0; q < np; ++q)
      if (score[q] != R(0)) unsafeAtomicAdd(p.grad + q * p.n_out + sd.out_off + off, val * static_cast<double>(score[q]));
  }
}

// Coherent and incoherent Stokes of a forward field (fx, fy) and a reverse
// field (rx, ry) projected on the lab axes: accumulate_stokes
// (detector.cpp:1194-1226).
template <class R>
__device__ __forceinline__ void cbs_stokes(cx<R> Efx, cx<R> Efy, cx<R> Erx, cx<R> Ery, double vals[8]) {
  const cx<R> Ex = cadd(Efx, Erx), Ey = cadd(Efy, Ery);
  const cx<R> xy = cmul(Ex, cx<R>{Ey.r, -Ey.i});
  const cx<R> fxy = cmul(Efx, cx<R>{Efy.r, -Efy.i});
  const cx<R> rxy = cmul(Erx, cx<R>{Ery.r, -Ery.i});
  vals[0] = static_cast<double>(cnorm(Ex) + cnorm(Ey));
  vals[1] = static_cast<double>(cnorm(Ex) - cnorm(Ey));
  vals[2] = static_cast<double>(R(2) * xy.r);
  vals[3] = static_cast<double>(R(2) * xy.i);
  vals[4] = static_cast<double>(cnorm(Efx) + cnorm(Efy) + cnorm(Erx) + cnorm(Ery));
  vals[5] = static_cast<double>(cnorm(Efx) - cnorm(Efy) + cnorm(Erx) - cnorm(Ery));
  vals[6] = static_cast<double>(R(2) * (fxy.r + rxy.r));
  vals[7] = static_cast<double>(R(2) * (fxy.i + rxy.i));
}

// FarFieldCBSSensor::process_hit (detector.cpp:1231-1310), transport side:
// time and angle bins, and either the forward-only deposit of a photon with
// fewer than two events (:1237-1265) or a queued record for the reverse-path
// kernel.  info.phase has unit modulus and cancels in every coherent and
// incoherent Stokes term, so it is not formed.
template <class R, class RNG, bool kAdj>
__device__ bool cbs_hit(const KParams<R>& p, const SceneView<R>& v, const SensorDev& sd, int si, Photon<R, RNG>& ph,
                        R gpl_hit, const R* score, unsigned& cbs_count) {
  const double dt = sd.g[5], tmax = sd.g[2];
  int t_idx = 0;
  if (dt > 0) {
    const double arr = static_cast<double>(ph.t0) + static_cast<double>(gpl_hit);
    if (arr < 0 || arr >= tmax) return false;
    t_idx = static_cast<int>(arr / dt) + 1;
    if (t_idx >= sd.n2) return false;
  }
  const V3<R> so = ph.P.s;
  const R theta = exit_theta(so);
  const R phi = exit_phi(so);
  const int it = static_cast<int>(theta / static_cast<R>(sd.g[3]));
  const int jp = static_cast<int>(phi / static_cast<R>(sd.g[4]));
  if (it < 0 || it >= sd.n0 || jp < 0 || jp >= sd.n1) return false;
  const int64_t bin = static_cast<int64_t>(it) * sd.n1 + jp;
  if (ph.events >= 2) {
    ++cbs_count;
    const unsigned long long slot = atomicAdd(&p.q_count[0], 1ull);
    if (static_cast<int64_t>(slot) >= p.q_cap) {
      atomicAdd(&p.q_count[1], 1ull);
      return true;
    }
    const int64_t c = p.q_cap;
    R* f = p.qf + slot;
    f[F_W * c] = ph.w;
    f[(F_EM + 0) * c] = ph.Em.r;
    f[(F_EM + 1) * c] = ph.Em.i;
    f[(F_EM + 2) * c] = ph.En.r;
    f[(F_EM + 3) * c] = ph.En.i;
#define LMC_PUT_FRAME(B, FR)          \
  f[(B + 0) * c] = FR.m.x;            \
  f[(B + 1) * c] = FR.m.y;            \
  f[(B + 2) * c] = FR.m.z;            \
  f[(B + 3) * c] = FR.n.x;            \
  f[(B + 4) * c] = FR.n.y;            \
  f[(B + 5) * c] = FR.n.z;            \
  f[(B + 6) * c] = FR.s.x;            \
  f[(B + 7) * c] = FR.s.y;            \
  f[(B + 8) * c] = FR.s.z;
    LMC_PUT_FRAME(F_P, ph.P)
    LMC_PUT_FRAME(F_P1, ph.P1)
    LMC_PUT_FRAME(F_PN1, ph.Pn1)
#undef LMC_PUT_FRAME
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      f[(F_T + 2 * k) * c] = ph.T[k].r;
      f[(F_T + 2 * k + 1) * c] = ph.T[k].i;
    }
    f[(F_D + 0) * c] = ph.rn.x - ph.r1.x;
    f[(F_D + 1) * c] = ph.rn.y - ph.r1.y;
    f[(F_D + 2) * c] = ph.rn.z - ph.r1.z;
    if constexpr (kAdj) {
#pragma unroll
      for (int q = 0; q < 3 * kMaxGradLayers; ++q) f[(F_SCORE + q) * c] = score[q];
    }
    int32_t* iq = p.qi + slot;
    iq[I_SENSOR * c] = si;
    iq[I_BIN * c] = static_cast<int32_t>(bin);
    iq[I_TIDX * c] = t_idx;
    iq[I_SP * c] = (ph.sp1 << 16) | ph.spn;
    return true;
  }
  // single scatter: its own reverse path, coh = incoh = |E_f|^2
  const R ws = M<R>::sqrt(ph.w);
  const cx<R> fm = cscale(ph.Em, ws), fn = cscale(ph.En, ws);
  const cx<R> Efx = cadd(cscale(fm, ph.P.m.x), cscale(fn, ph.P.n.x));
  const cx<R> Efy = cadd(cscale(fm, ph.P.m.y), cscale(fn, ph.P.n.y));
  double vals[8];
  cbs_stokes(Efx, Efy, cx<R>{R(0), R(0)}, cx<R>{R(0), R(0)}, vals);
  const int64_t tplane = static_cast<int64_t>(sd.n0) * sd.n1;
  const int64_t plane = tplane * sd.n2;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    deposit<R, RNG, kAdj>(p, sd, v.lgrid, q * plane + bin, vals[q], score);
    if (t_idx >= 1) deposit<R, RNG, kAdj>(p, sd, v.lgrid, q * plane + t_idx * tplane + bin, vals[q], score);
  }
  return true;
}

// PlanarFluenceSensor::process_hit (detector.cpp:780-841), lab x negated.
template <class R, class RNG, bool kAdj>
__device__ bool fluence_hit(const KParams<R>& p, const SceneView<R>& v, const SensorDev& sd, Photon<R, RNG>& ph,
                            V3<R> hp, R gpl_hit, const R* score) {
  const double len_x = sd.g[0], len_y = sd.g[1], len_t = sd.g[2], dx = sd.g[3], dy = sd.g[4], dt = sd.g[5];
  int t_idx = -1;
  if (dt > 0) {
    const double arr = static_cast<double>(ph.t0) + static_cast<double>(gpl_hit);
    if (arr < 0 || arr >= len_t) return false;
    t_idx = static_cast<int>(arr / dt) + 1;
    if (t_idx >= sd.n2) return false;
  }
  const int xi = static_cast<int>((static_cast<double>(hp.x) + 0.5 * len_x) / dx);
  const int yi = static_cast<int>((static_cast<double>(hp.y) + 0.5 * len_y) / dy);
  if (xi < 0 || xi >= sd.n0 || yi < 0 || yi >= sd.n1) return false;
  const R ws = M<R>::sqrt(ph.w);
  const cx<R> fm = cscale(ph.Em, ws), fn = cscale(ph.En, ws);
  const cx<R> Ex = cscale(cadd(cscale(fm, ph.P.m.x), cscale(fn, ph.P.n.x)), R(-1));
  const cx<R> Ey = cadd(cscale(fm, ph.P.m.y), cscale(fn, ph.P.n.y));
  const cx<R> xy = cmul(Ex, cx<R>{Ey.r, -Ey.i});
  const double vals[4] = {static_cast<double>(cnorm(Ex) + cnorm(Ey)), static_cast<double>(cnorm(Ex) - cnorm(Ey)),
                          static_cast<double>(R(2) * xy.r), static_cast<double>(R(2) * xy.i)};
  const int64_t tplane = static_cast<int64_t>(sd.n0) * sd.n1;
  const int64_t plane = tplane * sd.n2;
  const int64_t bin = static_cast<int64_t>(xi) * sd.n1 + yi;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (dt > 0) deposit<R, RNG, kAdj>(p, sd, v.lgrid, q * plane + t_idx * tplane + bin, vals[q], score);
    deposit<R, RNG, kAdj>(p, sd, v.lgrid, q * plane + bin, vals[q], score);
  }
  return true;
}

// PlanarFieldSensor::process_hit (detector.cpp:546-568).  The propagation
// phase exp(i k0 * optical_path) is formed from the FP64 optical path with an
// FP64 argument reduction.
template <class R, class RNG, bool kAdj>
__device__ bool field_hit(const KParams<R>& p, const SceneView<R>& v, const SensorDev& sd, Photon<R, RNG>& ph,
                          V3<R> hp, double opl_hit, const R* score) {
  const double len_x = sd.g[0], len_y = sd.g[1], dx = sd.g[2], dy = sd.g[3];
  const int xi = static_cast<int>((static_cast<double>(hp.x) + 0.5 * len_x) / dx);
  const int yi = static_cast<int>((static_cast<double>(hp.y) + 0.5 * len_y) / dy);
  if (xi < 0 || xi >= sd.n0 || yi < 0 || yi >= sd.n1) return false;
  const double ph_arg = static_cast<double>(p.k0) * opl_hit;
  double sp, cp;
  ::sincos(ph_arg, &sp, &cp);
  const double ws = ::sqrt(static_cast<double>(ph.w));
  const double exr = static_cast<double>(ph.Em.r * ph.P.m.x + ph.En.r * ph.P.n.x);
  const double exi = static_cast<double>(ph.Em.i * ph.P.m.x + ph.En.i * ph.P.n.x);
  const double eyr = static_cast<double>(ph.Em.r * ph.P.m.y + ph.En.r * ph.P.n.y);
  const double eyi = static_cast<double>(ph.Em.i * ph.P.m.y + ph.En.i * ph.P.n.y);
  const double vals[4] = {(exr * cp - exi * sp) * ws, (exr * sp + exi * cp) * ws, (eyr * cp - eyi * sp) * ws,
                          (eyr * sp + eyi * cp) * ws};
  const int64_t plane = static_cast<int64_t>(sd.n0) * sd.n1;
  const int64_t bin = static_cast<int64_t>(xi) * sd.n1 + yi;
  deposit<R, RNG, kAdj>(p, sd, v.lgrid, 2 * bin, vals[0], score);
  deposit<R, RNG, kAdj>(p, sd, v.lgrid, 2 * bin + 1, vals[1], score);
  deposit<R, RNG, kAdj>(p, sd, v.lgrid, 2 * (plane + bin), vals[2], score);
  deposit<R, RNG, kAdj>(p, sd, v.lgrid, 2 * (plane + bin) + 1, vals[3], score);
  return true;
}

// Far-field fluence (no reference counterpart): forward-field Stokes of the
// exit direction on the CBS angular binning.
template <class R, class RNG, bool kAdj>
__device__ bool fffluence_hit(const KParams<R>& p, const SceneView<R>& v, const SensorDev& sd, Photon<R, RNG>& ph,
                              const R* score) {
  const V3<R> so = ph.P.s;
  const R theta = exit_theta(so);
  const R phi = exit_phi(so);
  const int it = static_cast<int>(theta / static_cast<R>(sd.g[2]));
  const int jp = static_cast<int>(phi / static_cast<R>(sd.g[3]));
  if (it < 0 || it >= sd.n0 || jp < 0 || jp >= sd.n1) return false;
  const R ws = M<R>::sqrt(ph.w);
  const cx<R> fm = cscale(ph.Em, ws), fn = cscale(ph.En, ws);
  const cx<R> Ex = cadd(cscale(fm, ph.P.m.x), cscale(fn, ph.P.n.x));
  const cx<R> Ey = cadd(cscale(fm, ph.P.m.y), cscale(fn, ph.P.n.y));
  const cx<R> xy = cmul(Ex, cx<R>{Ey.r, -Ey.i});
  const double vals[4] = {static_cast<double>(cnorm(Ex) + cnorm(Ey)), static_cast<double>(cnorm(Ex) - cnorm(Ey)),
                          static_cast<double>(R(2) * xy.r), static_cast<double>(R(2) * xy.i)};
  const int64_t plane = static_cast<int64_t>(sd.n0) * sd.n1;
  const int64_t bin = static_cast<int64_t>(it) * sd.n1 + jp;
  for (int q = 0; q < 4; ++q) deposit<R, RNG, kAdj>(p, sd, v.lgrid, q * plane + bin, vals[q], score);
  return true;
}

// SensorsGroup::record_hit (detector.cpp:178-270) for the segment prev->pos.
// Returns true when an absorbing sensor accepted the photon.
template <class R, class RNG, bool kAdj>
__device__ bool record_hit(const KParams<R>& p, const SceneView<R>& v, Photon<R, RNG>& ph, V3<R> prev, unsigned& n_cross,
                           unsigned& n_cbs) {
  const R z1 = prev.z, z2 = ph.pos.z;
  if (M<R>::fabs(z2 - z1) < static_cast<R>(1e-12)) return false;
  const int cdir = (z2 > z1) ? LMC_DIR_FORWARD : LMC_DIR_BACKWARD;
  const R zmin = z1 < z2 ? z1 : z2, zmax = z1 < z2 ? z2 : z1;
  bool killed = false;
  for (int pl = 0; pl < p.n_planes; ++pl) {
    const R zp = static_cast<R>(v.planes[pl].z);
    if (!(zp >= zmin && zp <= zmax)) continue;
    ++n_cross;
    const V3<R> d{ph.pos.x - prev.x, ph.pos.y - prev.y, ph.pos.z - prev.z};
    const R t = (zp - prev.z) / d.z;
    const V3<R> hp{prev.x + d.x * t, prev.y + d.y * t, prev.z + d.z * t};
    const V3<R> ov{hp.x - ph.pos.x, hp.y - ph.pos.y, hp.z - ph.pos.z};
    const R corr = M<R>::sqrt(dot(ov, ov));
    const R nloc = v.L[ph.layer].n;
    R gpl_hit = ph.gpl;
    double opl_hit = ph.opl;
    if (corr > R(0)) {
      gpl_hit -= corr;
      opl_hit -= static_cast<double>(corr) * static_cast<double>(nloc);
    }
    R score[3 * kMaxGradLayers];
    if constexpr (kAdj) {
      // the overshoot beyond the plane is not travelled inside the sample
#pragma unroll
      for (int l = 0; l < kMaxGradLayers; ++l) {
        if (l == ph.layer) ph.aL[l] -= corr;
      }
      adj_scores(p, v, ph, score);
#pragma unroll
      for (int l = 0; l < kMaxGradLayers; ++l) {
        if (l == ph.layer) ph.aL[l] += corr;
      }
    }
    const int first = v.planes[pl].first, cnt = v.planes[pl].count;
    for (int si = first; si < first + cnt; ++si) {
      const SensorDev& sd = v.sens[si];
      // check_conditions (detector.cpp:374-417)
      if (sd.f_dir_en && sd.f_dir != LMC_DIR_BOTH && sd.f_dir != cdir) continue;
      if (sd.f_ev_en && (ph.events < sd.f_ev_min || ph.events > sd.f_ev_max)) continue;
      if (sd.f_th_en) {
        const R ct = -ph.P.s.z;
        if (ct < static_cast<R>(sd.cos_th_min) || ct > static_cast<R>(sd.cos_th_max)) continue;
      }
      if (sd.f_ph_en) {
        const R phi = exit_phi(ph.P.s);
        if (phi < static_cast<R>(sd.ph_min) || phi > static_cast<R>(sd.ph_max)) continue;
      }
      if (sd.f_pos_en) {
        if (hp.x < static_cast<R>(sd.x_min) || hp.x > static_cast<R>(sd.x_max)) continue;
        if (hp.y < static_cast<R>(sd.y_min) || hp.y > static_cast<R>(sd.y_max)) continue;
      }
      bool counted = false;
      switch (sd.kind) {
        case LMC_SENSOR_FARFIELD_CBS:
          counted = cbs_hit<R, RNG, kAdj>(p, v, sd, si, ph, gpl_hit, score, n_cbs);
          break;
        case LMC_SENSOR_PLANAR_FLUENCE:
          counted = fluence_hit<R, RNG, kAdj>(p, v, sd, ph, hp, gpl_hit, score);
          break;
        case LMC_SENSOR_PLANAR_FIELD:
          counted = field_hit<R, RNG, kAdj>(p, v, sd, ph, hp, opl_hit, score);
          break;
        case LMC_SENSOR_FARFIELD_FLUENCE:
          counted = fffluence_hit<R, RNG, kAdj>(p, v, sd, ph, score);
          break;
        case LMC_SENSOR_REFLECTANCE: {
          deposit<R, RNG, kAdj>(p, sd, v.lgrid, 0, static_cast<double>(ph.w), score);
          counted = true;
          break;
        }
        default:
          break;
      }
      if (counted) atomicAdd(&v.lhits[si], 1u);
      if (sd.absorb) killed = true;
    }
  }
  return killed;
}

// Laser::emit_photon (laser.cpp:77-99) and sample_emission_time (:42-75).
template <class R, class RNG>
__device__ __forceinline__ void launch(const KParams<R>& p, Photon<R, RNG>& ph, uint64_t idx) {
  ph.rng.init(p.seed, p.key1, idx);
  double x = p.lpos[0], y = p.lpos[1];
  const double z = p.lpos[2];
  const double kTwoPi = 2.0 * 3.14159265358979323846;
  double t0 = p.t_off * 1e-3;
  const double fwhm_to_sigma = 1.0 / (2.0 * 1.1774100225154747);  // 1/(2 sqrt(2 ln 2))
  if constexpr (RNG::kSequential) {
    if (p.source == LMC_SOURCE_UNIFORM) {
      const double th = ph.rng.uniform() * kTwoPi;
      const double r = ::sqrt(ph.rng.uniform()) * p.sigma;
      x += r * ::cos(th);
      y += r * ::sin(th);
    } else if (p.source == LMC_SOURCE_GAUSSIAN) {
      x = ph.rng.normal(x, p.sigma);
      y = ph.rng.normal(y, p.sigma);
    }
    switch (p.temporal) {
      case LMC_TIME_GAUSSIAN: t0 = ph.rng.normal(t0, p.pulse * fwhm_to_sigma * 1e-3); break;
      case LMC_TIME_TOPHAT: t0 = t0 + (ph.rng.uniform() - 0.5) * p.pulse * 1e-3; break;
      case LMC_TIME_EXPONENTIAL: t0 = t0 - ::log(ph.rng.uniform()) * p.pulse * 1e-3; break;
      case LMC_TIME_PULSE_TRAIN: if (p.rep > 0.0) t0 = t0 + ph.rng.uniform() * (1e9 / p.rep); break;
      case LMC_TIME_CW: if (p.pulse > 0.0) t0 = t0 + ph.rng.uniform() * p.pulse * 1e-3; break;
      default: break;
    }
  } else {
    if (p.source != LMC_SOURCE_POINT) {
      R a, b, c, d;
      ph.rng.emit_block(0u, a, b, c, d);
      if (p.source == LMC_SOURCE_UNIFORM) {
        const double r = ::sqrt(static_cast<double>(b)) * p.sigma;
        double s, co;
        ::sincos(static_cast<double>(a) * kTwoPi, &s, &co);
        x += r * co;
        y += r * s;
      } else {
        x += static_cast<double>(ph.rng.normal01(a, b)) * p.sigma;
        y += static_cast<double>(ph.rng.normal01(c, d)) * p.sigma;
      }
    }
    if (p.temporal != LMC_TIME_DELTA) {
      R a, b, c, d;
      ph.rng.emit_block(1u, a, b, c, d);
      const double ua = static_cast<double>(a);
      switch (p.temporal) {
        case LMC_TIME_GAUSSIAN:
          t0 = t0 + static_cast<double>(ph.rng.normal01(a, b)) * p.pulse * fwhm_to_sigma * 1e-3;
          break;
        case LMC_TIME_TOPHAT: t0 = t0 + (ua - 0.5) * p.pulse * 1e-3; break;
        case LMC_TIME_EXPONENTIAL: t0 = t0 - ::log(ua) * p.pulse * 1e-3; break;
        case LMC_TIME_PULSE_TRAIN: if (p.rep > 0.0) t0 = t0 + ua * (1e9 / p.rep); break;
        case LMC_TIME_CW: if (p.pulse > 0.0) t0 = t0 + ua * p.pulse * 1e-3; break;
        default: break;
      }
    }
  }
  ph.pos = {static_cast<R>(x), static_cast<R>(y), static_cast<R>(z)};
  ph.P.m = {p.lm[0], p.lm[1], p.lm[2]};
  ph.P.n = {p.ln[0], p.ln[1], p.ln[2]};
  ph.P.s = {p.ldir[0], p.ldir[1], p.ldir[2]};
  ph.Em = {p.pol[0], p.pol[1]};
  ph.En = {p.pol[2], p.pol[3]};
  ph.w = R(1);
  ph.gpl = R(0);
  ph.t0 = static_cast<R>(t0);
  ph.opl = 0.0;
  ph.events = 0;
  ph.layer = p.launch_layer;
  ph.alive = true;
  ph.P1 = ph.P;
  ph.Pn1 = ph.P;
  ph.r1 = ph.pos;
  ph.rn = ph.pos;
  ph.T[0] = {R(1), R(0)};
  ph.T[1] = {R(0), R(0)};
  ph.T[2] = {R(0), R(0)};
  ph.T[3] = {R(1), R(0)};
  ph.Tb[0] = ph.T[0];
  ph.Tb[1] = ph.T[1];
  ph.Tb[2] = ph.T[2];
  ph.Tb[3] = ph.T[3];
  ph.sp1 = 0;
  ph.spn = 0;
#pragma unroll
  for (int l = 0; l < kMaxGradLayers; ++l) {
    ph.aL[l] = R(0);
    ph.aN[l] = R(0);
    ph.aG[l] = R(0);
  }
}

template <class R, class RNG, bool kAdj>
__device__ __forceinline__ void add_path(Photon<R, RNG>& ph, R d) {
  if constexpr (kAdj) {
#pragma unroll
    for (int l = 0; l < kMaxGradLayers; ++l)
      if (l == ph.layer) ph.aL[l] += d;
  }
}

// One iteration of run_photon's loop (simulation.cpp:297-614).
template <class R, class RNG, bool kTrack, bool kAdj>
__device__ __forceinline__ void transport_event(const KParams<R>& p, const SceneView<R>& v, Photon<R, RNG>& ph,
                                                unsigned& n_trials, unsigned& n_cross, unsigned& n_cbs) {
  LayerDev<R> Ld = v.L[ph.layer];
  ph.rng.begin_event(static_cast<uint32_t>(ph.events));
  R step = -Ld.mfp * M<R>::log(static_cast<R>(ph.rng.free_path()));
  const V3<R> dir = ph.P.s;
  R new_z = ph.pos.z + dir.z * step;
  // interface crossings (simulation.cpp:317-383), with the layer index tracked
  // explicitly (the reference nudges z by 1e-12 and re-derives it)
  for (;;) {
    R zi = R(0);
    bool has = false;
    if (new_z > ph.pos.z) {
      if (ph.layer + 1 < p.n_layers) {
        zi = Ld.z_max;
        has = zi <= new_z;
      }
    } else if (new_z < ph.pos.z) {
      if (ph.layer > 0) {
        zi = Ld.z_min;
        has = zi >= new_z;
      }
    }
    if (!has) break;
    const R dzt = dir.z * step;
    if (M<R>::fabs(dzt) < static_cast<R>(1e-15)) break;
    const R t = (zi - ph.pos.z) / dzt;
    if (t < R(0) || t > R(1)) break;
    const R step_to = step * t;
    const R rem = (step - step_to) * Ld.mu_t;
    ph.gpl += step_to;
    ph.opl += static_cast<double>(Ld.n) * static_cast<double>(step_to);
    add_path<R, RNG, kAdj>(ph, step_to);
    const V3<R> prev = ph.pos;
    ph.pos.x += dir.x * step_to;
    ph.pos.y += dir.y * step_to;
    const int up = dir.z > R(0);
    if constexpr (sizeof(R) == 8)
      ph.pos.z = zi + (up ? 1e-12 : -1e-12);
    else
      ph.pos.z = zi;
    if (record_hit<R, RNG, kAdj>(p, v, ph, prev, n_cross, n_cbs)) {
      ph.alive = false;
      return;
    }
    ph.layer += up ? 1 : -1;
    Ld = v.L[ph.layer];
    step = Ld.mu_t > R(0) ? rem / Ld.mu_t : rem;
    new_z = ph.pos.z + dir.z * step;
  }
  ph.gpl += step;
  ph.opl += static_cast<double>(Ld.n) * static_cast<double>(step);
  add_path<R, RNG, kAdj>(ph, step);
  const V3<R> prev = ph.pos;
  ph.pos.x += dir.x * step;
  ph.pos.y += dir.y * step;
  ph.pos.z += dir.z * step;
  if (record_hit<R, RNG, kAdj>(p, v, ph, prev, n_cross, n_cbs)) {
    ph.alive = false;
    return;
  }
  // Sample::is_inside (sample.cpp:268-288)
  if (!(ph.pos.z >= R(0)) || (p.top_finite && ph.pos.z >= p.z_top) || !isfinite(ph.pos.x) ||
      !isfinite(ph.pos.y) || !isfinite(ph.pos.z)) {
    ph.alive = false;
    return;
  }
  // species of this vertex (MixtureLayer::select_scatter_medium, sample.cpp:96-104)
  int sp = Ld.sp_first;
  if (Ld.sp_count > 1) {
    const R u = static_cast<R>(ph.rng.species());
    const R* c = v.sel + Ld.sel_off;
    int lo = 0, hi = Ld.sp_count;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (c[mid] < u)
        lo = mid + 1;
      else
        hi = mid;
    }
    sp += lo < Ld.sp_count ? lo : Ld.sp_count - 1;
  }
  const SpeciesDev<R> S = v.S[sp];
  // theta (medium.cpp:51-58 -> phase.cpp)
  const R uth = static_cast<R>(ph.rng.theta());
  R theta, ct, st;
  if (S.kind == LMC_PHASE_HG) {  // phase.cpp:108-115
    const R g = S.hg_g, gg = g * g;
    const R frac = (R(1) - gg) / (R(1) - g + R(2) * g * uth);
    R c = (R(1) + gg - frac * frac) / (R(2) * g);
    if constexpr (sizeof(R) == 4) c = clamp1(c);
    theta = M<R>::acos(c);
    if constexpr (kAdj) {
#pragma unroll
      for (int l = 0; l < kMaxGradLayers; ++l)
        if (l == ph.layer) ph.aG[l] += -R(2) * g / (R(1) - gg) - R(3) * (g - c) / (R(1) + gg - R(2) * g * c);
    }
  } else if (S.kind == LMC_PHASE_TABLE_THETA) {
    theta = table_sample(v, S, uth);
  } else if (S.kind == LMC_PHASE_TABLE_COS) {
    theta = M<R>::acos(table_sample(v, S, uth));
  } else {
    theta = M<R>::acos(R(2) * uth - R(1));
  }
  M<R>::sincos(theta, &st, &ct);
  cx<R> s1, s2;
  s_lookup(v, S, theta, s1, s2);
  // azimuth by rejection (phase.cpp:27-49)
  const R s2sq = cnorm(s2), s1sq = cnorm(s1);
  const R e1sq = cnorm(ph.Em), e2sq = cnorm(ph.En);
  const R e12 = ph.Em.r * ph.En.r + ph.Em.i * ph.En.i;
  const R a = s2sq * e1sq + s1sq * e2sq;
  const R b = s1sq * e1sq + s2sq * e2sq;
  const R c = R(2) * (s2sq - s1sq) * e12;
  const R diff = a - b;
  const R Fmax = R(0.5) * (a + b) + R(0.5) * M<R>::sqrt(diff * diff + c * c);
  R cph = R(1), sph = R(0), F = R(0);
  for (int trial = 0; trial < 256; ++trial) {
    R uphi, uacc;
    ph.rng.phi_trial(uphi, uacc);
    ++n_trials;
    R sp_, cp_;
    M<R>::sincos_rev(uphi, &sp_, &cp_);
    const R Ft = a * cp_ * cp_ + b * sp_ * sp_ + c * cp_ * sp_;
    const R Fc = Ft > R(0) ? Ft : R(0);
    if (uacc * Fmax <= Fc) {
      cph = cp_;
      sph = sp_;
      F = Ft;
      break;
    }
  }
  // frame update P <- A(theta, phi) P (simulation.cpp:438-453)
  const Frame<R> Po = ph.P;
  const R ctc = ct * cph, cts = ct * sph, stc = st * cph, sts = st * sph;
  ph.P.m = {ctc * Po.m.x + cts * Po.n.x - st * Po.s.x, ctc * Po.m.y + cts * Po.n.y - st * Po.s.y,
            ctc * Po.m.z + cts * Po.n.z - st * Po.s.z};
  ph.P.n = {-sph * Po.m.x + cph * Po.n.x, -sph * Po.m.y + cph * Po.n.y, -sph * Po.m.z + cph * Po.n.z};
  ph.P.s = {stc * Po.m.x + sts * Po.n.x + ct * Po.s.x, stc * Po.m.y + sts * Po.n.y + ct * Po.s.y,
            stc * Po.m.z + sts * Po.n.z + ct * Po.s.z};
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
    if (evt == 1) {
      ph.r1 = ph.pos;
      ph.sp1 = sp;
      ph.P1 = ph.P;
    }
    ph.rn = ph.pos;
    ph.spn = sp;
    ph.Pn1 = Po;
    if (evt >= 2) {
      const cx<R> J0 = cscale(s2, cph), J1 = cscale(s2, sph), J2 = cscale(s1, -sph), J3 = cscale(s1, cph);
      if (evt > 2) {
        // T <- Tb * T, Frobenius-normalised
        const cx<R> t0 = cadd(cmul(ph.Tb[0], ph.T[0]), cmul(ph.Tb[1], ph.T[2]));
        const cx<R> t1 = cadd(cmul(ph.Tb[0], ph.T[1]), cmul(ph.Tb[1], ph.T[3]));
        const cx<R> t2 = cadd(cmul(ph.Tb[2], ph.T[0]), cmul(ph.Tb[3], ph.T[2]));
        const cx<R> t3 = cadd(cmul(ph.Tb[2], ph.T[1]), cmul(ph.Tb[3], ph.T[3]));
        const R fr2 = cnorm(t0) + cnorm(t1) + cnorm(t2) + cnorm(t3);
        const R inv = fr2 > R(0) ? M<R>::rsqrt(fr2) : R(1);
        ph.T[0] = cscale(t0, inv);
        ph.T[1] = cscale(t1, inv);
        ph.T[2] = cscale(t2, inv);
        ph.T[3] = cscale(t3, inv);
      }
      ph.Tb[0] = J0;
      ph.Tb[1] = J1;
      ph.Tb[2] = J2;
      ph.Tb[3] = J3;
    }
  }
  if constexpr (kAdj) {
#pragma unroll
    for (int l = 0; l < kMaxGradLayers; ++l)
      if (l == ph.layer) ph.aN[l] += R(1);
  }
  // weight, roulette, cap (simulation.cpp:578-613)
  ph.events++;
  if (Ld.mu_t > R(0)) ph.w *= Ld.albedo;
  if (ph.w < static_cast<R>(1e-4)) {
    if (static_cast<R>(ph.rng.roulette()) < static_cast<R>(0.1))
      ph.w /= static_cast<R>(0.1);
    else {
      ph.alive = false;
      return;
    }
  }
  if (ph.events > p.max_events) ph.alive = false;
}

template <class R, class RNG, bool kTrack, bool kLdsTables, bool kAdj>
__global__ void __launch_bounds__(256) k_transport(KParams<R> p) {
  extern __shared__ __align__(16) unsigned char smem[];
  // ---- stage scene constants (and tables) in LDS
  unsigned char* q = smem;
  auto carve = [&](size_t bytes) {
    unsigned char* r = q;
    q += (bytes + 15) & ~static_cast<size_t>(15);
    return r;
  };
  LayerDev<R>* sL = reinterpret_cast<LayerDev<R>*>(carve(sizeof(LayerDev<R>) * p.n_layers));
  SpeciesDev<R>* sS = reinterpret_cast<SpeciesDev<R>*>(carve(sizeof(SpeciesDev<R>) * p.n_species));
  SensorDev* sSen = reinterpret_cast<SensorDev*>(carve(sizeof(SensorDev) * p.n_sensors));
  PlaneDev* sPl = reinterpret_cast<PlaneDev*>(carve(sizeof(PlaneDev) * p.n_planes));
  unsigned* sHits = reinterpret_cast<unsigned*>(carve(sizeof(unsigned) * kMaxSensors));
  double* sGrid = reinterpret_cast<double*>(carve(sizeof(double) * p.lds_grid));
  const int tid = threadIdx.x, nt = blockDim.x;
  auto copy = [&](void* dst, const void* src, size_t bytes) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    for (size_t i = tid; i < bytes / 4; i += nt) d[i] = s[i];
  };
  copy(sL, p.layers, sizeof(LayerDev<R>) * p.n_layers);
  copy(sS, p.species, sizeof(SpeciesDev<R>) * p.n_species);
  copy(sSen, p.sensors, sizeof(SensorDev) * p.n_sensors);
  copy(sPl, p.planes, sizeof(PlaneDev) * p.n_planes);
  for (int i = tid; i < kMaxSensors; i += nt) sHits[i] = 0u;
  for (int i = tid; i < p.lds_grid; i += nt) sGrid[i] = 0.0;
  SceneView<R> v;
  v.L = sL;
  v.S = sS;
  v.sens = sSen;
  v.planes = sPl;
  v.lgrid = sGrid;
  v.lhits = sHits;
  if constexpr (kLdsTables) {
    R* tc = reinterpret_cast<R*>(carve(sizeof(R) * p.cdf_total));
    uint16_t* tg = reinterpret_cast<uint16_t*>(carve(sizeof(uint16_t) * p.guide_total + 4));
    STab<R>* ts = reinterpret_cast<STab<R>*>(carve(sizeof(STab<R>) * p.stab_total));
    R* tsel = reinterpret_cast<R*>(carve(sizeof(R) * p.sel_total));
    copy(tc, p.cdf, sizeof(R) * p.cdf_total);
    for (int i = tid; i < p.guide_total; i += nt) tg[i] = p.guide[i];
    copy(ts, p.stab, sizeof(STab<R>) * p.stab_total);
    copy(tsel, p.sel, sizeof(R) * p.sel_total);
    v.cdf = tc;
    v.guide = tg;
    v.stab = ts;
    v.sel = tsel;
  } else {
    v.cdf = p.cdf;
    v.guide = p.guide;
    v.stab = p.stab;
    v.sel = p.sel;
  }
  __syncthreads();

  // ---- persistent photon loop with wave-level refill
  Photon<R, RNG> ph;
  ph.alive = false;
  unsigned n_ph = 0, n_ev = 0, n_tr = 0, n_cr = 0, n_cbs = 0;
  uint64_t pool_next = 0, pool_end = 0;
  bool exhausted = false;
  const int lane = __lane_id();
  for (;;) {
    const uint64_t dead = __ballot(!ph.alive);
    if (dead != 0 && !exhausted) {
      if (pool_next >= pool_end) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(p.counter, static_cast<unsigned long long>(p.claim));
        base = __shfl(base, 0);
        if (base >= p.n_photons) {
          exhausted = true;
          pool_next = pool_end = p.n_photons;
        } else {
          pool_next = base;
          pool_end = base + p.claim < p.n_photons ? base + p.claim : p.n_photons;
        }
      }
      const uint64_t below = dead & ((1ull << lane) - 1ull);
      const uint64_t my = pool_next + static_cast<uint64_t>(__popcll(below));
      if (!ph.alive && my < pool_end) {
        launch<R, RNG>(p, ph, p.offset + my);
        ++n_ph;
      }
      const uint64_t nd = static_cast<uint64_t>(__popcll(dead));
      const uint64_t avail = pool_end - pool_next;
      pool_next += nd < avail ? nd : avail;
    }
    if (__ballot(ph.alive) == 0) {
      if (exhausted) break;
      continue;
    }
    if (ph.alive) {
      const int32_t ev0 = ph.events;
      transport_event<R, RNG, kTrack, kAdj>(p, v, ph, n_tr, n_cr, n_cbs);
      n_ev += static_cast<unsigned>(ph.events - ev0);
    }
  }
  // ---- flush counters and LDS accumulators
  atomicAdd(&p.stats[0], static_cast<unsigned long long>(n_ph));
  atomicAdd(&p.stats[1], static_cast<unsigned long long>(n_ev));
  atomicAdd(&p.stats[2], static_cast<unsigned long long>(n_tr));
  atomicAdd(&p.stats[3], static_cast<unsigned long long>(n_cr));
  atomicAdd(&p.stats[4], static_cast<unsigned long long>(n_cbs));
  __syncthreads();
  for (int i = tid; i < p.n_sensors; i += nt)
    if (sHits[i]) atomicAdd(&p.hits[i], static_cast<unsigned long long>(sHits[i]));
  for (int si = 0; si < p.n_sensors; ++si) {
    const SensorDev& sd = sSen[si];
    if (!sd.lds) continue;
    const int64_t len = (sd.kind == LMC_SENSOR_FARFIELD_CBS)
                            ? 8ll * sd.n0 * sd.n1 * sd.n2
                            : (sd.kind == LMC_SENSOR_REFLECTANCE ? 1ll : 4ll * sd.n0 * sd.n1);
    for (int64_t i = tid; i < len; i += nt) {
      const double val = sGrid[sd.lds_off + i];
      if (val != 0.0) unsafeAtomicAdd(p.out + sd.out_off + i, val);
    }
  }
}

// Reverse-path kernel: one thread per queued CBS detection.  Rebuilds the
// time-reversed partner's Jones vector, applies the CBS path phase
// exp(i k0 (s_out + s_in).(r_n - r_1)) (detector.cpp:1277-1310) and deposits
// the coherent/incoherent Stokes terms (accumulate_stokes, :1194-1226), into
// a per-workgroup LDS copy for small grids or straight to HBM otherwise.
template <class R, bool kAdj>
__global__ void __launch_bounds__(256) k_cbs_records(KParams<R> p) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* sgrid = reinterpret_cast<double*>(smem);
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < p.lds_grid; i += nt) sgrid[i] = 0.0;
  __syncthreads();
  const int64_t n_rec = static_cast<int64_t>(p.q_count[0]) < p.q_cap ? static_cast<int64_t>(p.q_count[0]) : p.q_cap;
  const int64_t c = p.q_cap;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * nt + tid; i < n_rec; i += static_cast<int64_t>(gridDim.x) * nt) {
    const R* f = p.qf + i;
    const int32_t* iq = p.qi + i;
    const int si = iq[I_SENSOR * c];
    const int64_t bin = iq[I_BIN * c];
    const int t_idx = iq[I_TIDX * c];
    const int spk = iq[I_SP * c];
    const SensorDev sd = p.sensors[si];
    const SpeciesDev<R> sp1 = p.species[spk >> 16];
    const SpeciesDev<R> spn = p.species[spk & 0xFFFF];
    Frame<R> P, P1, Pq;
#define LMC_GET_FRAME(B, FR)                                                          \
  FR.m = {f[(B + 0) * c], f[(B + 1) * c], f[(B + 2) * c]};                          \
  FR.n = {f[(B + 3) * c], f[(B + 4) * c], f[(B + 5) * c]};                          \
  FR.s = {f[(B + 6) * c], f[(B + 7) * c], f[(B + 8) * c]};
    LMC_GET_FRAME(F_P, P)
    LMC_GET_FRAME(F_P1, P1)
    LMC_GET_FRAME(F_PN1, Pq)
#undef LMC_GET_FRAME
    cx<R> T[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) T[k] = {f[(F_T + 2 * k) * c], f[(F_T + 2 * k + 1) * c]};
    const R w = f[F_W * c];
    const cx<R> Em{f[(F_EM + 0) * c], f[(F_EM + 1) * c]}, En{f[(F_EM + 2) * c], f[(F_EM + 3) * c]};
    const V3<R> d{f[(F_D + 0) * c], f[(F_D + 1) * c], f[(F_D + 2) * c]};
    R score[3 * kMaxGradLayers];
    if constexpr (kAdj) {
#pragma unroll
      for (int q = 0; q < 3 * kMaxGradLayers; ++q) score[q] = f[(F_SCORE + q) * c];
    }
    cx<R> rm, rn;
    reverse_field(p, p.stab, spn, sp1, P1, Pq, P, T, rm, rn);
    const R ws = M<R>::sqrt(w);
    const R arg = p.k0 * ((P.s.x + p.ldir[0]) * d.x + (P.s.y + p.ldir[1]) * d.y + (P.s.z + p.ldir[2]) * d.z);
    R sa, ca;
    M<R>::sincos(arg, &sa, &ca);
    const cx<R> pp{ca * ws, sa * ws};
    rm = cmul(rm, pp);
    rn = cmul(rn, pp);
    const cx<R> fm = cscale(Em, ws), fn = cscale(En, ws);
    const cx<R> Efx = cadd(cscale(fm, P.m.x), cscale(fn, P.n.x));
    const cx<R> Efy = cadd(cscale(fm, P.m.y), cscale(fn, P.n.y));
    const cx<R> Erx = cadd(cscale(rm, P.m.x), cscale(rn, P.n.x));
    const cx<R> Ery = cadd(cscale(rm, P.m.y), cscale(rn, P.n.y));
    double vals[8];
    cbs_stokes(Efx, Efy, Erx, Ery, vals);
    const int64_t tplane = static_cast<int64_t>(sd.n0) * sd.n1;
    const int64_t plane = tplane * sd.n2;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      deposit<R, PhiloxStream<R>, kAdj>(p, sd, sgrid, q * plane + bin, vals[q], score);
      if (t_idx >= 1) deposit<R, PhiloxStream<R>, kAdj>(p, sd, sgrid, q * plane + t_idx * tplane + bin, vals[q], score);
    }
  }
  __syncthreads();
  for (int si = 0; si < p.n_sensors; ++si) {
    const SensorDev sd = p.sensors[si];
    if (!sd.lds || sd.kind != LMC_SENSOR_FARFIELD_CBS) continue;
    const int64_t len = 8ll * sd.n0 * sd.n1 * sd.n2;
    for (int64_t i = tid; i < len; i += nt) {
      const double val = sgrid[sd.lds_off + i];
      if (val != 0.0) unsafeAtomicAdd(p.out + sd.out_off + i, val);
    }
  }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_philox_probe(int n, const uint4* ctr, uint32_t k0, uint32_t k1, uint4* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = philox4x32_10(ctr[i], k0, k1);
}

}  // namespace lmc

// =================================================================== host
namespace {

int g_device_init = -1;
int g_num_cu = 0;
size_t g_max_lds_block = 0;  // dynamic LDS one workgroup may use
size_t g_max_lds_cu = 0;     // LDS of one CU

int ensure_device() {
  if (g_device_init == 1) return LMC_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(LMC_ERR_DEVICE, "no HIP device available");
  hipDeviceProp_t prop;
  LMC_HIP(hipGetDeviceProperties(&prop, 0));
  g_num_cu = prop.multiProcessorCount;
  g_max_lds_block = prop.sharedMemPerBlock;
  g_max_lds_cu = prop.maxSharedMemoryPerMultiProcessor;
  if (g_max_lds_cu < g_max_lds_block) g_max_lds_cu = g_max_lds_block;
  g_device_init = 1;
  return LMC_OK;
}

template <class T> struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return LMC_OK;
    if (hipMalloc(&p, sizeof(T) * count) != hipSuccess) {
      p = nullptr;
      return fail(LMC_ERR_NOMEM, "hipMalloc failed");
    }
    return LMC_OK;
  }
  int upload(const T* h, size_t count) {
    int rc = alloc(count);
    if (rc) return rc;
    if (count && hipMemcpy(p, h, sizeof(T) * count, hipMemcpyHostToDevice) != hipSuccess)
      return fail(LMC_ERR_DEVICE, "hipMemcpy H2D failed");
    return LMC_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

// Precision-specific device copy of a scene's tables.
template <class R> struct DevTables {
  bool ready = false;
  DBuf<LayerDev<R>> layers;
  DBuf<SpeciesDev<R>> species;
  DBuf<R> cdf, sel;
  DBuf<uint16_t> guide;
  DBuf<STab<R>> stab;
  int cdf_total = 0, guide_total = 0, stab_total = 0, sel_total = 0;
  size_t table_bytes = 0;
};

}  // namespace

struct lmc_scene {
  // host copies of the descriptor
  std::vector<lmc_layer_desc> layers;
  std::vector<lmc_species_desc> species;
  std::vector<lmc_sensor_desc> sensors;
  std::vector<double> tables;
  lmc_laser_desc laser{};
  double n_medium = 1.0;
  int64_t n_out = 0;
  int launch_layer = 0;
  // device
  DevTables<float> t32;
  DevTables<double> t64;
  DBuf<SensorDev> sensors_dev;
  DBuf<PlaneDev> planes_dev;
  std::vector<SensorDev> sensors_host;
  std::vector<PlaneDev> planes_host;
  int lds_grid = 0;
  DBuf<double> out;
  DBuf<unsigned long long> hits;
  DBuf<unsigned long long> counter;  // [0] work counter, [1..5] stats
  DBuf<double> grad;
  DBuf<unsigned char> queue;        // CBS record queue (SoA)
  DBuf<unsigned long long> qcount;  // [0] written, [1] overflowed
  int64_t q_cap = 0;
  size_t q_rec_bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
};

namespace {

constexpr int kGuideN = 4096;

template <class R> int build_tables(lmc_scene* s, DevTables<R>& T) {
  if (T.ready) return LMC_OK;
  std::vector<SpeciesDev<R>> sp(s->species.size());
  std::vector<R> cdf;
  std::vector<uint16_t> guide;
  std::vector<STab<R>> stab;
  std::vector<R> sel;
  for (size_t i = 0; i < s->species.size(); ++i) {
    const lmc_species_desc& d = s->species[i];
    SpeciesDev<R>& o = sp[i];
    std::memset(&o, 0, sizeof(o));
    o.kind = d.phase_kind;
    o.cdf_min = static_cast<R>(d.cdf_min);
    o.cdf_step = static_cast<R>(d.cdf_step);
    o.hg_g = static_cast<R>(d.hg_g);
    if (d.phase_kind == LMC_PHASE_TABLE_THETA || d.phase_kind == LMC_PHASE_TABLE_COS) {
      if (d.cdf_n < 2 || d.cdf_n > 65535) return fail(LMC_ERR_INVALID, "cdf_n must be in [2, 65535]");
      o.cdf_n = d.cdf_n;
      o.cdf_off = static_cast<int32_t>(cdf.size());
      const double* c = s->tables.data() + d.cdf_offset;
      std::vector<R> cr(d.cdf_n);
      for (int j = 0; j < d.cdf_n; ++j) cr[j] = static_cast<R>(c[j]);
      cdf.insert(cdf.end(), cr.begin(), cr.end());
      // guide g[j] = lower_bound(cr, j/G) on the converted values; g[G] = n-1
      o.guide_off = static_cast<int32_t>(guide.size());
      o.guide_n = kGuideN;
      int idx = 0;
      for (int j = 0; j < kGuideN; ++j) {
        const R thr = static_cast<R>(static_cast<double>(j) / kGuideN);
        while (idx < d.cdf_n - 1 && cr[idx] < thr) ++idx;
        guide.push_back(static_cast<uint16_t>(idx));
      }
      guide.push_back(static_cast<uint16_t>(d.cdf_n - 1));
    }
    if (d.s_n < 2) return fail(LMC_ERR_INVALID, "species S table needs >= 2 nodes");
    o.s_n = d.s_n;
    o.s_off = static_cast<int32_t>(stab.size());
    o.inv_ds = static_cast<R>(1.0 / d.s_dtheta);
    for (int j = 0; j < d.s_n; ++j) {
      const double* a = s->tables.data() + d.s1_offset + 2 * j;
      const double* b = s->tables.data() + d.s2_offset + 2 * j;
      STab<R> e;
      if constexpr (sizeof(R) == 4) {
        e.v = make_float4(static_cast<float>(a[0]), static_cast<float>(a[1]), static_cast<float>(b[0]),
                          static_cast<float>(b[1]));
      } else {
        e.a = make_double2(a[0], a[1]);
        e.b = make_double2(b[0], b[1]);
      }
      stab.push_back(e);
    }
  }
  std::vector<LayerDev<R>> L(s->layers.size());
  for (size_t i = 0; i < s->layers.size(); ++i) {
    const lmc_layer_desc& d = s->layers[i];
    LayerDev<R>& o = L[i];
    std::memset(&o, 0, sizeof(o));
    o.z_min = static_cast<R>(d.z_min);
    o.z_max = static_cast<R>(d.z_max);
    o.mfp = static_cast<R>(d.mfp);
    o.mu_t = static_cast<R>(d.mu_t);
    o.albedo = static_cast<R>(d.mu_t > 0 ? d.mu_s / d.mu_t : 1.0);
    o.n = static_cast<R>(d.n);
    o.mu_s = static_cast<R>(d.mu_s);
    o.mu_a = static_cast<R>(d.mu_a);
    o.sp_first = d.species_first;
    o.sp_count = d.species_count;
    o.sel_off = static_cast<int32_t>(sel.size());
    if (d.species_count > 1)
      for (int j = 0; j < d.species_count; ++j) sel.push_back(static_cast<R>(s->tables[d.select_offset + j]));
  }
  int rc;
  if ((rc = T.layers.upload(L.data(), L.size()))) return rc;
  if ((rc = T.species.upload(sp.data(), sp.size()))) return rc;
  if (cdf.empty()) cdf.push_back(R(1));
  if (guide.empty()) guide.assign(2, 0);
  if (sel.empty()) sel.push_back(R(1));
  if ((rc = T.cdf.upload(cdf.data(), cdf.size()))) return rc;
  if ((rc = T.guide.upload(guide.data(), guide.size()))) return rc;
  if ((rc = T.stab.upload(stab.data(), stab.size()))) return rc;
  if ((rc = T.sel.upload(sel.data(), sel.size()))) return rc;
  T.cdf_total = static_cast<int>(cdf.size());
  T.guide_total = static_cast<int>(guide.size());
  T.stab_total = static_cast<int>(stab.size());
  T.sel_total = static_cast<int>(sel.size());
  T.table_bytes = sizeof(R) * cdf.size() + sizeof(uint16_t) * guide.size() + 4 + sizeof(STab<R>) * stab.size() +
                  sizeof(R) * sel.size() + 64;
  T.ready = true;
  return LMC_OK;
}

int validate(const lmc_scene_desc* d) {
  if (!d) return fail(LMC_ERR_INVALID, "null scene descriptor");
  if (d->n_layers <= 0 || d->n_layers > kMaxLayers)
    return fail(LMC_ERR_INVALID, "scene must hold 1..8 layers");
  if (d->n_species <= 0 || d->n_species > kMaxSpecies)
    return fail(LMC_ERR_INVALID, "scene must hold 1..64 species");
  if (d->n_sensors < 0 || d->n_sensors > kMaxSensors) return fail(LMC_ERR_INVALID, "scene must hold 0..8 sensors");
  if (!d->layers || !d->species || (d->n_sensors && !d->sensors) || (d->n_tables && !d->tables))
    return fail(LMC_ERR_INVALID, "null descriptor array");
  for (int i = 0; i < d->n_layers; ++i) {
    const lmc_layer_desc& L = d->layers[i];
    if (!(L.z_max > L.z_min)) return fail(LMC_ERR_INVALID, "layer z_max must exceed z_min");
    if (i == 0 && L.z_min != 0.0) return fail(LMC_ERR_INVALID, "first layer must start at z = 0");
    if (i > 0 && L.z_min != d->layers[i - 1].z_max) return fail(LMC_ERR_INVALID, "layers must be contiguous");
    if (i + 1 < d->n_layers && std::isinf(L.z_max)) return fail(LMC_ERR_INVALID, "only the last layer may be infinite");
    if (L.species_first < 0 || L.species_count < 1 || L.species_first + L.species_count > d->n_species)
      return fail(LMC_ERR_INVALID, "layer species range out of bounds");
    if (L.species_count > 1 && (L.select_offset < 0 || L.select_offset + L.species_count > d->n_tables))
      return fail(LMC_ERR_INVALID, "mixture selection CDF out of bounds");
  }
  for (int i = 0; i < d->n_species; ++i) {
    const lmc_species_desc& S = d->species[i];
    const bool tab = S.phase_kind == LMC_PHASE_TABLE_THETA || S.phase_kind == LMC_PHASE_TABLE_COS;
    if (S.phase_kind < 0 || S.phase_kind > LMC_PHASE_UNIFORM) return fail(LMC_ERR_INVALID, "unknown phase kind");
    if (tab && (S.cdf_offset < 0 || S.cdf_offset + S.cdf_n > d->n_tables))
      return fail(LMC_ERR_INVALID, "species CDF out of bounds");
    if (S.phase_kind == LMC_PHASE_HG && S.hg_g == 0.0)
      return fail(LMC_ERR_INVALID, "Henyey-Greenstein sampling needs g != 0");
    if (S.s1_offset < 0 || S.s1_offset + 2 * S.s_n > d->n_tables || S.s2_offset < 0 ||
        S.s2_offset + 2 * S.s_n > d->n_tables || !(S.s_dtheta > 0))
      return fail(LMC_ERR_INVALID, "species S table out of bounds");
  }
  for (int i = 0; i < d->n_sensors; ++i) {
    const lmc_sensor_desc& s = d->sensors[i];
    if (s.out_offset < 0 || s.out_offset + s.out_len > d->n_out)
      return fail(LMC_ERR_INVALID, "sensor accumulators out of bounds");
    int64_t need = 0;
    switch (s.kind) {
      case LMC_SENSOR_FARFIELD_CBS: need = 8ll * s.n[0] * s.n[1] * s.n[2]; break;
      case LMC_SENSOR_PLANAR_FLUENCE: need = 4ll * s.n[0] * s.n[1] * s.n[2]; break;
      case LMC_SENSOR_PLANAR_FIELD: need = 4ll * s.n[0] * s.n[1]; break;
      case LMC_SENSOR_FARFIELD_FLUENCE: need = 4ll * s.n[0] * s.n[1]; break;
      case LMC_SENSOR_REFLECTANCE: need = 1; break;
      default: return fail(LMC_ERR_INVALID, "unknown sensor kind");
    }
    if (need != s.out_len) return fail(LMC_ERR_INVALID, "sensor out_len does not match its grid");
    if (s.n[0] < 0 || s.n[1] < 0 || s.n[2] < 0) return fail(LMC_ERR_INVALID, "negative grid size");
  }
  return LMC_OK;
}

int launch_layer_of(const lmc_scene_desc* d, double z) {
  if (z < 0.0) return -1;
  const double ztop = d->layers[d->n_layers - 1].z_max;
  if (!std::isinf(ztop) && z >= ztop) return -1;
  int idx = 0;
  for (int i = 1; i < d->n_layers; ++i)
    if (d->layers[i].z_min <= z) idx = i;
  return idx;
}

template <class R> void fill_laser(const lmc_scene* s, KParams<R>& p) {
  const lmc_laser_desc& L = s->laser;
  for (int j = 0; j < 3; ++j) {
    p.lpos[j] = L.pos[j];
    p.ldir[j] = static_cast<R>(L.dir[j]);
    p.lm[j] = static_cast<R>(L.m[j]);
    p.ln[j] = static_cast<R>(L.n[j]);
  }
  for (int j = 0; j < 4; ++j) p.pol[j] = static_cast<R>(L.pol[j]);
  p.k0 = static_cast<R>(2.0 * M_PI / L.wavelength);
  p.sigma = L.sigma;
  p.source = L.source;
  p.temporal = L.temporal;
  p.pulse = L.pulse_duration;
  p.rep = L.repetition_rate;
  p.t_off = L.time_offset;
}

template <class K> int raise_lds_limit(K kfn, size_t lds) {
  if (lds > 65536)
    LMC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(lds)));
  return LMC_OK;
}

// Records per transport launch: the queue holds one record per photon and CBS
// sensor in the worst case, so a run is split into chunks of
// q_cap / n_cbs photons.
constexpr int64_t kQueueRecords = int64_t(1) << 22;

template <class R> int ensure_queue(lmc_scene* s, int64_t cap) {
  const size_t rec = sizeof(R) * kRecF + sizeof(int32_t) * kRecI;
  if (s->queue.p && s->q_cap >= cap && s->q_rec_bytes == rec) return LMC_OK;
  int rc = s->queue.alloc(static_cast<size_t>(cap) * rec + 256);
  if (rc) return rc;
  s->q_cap = cap;
  s->q_rec_bytes = rec;
  return LMC_OK;
}

template <class R, class RNG, bool kAdj>
int run_impl(lmc_scene* s, DevTables<R>& T, const lmc_run_desc* r, lmc_stats* st) {
  KParams<R> p;
  std::memset(&p, 0, sizeof(p));
  p.layers = T.layers.p;
  p.species = T.species.p;
  p.cdf = T.cdf.p;
  p.guide = T.guide.p;
  p.stab = T.stab.p;
  p.sel = T.sel.p;
  p.sensors = s->sensors_dev.p;
  p.planes = s->planes_dev.p;
  p.n_layers = static_cast<int32_t>(s->layers.size());
  p.n_species = static_cast<int32_t>(s->species.size());
  p.n_sensors = static_cast<int32_t>(s->sensors.size());
  p.n_planes = static_cast<int32_t>(s->planes_host.size());
  p.cdf_total = T.cdf_total;
  p.guide_total = T.guide_total;
  p.stab_total = T.stab_total;
  p.sel_total = T.sel_total;
  p.lds_grid = s->lds_grid;
  p.launch_layer = s->launch_layer;
  fill_laser(s, p);
  p.seed = r->seed;
  p.key1 = static_cast<uint32_t>(r->seed >> 32);
  p.max_events = r->max_events;
  p.claim = r->claim > 0 ? r->claim : 256;
  const double ztop = s->layers.back().z_max;
  p.top_finite = std::isinf(ztop) ? 0 : 1;
  p.z_top = static_cast<R>(p.top_finite ? ztop : 0.0);
  p.out = s->out.p;
  p.hits = s->hits.p;
  p.counter = s->counter.p;
  p.stats = s->counter.p + 1;
  p.grad = s->grad.p;
  p.n_out = s->n_out;
  const int block = r->block_size > 0 ? r->block_size : 256;
  if (block % 64 != 0 || block > 256) return fail(LMC_ERR_INVALID, "block_size must be a multiple of 64, <= 256");
  const int bpc = r->blocks_per_cu > 0 ? r->blocks_per_cu : 8;
  const int grid = g_num_cu * bpc;
  size_t base_lds = 0;
  auto add = [&](size_t b) { base_lds += (b + 15) & ~static_cast<size_t>(15); };
  add(sizeof(LayerDev<R>) * p.n_layers);
  add(sizeof(SpeciesDev<R>) * p.n_species);
  add(sizeof(SensorDev) * p.n_sensors);
  add(sizeof(PlaneDev) * p.n_planes);
  add(sizeof(unsigned) * kMaxSensors);
  add(sizeof(double) * p.lds_grid);
  const size_t tab_lds = ((sizeof(R) * T.cdf_total + 15) & ~15ull) +
                         ((sizeof(uint16_t) * T.guide_total + 4 + 15) & ~15ull) +
                         ((sizeof(STab<R>) * T.stab_total + 15) & ~15ull) + ((sizeof(R) * T.sel_total + 15) & ~15ull);
  // stage the tables in LDS when `bpc` workgroups still fit on one CU
  const bool lds_tables = (base_lds + tab_lds) * static_cast<size_t>(bpc) <= g_max_lds_cu &&
                          base_lds + tab_lds <= g_max_lds_block;
  const size_t lds = lds_tables ? base_lds + tab_lds : base_lds;
  if (lds > g_max_lds_block) return fail(LMC_ERR_UNSUPPORTED, "scene constants exceed LDS");
  int n_cbs = 0;
  for (auto& sd : s->sensors) n_cbs += sd.kind == LMC_SENSOR_FARFIELD_CBS;
  const bool track = r->track_reverse_paths != 0;
  int64_t chunk = static_cast<int64_t>(r->n_photons);
  if (n_cbs > 0 && track) {
    const int64_t cap = std::min<int64_t>(kQueueRecords, static_cast<int64_t>(r->n_photons) * n_cbs);
    int rc = ensure_queue<R>(s, cap);
    if (rc) return rc;
    chunk = std::max<int64_t>(1, s->q_cap / n_cbs);
    p.qf = reinterpret_cast<R*>(s->queue.p);
    p.qi = reinterpret_cast<int32_t*>(s->queue.p + sizeof(R) * kRecF * s->q_cap);
    p.q_cap = s->q_cap;
  }
  p.q_count = s->qcount.p;
  void (*kt)(KParams<R>) = nullptr;
  if (track)
    kt = lds_tables ? k_transport<R, RNG, true, true, kAdj> : k_transport<R, RNG, true, false, kAdj>;
  else
    kt = lds_tables ? k_transport<R, RNG, false, true, kAdj> : k_transport<R, RNG, false, false, kAdj>;
  int rc = raise_lds_limit(kt, lds);
  if (rc) return rc;
  void (*kr)(KParams<R>) = k_cbs_records<R, kAdj>;
  const size_t lds_r = sizeof(double) * std::max(1, p.lds_grid);
  if ((rc = raise_lds_limit(kr, lds_r))) return rc;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->counter.p, 8);
  LMC_HIP(hipGetLastError());
  float ms_t = 0.f, ms_r = 0.f;
  uint64_t done = 0, launches = 0;
  while (done < r->n_photons) {
    ++launches;
    const uint64_t n = std::min<uint64_t>(static_cast<uint64_t>(chunk), r->n_photons - done);
    p.n_photons = n;
    p.offset = r->photon_offset + done;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->counter.p, 1);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->qcount.p, 2);
    LMC_HIP(hipGetLastError());
    LMC_HIP(hipEventRecord(s->ev0, s->stream));
    hipLaunchKernelGGL(kt, dim3(grid), dim3(block), lds, s->stream, p);
    LMC_HIP(hipGetLastError());
    LMC_HIP(hipEventRecord(s->ev1, s->stream));
    if (p.q_cap > 0) {
      hipLaunchKernelGGL(kr, dim3(g_num_cu * 4), dim3(256), lds_r, s->stream, p);
      LMC_HIP(hipGetLastError());
    }
    LMC_HIP(hipEventRecord(s->ev2, s->stream));
    LMC_HIP(hipEventSynchronize(s->ev2));
    float a = 0.f, b = 0.f;
    LMC_HIP(hipEventElapsedTime(&a, s->ev0, s->ev1));
    LMC_HIP(hipEventElapsedTime(&b, s->ev1, s->ev2));
    ms_t += a;
    ms_r += b;
    if (p.q_cap > 0) {
      unsigned long long qc[2];
      LMC_HIP(hipMemcpy(qc, s->qcount.p, sizeof(qc), hipMemcpyDeviceToHost));
      if (qc[1] != 0) return fail(LMC_ERR_DEVICE, "CBS record queue overflowed");
    }
    done += n;
  }
  if (st) {
    unsigned long long c[8];
    LMC_HIP(hipMemcpy(c, s->counter.p, sizeof(c), hipMemcpyDeviceToHost));
    st->photons = c[1];
    st->events = c[2];
    st->phi_trials = c[3];
    st->crossings = c[4];
    st->cbs_detections = c[5];
    st->kernel_ms = ms_t + ms_r;
    st->transport_ms = ms_t;
    st->records_ms = ms_r;
    st->transport_launches = launches;
  }
  return LMC_OK;
}

template <bool kAdj> int run_dispatch(lmc_scene* s, const lmc_run_desc* r, lmc_stats* st) {
  if (!s || !r) return fail(LMC_ERR_INVALID, "null argument");
  if (r->max_events < 0) return fail(LMC_ERR_INVALID, "max_events must be >= 0");
  if (r->n_photons == 0) {
    if (st) std::memset(st, 0, sizeof(*st));
    return LMC_OK;
  }
  if (r->precision == LMC_PREC_FP32) {
    if (r->rng != LMC_RNG_PHILOX) return fail(LMC_ERR_UNSUPPORTED, "FP32 transport supports the Philox stream only");
    int rc = build_tables<float>(s, s->t32);
    if (rc) return rc;
    return run_impl<float, PhiloxStream<float>, kAdj>(s, s->t32, r, st);
  }
  if (r->precision == LMC_PREC_FP64) {
    int rc = build_tables<double>(s, s->t64);
    if (rc) return rc;
    if (r->rng == LMC_RNG_XOSHIRO) return run_impl<double, XoshiroStream, kAdj>(s, s->t64, r, st);
    return run_impl<double, PhiloxStream<double>, kAdj>(s, s->t64, r, st);
  }
  return fail(LMC_ERR_INVALID, "unknown precision");
}

}  // namespace

extern "C" {

int lmc_abi_version(void) { return LMC_ABI_VERSION; }
const char* lmc_last_error(void) { return g_last_error.c_str(); }

int lmc_get_device_info(lmc_device_info* info) {
  if (!info) return fail(LMC_ERR_INVALID, "null info");
  int rc = ensure_device();
  if (rc) return rc;
  hipDeviceProp_t prop;
  LMC_HIP(hipGetDeviceProperties(&prop, 0));
  std::memset(info, 0, sizeof(*info));
  info->n_cu = prop.multiProcessorCount;
  info->clock_khz = prop.clockRate;
  info->lds_per_block = static_cast<int64_t>(prop.sharedMemPerBlock);
  info->lds_per_cu = static_cast<int64_t>(prop.maxSharedMemoryPerMultiProcessor);
  info->global_mem = static_cast<int64_t>(prop.totalGlobalMem);
  std::snprintf(info->name, sizeof(info->name), "%s", prop.gcnArchName);
  return LMC_OK;
}

int lmc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int lmc_scene_create(const lmc_scene_desc* d, lmc_scene** out) {
  if (!out) return fail(LMC_ERR_INVALID, "null output pointer");
  *out = nullptr;
  int rc = validate(d);
  if (rc) return rc;
  if ((rc = ensure_device())) return rc;
  const int ll = launch_layer_of(d, d->laser.pos[2]);
  if (ll < 0) return fail(LMC_ERR_INVALID, "laser position lies outside the sample");
  lmc_scene* s = new lmc_scene();
  s->layers.assign(d->layers, d->layers + d->n_layers);
  s->species.assign(d->species, d->species + d->n_species);
  if (d->n_sensors) s->sensors.assign(d->sensors, d->sensors + d->n_sensors);
  if (d->n_tables) s->tables.assign(d->tables, d->tables + d->n_tables);
  s->laser = d->laser;
  s->n_medium = d->n_medium;
  s->n_out = d->n_out;
  s->launch_layer = ll;
  // group sensor planes with the reference's 1e-9 key tolerance
  // (detector.hpp:46-49), ascending z; sensors reordered by plane
  std::vector<double> zs;
  for (auto& sd : s->sensors) {
    bool found = false;
    for (double z : zs)
      if (!(sd.z < z - 1e-9) && !(z < sd.z - 1e-9)) found = true;
    if (!found) zs.push_back(sd.z);
  }
  std::sort(zs.begin(), zs.end());
  int lds_grid = 0;
  for (size_t pl = 0; pl < zs.size(); ++pl) {
    PlaneDev P;
    P.z = zs[pl];
    P.first = static_cast<int32_t>(s->sensors_host.size());
    P.count = 0;
    for (size_t i = 0; i < s->sensors.size(); ++i) {
      const lmc_sensor_desc& sd = s->sensors[i];
      if ((sd.z < zs[pl] - 1e-9) || (zs[pl] < sd.z - 1e-9)) continue;
      SensorDev o;
      std::memset(&o, 0, sizeof(o));
      o.kind = sd.kind;
      o.absorb = sd.absorb;
      o.plane = static_cast<int32_t>(pl);
      o.id = static_cast<int32_t>(i);
      o.f_dir_en = sd.f_dir_en;
      o.f_dir = sd.f_dir;
      o.f_ev_en = sd.f_ev_en;
      o.f_ev_min = sd.f_ev_min;
      o.f_ev_max = sd.f_ev_max;
      o.f_th_en = sd.f_th_en;
      o.f_ph_en = sd.f_ph_en;
      o.f_pos_en = sd.f_pos_en;
      o.cos_th_min = static_cast<float>(sd.cos_th_min);
      o.cos_th_max = static_cast<float>(sd.cos_th_max);
      o.ph_min = static_cast<float>(sd.ph_min);
      o.ph_max = static_cast<float>(sd.ph_max);
      o.x_min = static_cast<float>(sd.x_min);
      o.x_max = static_cast<float>(sd.x_max);
      o.y_min = static_cast<float>(sd.y_min);
      o.y_max = static_cast<float>(sd.y_max);
      for (int j = 0; j < 8; ++j) o.g[j] = sd.g[j];
      o.n0 = sd.n[0];
      o.n1 = sd.n[1];
      o.n2 = sd.n[2];
      o.out_off = sd.out_offset;
      // privatise small angular grids per workgroup (<= 16 KiB)
      const bool small = (sd.kind == LMC_SENSOR_FARFIELD_CBS || sd.kind == LMC_SENSOR_FARFIELD_FLUENCE ||
                          sd.kind == LMC_SENSOR_REFLECTANCE) &&
                         (lds_grid + sd.out_len) * 8 <= 16384;
      if (small) {
        o.lds = 1;
        o.lds_off = lds_grid;
        lds_grid += static_cast<int>(sd.out_len);
      }
      s->sensors_host.push_back(o);
      P.count++;
    }
    s->planes_host.push_back(P);
  }
  // hit counters are indexed by position in sensors_host; keep a map back
  s->lds_grid = lds_grid;
  rc = s->sensors_dev.upload(s->sensors_host.data(), s->sensors_host.size());
  if (!rc) rc = s->planes_dev.upload(s->planes_host.data(), s->planes_host.size());
  if (!rc) rc = s->out.alloc(std::max<int64_t>(1, s->n_out));
  if (!rc) rc = s->hits.alloc(kMaxSensors);
  if (!rc) rc = s->counter.alloc(8);
  if (!rc) rc = s->qcount.alloc(2);
  if (rc) {
    delete s;
    return rc;
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess ||
      hipEventCreate(&s->ev2) != hipSuccess) {
    delete s;
    return fail(LMC_ERR_DEVICE, "stream/event creation failed");
  }
  rc = lmc_scene_zero(s);
  if (rc) {
    lmc_scene_destroy(s);
    return rc;
  }
  *out = s;
  return LMC_OK;
}

void lmc_scene_destroy(lmc_scene* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->ev2) (void)hipEventDestroy(s->ev2);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int lmc_scene_zero(lmc_scene* s) {
  if (!s) return fail(LMC_ERR_INVALID, "null scene");
  LMC_HIP(hipMemsetAsync(s->out.p, 0, sizeof(double) * s->out.n, s->stream));
  LMC_HIP(hipMemsetAsync(s->hits.p, 0, sizeof(unsigned long long) * s->hits.n, s->stream));
  if (s->grad.p) LMC_HIP(hipMemsetAsync(s->grad.p, 0, sizeof(double) * s->grad.n, s->stream));
  LMC_HIP(hipStreamSynchronize(s->stream));
  return LMC_OK;
}

int lmc_scene_run(lmc_scene* s, const lmc_run_desc* r, lmc_stats* st) { return run_dispatch<false>(s, r, st); }

int lmc_scene_read(lmc_scene* s, double* out, int64_t n, uint64_t* hits) {
  if (!s) return fail(LMC_ERR_INVALID, "null scene");
  if (out) {
    if (n != s->n_out) return fail(LMC_ERR_INVALID, "read size does not match the scene's accumulators");
    std::vector<double> tmp(s->n_out);
    if (s->n_out) LMC_HIP(hipMemcpy(tmp.data(), s->out.p, sizeof(double) * s->n_out, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < s->n_out; ++i) out[i] += tmp[i];
  }
  if (hits) {
    unsigned long long h[kMaxSensors];
    LMC_HIP(hipMemcpy(h, s->hits.p, sizeof(h), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < s->sensors_host.size(); ++k) hits[s->sensors_host[k].id] += h[k];
  }
  return LMC_OK;
}

int lmc_scene_run_adjoint(lmc_scene* s, const lmc_run_desc* r, double* grad, int64_t n_grad, lmc_stats* st) {
  if (!s || !r) return fail(LMC_ERR_INVALID, "null argument");
  const int nl = static_cast<int>(s->layers.size());
  if (nl > kMaxGradLayers) return fail(LMC_ERR_UNSUPPORTED, "adjoint supports at most 4 layers");
  for (auto& L : s->layers)
    if (!(L.mu_t > 0) || std::fabs(L.mfp * L.mu_t - 1.0) > 1e-9)
      return fail(LMC_ERR_INVALID, "adjoint requires mfp == 1/mu_t in every layer");
  for (auto& L : s->layers)
    if (L.species_count != 1) return fail(LMC_ERR_UNSUPPORTED, "adjoint supports homogeneous layers only");
  const int64_t need = 3ll * nl * s->n_out;
  if (grad && n_grad != need) return fail(LMC_ERR_INVALID, "grad buffer must hold 3*n_layers*n_out doubles");
  if (!s->grad.p || static_cast<int64_t>(s->grad.n) != need) {
    int rc = s->grad.alloc(need);
    if (rc) return rc;
    LMC_HIP(hipMemset(s->grad.p, 0, sizeof(double) * need));
  }
  LMC_HIP(hipMemset(s->grad.p, 0, sizeof(double) * need));
  int rc = run_dispatch<true>(s, r, st);
  if (rc) return rc;
  if (grad) {
    std::vector<double> tmp(need);
    LMC_HIP(hipMemcpy(tmp.data(), s->grad.p, sizeof(double) * need, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < need; ++i) grad[i] += tmp[i];
  }
  return LMC_OK;
}

struct lmc_batch;
int lmc_batch_create(int32_t, const lmc_scene_desc*, lmc_batch** out) {
  if (out) *out = nullptr;
  return fail(LMC_ERR_UNSUPPORTED, "batched launch not built yet");
}
void lmc_batch_destroy(lmc_batch*) {}
int lmc_batch_run(lmc_batch*, const uint64_t*, const lmc_run_desc*, int32_t, lmc_stats*) {
  return fail(LMC_ERR_UNSUPPORTED, "batched launch not built yet");
}
int lmc_batch_read(lmc_batch*, int32_t, double*, int64_t, uint64_t*) {
  return fail(LMC_ERR_UNSUPPORTED, "batched launch not built yet");
}
int lmc_batch_zero(lmc_batch*) { return fail(LMC_ERR_UNSUPPORTED, "batched launch not built yet"); }

int lmc_probe_philox(int32_t n, const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out) {
  if (n <= 0 || !ctr || !out) return fail(LMC_ERR_INVALID, "bad probe arguments");
  int rc = ensure_device();
  if (rc) return rc;
  uint4 *dc = nullptr, *dout = nullptr;
  LMC_HIP(hipMalloc(&dc, sizeof(uint4) * n));
  LMC_HIP(hipMalloc(&dout, sizeof(uint4) * n));
  LMC_HIP(hipMemcpy(dc, ctr, sizeof(uint4) * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_philox_probe, dim3((n + 255) / 256), dim3(256), 0, 0, n, dc, k0, k1, dout);
  LMC_HIP(hipGetLastError());
  LMC_HIP(hipMemcpy(out, dout, sizeof(uint4) * n, hipMemcpyDeviceToHost));
  LMC_HIP(hipFree(dc));
  LMC_HIP(hipFree(dout));
  return LMC_OK;
}

}  // extern "C"
