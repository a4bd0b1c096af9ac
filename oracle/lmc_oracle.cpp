// lmc_oracle.cpp -- CPU restatement of luminis-mc's photon-transport hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the MI355X library is
// checked against; it is never part of the product path.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// It restates, in FP64 with the reference's own RNG (xoshiro256++ re-seeded
// per photon with mix_seed) and its std::thread clone/merge driver, the
// algorithm of these reference files (all paths relative to the reference
// repository root):
//   src/simulation.cpp:105-249   run_simulation_parallel  -> oracle_run
//   src/simulation.cpp:255-617   run_photon               -> run_photon
//   src/detector.cpp:178-270     SensorsGroup::record_hit -> record_hit
//   src/detector.cpp:374-417     Sensor::check_conditions -> check_conditions
//   src/detector.cpp:546-568     PlanarFieldSensor::process_hit
//   src/detector.cpp:780-841     PlanarFluenceSensor::process_hit
//   src/detector.cpp:1080-1310   reverse_field_prefix/suffix, time_bin,
//                                accumulate_stokes, FarFieldCBSSensor::process_hit
//   src/detector.cpp:36-161,1012-1044,1315-1566  the CBS estimator
//                                (process_estimation and its helpers)
//   src/detector.cpp:1703-1800   StatisticsSensor::process_hit (and the
//                                penetration depth, simulation.cpp:396-400)
//   src/absortion.cpp:47-77      Absorption::record_absorption
//   src/sample.cpp:90-104,221-288 layer navigation, mixture species draw
//   src/medium.cpp:123-126,175-191 free path, S(theta) lookup
//   src/phase.cpp:27-49,108-115  azimuth rejection sampler, HG inversion
//   src/table.cpp:6-66,100-118   SamplingTable / DataTable
//   src/laser.cpp:11-99          photon emission
//   src/rng.cpp:5-28, include/luminis/math/rng.hpp:59-85  RNG
//   src/vec.cpp:11-112, src/utils.cpp:25-69  small linear algebra helpers
// plus two sensors with no reference counterpart (far-field fluence and total
// reflectance), defined in the same arithmetic as the CBS forward-only
// deposit (src/detector.cpp:1237-1265).
//
// It also holds an independent Mie series (Bohren & Huffman 1983, BHMIE
// recurrences, in MIEV0's sign convention) and the SamplingTable CDF builder,
// used to check the library's own host table builders.
//
// Built with -ffp-contract=off so its arithmetic does not depend on the host
// (liblmc_oracle.so, the parity checker).  The same source also builds
// liblmc_oracle_fast.so at release flags (the CPU baseline's timing build) and
// liblmc_oracle_plant.so with -DLMC_ORACLE_PLANTS: there oracle_set_plant(k)
// switches on one planted mistake (enum Plant) so the tests can show that each
// of their checks fails for a plausible slip.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../include/luminis_hip.h"

namespace {

using cplx = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;

// Planted mistakes (sensitivity checks of the tests; plant build only).
enum Plant {
  P_NONE = 0,
  P_XO_ROT,               // xoshiro256++ state rotation 44 instead of 45
  P_TABLE_T,              // SamplingTable lerp fraction measured from the upper node
  P_DATA_T,               // DataTable lerp fraction measured from the upper node
  P_PHI_C_SIGN,           // azimuth density with the sin*cos term's sign flipped
  P_HG_FRAC,              // HG inversion with the fraction not squared
  P_RESCALE_OLD_MU,       // interface rescale with the old layer's mu_t
  P_FRAME_TRANSPOSE,      // frame update with A^T
  P_JONES_SWAP,           // Jones update with S1 and S2 exchanged
  P_DISK_NO_SQRT,         // uniform disk source radius sigma*u instead of sigma*sqrt(u)
  P_PULSE_FWHM_AS_SIGMA,  // Gaussian pulse: FWHM used as the standard deviation
  P_FLUENCE_POS_NEG,      // planar fluence binning at -x
  P_FLUENCE_NO_XNEG,      // planar fluence lab field without the x negation
  P_FIELD_PHASE_GEOM,     // field phase from the geometric instead of the optical path
  P_ISO_THETA,            // isotropic theta as acos(u) instead of acos(2u - 1)
};
#ifdef LMC_ORACLE_PLANTS
int g_plant = P_NONE;
#define PLANT(k) (g_plant == (k))
#else
#define PLANT(k) false
#endif

// ---------------------------------------------------------------- RNG ----
// mix_seed: src/rng.cpp:5-10.  xoshiro256++ and SplitMix64 seeding:
// include/luminis/math/rng.hpp:59-85.  uniform: src/rng.cpp:12-16.
// normal (Box-Muller, cos branch): src/rng.cpp:18-28.
uint64_t mix_seed(uint64_t base, uint64_t tid) {
  uint64_t z = base + 0x9E3779B97F4A7C15ULL * (tid + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Xoshiro {
  uint64_t s[4];
  void seed(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9E3779B97F4A7C15ULL;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
      s[i] = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], PLANT(P_XO_ROT) ? 44 : 45);
    return r;
  }
  double uniform() { return (next() >> 11) * 0x1.0p-53; }
  double normal(double mean, double sd) {
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    const double z0 = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    return z0 * sd + mean;
  }
};

// ------------------------------------------------------- small algebra ----
struct V3 {
  double x{0}, y{0}, z{0};
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
// cross() normalises its result (src/vec.cpp:15-28).
inline V3 cross_unit(V3 a, V3 b) {
  V3 r{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  const double nf = norm(r);
  if (nf > 0.0) r = r * (1.0 / nf);
  return r;
}
inline V3 safe_unit(V3 v, V3 fb, double eps = 1e-20) {  // src/utils.cpp:39-51
  const double n2 = dot(v, v);
  if (n2 < eps) {
    const double f2 = dot(fb, fb);
    if (f2 < eps) return {1, 0, 0};
    return fb * (1.0 / std::sqrt(f2));
  }
  return v * (1.0 / std::sqrt(n2));
}
inline double clamp_pm1(double x) { return x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x); }

struct M3 {  // rows m, n, s
  double a[9];
  double& operator()(int i, int j) { return a[i * 3 + j]; }
  double operator()(int i, int j) const { return a[i * 3 + j]; }
  V3 row(int r) const { return {a[r * 3], a[r * 3 + 1], a[r * 3 + 2]}; }
};
// matmul (src/vec.cpp:76-112): accumulation order k = 0,1,2 from 0.0.
inline M3 matmul(const M3& A, const M3& B) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) acc += A(i, k) * B(k, j);
      C(i, j) = acc;
    }
  return C;
}
struct C2 {  // 2x2 complex, row-major
  cplx a[4];
  cplx& operator()(int i, int j) { return a[i * 2 + j]; }
  cplx operator()(int i, int j) const { return a[i * 2 + j]; }
  static C2 identity() {
    C2 r;
    r.a[0] = 1.0;
    r.a[1] = 0.0;
    r.a[2] = 0.0;
    r.a[3] = 1.0;
    return r;
  }
};
inline C2 matcmul(const C2& A, const C2& B) {  // src/vec.cpp:34-74
  C2 C;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      cplx acc(0.0, 0.0);
      for (int k = 0; k < 2; ++k) acc += A(i, k) * B(k, j);
      C(i, j) = acc;
    }
  return C;
}
struct CV {
  cplx m, n;
};
inline CV apply2(const C2& A, const CV& v) {  // src/utils.cpp:63-69
  return {A(0, 0) * v.m + A(0, 1) * v.n, A(1, 0) * v.m + A(1, 1) * v.n};
}
inline C2 rot2(V3 m_to, V3 n_to, V3 m_from, V3 n_from) {  // src/utils.cpp:53-61
  C2 R;
  R(0, 0) = cplx(dot(m_to, m_from), 0.0);
  R(0, 1) = cplx(dot(m_to, n_from), 0.0);
  R(1, 0) = cplx(dot(n_to, m_from), 0.0);
  R(1, 1) = cplx(dot(n_to, n_from), 0.0);
  return R;
}

// -------------------------------------------------------------- scene ----
struct Species {
  lmc_species_desc d;
  std::vector<double> cdf;  // SamplingTable cdf (values are implicit)
  std::vector<cplx> s1, s2;
  double inv_ds{0};
  double i_norm{0};  // compute_I_norm (detector.cpp:105-118)
};

// SamplingTable::Sample (src/table.cpp:50-66): lower_bound, then lerp.
double table_sample(const Species& sp, double u) {
  const auto& c = sp.cdf;
  auto it = std::lower_bound(c.begin(), c.end(), u);
  const long idx = static_cast<long>(it - c.begin());
  const double v0 = sp.d.cdf_min;
  const double st = sp.d.cdf_step;
  if (idx == 0) return v0;
  if (idx >= static_cast<long>(c.size())) return v0 + (c.size() - 1) * st;
  const double xa = v0 + (idx - 1) * st;
  const double xb = v0 + idx * st;
  const double t = PLANT(P_TABLE_T) ? (c[idx] - u) / (c[idx] - c[idx - 1]) : (u - c[idx - 1]) / (c[idx] - c[idx - 1]);
  return xa + t * (xb - xa);
}

// DataTable::Sample, uniform grid (src/table.cpp:100-118).
cplx data_sample(const std::vector<cplx>& y, double inv_dx, double x) {
  const double pos = x * inv_dx;
  if (pos <= 0.0) return y.front();
  const size_t last = y.size() - 1;
  if (pos >= static_cast<double>(last)) return y.back();
  const size_t i = static_cast<size_t>(pos);
  const double t = PLANT(P_DATA_T) ? 1.0 - (pos - static_cast<double>(i)) : pos - static_cast<double>(i);
  return y[i] + t * (y[i + 1] - y[i]);
}

// ScatteringMedium::sample_scattering_angle -> phase function (medium.cpp:51-58).
double sample_theta(const Species& sp, double u) {
  switch (sp.d.phase_kind) {
    case LMC_PHASE_TABLE_THETA:
      return table_sample(sp, u);
    case LMC_PHASE_TABLE_COS:
      return std::acos(table_sample(sp, u));
    case LMC_PHASE_HG: {  // src/phase.cpp:108-115
      const double g = sp.d.hg_g;
      const double gg = g * g;
      const double frac = (1.0 - gg) / (1.0 - g + 2.0 * g * u);
      return std::acos((1.0 + gg - (PLANT(P_HG_FRAC) ? frac : frac * frac)) / (2.0 * g));
    }
    default:  // uniform, src/phase.cpp:78-83
      return std::acos(PLANT(P_ISO_THETA) ? u : 2.0 * u - 1.0);
  }
}

// S(theta) = diag(s2, s1) (medium.cpp:138-148,181-191).
inline void smatrix(const Species& sp, double th, cplx& s2, cplx& s1) {
  s1 = data_sample(sp.s1, sp.inv_ds, th);
  s2 = data_sample(sp.s2, sp.inv_ds, th);
}

struct Layer {
  lmc_layer_desc d;
  std::vector<double> select_cdf;
};

struct Sensor {
  lmc_sensor_desc d;
  std::vector<double> stat;  // STATISTICS parameter block (27 values)
  // FarFieldCBS bin geometry (detector.cpp:884-907, 909-943)
  std::vector<double> cos_th, sin_th, domega, cos_ph, sin_ph;
};

struct Scene {
  std::vector<Layer> layers;
  std::vector<double> interfaces;  // z_min of layers 1..L-1 (sample.cpp:155-158)
  std::vector<Species> species;
  std::vector<Sensor> sensors;
  // sensor planes grouped with the reference's 1e-9 key tolerance
  // (detector.hpp:46-49), in increasing z.
  std::vector<double> plane_z;
  std::vector<std::vector<int>> plane_sensors;
  lmc_laser_desc laser;
  double light_speed{1.0};
  int64_t n_out{0};
  lmc_absorption_desc absorption{};
};


size_t layer_index_at(const Scene& S, double z) {  // sample.cpp:221-235
  const size_t L = S.layers.size();
  if (L == 0 || z < 0.0) return L;
  const double ztop = S.layers.back().d.z_max;
  if (!std::isinf(ztop) && z >= ztop) return L;
  auto it = std::upper_bound(S.interfaces.begin(), S.interfaces.end(), z);
  const size_t idx = static_cast<size_t>(it - S.interfaces.begin());
  return idx < L ? idx : L;
}

bool next_interface(const Scene& S, double zc, double zt, double& out) {  // sample.cpp:237-266
  const auto& I = S.interfaces;
  if (I.empty()) return false;
  if (zt > zc) {
    auto it = std::upper_bound(I.begin(), I.end(), zc);
    if (it != I.end() && *it <= zt) {
      out = *it;
      return true;
    }
  } else if (zt < zc) {
    auto it = std::lower_bound(I.begin(), I.end(), zc);
    if (it != I.begin()) {
      --it;
      if (*it >= zt) {
        out = *it;
        return true;
      }
    }
  }
  return false;
}

bool is_inside(const Scene& S, V3 p) {  // sample.cpp:268-288
  if (S.layers.empty()) return false;
  if (std::isnan(p.x) || std::isnan(p.y) || std::isnan(p.z)) return false;
  if (std::isinf(p.x) || std::isinf(p.y) || std::isinf(p.z)) return false;
  if (p.z < 0.0) return false;
  const double ztop = S.layers.back().d.z_max;
  if (!std::isinf(ztop) && p.z >= ztop) return false;
  return true;
}

// ------------------------------------------------------------- photon ----
struct Photon {
  V3 prev_pos, pos;
  unsigned events{0};
  bool alive{true};
  double k{0};
  double opticalpath{0};   // geometric path (time)
  double optical_path{0};  // sum n*d (phase)
  double launch_time{0};
  double velocity{1.0};
  double weight{1.0};
  double penetration_depth{0.0};  // maximum z reached (simulation.cpp:396-400)
  size_t layer{0};
  int active{-1}, first_sp{-1}, last_sp{-1};
  M3 P;
  CV E;
  CV E0;
  M3 P0, P1, Pn1, Pn;
  V3 r1, rn;
  C2 Traw, Traw_buf;
  bool has_T_prev{false};
  CV E_rev{cplx(0.0, 0.0), cplx(0.0, 0.0)};  // polarization_reverse (photon.hpp), set by CBS process_hit
};

// Absorption::record_absorption (src/absortion.cpp:47-77).
void record_absorption(const Scene& S, const Photon& ph, double d_weight, double* out) {
  const lmc_absorption_desc& A = S.absorption;
  const double r = std::sqrt(ph.pos.x * ph.pos.x + ph.pos.y * ph.pos.y), z = ph.pos.z;
  if (r > A.radius || z < 0.0 || z > A.depth) return;
  const size_t i_r = static_cast<size_t>(std::floor(r / A.d_r));
  const size_t i_z = static_cast<size_t>(std::floor(z / A.d_z));
  if (i_r >= static_cast<size_t>(A.n_r) || i_z >= static_cast<size_t>(A.n_z)) return;
  double* g = out + A.out_offset;
  g[i_r * A.n_z + i_z] += d_weight;
  if (A.d_t > 0.0) {
    const double time = ph.launch_time + (ph.opticalpath / ph.velocity);
    const int k = static_cast<int>(std::floor(time / A.d_t)) + 1;
    if (k >= 1 && k < A.n_t) g[(static_cast<size_t>(k) * A.n_r + i_r) * A.n_z + i_z] += d_weight;
  }
}

// Per-thread accumulators (the clone of the SensorsGroup, detector.cpp:299-307).
struct Accum {
  std::vector<double> out;
  std::vector<uint64_t> hits;
  std::vector<double> records;  // PhotonRecordSensor rows, LMC_RECORD_FIELDS each
  uint64_t events{0}, trials{0};
};

// Records of the last oracle_run (all threads, thread order).
std::vector<double> g_records;

// Sensor::check_conditions (detector.cpp:374-417).
bool check_conditions(const lmc_sensor_desc& d, V3 hit, V3 dir, int cross_dir, int events) {
  if (d.f_dir_en) {
    if (d.f_dir != LMC_DIR_BOTH && d.f_dir != cross_dir) return false;
  }
  if (d.f_ev_en) {
    if (events < d.f_ev_min || events > d.f_ev_max) return false;
  }
  if (d.f_th_en) {
    const double ct = -1 * dir.z;
    if (ct < d.cos_th_min || ct > d.cos_th_max) return false;
  }
  if (d.f_ph_en) {
    double phi = std::atan2(dir.y, dir.x);
    if (phi < 0) phi += 2.0 * kPi;
    if (phi < d.ph_min || phi > d.ph_max) return false;
  }
  if (d.f_pos_en) {
    if (hit.x < d.x_min || hit.x > d.x_max) return false;
    if (hit.y < d.y_min || hit.y > d.y_max) return false;
  }
  return true;
}

// reverse_field_prefix (detector.cpp:1080-1121): stages A and B.
CV reverse_prefix(const Scene& S, const M3& P0, const M3& Pnm1, const C2& Traw, int sp_n, const CV& E_in) {
  const V3 s0 = P0.row(2), snm1 = Pnm1.row(2);
  const V3 m0 = P0.row(0), n0 = P0.row(1);
  const V3 mnm1 = Pnm1.row(0), nnm1 = Pnm1.row(1);
  C2 Q;
  Q(0, 0) = 1;
  Q(0, 1) = 0;
  Q(1, 0) = 0;
  Q(1, 1) = -1;
  // Stage A: scatter at r_n, s0 -> -s_{n-1}
  const V3 s_in_a = s0;
  const V3 s_out_a = snm1 * (-1.0);
  const double th_a = std::acos(clamp_pm1(dot(s_in_a, s_out_a)));
  C2 Sa;
  {
    cplx s2, s1;
    smatrix(S.species[sp_n], th_a, s2, s1);
    Sa(0, 0) = s2;
    Sa(0, 1) = 0;
    Sa(1, 0) = 0;
    Sa(1, 1) = s1;
  }
  const V3 nprime = safe_unit(cross_unit(s_in_a, s_out_a), n0);
  const V3 mprime_in = safe_unit(cross_unit(nprime, s_in_a), m0);
  const V3 mprime_out = safe_unit(cross_unit(nprime, s_out_a), mnm1);
  C2 SRa = matcmul(Sa, rot2(mprime_in, nprime, m0, n0));
  CV E = apply2(SRa, E_in);
  E = apply2(rot2(mnm1, nnm1 * (-1.0), mprime_out, nprime), E);
  // Stage B: E <- Q T^T Q E
  C2 Tt;
  Tt(0, 0) = Traw(0, 0);
  Tt(0, 1) = Traw(1, 0);
  Tt(1, 0) = Traw(0, 1);
  Tt(1, 1) = Traw(1, 1);
  E = apply2(Q, E);
  E = apply2(Tt, E);
  E = apply2(Q, E);
  return E;
}

// reverse_field_suffix (detector.cpp:1123-1161): stage C and normalisation.
CV reverse_suffix(const Scene& S, const M3& P1, const M3& Pn, int sp_1, CV E) {
  // Stage C: scatter at r_1, -s1 -> s_n
  const V3 s1 = P1.row(2), sn = Pn.row(2);
  const V3 m1 = P1.row(0), n1 = P1.row(1);
  const V3 mn = Pn.row(0), nn = Pn.row(1);
  const V3 s_in_c = s1 * (-1.0);
  const V3 s_out_c = sn;
  const double th_c = std::acos(clamp_pm1(dot(s_in_c, s_out_c)));
  C2 Sc;
  {
    cplx s2, s1v;
    smatrix(S.species[sp_1], th_c, s2, s1v);
    Sc(0, 0) = s2;
    Sc(0, 1) = 0;
    Sc(1, 0) = 0;
    Sc(1, 1) = s1v;
  }
  const V3 npp = safe_unit(cross_unit(s_in_c, s_out_c), n1 * (-1.0));
  const V3 mpp_in = safe_unit(cross_unit(npp, s_in_c), m1);
  const V3 mpp_out = safe_unit(cross_unit(npp, s_out_c), mn);
  C2 SRc = matcmul(Sc, rot2(mpp_in, npp, m1, n1 * (-1.0)));
  E = apply2(SRc, E);
  E = apply2(rot2(mn, nn, mpp_out, npp), E);
  const double nrm = std::sqrt(std::norm(E.m) + std::norm(E.n));
  if (nrm > 1e-300) {
    E.m /= nrm;
    E.n /= nrm;
  } else {
    E.m = 0;
    E.n = 0;
  }
  return E;
}

// reverse_field (detector.cpp:1165-1173) for a detected photon.
CV reverse_field(const Scene& S, const Photon& ph) {
  return reverse_suffix(S, ph.P1, ph.Pn, ph.first_sp, reverse_prefix(S, ph.P0, ph.Pn1, ph.Traw, ph.last_sp, ph.E0));
}

inline void project_to_lab(const CV& E, const M3& P, cplx& Ex, cplx& Ey) {  // :1048-1053
  Ex = E.m * P(0, 0) + E.n * P(1, 0);
  Ey = E.m * P(0, 1) + E.n * P(1, 1);
}

// FarFieldCBSSensor::accumulate_stokes (detector.cpp:1194-1226).
void accumulate_stokes(double* g, int64_t plane, int64_t bin, int t_idx, cplx Efx, cplx Efy,
                       cplx Erx, cplx Ery) {
  const cplx Ex = Efx + Erx;
  const cplx Ey = Efy + Ery;
  const double v[8] = {
      std::norm(Ex) + std::norm(Ey),
      std::norm(Ex) - std::norm(Ey),
      2.0 * std::real(Ex * std::conj(Ey)),
      2.0 * std::imag(Ex * std::conj(Ey)),
      std::norm(Efx) + std::norm(Efy) + std::norm(Erx) + std::norm(Ery),
      std::norm(Efx) - std::norm(Efy) + std::norm(Erx) - std::norm(Ery),
      2.0 * (std::real(Efx * std::conj(Efy)) + std::real(Erx * std::conj(Ery))),
      2.0 * (std::imag(Efx * std::conj(Efy)) + std::imag(Erx * std::conj(Ery)))};
  // grid q of the sensor is at q*plane; time bin t at t*(N_theta*N_phi)
  for (int q = 0; q < 8; ++q) g[q * plane + bin] += v[q];
  (void)t_idx;
}

int cbs_time_bin(const lmc_sensor_desc& d, double arrival) {  // detector.cpp:1182-1190
  const double dt = d.g[5], tmax = d.g[2];
  if (dt <= 0) return 0;
  if (arrival < 0 || arrival >= tmax) return -1;
  const int idx = static_cast<int>(arrival / dt) + 1;
  return idx < d.n[2] ? idx : -1;
}

struct HitInfo {
  V3 point;
  cplx phase;
  int dir;
  double opticalpath_at_hit;
};

// FarFieldCBSSensor::process_hit (detector.cpp:1231-1310).
void cbs_process_hit(const Scene& S, const lmc_sensor_desc& d, Photon& ph, const HitInfo& info,
                     double* g, uint64_t& hits) {
  const int t_idx = cbs_time_bin(d, ph.launch_time + info.opticalpath_at_hit / ph.velocity);
  if (t_idx < 0) return;
  const M3& P = ph.P;
  const V3 s_out = P.row(2);
  const double theta = std::acos(-s_out.z);
  double phi = std::atan2(s_out.y, s_out.x);
  if (phi < 0) phi += 2.0 * kPi;
  const int it = static_cast<int>(theta / d.g[3]);
  const int jp = static_cast<int>(phi / d.g[4]);
  const int Nth = d.n[0], Nph = d.n[1];
  if (it < 0 || it >= Nth || jp < 0 || jp >= Nph) return;
  const int64_t tplane = static_cast<int64_t>(Nth) * Nph;
  const int64_t plane = tplane * d.n[2];
  const int64_t bin0 = static_cast<int64_t>(it) * Nph + jp;
  if (ph.events < 2) {
    const double w_sqrt = std::sqrt(ph.weight);
    const CV Ef{ph.E.m * info.phase * w_sqrt, ph.E.n * info.phase * w_sqrt};
    cplx Efx, Efy;
    project_to_lab(Ef, P, Efx, Efy);
    accumulate_stokes(g, plane, bin0, 0, Efx, Efy, 0.0, 0.0);
    if (t_idx >= 1) accumulate_stokes(g, plane, t_idx * tplane + bin0, t_idx, Efx, Efy, 0.0, 0.0);
    hits += 1;
    return;
  }
  const CV Er_loc = reverse_field(S, ph);
  ph.E_rev = Er_loc;
  const V3 s_in = ph.P0.row(2);
  const V3 qb = (s_out + s_in) * ph.k;
  const cplx path_phase = std::exp(cplx(0.0, dot(qb, ph.rn - ph.r1)));
  const cplx amp = info.phase * std::sqrt(ph.weight);
  const CV Ef{ph.E.m * amp, ph.E.n * amp};
  const CV Er{Er_loc.m * amp * path_phase, Er_loc.n * amp * path_phase};
  cplx Efx, Efy, Erx, Ery;
  project_to_lab(Ef, P, Efx, Efy);
  project_to_lab(Er, P, Erx, Ery);
  accumulate_stokes(g, plane, bin0, 0, Efx, Efy, Erx, Ery);
  if (t_idx >= 1) accumulate_stokes(g, plane, t_idx * tplane + bin0, t_idx, Efx, Efy, Erx, Ery);
  hits += 1;
}

// ---------------------------------------------------------- estimator ----
// optical_depth_to_detector (detector.cpp:53-101): sum of mu_t,i L_i along the
// straight ray from `start` over the distance L_total.
double optical_depth_to_detector(const Scene& S, V3 start, V3 dir, double L_total) {
  double tau = 0.0, traveled = 0.0;
  V3 p = start;
  size_t layer = layer_index_at(S, p.z);
  if (layer >= S.layers.size()) return 0.0;
  if (S.layers.size() == 1) return S.layers[layer].d.mu_t * L_total;
  const double dz = dir.z;
  while (traveled < L_total - 1e-15) {
    const double remaining = L_total - traveled;
    const double mu = S.layers[layer].d.mu_t;
    double seg = remaining;
    if (std::abs(dz) > 1e-15) {
      double zi;
      if (next_interface(S, p.z, p.z + dz * remaining, zi)) {
        const double sd = (zi - p.z) / dz;
        if (sd > 1e-15 && sd < remaining) seg = sd;
      }
    }
    tau += mu * seg;
    p.x += dir.x * seg;
    p.y += dir.y * seg;
    p.z += dir.z * seg;
    traveled += seg;
    if (seg >= remaining - 1e-15) break;
    const double nudge = (dz > 0) ? 1e-12 : -1e-12;
    const size_t next = layer_index_at(S, p.z + nudge);
    if (next >= S.layers.size()) break;
    layer = next;
  }
  return tau;
}

// phase_F (detector.cpp:1012-1023).
double phase_F(cplx s2, cplx s1, double c, double sn, const CV& E) {
  const double s22 = std::norm(s2), s11 = std::norm(s1);
  const double Emm = std::norm(E.m), Enn = std::norm(E.n);
  const double c2 = c * c, s2q = sn * sn;
  return Emm * (s22 * c2 + s11 * s2q) + Enn * (s22 * s2q + s11 * c2) +
         2.0 * std::real(E.m * std::conj(E.n)) * (s22 - s11) * sn * c;
}

// apply_scatter_normalized (detector.cpp:121-161).
CV apply_scatter_normalized(cplx s2, cplx s1, double c, double sn, const CV& Ein) {
  C2 Sm;
  Sm(0, 0) = s2;
  Sm(0, 1) = 0;
  Sm(1, 0) = 0;
  Sm(1, 1) = s1;
  C2 R;
  R(0, 0) = c;
  R(0, 1) = sn;
  R(1, 0) = -sn;
  R(1, 1) = c;
  const C2 T = matcmul(Sm, R);
  const double F = phase_F(s2, s1, c, sn, Ein);
  if (F < 1e-300) return CV{0.0, 0.0};
  const double inv = 1.0 / std::sqrt(F);
  return CV{inv * (T(0, 0) * Ein.m + T(0, 1) * Ein.n), inv * (T(1, 0) * Ein.m + T(1, 1) * Ein.n)};
}

// scatter_frame (detector.cpp:1027-1044).
M3 scatter_frame(const M3& P, double theta, double c, double sn) {
  const double ct = std::cos(theta), st = std::sin(theta);
  M3 A;
  A(0, 0) = ct * c;
  A(0, 1) = ct * sn;
  A(0, 2) = -st;
  A(1, 0) = -sn;
  A(1, 1) = c;
  A(1, 2) = 0;
  A(2, 0) = st * c;
  A(2, 1) = st * sn;
  A(2, 2) = ct;
  return matmul(A, P);
}

// FarFieldCBSSensor::process_estimation (detector.cpp:1315-1566): at the
// vertex `ph` (species selected, not yet scattered) force the next scatter
// into every (theta, phi) bin around the backscatter axis and deposit the
// forward field and, from the second order on, the reverse-path partner.
void cbs_process_estimation(const Scene& S, const Sensor& sen, const Photon& ph, double* g, uint64_t& hits) {
  const lmc_sensor_desc& d = sen.d;
  if (d.f_ev_en && (static_cast<int>(ph.events) < d.f_ev_min || static_cast<int>(ph.events) > d.f_ev_max)) return;
  const Species& med = S.species[ph.active];
  const Layer& lay = S.layers[ph.layer];
  if (med.i_norm < 1e-300) return;
  const V3 e1 = ph.P0.row(0), e2 = ph.P0.row(1), s_in = ph.P0.row(2);
  const V3 e3 = s_in * (-1.0);
  const V3 m_cur = ph.P.row(0), n_cur = ph.P.row(1), s_cur = ph.P.row(2);
  const bool single = ph.events == 0;
  const double w_scatter = single ? ph.weight * (lay.d.mu_s / lay.d.mu_t)
                                  : ph.weight * (med.d.mu_s / (med.d.mu_a + med.d.mu_s));
  if (w_scatter < 1e-300) return;
  const double theta_max = d.g[0], dtheta = d.g[3];
  const int Nth = d.n[0], Nph = d.n[1];
  const double th_cap = (d.theta_pp_max > 0.0) ? std::min(d.theta_pp_max, theta_max) : theta_max;
  const int i_max = std::min(Nth, static_cast<int>(th_cap / dtheta));
  const double dphi = d.g[4];
  const int64_t tplane = static_cast<int64_t>(Nth) * Nph;
  const int64_t plane = tplane * d.n[2];
  CV pre{0.0, 0.0};
  if (!single) {
    C2 Tmid = ph.Traw;
    if (ph.events >= 2 && ph.has_T_prev) Tmid = matcmul(ph.Traw_buf, ph.Traw);
    pre = reverse_prefix(S, ph.P0, ph.P, Tmid, ph.active, ph.E0);
  }
  for (int it = 0; it < i_max; ++it) {
    const double dOmega_band = sen.domega[it], s_th = sen.sin_th[it], c_th = sen.cos_th[it];
    for (int jp = 0; jp < Nph; ++jp) {
      const double c_ph = sen.cos_ph[jp], s_ph = sen.sin_ph[jp];
      const V3 s_out = e1 * (s_th * c_ph) + e2 * (s_th * s_ph) + e3 * c_th;
      // intersect_plane (detector.cpp:36-47) with the plane z = d.z, normal +z
      const double denom = s_out.z;
      if (std::abs(denom) < 1e-14) continue;
      const double L = (d.z - ph.pos.z) / denom;
      if (L <= 0.0) continue;
      const double Tr = std::exp(-optical_depth_to_detector(S, ph.pos, s_out, L));
      if (Tr < 1e-20) continue;
      const V3 pv = cross_unit(s_cur, s_out);
      const double pn = norm(pv);
      if (pn < 1e-12) continue;
      const V3 p_hat = pv * (1.0 / pn);
      const double sin_phi = -dot(m_cur, p_hat);
      const double cos_phi = dot(n_cur, p_hat);
      const double th_scat = std::acos(clamp_pm1(dot(s_cur, s_out)));
      cplx s2, s1;
      smatrix(med, th_scat, s2, s1);
      const double F = phase_F(s2, s1, cos_phi, sin_phi, ph.E);
      if (F < 1e-300) continue;
      const double prob_bin = (F / (kPi * med.i_norm)) * (dOmega_band * dphi);
      const double w_bin = w_scatter * Tr * prob_bin;
      if (w_bin < 1e-300) continue;
      const cplx amp = std::sqrt(w_bin) * std::exp(cplx(0.0, ph.k * L));
      const M3 P_exit = scatter_frame(ph.P, th_scat, cos_phi, sin_phi);
      const CV Ef_loc = apply_scatter_normalized(s2, s1, cos_phi, sin_phi, ph.E);
      const int t_idx = cbs_time_bin(d, ph.launch_time + (ph.opticalpath + L) / ph.velocity);
      if (t_idx < 0) continue;
      const CV Ef{Ef_loc.m * amp, Ef_loc.n * amp};
      cplx Efx, Efy, Erx(0.0, 0.0), Ery(0.0, 0.0);
      project_to_lab(Ef, P_exit, Efx, Efy);
      if (!single) {
        const CV Er_loc = reverse_suffix(S, ph.P1, P_exit, ph.first_sp, pre);
        const V3 qb = (s_out + s_in) * ph.k;
        const cplx path_phase = std::exp(cplx(0.0, dot(qb, ph.pos - ph.r1)));
        const CV Er{Er_loc.m * amp * path_phase, Er_loc.n * amp * path_phase};
        project_to_lab(Er, P_exit, Erx, Ery);
      }
      const int64_t bin0 = static_cast<int64_t>(it) * Nph + jp;
      accumulate_stokes(g, plane, bin0, 0, Efx, Efy, Erx, Ery);
      if (t_idx >= 1) accumulate_stokes(g, plane, t_idx * tplane + bin0, t_idx, Efx, Efy, Erx, Ery);
    }
  }
  hits += 1;
}

// PlanarFluenceSensor::process_hit (detector.cpp:780-841).
void fluence_process_hit(const lmc_sensor_desc& d, Photon& ph, const HitInfo& info, double* g,
                         uint64_t& hits) {
  const double len_x = d.g[0], len_y = d.g[1], len_t = d.g[2];
  const double dx = d.g[3], dy = d.g[4], dt = d.g[5];
  const int Nx = d.n[0], Ny = d.n[1], Nt = d.n[2];
  int t_idx = -1;
  if (dt > 0) {
    const double arrival = ph.launch_time + (info.opticalpath_at_hit / ph.velocity);
    if (arrival < 0 || arrival >= len_t) return;
    t_idx = static_cast<int>(arrival / dt) + 1;
    if (t_idx >= Nt) return;
  }
  const int xi = static_cast<int>(((PLANT(P_FLUENCE_POS_NEG) ? -info.point.x : info.point.x) + 0.5 * len_x) / dx);
  const int yi = static_cast<int>((info.point.y + 0.5 * len_y) / dy);
  if (xi < 0 || xi >= Nx || yi < 0 || yi >= Ny) return;
  const double w_sqrt = std::sqrt(ph.weight);
  const cplx Em = ph.E.m * info.phase * w_sqrt;
  const cplx En = ph.E.n * info.phase * w_sqrt;
  const M3& P = ph.P;
  const cplx Ex = (Em * P(0, 0) + En * P(1, 0)) * (PLANT(P_FLUENCE_NO_XNEG) ? 1.0 : -1.0);
  const cplx Ey = (Em * P(0, 1) + En * P(1, 1));
  const double v[4] = {std::norm(Ex) + std::norm(Ey), std::norm(Ex) - std::norm(Ey),
                       2.0 * std::real(Ex * std::conj(Ey)), 2.0 * std::imag(Ex * std::conj(Ey))};
  const int64_t tplane = static_cast<int64_t>(Nx) * Ny;
  const int64_t plane = tplane * Nt;
  const int64_t bin = static_cast<int64_t>(xi) * Ny + yi;
  if (dt > 0)
    for (int q = 0; q < 4; ++q) g[q * plane + t_idx * tplane + bin] += v[q];
  for (int q = 0; q < 4; ++q) g[q * plane + bin] += v[q];
  hits += 1;
}

// PlanarFieldSensor::process_hit (detector.cpp:546-568).
void field_process_hit(const lmc_sensor_desc& d, Photon& ph, const HitInfo& info, double* g,
                       uint64_t& hits) {
  const double len_x = d.g[0], len_y = d.g[1], dx = d.g[2], dy = d.g[3];
  const int Nx = d.n[0], Ny = d.n[1];
  const int xi = static_cast<int>((info.point.x + 0.5 * len_x) / dx);
  const int yi = static_cast<int>((info.point.y + 0.5 * len_y) / dy);
  if (xi < 0 || xi >= Nx || yi < 0 || yi >= Ny) return;
  const double w_sqrt = std::sqrt(ph.weight);
  const M3& P = ph.P;
  const cplx Ex = (ph.E.m * P(0, 0) + ph.E.n * P(1, 0)) * info.phase * w_sqrt;
  const cplx Ey = (ph.E.m * P(0, 1) + ph.E.n * P(1, 1)) * info.phase * w_sqrt;
  const int64_t bin = static_cast<int64_t>(xi) * Ny + yi;
  const int64_t plane = static_cast<int64_t>(Nx) * Ny;
  g[2 * bin] += Ex.real();
  g[2 * bin + 1] += Ex.imag();
  g[2 * (plane + bin)] += Ey.real();
  g[2 * (plane + bin) + 1] += Ey.imag();
  hits += 1;
}

// Far-field fluence (no reference counterpart): forward-field Stokes of the
// exit direction, binned like the CBS sensor (detector.cpp:1237-1265 with Er=0).
void fffluence_process_hit(const lmc_sensor_desc& d, Photon& ph, const HitInfo& info, double* g,
                           uint64_t& hits) {
  const V3 s_out = ph.P.row(2);
  const double theta = std::acos(-s_out.z);
  double phi = std::atan2(s_out.y, s_out.x);
  if (phi < 0) phi += 2.0 * kPi;
  const int it = static_cast<int>(theta / d.g[2]);
  const int jp = static_cast<int>(phi / d.g[3]);
  const int Nth = d.n[0], Nph = d.n[1];
  if (it < 0 || it >= Nth || jp < 0 || jp >= Nph) return;
  const double w_sqrt = std::sqrt(ph.weight);
  const CV Ef{ph.E.m * w_sqrt, ph.E.n * w_sqrt};
  cplx Ex, Ey;
  project_to_lab(Ef, ph.P, Ex, Ey);
  const double v[4] = {std::norm(Ex) + std::norm(Ey), std::norm(Ex) - std::norm(Ey),
                       2.0 * std::real(Ex * std::conj(Ey)), 2.0 * std::imag(Ex * std::conj(Ey))};
  const int64_t plane = static_cast<int64_t>(Nth) * Nph;
  const int64_t bin = static_cast<int64_t>(it) * Nph + jp;
  for (int q = 0; q < 4; ++q) g[q * plane + bin] += v[q];
  (void)info;
  hits += 1;
}

// StatisticsSensor::process_hit (detector.cpp:1703-1800): counts in the set

// PhotonRecordSensor::process_hit (detector.cpp:463-485): one row of the
// record layout (luminis_hip.h LMC_RECORD_FIELDS) per accepted crossing.
void record_process_hit(int si, const Photon& ph, const HitInfo& info, Accum& acc) {
  const V3 d = ph.P.row(2), m = ph.P.row(0), n = ph.P.row(1);
  const double row[LMC_RECORD_FIELDS] = {
      static_cast<double>(si), static_cast<double>(ph.events), ph.penetration_depth, ph.launch_time,
      ph.launch_time + info.opticalpath_at_hit / ph.velocity, info.opticalpath_at_hit, ph.weight, ph.k,
      ph.r1.x, ph.r1.y, ph.r1.z, ph.rn.x, ph.rn.y, ph.rn.z, info.point.x, info.point.y, info.point.z,
      d.x, d.y, d.z, m.x, m.y, m.z, n.x, n.y, n.z,
      ph.E.m.real(), ph.E.m.imag(), ph.E.n.real(), ph.E.n.imag(),
      ph.E_rev.m.real(), ph.E_rev.m.imag(), ph.E_rev.n.real(), ph.E_rev.n.imag()};
  acc.records.insert(acc.records.end(), row, row + LMC_RECORD_FIELDS);
  acc.hits[si] += 1;
}
// histograms; events, theta, phi and depth also per time window.
void stats_process_hit(const Sensor& sen, Photon& ph, const HitInfo& info, double* g, uint64_t& hits) {
  const std::vector<double>& a = sen.stat;
  const int Nt = static_cast<int>(a[24]);
  const double dt = a[25], tmax = a[26];
  const double arrival = ph.launch_time + (info.opticalpath_at_hit / ph.velocity);
  int t_idx = -1;
  if (Nt > 1 && dt > 0.0 && arrival >= 0.0 && arrival < tmax) {
    t_idx = static_cast<int>(arrival / dt) + 1;
    if (t_idx >= Nt) t_idx = -1;
  }
  int64_t base = 0;
  auto temporal = [&](int n, int idx) {
    g[base + idx] += 1.0;
    if (t_idx >= 1 && t_idx < Nt) g[base + static_cast<int64_t>(t_idx) * n + idx] += 1.0;
  };
  if (a[0] != 0.0) {
    const int me = static_cast<int>(a[6]);
    const int ev = static_cast<int>(ph.events);
    if (ev >= 0 && ev < me) temporal(me, ev);
    base += static_cast<int64_t>(Nt) * me;
  }
  if (a[1] != 0.0) {
    const int n = static_cast<int>(a[9]);
    const double theta = std::acos(-ph.P(2, 2));
    if (theta >= a[7] && theta < a[8]) temporal(n, std::min(n - 1, static_cast<int>((theta - a[7]) / a[10])));
    base += static_cast<int64_t>(Nt) * n;
  }
  if (a[2] != 0.0) {
    const int n = static_cast<int>(a[13]);
    double phi = std::atan2(ph.P(2, 1), ph.P(2, 0));
    if (phi < 0) phi += 2.0 * kPi;
    if (phi >= a[11] && phi < a[12]) temporal(n, std::min(n - 1, static_cast<int>((phi - a[11]) / a[14])));
    base += static_cast<int64_t>(Nt) * n;
  }
  if (a[3] != 0.0) {
    const int n = static_cast<int>(a[16]);
    const double depth = ph.penetration_depth;
    if (depth >= 0 && depth < a[15]) temporal(n, std::min(n - 1, static_cast<int>(depth / a[17])));
    base += static_cast<int64_t>(Nt) * n;
  }
  if (a[4] != 0.0) {
    const int n = static_cast<int>(a[19]);
    const int idx = static_cast<int>(arrival / a[20]);
    if (idx >= 1 && idx < n) g[base + idx] += 1.0;
    base += n;
  }
  if (a[5] != 0.0) {
    const int n = static_cast<int>(a[22]);
    const double w = ph.weight;
    if (w >= 0 && w < a[21]) g[base + std::min(n - 1, static_cast<int>(w / a[23]))] += 1.0;
  }
  hits += 1;
}

// SensorsGroup::record_hit (detector.cpp:178-270).
bool record_hit(const Scene& S, Photon& ph, Accum& acc) {
  bool killed = false;
  const double z1 = ph.prev_pos.z, z2 = ph.pos.z;
  if (std::abs(z2 - z1) < 1e-12) return false;
  const int cdir = (z2 > z1) ? LMC_DIR_FORWARD : LMC_DIR_BACKWARD;
  const double zmin = std::min(z1, z2), zmax = std::max(z1, z2);
  for (size_t p = 0; p < S.plane_z.size(); ++p) {
    const double zp = S.plane_z[p];
    // map range [lower_bound(zmin), upper_bound(zmax)) with the 1e-9 comparator
    if (zp < zmin - 1e-9 || zp > zmax + 1e-9) continue;
    if (!((zp >= zmin) && (zp <= zmax))) continue;
    const V3 xn = ph.prev_pos, xf = ph.pos;
    const V3 d = xf - xn;
    const double denom = d.z;
    const double t = (zp - xn.z) / denom;
    const V3 hp = xn + d * t;
    const double corr = norm(hp - xf);
    const double n_local = 1.0 / S.light_speed;
    double opl_corr = ph.optical_path;
    double gpl_hit = ph.opticalpath;
    if (corr > 0) {
      opl_corr -= corr * n_local;
      gpl_hit -= corr;
    }
    HitInfo info;
    info.point = hp;
    info.phase = std::exp(cplx(0, ph.k * (PLANT(P_FIELD_PHASE_GEOM) ? gpl_hit : opl_corr)));
    info.dir = cdir;
    info.opticalpath_at_hit = gpl_hit;
    for (int si : S.plane_sensors[p]) {
      const lmc_sensor_desc& sd = S.sensors[si].d;
      if (sd.estimator) continue;  // detected only through the estimator (detector.cpp:246-253)
      const V3 dir = ph.P.row(2);
      if (!check_conditions(sd, hp, dir, cdir, static_cast<int>(ph.events))) continue;
      double* g = acc.out.data() + sd.out_offset;
      switch (sd.kind) {
        case LMC_SENSOR_FARFIELD_CBS:
          cbs_process_hit(S, sd, ph, info, g, acc.hits[si]);
          break;
        case LMC_SENSOR_PLANAR_FLUENCE:
          fluence_process_hit(sd, ph, info, g, acc.hits[si]);
          break;
        case LMC_SENSOR_PLANAR_FIELD:
          field_process_hit(sd, ph, info, g, acc.hits[si]);
          break;
        case LMC_SENSOR_FARFIELD_FLUENCE:
          fffluence_process_hit(sd, ph, info, g, acc.hits[si]);
          break;
        case LMC_SENSOR_REFLECTANCE:
          g[0] += ph.weight;
          acc.hits[si] += 1;
          break;
        case LMC_SENSOR_STATISTICS:
          stats_process_hit(S.sensors[si], ph, info, g, acc.hits[si]);
          break;
        case LMC_SENSOR_PHOTON_RECORD:
          record_process_hit(si, ph, info, acc);
          break;
        default:
          break;
      }
      if (sd.absorb) killed = true;
    }
  }
  return killed;
}

// Laser::emit_photon (laser.cpp:77-99) + sample_emission_time (:42-75).
void emit(const Scene& S, Xoshiro& rng, Photon& ph) {
  const lmc_laser_desc& L = S.laser;
  V3 pos{L.pos[0], L.pos[1], L.pos[2]};
  if (L.source == LMC_SOURCE_UNIFORM) {
    const double th = rng.uniform() * 2.0 * kPi;
    const double u = rng.uniform();
    const double r = (PLANT(P_DISK_NO_SQRT) ? u : std::sqrt(u)) * L.sigma;
    pos = {L.pos[0] + r * std::cos(th), L.pos[1] + r * std::sin(th), L.pos[2]};
  } else if (L.source == LMC_SOURCE_GAUSSIAN) {
    const double gx = rng.normal(L.pos[0], L.sigma);
    const double gy = rng.normal(L.pos[1], L.sigma);
    pos = {gx, gy, L.pos[2]};
  }
  ph = Photon();
  ph.pos = pos;
  ph.prev_pos = pos;
  ph.k = 2.0 * kPi / L.wavelength;
  for (int j = 0; j < 3; ++j) {
    ph.P(0, j) = L.m[j];
    ph.P(1, j) = L.n[j];
    ph.P(2, j) = L.dir[j];
  }
  ph.E = {cplx(L.pol[0], L.pol[1]), cplx(L.pol[2], L.pol[3])};
  const double PS_TO_NS = 1e-3;
  const double FWHM_TO_SIGMA = 1.0 / (2.0 * std::sqrt(2.0 * std::log(2.0)));
  const double t0 = L.time_offset * PS_TO_NS;
  double t = t0;
  switch (L.temporal) {
    case LMC_TIME_GAUSSIAN:
      t = rng.normal(t0, L.pulse_duration * (PLANT(P_PULSE_FWHM_AS_SIGMA) ? 1.0 : FWHM_TO_SIGMA) * PS_TO_NS);
      break;
    case LMC_TIME_TOPHAT:
      t = t0 + (rng.uniform() - 0.5) * L.pulse_duration * PS_TO_NS;
      break;
    case LMC_TIME_EXPONENTIAL:
      t = t0 - std::log(rng.uniform()) * L.pulse_duration * PS_TO_NS;
      break;
    case LMC_TIME_PULSE_TRAIN:
      t = (L.repetition_rate <= 0.0) ? t0 : t0 + rng.uniform() * (1e9 / L.repetition_rate);
      break;
    case LMC_TIME_CW:
      t = (L.pulse_duration <= 0.0) ? t0 : t0 + rng.uniform() * L.pulse_duration * PS_TO_NS;
      break;
    default:
      break;
  }
  ph.launch_time = t;
}

// PhaseFunction::sample_phi_conditional (phase.cpp:27-49): phi by rejection
// against F(phi) = a cos^2 + b sin^2 + c sin cos with the closed-form maximum
// Fmax = (a+b)/2 + sqrt((a-b)^2 + c^2)/2, two draws per trial.
double phi_rejection(cplx s2, cplx s1, const CV& E, Xoshiro& rng, uint64_t& trials) {
  const double s2sq = std::norm(s2), s1sq = std::norm(s1);
  const double e1sq = std::norm(E.m), e2sq = std::norm(E.n);
  const double e12 = std::real(E.m * std::conj(E.n));
  const double a = s2sq * e1sq + s1sq * e2sq;
  const double b = s1sq * e1sq + s2sq * e2sq;
  const double c = 2.0 * (s2sq - s1sq) * e12 * (PLANT(P_PHI_C_SIGN) ? -1.0 : 1.0);
  const double diff = (a - b);
  const double Fmax = 0.5 * (a + b) + 0.5 * std::sqrt(diff * diff + c * c);
  while (true) {
    const double p = 2.0 * kPi * rng.uniform();
    const double cp = std::cos(p), spn = std::sin(p);
    const double F = a * cp * cp + b * spn * spn + c * cp * spn;
    const double Fc = std::max(F, 0.0);
    trials++;
    if (rng.uniform() * Fmax <= Fc) return p;
  }
}

// One scattering of the frame and the Jones vector (simulation.cpp:438-492):
// P <- A(theta, phi) P with A's rows m' = cos(theta) (cos phi m + sin phi n) -
// sin(theta) s, n' = -sin phi m + cos phi n, s' = sin(theta) (cos phi m +
// sin phi n) + cos(theta) s; E <- S R(phi) E / sqrt(F), S = diag(s2, s1),
// R(phi) = [[cos phi, sin phi], [-sin phi, cos phi]], F = |S R(phi) E|^2.
void scatter_step(M3& P, CV& E, double theta, double phi, cplx s2, cplx s1) {
  const double ct = std::cos(theta), st = std::sin(theta);
  const double cph = std::cos(phi), sph = std::sin(phi);
  M3 A;
  A(0, 0) = ct * cph;
  A(0, 1) = ct * sph;
  A(0, 2) = -st;
  A(1, 0) = -sph;
  A(1, 1) = cph;
  A(1, 2) = 0;
  A(2, 0) = st * cph;
  A(2, 1) = st * sph;
  A(2, 2) = ct;
  if (PLANT(P_FRAME_TRANSPOSE)) {
    M3 At;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) At(i, j) = A(j, i);
    A = At;
  }
  P = matmul(A, P);
  if (PLANT(P_JONES_SWAP)) std::swap(s1, s2);
  C2 Sm;
  Sm(0, 0) = s2;
  Sm(0, 1) = 0;
  Sm(1, 0) = 0;
  Sm(1, 1) = s1;
  C2 R;
  R(0, 0) = cph;
  R(0, 1) = sph;
  R(1, 0) = -sph;
  R(1, 1) = cph;
  C2 T = matcmul(Sm, R);
  const cplx Em = E.m, En = E.n;
  const double Emm = std::norm(Em), Enn = std::norm(En);
  const double s22 = std::norm(s2), s11 = std::norm(s1);
  const double pc = cph * cph, ps = sph * sph;
  const double F = Emm * (s22 * pc + s11 * ps) + Enn * (s22 * ps + s11 * pc) +
                   2.0 * std::real(Em * std::conj(En)) * (s22 - s11) * sph * cph;
  const double Finv = 1.0 / std::sqrt(F);
  for (auto& v : T.a) v *= Finv;
  E.m = (T(0, 0) * Em + T(0, 1) * En);
  E.n = (T(1, 0) * Em + T(1, 1) * En);
}

// run_photon (simulation.cpp:255-617).
void run_photon(const Scene& S, Photon& ph, Xoshiro& rng, Accum& acc, bool track, int max_events) {
  ph.r1 = ph.pos;
  ph.rn = ph.pos;
  ph.P0 = ph.P;
  ph.P1 = ph.P;
  ph.Pn1 = ph.P;
  ph.Pn = ph.P;
  ph.E0 = ph.E;
  ph.Traw = C2::identity();
  ph.Traw_buf = C2::identity();
  ph.has_T_prev = false;
  ph.active = ph.first_sp = ph.last_sp = -1;
  const double n_step = 1.0 / S.light_speed;
  while (ph.alive) {
    const Layer& cur = S.layers.at(ph.layer);
    double step = -1 * cur.d.mfp * std::log(rng.uniform());
    const double dx = ph.P(2, 0), dy = ph.P(2, 1), dz = ph.P(2, 2);
    double new_z = ph.pos.z + dz * step;
    {
      double zi;
      bool has = next_interface(S, ph.pos.z, new_z, zi);
      while (has) {
        const double dz_total = dz * step;
        if (std::abs(dz_total) < 1e-15) break;
        const double t = (zi - ph.pos.z) / dz_total;
        if (t < 0.0 || t > 1.0) break;
        const double step_to = step * t;
        const double rem_opt = (step - step_to) * S.layers.at(ph.layer).d.mu_t;
        ph.opticalpath += step_to;
        ph.optical_path += n_step * step_to;
        ph.prev_pos = ph.pos;
        ph.pos.x += dx * step_to;
        ph.pos.y += dy * step_to;
        const double nudge = (dz > 0) ? 1e-12 : -1e-12;
        ph.pos.z = zi + nudge;
        if (record_hit(S, ph, acc)) {
          ph.alive = false;
          break;
        }
        const size_t nl = layer_index_at(S, ph.pos.z);
        if (nl >= S.layers.size()) {
          ph.alive = false;
          break;
        }
        const size_t old_layer = ph.layer;
        ph.layer = nl;
        const double new_mu_t = S.layers[PLANT(P_RESCALE_OLD_MU) ? old_layer : nl].d.mu_t;
        step = new_mu_t > 0 ? rem_opt / new_mu_t : rem_opt;
        new_z = ph.pos.z + dz * step;
        has = next_interface(S, ph.pos.z, new_z, zi);
      }
    }
    if (!ph.alive) break;
    ph.opticalpath += step;
    ph.optical_path += (1.0 / S.light_speed) * step;
    ph.prev_pos = ph.pos;
    ph.pos.x += dx * step;
    ph.pos.y += dy * step;
    ph.pos.z += dz * step;
    if (ph.pos.z > ph.penetration_depth) ph.penetration_depth = ph.pos.z;
    if (record_hit(S, ph, acc)) {
      ph.alive = false;
      break;
    }
    if (!is_inside(S, ph.pos)) {
      ph.alive = false;
      break;
    }
    ph.layer = layer_index_at(S, ph.pos.z);
    const Layer& sl = S.layers.at(ph.layer);
    // select_scatter_medium (sample.cpp:96-104; Homogeneous draws nothing)
    int sp = sl.d.species_first;
    if (sl.d.species_count > 1) {
      const double u = rng.uniform();
      auto it = std::lower_bound(sl.select_cdf.begin(), sl.select_cdf.end(), u);
      size_t idx = static_cast<size_t>(it - sl.select_cdf.begin());
      if (idx >= static_cast<size_t>(sl.d.species_count)) idx = sl.d.species_count - 1;
      sp += static_cast<int>(idx);
    }
    ph.active = sp;
    // SensorsGroup::run_estimators (detector.cpp:272-281)
    for (size_t si = 0; si < S.sensors.size(); ++si) {
      const Sensor& sen = S.sensors[si];
      if (sen.d.estimator && sen.d.kind == LMC_SENSOR_FARFIELD_CBS)
        cbs_process_estimation(S, sen, ph, acc.out.data() + sen.d.out_offset, acc.hits[si]);
    }
    const Species& med = S.species[sp];
    const double theta = sample_theta(med, rng.uniform());
    cplx s2, s1;
    smatrix(med, theta, s2, s1);
    // sample_phi_conditional (phase.cpp:27-49)
    const double phi = phi_rejection(s2, s1, ph.E, rng, acc.trials);
    const double cph = std::cos(phi), sph = std::sin(phi);
    scatter_step(ph.P, ph.E, theta, phi, s2, s1);
    {
      if (track) {  // simulation.cpp:495-567
        C2 Tr;
        Tr(0, 0) = s2 * cph;
        Tr(0, 1) = s2 * sph;
        Tr(1, 0) = -s1 * sph;
        Tr(1, 1) = s1 * cph;
        const unsigned evt = ph.events + 1;
        if (evt == 1) {
          ph.r1 = ph.pos;
          ph.first_sp = ph.active;
          ph.P1 = ph.P;
        }
        ph.rn = ph.pos;
        ph.last_sp = ph.active;
        ph.Pn1 = ph.Pn;
        ph.Pn = ph.P;
        if (evt >= 2) {
          if (!ph.has_T_prev) {
            ph.Traw_buf = Tr;
            ph.has_T_prev = true;
          } else {
            C2 t = matcmul(ph.Traw_buf, ph.Traw);
            const double frob = std::sqrt(std::norm(t(0, 0)) + std::norm(t(0, 1)) +
                                          std::norm(t(1, 0)) + std::norm(t(1, 1)));
            if (frob > 1e-300)
              for (auto& v : t.a) v /= frob;
            ph.Traw = t;
            ph.Traw_buf = Tr;
          }
        }
      }
    }
    ph.events++;
    acc.events++;
    const double mu_t = sl.d.mu_t;
    double d_weight = 1.0;  // simulation.cpp:577-592
    if (mu_t > 0) {
      const double absorption_prob = sl.d.mu_a / mu_t;
      const double scat = sl.d.mu_s / mu_t;
      d_weight = ph.weight * absorption_prob;
      ph.weight *= scat;
    }
    if (S.absorption.enabled) record_absorption(S, ph, d_weight, acc.out.data());
    if (ph.weight < 1e-4) {
      if (rng.uniform() < 0.1) {
        ph.weight /= 0.1;
      } else {
        ph.alive = false;
        break;
      }
    }
    if (static_cast<int>(ph.events) > max_events) {
      ph.alive = false;
      break;
    }
  }
}

Scene build_scene(const lmc_scene_desc* d) {
  Scene S;
  if (d->n_layers <= 0) throw std::invalid_argument("scene has no layers");
  for (int i = 0; i < d->n_species; ++i) {
    Species sp;
    sp.d = d->species[i];
    if (sp.d.phase_kind == LMC_PHASE_TABLE_THETA || sp.d.phase_kind == LMC_PHASE_TABLE_COS)
      sp.cdf.assign(d->tables + sp.d.cdf_offset, d->tables + sp.d.cdf_offset + sp.d.cdf_n);
    sp.s1.resize(sp.d.s_n);
    sp.s2.resize(sp.d.s_n);
    for (int j = 0; j < sp.d.s_n; ++j) {
      sp.s1[j] = cplx(d->tables[sp.d.s1_offset + 2 * j], d->tables[sp.d.s1_offset + 2 * j + 1]);
      sp.s2[j] = cplx(d->tables[sp.d.s2_offset + 2 * j], d->tables[sp.d.s2_offset + 2 * j + 1]);
    }
    sp.inv_ds = 1.0 / sp.d.s_dtheta;
    {  // compute_I_norm (detector.cpp:105-118): midpoint rule, 2048 points
      const int n = 2048;
      const double dth = kPi / n;
      double acc = 0.0;
      for (int j = 0; j < n; ++j) {
        const double th = (j + 0.5) * dth;
        cplx s2, s1;
        smatrix(sp, th, s2, s1);
        acc += (std::norm(s1) + std::norm(s2)) * std::sin(th);
      }
      sp.i_norm = acc * dth;
    }
    S.species.push_back(std::move(sp));
  }
  for (int i = 0; i < d->n_layers; ++i) {
    Layer L;
    L.d = d->layers[i];
    if (L.d.species_count > 1)
      L.select_cdf.assign(d->tables + L.d.select_offset,
                          d->tables + L.d.select_offset + L.d.species_count);
    if (i > 0) S.interfaces.push_back(L.d.z_min);
    S.layers.push_back(std::move(L));
  }
  for (int i = 0; i < d->n_sensors; ++i) {
    Sensor s;
    s.d = d->sensors[i];
    if (s.d.kind == LMC_SENSOR_STATISTICS) s.stat.assign(d->tables + s.d.stat_offset, d->tables + s.d.stat_offset + 27);
    if (s.d.kind == LMC_SENSOR_FARFIELD_CBS) {  // detector.cpp:884-907, 909-943
      const double dth = s.d.g[3], dph = s.d.g[4];
      for (int it = 0; it < s.d.n[0]; ++it) {
        const double th = (it + 0.5) * dth;
        s.cos_th.push_back(std::cos(th));
        s.sin_th.push_back(std::sin(th));
        s.domega.push_back(std::cos(it * dth) - std::cos((it + 1) * dth));
      }
      for (int jp = 0; jp < s.d.n[1]; ++jp) {
        const double ph = s.d.n_phi_explicit > 0 ? d->tables[s.d.phi_offset + jp] : (jp + 0.5) * dph;
        s.cos_ph.push_back(std::cos(ph));
        s.sin_ph.push_back(std::sin(ph));
      }
    }
    S.sensors.push_back(s);
    // z_layers map keyed with the 1e-9 comparator (detector.cpp:166-176)
    bool placed = false;
    for (size_t p = 0; p < S.plane_z.size(); ++p)
      if (!(s.d.z < S.plane_z[p] - 1e-9) && !(S.plane_z[p] < s.d.z - 1e-9)) {
        S.plane_sensors[p].push_back(i);
        placed = true;
        break;
      }
    if (!placed) {
      S.plane_z.push_back(s.d.z);
      S.plane_sensors.push_back({i});
    }
  }
  // planes in increasing z (std::map order)
  std::vector<size_t> ord(S.plane_z.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return S.plane_z[a] < S.plane_z[b]; });
  std::vector<double> pz;
  std::vector<std::vector<int>> ps;
  for (size_t i : ord) {
    pz.push_back(S.plane_z[i]);
    ps.push_back(S.plane_sensors[i]);
  }
  S.plane_z = pz;
  S.plane_sensors = ps;
  S.laser = d->laser;
  S.light_speed = 1.0 / d->n_medium;
  S.n_out = d->n_out;
  S.absorption = d->absorption;
  return S;
}

thread_local std::string g_err;

// ------------------------------------------------------------- Mie ------
// Independent Mie series: Bohren & Huffman (1983), App. A (BHMIE), with the
// logarithmic derivative D_n(mx) by downward recurrence and psi/chi by upward
// recurrence; term count and sign convention as MIEV0 (MIEV0noP.f:160-161,
// 190-206, 472-485): BH's e^{-iwt} amplitudes are conjugated for Im(m) <= 0.
void mie_bh(double x, cplx m_in, int nmu, const double* mu, cplx* S1, cplx* S2, double* qext,
            double* qsca, double* gqsc) {
  const bool conj_out = !(m_in.imag() > 0.0);
  const cplx m(m_in.real(), std::abs(m_in.imag()));  // BH: absorbing has Im > 0
  int nstop;
  if (x <= 8.0)
    nstop = static_cast<int>(x + 4.0 * std::cbrt(x) + 1.0);
  else if (x < 4200.0)
    nstop = static_cast<int>(x + 4.05 * std::cbrt(x) + 2.0);
  else
    nstop = static_cast<int>(x + 4.0 * std::cbrt(x) + 2.0);
  const cplx y = m * x;
  const int nmx = static_cast<int>(std::max(static_cast<double>(nstop), std::abs(y))) + 15;
  std::vector<cplx> D(nmx + 1);
  D[nmx] = cplx(0, 0);
  for (int n = nmx; n >= 1; --n) {
    const double en = n;
    D[n - 1] = en / y - 1.0 / (D[n] + en / y);
  }
  double psi0 = std::cos(x), psi1 = std::sin(x);
  double chi0 = -std::sin(x), chi1 = std::cos(x);
  cplx xi1(psi1, -chi1);
  std::vector<double> pi0(nmu, 0.0), pi1(nmu, 1.0);
  for (int j = 0; j < nmu; ++j) {
    S1[j] = 0;
    S2[j] = 0;
  }
  double qs = 0, gq = 0;
  cplx an_prev, bn_prev;
  cplx sforw(0, 0);
  for (int n = 1; n <= nstop; ++n) {
    const double en = n;
    const double fn = (2.0 * en + 1.0) / (en * (en + 1.0));
    const double psi = (2.0 * en - 1.0) * psi1 / x - psi0;
    const double chi = (2.0 * en - 1.0) * chi1 / x - chi0;
    const cplx xi(psi, -chi);
    const cplx an = ((D[n] / m + en / x) * psi - psi1) / ((D[n] / m + en / x) * xi - xi1);
    const cplx bn = ((D[n] * m + en / x) * psi - psi1) / ((D[n] * m + en / x) * xi - xi1);
    qs += (2.0 * en + 1.0) * (std::norm(an) + std::norm(bn));
    gq += (2.0 * en + 1.0) / (en * (en + 1.0)) * std::real(an * std::conj(bn));
    if (n > 1)
      gq += (en - 1.0) * (en + 1.0) / en * std::real(an_prev * std::conj(an) + bn_prev * std::conj(bn));
    sforw += (2.0 * en + 1.0) * (an + bn);
    for (int j = 0; j < nmu; ++j) {
      const double p = pi1[j];
      const double t = en * mu[j] * p - (en + 1.0) * pi0[j];
      S1[j] += fn * (an * p + bn * t);
      S2[j] += fn * (an * t + bn * p);
    }
    for (int j = 0; j < nmu; ++j) {
      const double p = pi1[j];
      const double pnew = ((2.0 * en + 1.0) * mu[j] * p - (en + 1.0) * pi0[j]) / en;
      pi0[j] = p;
      pi1[j] = pnew;
    }
    an_prev = an;
    bn_prev = bn;
    psi0 = psi1;
    psi1 = psi;
    chi0 = chi1;
    chi1 = chi;
    xi1 = cplx(psi1, -chi1);
  }
  // gq above uses the BH form sum_n [n(n+2)/(n+1)] Re(a_n a*_{n+1} + b_n b*_{n+1})
  // + sum_n (2n+1)/(n(n+1)) Re(a_n b*_n); the index shift is applied via an_prev.
  if (qsca) *qsca = 2.0 / (x * x) * qs;
  if (qext) *qext = 2.0 / (x * x) * sforw.real();
  if (gqsc) *gqsc = 4.0 / (x * x) * gq;
  if (conj_out)
    for (int j = 0; j < nmu; ++j) {
      S1[j] = std::conj(S1[j]);
      S2[j] = std::conj(S2[j]);
    }
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

// run_simulation_parallel (simulation.cpp:105-249): n_threads workers claim
// batches of 256 photon indices from an atomic counter; every photon's RNG is
// re-seeded with mix_seed(seed, index); per-thread accumulators are merged
// additively after join.  Adds into out[n_out] and hits[n_sensors].
int oracle_run(const lmc_scene_desc* desc, uint64_t n_photons, uint64_t photon_offset,
               uint64_t seed, int max_events, int track, int n_threads, double* out,
               uint64_t* hits, uint64_t* events_out, uint64_t* trials_out) {
  try {
    const Scene S = build_scene(desc);
    unsigned nt = n_threads > 0 ? static_cast<unsigned>(n_threads) : std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > n_photons) nt = static_cast<unsigned>(std::max<uint64_t>(1, n_photons));
    const uint64_t B = 256;
    std::atomic<uint64_t> next{0};
    std::vector<Accum> acc(nt);
    for (auto& a : acc) {
      a.out.assign(S.n_out, 0.0);
      a.hits.assign(S.sensors.size(), 0);
    }
    std::vector<std::thread> th;
    std::mutex emu;
    std::string err;
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t]() {
        try {
          Xoshiro rng;
          for (;;) {
            const uint64_t start = next.fetch_add(B, std::memory_order_relaxed);
            if (start >= n_photons) break;
            const uint64_t end = std::min(start + B, n_photons);
            for (uint64_t i = start; i < end; ++i) {
              const uint64_t idx = photon_offset + i;
              rng.seed(mix_seed(seed, idx));
              Photon ph;
              emit(S, rng, ph);
              ph.velocity = 1.0;
              ph.layer = layer_index_at(S, ph.pos.z);
              // the reference then fails in Sample::get_layer (layers.at)
              if (ph.layer >= S.layers.size()) throw std::out_of_range("photon launched outside the sample");
              run_photon(S, ph, rng, acc[t], track != 0, max_events);
            }
          }
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> lk(emu);
          if (err.empty()) err = e.what();
        }
      });
    for (auto& x : th) x.join();
    if (!err.empty()) {
      g_err = err;
      return LMC_ERR_INVALID;
    }
    uint64_t ev = 0, tr = 0;
    g_records.clear();
    for (auto& a : acc) g_records.insert(g_records.end(), a.records.begin(), a.records.end());
    for (auto& a : acc) {
      for (int64_t i = 0; i < S.n_out; ++i) out[i] += a.out[i];
      if (hits)
        for (size_t s = 0; s < S.sensors.size(); ++s) hits[s] += a.hits[s];
      ev += a.events;
      tr += a.trials;
    }
    if (events_out) *events_out += ev;
    if (trials_out) *trials_out += tr;
    return LMC_OK;
  } catch (const std::exception& e) {
    g_err = e.what();
    return LMC_ERR_INVALID;
  }
}

// PhotonRecordSensor rows of the last oracle_run: count into *n, up to
// max_rows rows into out (may be null).
int oracle_records(double* out, int64_t max_rows, int64_t* n) {
  const int64_t rows = static_cast<int64_t>(g_records.size() / LMC_RECORD_FIELDS);
  if (n) *n = rows;
  if (out && max_rows > 0)
    std::memcpy(out, g_records.data(), sizeof(double) * LMC_RECORD_FIELDS * std::min(rows, max_rows));
  return LMC_OK;
}

// Mie amplitudes (MIEV0 convention) and efficiencies, independent series.
int oracle_mie(double x, double m_re, double m_im, int nmu, const double* mu, double* s1,
               double* s2, double* qext, double* qsca, double* gqsc) {
  std::vector<cplx> a(std::max(nmu, 1)), b(std::max(nmu, 1));
  mie_bh(x, cplx(m_re, m_im), nmu, mu, a.data(), b.data(), qext, qsca, gqsc);
  for (int j = 0; j < nmu; ++j) {
    s1[2 * j] = a[j].real();
    s1[2 * j + 1] = a[j].imag();
    s2[2 * j] = b[j].real();
    s2[2 * j + 1] = b[j].imag();
  }
  return LMC_OK;
}

// SamplingTable::initialize (src/table.cpp:6-48), pdf sized nDiv+1 (the
// reference's pdf vector is one short, src/table.cpp:23 vs :27-31).
int oracle_sampling_cdf(int n_div, double min_val, double max_val, const double* pdf_in,
                        double* cdf) {
  if (n_div <= 0 || !(min_val < max_val)) return LMC_ERR_INVALID;
  const double step = (max_val - min_val) / n_div;
  std::vector<double> pdf(pdf_in, pdf_in + n_div + 1);
  double sum = 0.0;
  for (int i = 0; i < n_div + 1; i++) sum += pdf[i] * step;
  for (int i = 0; i < n_div + 1; i++) pdf[i] /= sum;
  cdf[0] = pdf[0] * step;
  for (int i = 1; i < n_div + 1; i++) cdf[i] = cdf[i - 1] + pdf[i] * step;
  const double last = cdf[n_div];
  for (int i = 0; i < n_div + 1; i++) cdf[i] /= last;
  return LMC_OK;
}

// Sequential xoshiro probe: out[i] = i-th uniform of mix_seed(seed, idx).
void oracle_xoshiro_uniforms(uint64_t seed, uint64_t idx, int n, double* out) {
  Xoshiro r;
  r.seed(mix_seed(seed, idx));
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
}
void oracle_xoshiro_u64(uint64_t state_seed, int n, uint64_t* out) {
  Xoshiro r;
  r.seed(state_seed);
  for (int i = 0; i < n; ++i) out[i] = r.next();
}
uint64_t oracle_mix_seed(uint64_t base, uint64_t tid) { return mix_seed(base, tid); }

// Philox4x32-10 (Salmon et al., SC'11), host restatement for known answers.
// reverse_field on explicit inputs (test hook): frames P0, P1, P_{n-1}, P_n
// (rows m, n, s; 9 doubles each), raw interior product T (4 complex, row
// major, re/im interleaved), launch Jones vector E0 (re/im of m, n), and the
// S1/S2 tables (n_s nodes on theta_i = i*dtheta, re/im interleaved) of the
// last- and first-vertex species.  out: the normalised reverse field (re/im of
// m, n) in the exit basis.
void oracle_reverse_field(const double* P0, const double* P1, const double* Pnm1, const double* Pn,
                          const double* T, const double* E0, int n_s, double dtheta, const double* s1_last,
                          const double* s2_last, const double* s1_first, const double* s2_first, double* out) {
  Scene S;
  S.species.resize(2);
  for (int k = 0; k < 2; ++k) {
    Species& sp = S.species[k];
    std::memset(&sp.d, 0, sizeof(sp.d));
    sp.inv_ds = 1.0 / dtheta;
    const double* a = k == 0 ? s1_last : s1_first;
    const double* b = k == 0 ? s2_last : s2_first;
    for (int i = 0; i < n_s; ++i) {
      sp.s1.push_back(cplx(a[2 * i], a[2 * i + 1]));
      sp.s2.push_back(cplx(b[2 * i], b[2 * i + 1]));
    }
  }
  Photon ph;
  for (int i = 0; i < 9; ++i) {
    ph.P0.a[i] = P0[i];
    ph.P1.a[i] = P1[i];
    ph.Pn1.a[i] = Pnm1[i];
    ph.Pn.a[i] = Pn[i];
  }
  for (int i = 0; i < 4; ++i) ph.Traw.a[i] = cplx(T[2 * i], T[2 * i + 1]);
  ph.E0 = {cplx(E0[0], E0[1]), cplx(E0[2], E0[3])};
  ph.last_sp = 0;
  ph.first_sp = 1;
  const CV E = reverse_field(S, ph);
  out[0] = E.m.real();
  out[1] = E.m.imag();
  out[2] = E.n.real();
  out[3] = E.n.imag();
}

// compute_I_norm of species `sp` of a scene (detector.cpp:105-118).
double oracle_species_i_norm(const lmc_scene_desc* desc, int sp) {
  try {
    const Scene S = build_scene(desc);
    return S.species.at(sp).i_norm;
  } catch (...) {
    return -1.0;
  }
}

void oracle_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}


// ---- test hooks on the restatement's building blocks ----------------------
// Switch on planted mistake k (0: none).  Returns 1 in the plant build, 0 (and
// does nothing) in the others.
int oracle_set_plant(int k) {
#ifdef LMC_ORACLE_PLANTS
  g_plant = k;
  return 1;
#else
  (void)k;
  return 0;
#endif
}

// xoshiro256++ from a raw state s_in[4]: n outputs, and the state after each
// step (4n words) when states is non-null.
void oracle_xoshiro_state_seq(const uint64_t* s_in, int n, uint64_t* outputs, uint64_t* states) {
  Xoshiro r;
  for (int i = 0; i < 4; ++i) r.s[i] = s_in[i];
  for (int i = 0; i < n; ++i) {
    outputs[i] = r.next();
    if (states)
      for (int j = 0; j < 4; ++j) states[4 * i + j] = r.s[j];
  }
}

// SamplingTable::Sample on a CDF of n nodes x_i = min + i*step, at m uniforms.
void oracle_table_sample(const double* cdf, int n, double min_val, double step, int m, const double* u,
                         double* out) {
  Species sp;
  std::memset(&sp.d, 0, sizeof(sp.d));
  sp.d.cdf_min = min_val;
  sp.d.cdf_step = step;
  sp.cdf.assign(cdf, cdf + n);
  for (int i = 0; i < m; ++i) out[i] = table_sample(sp, u[i]);
}

// DataTable::Sample on the uniform grid x_i = i*dx of n complex values
// (re/im interleaved), at m points; out re/im interleaved.
void oracle_data_sample(const double* y, int n, double dx, int m, const double* x, double* out) {
  std::vector<cplx> v(n);
  for (int i = 0; i < n; ++i) v[i] = cplx(y[2 * i], y[2 * i + 1]);
  for (int i = 0; i < m; ++i) {
    const cplx r = data_sample(v, 1.0 / dx, x[i]);
    out[2 * i] = r.real();
    out[2 * i + 1] = r.imag();
  }
}

// n azimuths of the rejection sampler for amplitudes (s2, s1) and Jones
// vector E (re/im of s2, s1, E.m, E.n in amp[8]) on the xoshiro stream of
// state seed `seed`; total trials into *trials.
void oracle_phi_rejection(const double* amp, uint64_t seed, int n, double* phi, uint64_t* trials) {
  Xoshiro r;
  r.seed(seed);
  const cplx s2(amp[0], amp[1]), s1(amp[2], amp[3]);
  const CV E{cplx(amp[4], amp[5]), cplx(amp[6], amp[7])};
  uint64_t t = 0;
  for (int i = 0; i < n; ++i) phi[i] = phi_rejection(s2, s1, E, r, t);
  if (trials) *trials = t;
}

// The theta sampler of a closed-form phase function (kind HG or UNIFORM;
// g for HG) at m uniforms.
void oracle_sample_theta(int kind, double g, int m, const double* u, double* out) {
  Species sp;
  std::memset(&sp.d, 0, sizeof(sp.d));
  sp.d.phase_kind = kind;
  sp.d.hg_g = g;
  for (int i = 0; i < m; ++i) out[i] = sample_theta(sp, u[i]);
}

// One scatter_step: frame P (rows m, n, s; 9 doubles) and Jones vector E
// (re/im of m, n) in place, with amplitudes s2, s1 (re/im each in amp[4]).
void oracle_scatter_step(double* P, double* E, double theta, double phi, const double* amp) {
  M3 M;
  for (int i = 0; i < 9; ++i) M.a[i] = P[i];
  CV e{cplx(E[0], E[1]), cplx(E[2], E[3])};
  scatter_step(M, e, theta, phi, cplx(amp[0], amp[1]), cplx(amp[2], amp[3]));
  for (int i = 0; i < 9; ++i) P[i] = M.a[i];
  E[0] = e.m.real();
  E[1] = e.m.imag();
  E[2] = e.n.real();
  E[3] = e.n.imag();
}

// Emission of photons [offset, offset + n) of a scene (laser.cpp:77-99 on
// each photon's stream mix_seed(seed, index)): x, y, z, launch time each.
int oracle_emit(const lmc_scene_desc* desc, uint64_t seed, uint64_t offset, int64_t n, double* out) {
  try {
    const Scene S = build_scene(desc);
    Xoshiro rng;
    for (int64_t i = 0; i < n; ++i) {
      rng.seed(mix_seed(seed, offset + static_cast<uint64_t>(i)));
      Photon ph;
      emit(S, rng, ph);
      out[4 * i] = ph.pos.x;
      out[4 * i + 1] = ph.pos.y;
      out[4 * i + 2] = ph.pos.z;
      out[4 * i + 3] = ph.launch_time;
    }
    return LMC_OK;
  } catch (const std::exception& e) {
    g_err = e.what();
    return LMC_ERR_INVALID;
  }
}

}  // extern "C"
