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
#include <limits>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <vector>

#include "lmc_device.h"

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

#include "lmc_launch.h"

using namespace lmc;

// =================================================================== host
namespace {

int g_device_init = -1;
int g_num_cu = 0;
size_t g_max_lds_block = 0;  // dynamic LDS one workgroup may use
size_t g_max_lds_cu = 0;     // LDS of one CU

std::mutex g_device_mu;  // first use from several host threads

int ensure_device() {
  std::lock_guard<std::mutex> lk(g_device_mu);
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

// Precision-specific device copy of a scene set's tables.
template <class R> struct DevTables {
  bool ready = false;
  DBuf<LayerDev<R>> layers;
  DBuf<SpeciesDev<R>> species;
  DBuf<unsigned char> tab;  // CDFs, guide tables, S1/S2 nodes, selection CDFs (KParams::tab)
  DBuf<LaunchConst<R>> lcs;
  DBuf<R> aux;
  DBuf<AbsDev<R>> abs;  // per-scene Absorption grids (batched runs)
  size_t tab_bytes = 0;
};

// Per-scene bookkeeping of a scene set.
struct SceneInfo {
  int layer_base = 0, n_layers = 0;
  int species_base = 0, n_species = 0;
  int plane_first = 0, plane_count = 0;
  int launch_layer = 0;        // absolute
  int64_t out_base = 0, n_out = 0;
  int n_sensors = 0;
  int n_single_planes = 0;     // planes a photon crosses at most once (records <= 1)
  int n_multi_planes = 0;      // interior non-absorbing planes (records <= max_events + 2)
  double k_wave = 0.0;         // 2 pi / wavelength of the scene's laser (records)
  lmc_laser_desc laser{};
  lmc_absorption_desc absorption{};  // out_offset relative to out_base
};

}  // namespace

// A scene set: one scene (lmc_scene) or many independent scenes run in one
// launch (lmc_batch).  Layers, species, tables, sensors and accumulators of
// all scenes are concatenated with absolute indices.
struct lmc_scene {
  // merged host copies of the descriptors
  std::vector<lmc_layer_desc> layers;
  std::vector<lmc_species_desc> species;
  std::vector<double> tables;
  std::vector<int> layer_scene;
  std::vector<SceneInfo> scenes;
  int64_t n_out = 0;
  // device
  DevTables<float> t32;
  DevTables<double> t64;
  DBuf<SensorDev> sensors_dev;
  DBuf<PlaneDev> planes_dev;
  std::vector<SensorDev> sensors_host;
  std::vector<int> sensor_scene;
  std::vector<PlaneDev> planes_host;
  int lds_grid = 0;
  // estimator-mode CBS sensors: bin geometry (double, converted per precision)
  // and the doubles of their grids staged in the transport kernel's LDS
  std::vector<double> aux_host;
  int est_lds_n = 0;
  bool has_est = false;
  // a sensor plane that a flight inside the sample can cross: strictly inside
  // some scene's stack, or through its launch point
  bool flight_planes = false;
  // PhotonRecordSensor records: a device list per sensor-kernel launch,
  // gathered on the host into one list per scene; rec_per_crossing = most
  // record sensors on one plane
  int rec_per_crossing = 0;
  DBuf<double> rec;
  DBuf<unsigned long long> rec_count;  // [0] written, [1] no room
  std::vector<std::vector<double>> rec_host;
  DBuf<double> out;
  DBuf<unsigned long long> hits;
  DBuf<unsigned long long> counter;  // [0] work counter, [1..5] stats
  DBuf<double> grad;
  DBuf<unsigned char> queue;        // crossing record queue (SoA)
  DBuf<unsigned long long> qcount;  // [0] written, [1] overflowed
  // binned deposits of the sensor kernel (forward runs): entries, per-tile
  // per-workgroup positions, tile starts, accumulation work items
  DBuf<uint32_t> bin_keys;
  DBuf<double> bin_vals;
  DBuf<uint32_t> bin_pos;
  DBuf<uint32_t> bin_tile;
  DBuf<uint4> bin_items;
  DBuf<uint32_t> bin_n_items;
  DBuf<uint64_t> prefix;            // batched runs: photon-index prefix over scenes
  int64_t q_cap = 0;
  size_t q_rec_bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  // captured batched run, replayed while its key is unchanged
  hipGraphExec_t graph = nullptr;
  std::vector<uint64_t> graph_key;
};

struct lmc_batch {
  lmc_scene set;
};

namespace {

constexpr int kGuideN = 4096;
constexpr int kGuideFineN = 65536;
constexpr int64_t kGuideBudget = int64_t(1) << 21;  // bytes of guide tables a single scene keeps
constexpr int kMaxBatchScenes = 4096;

// compute_I_norm (detector.cpp:105-118): integral of (|S1|^2 + |S2|^2) sin(theta)
// by the midpoint rule on 2048 points, S from the species' uniform-theta
// tables as DataTable::Sample reads them (table.cpp:100-118).
double i_norm_of(const lmc_species_desc& d, const double* tables) {
  const int n = 2048;
  const double dth = M_PI / n, inv_dx = 1.0 / d.s_dtheta;
  const double* a = tables + d.s1_offset;
  const double* b = tables + d.s2_offset;
  double acc = 0.0;
  for (int j = 0; j < n; ++j) {
    const double th = (j + 0.5) * dth;
    const double pos = th * inv_dx;
    double s1r, s1i, s2r, s2i;
    if (pos <= 0.0) {
      s1r = a[0], s1i = a[1], s2r = b[0], s2i = b[1];
    } else if (pos >= static_cast<double>(d.s_n - 1)) {
      const int l = 2 * (d.s_n - 1);
      s1r = a[l], s1i = a[l + 1], s2r = b[l], s2i = b[l + 1];
    } else {
      const int i = static_cast<int>(pos);
      const double t = pos - static_cast<double>(i);
      s1r = a[2 * i] + t * (a[2 * i + 2] - a[2 * i]);
      s1i = a[2 * i + 1] + t * (a[2 * i + 3] - a[2 * i + 1]);
      s2r = b[2 * i] + t * (b[2 * i + 2] - b[2 * i]);
      s2i = b[2 * i + 1] + t * (b[2 * i + 3] - b[2 * i + 1]);
    }
    acc += (s1r * s1r + s1i * s1i + s2r * s2r + s2i * s2i) * std::sin(th);
  }
  return acc * dth;
}

template <class R> int build_tables(lmc_scene* s, DevTables<R>& T) {
  if (T.ready) return LMC_OK;
  // Guide entries per CDF.  The guides are read from global memory (cache
  // resident) in every launch, also when the CDFs and S tables sit in LDS.  A
  // 65536-entry guide brackets about one node, so a theta sample takes fewer
  // dependent loads and the wave's search loop runs fewer rounds than with
  // 4096 entries (the wave waits for its widest bracket: about 2 rounds instead
  // of 4-6 on 10^4-node Mie and RDG tables): a batch ran 6 % faster with it,
  // and a batch keeps only a few scenes' guides in flight.  A single scene's
  // guides all stay in use, so they get the largest power of two entries
  // (4096 to 65536) whose guides fit kGuideBudget bytes in all: 65536 for up
  // to 16 distinct CDFs, 32768 for a 32-species mixture (2 MB: 2.5 % faster
  // than 16384 there; its 65536-entry guides, 4 MB, no longer stayed in L2
  // beside the CDFs and S tables: 6 % slower than 4096).
  int gN = kGuideFineN;
  if (s->scenes.size() == 1) {
    std::set<int64_t> offs;
    for (const lmc_species_desc& d : s->species)
      if (d.phase_kind == LMC_PHASE_TABLE_THETA || d.phase_kind == LMC_PHASE_TABLE_COS) offs.insert(d.cdf_offset);
    const int64_t n_cdfs = std::max<int64_t>(1, static_cast<int64_t>(offs.size()));
    while (gN > kGuideN && n_cdfs * gN * static_cast<int64_t>(sizeof(uint16_t)) > kGuideBudget) gN >>= 1;
  }
  std::vector<SpeciesDev<R>> sp(s->species.size());
  std::vector<R> cdf;
  std::vector<uint16_t> guide;
  std::vector<STab<R>> stab;
  std::vector<R> sel;
  // species with identical tables (e.g. equal media in several layers) share one copy
  std::map<std::vector<double>, std::pair<int32_t, int32_t>> cdf_seen;  // -> (cdf_off, guide_off)
  std::map<std::vector<double>, int32_t> s_seen;                       // -> s_off
  for (size_t i = 0; i < s->species.size(); ++i) {
    const lmc_species_desc& d = s->species[i];
    SpeciesDev<R>& o = sp[i];
    std::memset(&o, 0, sizeof(o));
    o.kind = d.phase_kind;
    o.cdf_min = static_cast<R>(d.cdf_min);
    o.cdf_step = static_cast<R>(d.cdf_step);
    o.hg_g = static_cast<R>(d.hg_g);
    if (d.phase_kind == LMC_PHASE_TABLE_THETA || d.phase_kind == LMC_PHASE_TABLE_COS) {
      if (d.cdf_n < 2 || d.cdf_n > (1 << 24)) return fail(LMC_ERR_INVALID, "cdf_n must be in [2, 2^24]");
      o.cdf_n = d.cdf_n;
      o.guide_n = gN;
      // u16 guide entries: indices of tables beyond 65536 nodes are stored
      // shifted right, and the search widens each bracket by 2^shift - 1
      o.guide_shift = 0;
      while (((d.cdf_n - 1) >> o.guide_shift) > 65535) ++o.guide_shift;
      const double* c = s->tables.data() + d.cdf_offset;
      std::vector<double> key(c, c + d.cdf_n);
      auto hit = cdf_seen.find(key);
      if (hit != cdf_seen.end()) {
        o.cdf_off = hit->second.first;
        o.guide_off = hit->second.second;
      } else {
        o.cdf_off = static_cast<int32_t>(cdf.size());
        o.guide_off = static_cast<int32_t>(guide.size());
        cdf_seen.emplace(std::move(key), std::make_pair(o.cdf_off, o.guide_off));
        std::vector<R> cr(d.cdf_n);
        for (int j = 0; j < d.cdf_n; ++j) cr[j] = static_cast<R>(c[j]);
        cdf.insert(cdf.end(), cr.begin(), cr.end());
        // guide g[j] = lower_bound(cr, j/G) on the converted values; g[G] = n-1
        int idx = 0;
        for (int j = 0; j < gN; ++j) {
          const R thr = static_cast<R>(static_cast<double>(j) / gN);
          while (idx < d.cdf_n - 1 && cr[idx] < thr) ++idx;
          guide.push_back(static_cast<uint16_t>(idx >> o.guide_shift));
        }
        guide.push_back(static_cast<uint16_t>((d.cdf_n - 1) >> o.guide_shift));
      }
    }
    if (d.s_n < 2) return fail(LMC_ERR_INVALID, "species S table needs >= 2 nodes");
    o.s_n = d.s_n;
    o.inv_ds = static_cast<R>(1.0 / d.s_dtheta);
    o.pi_inorm = static_cast<R>(M_PI * i_norm_of(d, s->tables.data()));
    o.alb = static_cast<R>(d.mu_s / (d.mu_a + d.mu_s));
    {
      std::vector<double> key(s->tables.data() + d.s1_offset, s->tables.data() + d.s1_offset + 2 * d.s_n);
      key.insert(key.end(), s->tables.data() + d.s2_offset, s->tables.data() + d.s2_offset + 2 * d.s_n);
      auto hit = s_seen.find(key);
      if (hit != s_seen.end()) {
        o.s_off = hit->second;
        continue;
      }
      s_seen.emplace(std::move(key), static_cast<int32_t>(stab.size()));
    }
    o.s_off = static_cast<int32_t>(stab.size());
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
  if (cdf.size() > 0x7FFFFFFFull || stab.size() > 0x7FFFFFFFull || guide.size() > 0x7FFFFFFFull)
    return fail(LMC_ERR_INVALID, "tables too large");
  std::vector<LayerDev<R>> L(s->layers.size());
  for (size_t i = 0; i < s->layers.size(); ++i) {
    const lmc_layer_desc& d = s->layers[i];
    const SceneInfo& sc = s->scenes[s->layer_scene[i]];
    const int li = static_cast<int>(i) - sc.layer_base;
    LayerDev<R>& o = L[i];
    std::memset(&o, 0, sizeof(o));
    o.z_min = static_cast<R>(d.z_min);
    o.z_max = static_cast<R>(d.z_max);
    o.mfp = static_cast<R>(d.mfp);
    o.mu_t = static_cast<R>(d.mu_t);
    o.albedo = static_cast<R>(d.mu_t > 0 ? d.mu_s / d.mu_t : 1.0);
    o.scene = s->layer_scene[i];
    o.n64 = d.n;
    o.mu_s = static_cast<R>(d.mu_s);
    o.mu_a = static_cast<R>(d.mu_a);
    o.sp_first = d.species_first;
    o.sp_count = d.species_count;
    o.sel_off = static_cast<int32_t>(sel.size());
    o.nbr = static_cast<int16_t>((li + 1 < sc.n_layers ? 1 : 0) | (li > 0 ? 2 : 0));
    o.pl_first = sc.plane_first;
    o.pl_count = static_cast<int16_t>(sc.plane_count);
    o.z_top = static_cast<R>(s->layers[sc.layer_base + sc.n_layers - 1].z_max);  // +inf when semi-infinite
    if (d.species_count > 1) {
      // the selection CDF in R, then its guide: g[j] = lower_bound(cdf, j / G)
      // on the converted values (the kernel compares in R), g[G] = K, as u16
      // packed into R slots
      const int K = d.species_count;
      std::vector<R> cr(K);
      for (int j = 0; j < K; ++j) cr[j] = static_cast<R>(s->tables[d.select_offset + j]);
      sel.insert(sel.end(), cr.begin(), cr.end());
      std::vector<uint16_t> g(kSelGuide + 1);
      for (int j = 0; j < kSelGuide; ++j) {
        const R thr = static_cast<R>(static_cast<double>(j) / kSelGuide);
        g[j] = static_cast<uint16_t>(std::lower_bound(cr.begin(), cr.end(), thr) - cr.begin());
      }
      g[kSelGuide] = static_cast<uint16_t>(K);
      const size_t slots = (sizeof(uint16_t) * g.size() + sizeof(R) - 1) / sizeof(R);
      const size_t at = sel.size();
      sel.resize(at + slots, R(0));
      std::memcpy(sel.data() + at, g.data(), sizeof(uint16_t) * g.size());
    }
  }
  std::vector<LaunchConst<R>> lc(s->scenes.size());
  for (size_t k = 0; k < s->scenes.size(); ++k) {
    const lmc_laser_desc& Ls = s->scenes[k].laser;
    LaunchConst<R>& c = lc[k];
    std::memset(&c, 0, sizeof(c));
    c.lpos_x = Ls.pos[0];
    c.lpos_y = Ls.pos[1];
    c.lpos_z = static_cast<R>(Ls.pos[2]);
    for (int j = 0; j < 3; ++j) {
      c.ldir[j] = static_cast<R>(Ls.dir[j]);
      c.lm[j] = static_cast<R>(Ls.m[j]);
      c.ln[j] = static_cast<R>(Ls.n[j]);
    }
    for (int j = 0; j < 4; ++j) c.pol[j] = static_cast<R>(Ls.pol[j]);
    // handedness of the launch frame, (m x n) . s: every photon frame keeps it
    const double hx = Ls.m[1] * Ls.n[2] - Ls.m[2] * Ls.n[1], hy = Ls.m[2] * Ls.n[0] - Ls.m[0] * Ls.n[2],
                 hz = Ls.m[0] * Ls.n[1] - Ls.m[1] * Ls.n[0];
    c.hand = (hx * Ls.dir[0] + hy * Ls.dir[1] + hz * Ls.dir[2]) < 0.0 ? R(-1) : R(1);
    c.k0 = static_cast<R>(2.0 * M_PI / Ls.wavelength);
    c.k0d = 2.0 * M_PI / Ls.wavelength;
    c.sigma = Ls.sigma;
    c.source = Ls.source;
    c.temporal = Ls.temporal;
    c.pulse = Ls.pulse_duration;
    c.rep = Ls.repetition_rate;
    c.t_off = Ls.time_offset;
    c.launch_layer = s->scenes[k].launch_layer;
  }
  std::vector<AbsDev<R>> ab(s->scenes.size());
  for (size_t k = 0; k < s->scenes.size(); ++k) {
    const lmc_absorption_desc& A = s->scenes[k].absorption;
    AbsDev<R>& a = ab[k];
    std::memset(&a, 0, sizeof(a));
    if (!A.enabled) continue;
    a.on = 1;
    a.nr = A.n_r;
    a.nz = A.n_z;
    a.nt = A.n_t;
    a.radius = static_cast<R>(A.radius);
    a.depth = static_cast<R>(A.depth);
    a.dr = static_cast<R>(A.d_r);
    a.dz = static_cast<R>(A.d_z);
    a.dt = static_cast<R>(A.d_t);
    a.off = s->scenes[k].out_base + A.out_offset;
  }
  int rc;
  if ((rc = T.abs.upload(ab.data(), ab.size()))) return rc;
  if ((rc = T.lcs.upload(lc.data(), lc.size()))) return rc;
  {
    std::vector<R> eg(s->aux_host.begin(), s->aux_host.end());
    if (eg.empty()) eg.push_back(R(0));
    if ((rc = T.aux.upload(eg.data(), eg.size()))) return rc;
  }
  // one blob, sections 32-byte aligned (an STab<double> is 32 bytes), the
  // guides last (never copied to LDS); the species' and layers' offsets move
  // by their section's start
  auto al32 = [](size_t b) { return (b + 31) & ~static_cast<size_t>(31); };
  const size_t off_s = al32(sizeof(R) * cdf.size());
  const size_t off_l = off_s + al32(sizeof(STab<R>) * stab.size());
  const size_t off_g = off_l + al32(sizeof(R) * sel.size() + 4);
  const size_t total = off_g + al32(sizeof(uint16_t) * guide.size() + 4);
  if (total > 0x7FFFFFF0ull) return fail(LMC_ERR_INVALID, "tables too large");
  for (SpeciesDev<R>& o : sp) {
    o.guide_off += static_cast<int32_t>(off_g / sizeof(uint16_t));
    o.s_off += static_cast<int32_t>(off_s / sizeof(STab<R>));
  }
  for (LayerDev<R>& o : L) o.sel_off += static_cast<int32_t>(off_l / sizeof(R));
  if ((rc = T.layers.upload(L.data(), L.size()))) return rc;
  if ((rc = T.species.upload(sp.data(), sp.size()))) return rc;
  std::vector<unsigned char> blob(total, 0);
  if (!cdf.empty()) std::memcpy(blob.data(), cdf.data(), sizeof(R) * cdf.size());
  if (!guide.empty()) std::memcpy(blob.data() + off_g, guide.data(), sizeof(uint16_t) * guide.size());
  if (!stab.empty()) std::memcpy(blob.data() + off_s, stab.data(), sizeof(STab<R>) * stab.size());
  if (!sel.empty()) std::memcpy(blob.data() + off_l, sel.data(), sizeof(R) * sel.size());
  if ((rc = T.tab.upload(blob.data(), blob.size()))) return rc;
  T.tab_bytes = off_g;  // the LDS part: CDFs, S tables, selection CDFs
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
      case LMC_SENSOR_STATISTICS: {
        if (s.stat_offset < 0 || s.stat_offset + 27 > d->n_tables)
          return fail(LMC_ERR_INVALID, "statistics parameters out of bounds");
        const double* a = d->tables + s.stat_offset;
        const int64_t nt = static_cast<int64_t>(a[24]);
        const double sizes[6] = {nt * a[6], nt * a[9], nt * a[13], nt * a[16], a[19], a[22]};
        for (int k = 0; k < 6; ++k) {
          if (a[k] == 0.0) continue;
          if (!(sizes[k] >= 1) || sizes[k] > 1e9) return fail(LMC_ERR_INVALID, "statistics histogram size");
          need += static_cast<int64_t>(sizes[k]);
        }
        if (nt < 1) return fail(LMC_ERR_INVALID, "statistics N_t must be >= 1");
        break;
      }
      case LMC_SENSOR_PHOTON_RECORD:
        if (s.estimator) return fail(LMC_ERR_INVALID, "a PhotonRecordSensor has no estimator mode");
        break;  // no accumulators: its records go to the scene's record list
      default: return fail(LMC_ERR_INVALID, "unknown sensor kind");
    }
    if (need != s.out_len) return fail(LMC_ERR_INVALID, "sensor out_len does not match its grid");
    if (s.n[0] < 0 || s.n[1] < 0 || s.n[2] < 0) return fail(LMC_ERR_INVALID, "negative grid size");
    if (s.n_phi_explicit != 0 &&
        (s.kind != LMC_SENSOR_FARFIELD_CBS || s.n_phi_explicit != s.n[1] || s.phi_offset < 0 ||
         s.phi_offset + s.n_phi_explicit > d->n_tables))
      return fail(LMC_ERR_INVALID, "explicit phi slices out of bounds");
    if (s.estimator && s.kind == LMC_SENSOR_FARFIELD_CBS && s.z != 0.0)
      return fail(LMC_ERR_INVALID, "the CBS estimator is defined for a sensor at z = 0");
  }
  const lmc_absorption_desc& A = d->absorption;
  if (A.enabled) {
    if (!(A.d_r > 0) || !(A.d_z > 0) || !(A.radius >= 0) || !(A.depth >= 0) || A.n_r < 1 || A.n_z < 1 || A.n_t < 1)
      return fail(LMC_ERR_INVALID, "absorption grid needs positive bin sizes");
    const int64_t n = static_cast<int64_t>(A.n_r) * A.n_z * A.n_t;
    if (A.out_offset < 0 || A.out_offset + n > d->n_out)
      return fail(LMC_ERR_INVALID, "absorption grid accumulators out of bounds");
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

template <class K> int raise_lds_limit(K kfn, size_t lds) {
  if (lds > 65536)
    LMC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kfn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(lds)));
  return LMC_OK;
}

// Crossing records per transport launch.  A run is split into launches of
// q_cap / (worst-case records per photon) photons so the queue never fills:
// one record per plane a photon can cross only when leaving the sample (or
// whose sensors all absorb), max_events + 2 for other interior planes.  Large
// launches keep the drain tail of the persistent kernel (the last, longest
// paths) short against the launch; the queue takes at most a quarter of the
// free device memory.
constexpr int64_t kQueueRecords = int64_t(1) << 26;
constexpr int kWaveQueueChunk = 128;  // queue slots a wave takes from the global counter at a time

// Crossing queue of lmc_scene_run_noalloc: a file-scope device array instead
// of a hipMalloc'd buffer, for callers that cannot allocate.  64 MiB holds
// 284k FP32 records (236 B each) or 151k FP64 ones; waves take 16 slots at a
// time there (a smaller per-wave reserve: 16 slots x 16 waves x 256 CUs), so a
// launch carries ~219k FP32 photons of a scene with one sensor plane.  Each
// device holds its own copy for the life of the process.  hipGetSymbolAddress
// gives its device address once per device (g_scratch_ptr); host code uses
// only that address, never the array's name.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kScratchQueueChunk = 16;
constexpr int kMaxDevices = 64;

void drop_graph(lmc_scene* s) {
  if (s->graph) (void)hipGraphExecDestroy(s->graph);
  s->graph = nullptr;
  s->graph_key.clear();
}

template <class R> int ensure_queue(lmc_scene* s, int64_t cap) {
  const size_t rec = sizeof(R) * 4 * kRecQ + sizeof(double) + sizeof(int32_t) * kRecI;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
    const int64_t fit = static_cast<int64_t>((free_b + (s->queue.p ? s->queue.n : 0)) / 4 / rec);
    cap = std::min(cap, std::max<int64_t>(fit, 1 << 16));
  }
  cap = (cap + 1) & ~int64_t(1);
  if (s->queue.p && s->q_cap >= cap && s->q_rec_bytes == rec) return LMC_OK;
  drop_graph(s);  // a captured graph holds the old queue's address
  s->queue.release();
  int rc = s->queue.alloc(static_cast<size_t>(cap) * rec + 256);
  if (rc) return rc;
  s->q_cap = cap;
  s->q_rec_bytes = rec;
  return LMC_OK;
}

// Binned deposits: 2^13 accumulators (64 KiB of LDS) per tile, at most 16384
// tiles (the sensor kernel's per-tile counters in LDS), at most 2^28 entries
// (3 GiB) per sensor-kernel pass, work items of at most 65536 entries.
constexpr int kBinShift = 13;
constexpr int64_t kMaxBinTiles = 16384;
constexpr int64_t kMaxBinEntries = int64_t(1) << 28;
constexpr int kBinChunk = 65536;
// Which grids take binned deposits.  Measured on the sensor kernel (ms per 5e7
// photons): the multilayer run, FP64 atomics everywhere 21.3; its 136 MB planar
// fluence grid binned 15.9; with its 1 MB CBS grid binned too 14.0, with the
// 518 KB far-field grid as well 14.1.  The polydisperse run, atomics 7.9; its
// 32 MB field and 4 MB CBS grids binned 9.3, the CBS grid alone 13.6.  The
// batch's many small CBS grids, 0.96 with atomics and ten times that binned.
// So a set bins once one of its grids reaches kBinMinBytes (the counting and
// summing passes then run anyway), and then bins all its grids in HBM.
constexpr int64_t kBinMinBytes = int64_t(64) << 20;

bool bin_eligible(const SensorDev& o) {
  return !o.lds && !o.est &&
         (o.kind == LMC_SENSOR_FARFIELD_CBS || o.kind == LMC_SENSOR_PLANAR_FLUENCE ||
          o.kind == LMC_SENSOR_FARFIELD_FLUENCE || o.kind == LMC_SENSOR_PLANAR_FIELD);
}

// Binned HBM deposits one accepted crossing makes into a sensor's grids (only
// sensors marked for binned deposits have any).
int hbm_deposits(const SensorDev& sd, bool direct_run) {
  if (sd.lds || sd.est || !(direct_run ? sd.bin_d : sd.bin)) return 0;
  switch (sd.kind) {
    case LMC_SENSOR_FARFIELD_CBS: return 8;
    case LMC_SENSOR_PLANAR_FLUENCE:
    case LMC_SENSOR_PLANAR_FIELD:
    case LMC_SENSOR_FARFIELD_FLUENCE: return 4;
    case LMC_SENSOR_REFLECTANCE: return 1;
    default: return 0;  // StatisticsSensor deposits directly in the second pass
  }
}

// Grow a device buffer to at least n elements (a captured graph holds the old
// address, so it is dropped).
template <class T> int ensure_buf(lmc_scene* s, DBuf<T>& b, size_t n) {
  if (b.p && b.n >= n) return LMC_OK;
  drop_graph(s);
  return b.alloc(n);
}

int64_t records_per_photon(const SceneInfo& sc, int max_events) {
  return sc.n_single_planes + static_cast<int64_t>(sc.n_multi_planes) * (static_cast<int64_t>(max_events) + 2);
}


// One run over a scene set.  budgets == nullptr: the set's single scene, photons
// [photon_offset, photon_offset + n_photons).  Otherwise budgets[i] photons of
// scene i, photon indices photon_offset + [0, budgets[i]) within each scene, in
// launches over the concatenated index space; use_graph captures the launch
// sequence once into a hipGraph and replays it while (budgets, run) repeat.
// PhotonRecordSensor records of one launch: once the transport kernel has
// filled the crossing queue, size the device list for every queued crossing
// times the record sensors of one plane (so no record can miss a slot), and
// clear its counters; after the sensor kernel, append each record to its
// scene's host list.  The device writes the record's absolute sensor index
// into field 7; the host maps it to the scene and puts the scene's wave
// number there.
template <class R> int prepare_records(lmc_scene* s, KParams<R>& p) {
  LMC_HIP(hipEventSynchronize(s->ev1));
  unsigned long long nq = 0;
  LMC_HIP(hipMemcpy(&nq, s->qcount.p, sizeof(nq), hipMemcpyDeviceToHost));
  const int64_t need = std::max<int64_t>(1, std::min<int64_t>(static_cast<int64_t>(nq), p.q_cap)) * s->rec_per_crossing;
  if (static_cast<int64_t>(s->rec.n) < need * LMC_RECORD_FIELDS) {
    const int64_t cap = need + need / 4 + 1024;
    if (s->rec.alloc(static_cast<size_t>(cap) * LMC_RECORD_FIELDS)) return fail(LMC_ERR_NOMEM, "PhotonRecordSensor list: hipMalloc failed");
  }
  p.rec = s->rec.p;
  p.rec_cap = static_cast<int64_t>(s->rec.n / LMC_RECORD_FIELDS);
  p.rec_count = s->rec_count.p;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->rec_count.p, 2);
  LMC_HIP(hipGetLastError());
  return LMC_OK;
}

int gather_records(lmc_scene* s) {
  unsigned long long cnt[2];
  LMC_HIP(hipMemcpy(cnt, s->rec_count.p, sizeof(cnt), hipMemcpyDeviceToHost));
  if (cnt[1] != 0) return fail(LMC_ERR_DEVICE, "PhotonRecordSensor list overflowed");
  const size_t n = static_cast<size_t>(cnt[0]);
  if (n == 0) return LMC_OK;
  std::vector<double> tmp(n * LMC_RECORD_FIELDS);
  LMC_HIP(hipMemcpy(tmp.data(), s->rec.p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
  s->rec_host.resize(s->scenes.size());
  for (size_t i = 0; i < n; ++i) {
    double* rec = tmp.data() + i * LMC_RECORD_FIELDS;
    const double sid = rec[7];
    if (!(sid >= 0.0 && sid < static_cast<double>(s->sensor_scene.size())))
      return fail(LMC_ERR_DEVICE, "PhotonRecordSensor record with a bad sensor index");
    const int k = s->sensor_scene[static_cast<size_t>(sid)];
    rec[7] = s->scenes[k].k_wave;
    s->rec_host[k].insert(s->rec_host[k].end(), rec, rec + LMC_RECORD_FIELDS);
  }
  return LMC_OK;
}

// scratch != nullptr: the no-allocation run (lmc_scene_run_noalloc): the
// crossing queue lives at `scratch` (the device address of the static
// array, kScratchBytes), nothing is allocated, PhotonRecordSensor records
// (a host-gathered list) are refused.
template <class R, bool kXo, bool kAdj>
int run_impl(lmc_scene* s, DevTables<R>& T, const lmc_run_desc* r, const uint64_t* budgets, bool use_graph,
             lmc_stats* st, unsigned char* scratch = nullptr) {
  const bool batch = budgets != nullptr;
  if (scratch && (batch || use_graph)) return fail(LMC_ERR_INVALID, "no-alloc runs take one scene, eagerly");
  const int n_sc = static_cast<int>(s->scenes.size());
  KParams<R> p;
  std::memset(&p, 0, sizeof(p));
  p.layers = T.layers.p;
  p.species = T.species.p;
  p.tab = T.tab.p;
  p.sensors = s->sensors_dev.p;
  p.planes = s->planes_dev.p;
  p.lcs = T.lcs.p;
  p.aux = T.aux.p;  // estimator bin geometry, statistics histogram parameters
  p.abs_scenes = T.abs.p;
  if (batch) {
    for (const SceneInfo& sc : s->scenes) p.abs_on |= sc.absorption.enabled ? 1 : 0;
  } else if (!s->scenes.empty() && s->scenes[0].absorption.enabled) {
    const lmc_absorption_desc& A = s->scenes[0].absorption;
    p.abs_on = 1;
    p.abs_nr = A.n_r;
    p.abs_nz = A.n_z;
    p.abs_nt = A.n_t;
    p.abs_radius = static_cast<R>(A.radius);
    p.abs_depth = static_cast<R>(A.depth);
    p.abs_dr = static_cast<R>(A.d_r);
    p.abs_dz = static_cast<R>(A.d_z);
    p.abs_dt = static_cast<R>(A.d_t);
    p.abs_off = s->scenes[0].out_base + A.out_offset;
  }
  p.n_layers = static_cast<int32_t>(s->layers.size());
  p.n_species = static_cast<int32_t>(s->species.size());
  p.n_sensors = static_cast<int32_t>(s->sensors_host.size());
  p.n_planes = static_cast<int32_t>(s->planes_host.size());
  p.pz_lo = std::numeric_limits<R>::infinity();
  p.pz_hi = -std::numeric_limits<R>::infinity();
  for (const PlaneDev& P : s->planes_host) {
    p.pz_lo = std::min(p.pz_lo, static_cast<R>(P.z));
    p.pz_hi = std::max(p.pz_hi, static_cast<R>(P.z));
  }
  p.n_scenes = n_sc;
  p.tab_bytes = static_cast<int32_t>(T.tab_bytes);
  p.lds_grid = s->lds_grid;
  p.key0 = static_cast<uint32_t>(r->seed);
  p.key1 = static_cast<uint32_t>(r->seed >> 32);
  p.max_events = r->max_events;
  // photons a wave claims at a time: small claims keep the launch's tail
  // short (an estimator photon costs hundreds of plain ones; a single scene
  // ran 0.7 % faster at 64 than at 256 over 6e7 photons); a batch claims 256
  // so a wave's photons stay in one scene
  p.claim = r->claim > 0 ? r->claim : (s->has_est ? 8 : (s->scenes.size() > 1 ? 256 : 64));
  p.out = s->out.p;
  p.hits = s->hits.p;
  p.counter = s->counter.p;
  p.stats = s->counter.p + 1;
  p.grad = s->grad.p;
  p.n_out = s->n_out;
  uint64_t total = r->n_photons;
  int64_t per_photon = 0;
  std::vector<uint64_t> pre;
  if (batch) {
    pre.assign(n_sc + 1, 0);
    for (int k = 0; k < n_sc; ++k) {
      if (budgets[k] > 0 && r->photon_offset + budgets[k] > (uint64_t(1) << 32))
        return fail(LMC_ERR_INVALID, "batched runs index photons of a scene below 2^32");
      pre[k + 1] = pre[k] + budgets[k];
      if (budgets[k] > 0) per_photon = std::max(per_photon, records_per_photon(s->scenes[k], r->max_events));
    }
    total = pre[n_sc];
    if (s->prefix.n != pre.size()) {  // a captured graph holds the buffer's address
      drop_graph(s);
      int rc = s->prefix.alloc(pre.size());
      if (rc) return rc;
    }
    LMC_HIP(hipMemcpy(s->prefix.p, pre.data(), sizeof(uint64_t) * pre.size(), hipMemcpyHostToDevice));
    p.prefix = s->prefix.p;
    p.local_offset = r->photon_offset;
  } else {
    per_photon = records_per_photon(s->scenes[0], r->max_events);
  }
  if (st) std::memset(st, 0, sizeof(*st));
  if (total == 0) return LMC_OK;
  // LDS of the transport kernel: the fixed scene header (LdsHdr), then the
  // photon-index prefix of a batch or, when they fit, the staged tables
  const size_t base_lds = batch ? LdsHdr<R>::common_end + al16(sizeof(uint64_t) * (n_sc + 1)) : LdsHdr<R>::end;
  if (p.n_planes > kMaxSensors && !batch) return fail(LMC_ERR_INVALID, "too many sensor planes");
  const size_t tab_lds = T.tab_bytes;
  // A batch keeps the tables in HBM but can stage the per-event scene arrays
  // (layers, species, planes of all its scenes) in LDS.
  const size_t scene_lds = batch ? al16(sizeof(LayerDev<R>) * p.n_layers) + al16(sizeof(SpeciesDev<R>) * p.n_species) +
                                       al16(sizeof(PlaneDev) * p.n_planes)
                                 : 0;
  // Launch shape: 16 resident waves per CU (4 per SIMD at <= 128 VGPRs).  The
  // staged data (a single scene's tables, a batch's scene arrays) go to LDS
  // when two 512-thread workgroups per CU can hold them, or else (FP32
  // forward runs) when one 1024-thread workgroup can; otherwise 256-thread
  // workgroups read them through the caches.
  const size_t need = base_lds + (batch ? scene_lds : tab_lds);
  const bool est = s->has_est;
  if (est && (batch || kAdj))
    return fail(LMC_ERR_UNSUPPORTED, "the CBS estimator runs single scenes without gradients");
  if (est && !r->track_reverse_paths)
    return fail(LMC_ERR_INVALID, "the CBS estimator needs track_reverse_paths (reverse path of every forced scatter)");
  const bool abs_run = p.abs_on != 0;
  const bool rec_run = s->rec_per_crossing > 0;
  if (rec_run && scratch)
    return fail(LMC_ERR_UNSUPPORTED, "no-alloc runs keep no PhotonRecordSensor list (it is sized per launch): "
                                     "use lmc_scene_run");
  // the record list is sized and gathered on the host around every sensor
  // launch, so a set with record sensors runs eagerly
  if (rec_run) use_graph = false;
  const bool track = r->track_reverse_paths != 0;
  bool has_stats = false;
  for (const SensorDev& sd : s->sensors_host) has_stats = has_stats || sd.kind == LMC_SENSOR_STATISTICS;
  // kernels without the penetration depth only where nothing reads it, and
  // without plane code in the flight loop only where a flight crosses a plane
  // only when it leaves the sample
  const int aux_mode =
      abs_run ? 2 : ((has_stats || rec_run || est || s->flight_planes || sizeof(R) != 4 || kXo) ? 1 : 0);
  // the light sensors' crossings: deposited by the transport kernel or queued
  p.direct = deposits_direct(sizeof(R) == 8, kXo, kAdj, aux_mode, batch) ? 1 : 0;
  auto pick = [&](bool wide, bool lds_t) -> TransportFn<R> {
    return transport_kernel<R, kXo, kAdj>(aux_mode, est, wide, lds_t, batch, track);
  };
  const bool fit512 = 2 * need <= g_max_lds_cu && need <= g_max_lds_block;
  const bool wide_ok = !kAdj && sizeof(R) == 4 && need <= g_max_lds_block && need <= g_max_lds_cu;
  int block = r->block_size > 0 ? r->block_size : (fit512 ? 512 : (wide_ok ? 1024 : 256));
  if (est) block = 256;
  if (block % 64 != 0 || block > 1024) return fail(LMC_ERR_INVALID, "block_size must be a multiple of 64, <= 1024");
  if (block > 512 && !wide_ok)
    return fail(LMC_ERR_UNSUPPORTED, "1024-thread workgroups need FP32 forward runs with LDS-resident tables");
  int bpc = r->blocks_per_cu > 0 ? r->blocks_per_cu : std::max(1, 1024 / block);
  bool lds_tables =
      block > 512 ||
      (need * static_cast<size_t>(std::min(bpc, std::max(1, 1024 / block))) <= g_max_lds_cu && need <= g_max_lds_block);
  size_t lds = lds_tables ? need : base_lds;
  void (*kt_est)(KParams<R>) = nullptr;
  if (est) {
    // estimator runs: four 256-thread workgroups per CU (the kernel's
    // register budget), the estimator grids in LDS, the tables beside them
    // when all four copies fit
    const size_t est_need = base_lds + al16(sizeof(double) * s->est_lds_n);
    lds_tables = 4 * (est_need + tab_lds) <= g_max_lds_cu && est_need + tab_lds <= g_max_lds_block;
    lds = lds_tables ? est_need + tab_lds : est_need;
    kt_est = pick(false, lds_tables);
    if (!kt_est) return fail(LMC_ERR_UNSUPPORTED, "no estimator kernel for this run");
    int rcx = raise_lds_limit(kt_est, lds);
    if (rcx) return rcx;
    int per_cu = 0;
    LMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kt_est, block, lds));
    bpc = std::max(1, per_cu);
    p.est_lds_n = s->est_lds_n;
  }
  const int grid = g_num_cu * bpc;
  if (lds > g_max_lds_block) return fail(LMC_ERR_UNSUPPORTED, "scene constants exceed LDS");
  int64_t chunk = static_cast<int64_t>(total);
  p.q_chunk = scratch ? kScratchQueueChunk : kWaveQueueChunk;
  // slots a launch can leave unused: the rest of every wave's last chunk
  const int64_t reserve = static_cast<int64_t>(grid) * ((block + 63) / 64) * p.q_chunk;
  if (per_photon > 0 && scratch) {
    const size_t rec = sizeof(R) * 4 * kRecQ + sizeof(double) + sizeof(int32_t) * kRecI;
    const int64_t cap = static_cast<int64_t>(kScratchBytes / rec) & ~int64_t(1);
    if (cap - reserve < per_photon)
      return fail(LMC_ERR_INVALID,
                  "no-alloc run: the crossing queue needs " + std::to_string((reserve + per_photon) * rec) +
                      " bytes for one photon per launch (records per photon " + std::to_string(per_photon) +
                      "), the static scratch holds " + std::to_string(kScratchBytes) + ": use lmc_scene_run");
    chunk = std::max<int64_t>(1, (cap - reserve) / per_photon);
    p.qf = reinterpret_cast<R*>(scratch);
    p.qd = reinterpret_cast<double*>(scratch + sizeof(R) * 4 * kRecQ * cap);
    p.qi = reinterpret_cast<int32_t*>(scratch + (sizeof(R) * 4 * kRecQ + sizeof(double)) * cap);
    p.q_cap = cap;
  } else if (per_photon > 0) {
    const int64_t want = static_cast<int64_t>(std::min<uint64_t>(total, uint64_t(1) << 40)) * per_photon;
    const int64_t cap = std::min<int64_t>(kQueueRecords, want) + reserve;
    int rc = ensure_queue<R>(s, cap);
    if (rc) return rc;
    if (s->q_cap - reserve < per_photon) return fail(LMC_ERR_NOMEM, "crossing record queue too small");
    chunk = std::max<int64_t>(1, (s->q_cap - reserve) / per_photon);
    p.qf = reinterpret_cast<R*>(s->queue.p);
    p.qd = reinterpret_cast<double*>(s->queue.p + sizeof(R) * 4 * kRecQ * s->q_cap);
    p.qi = reinterpret_cast<int32_t*>(s->queue.p + (sizeof(R) * 4 * kRecQ + sizeof(double)) * s->q_cap);
    p.q_cap = s->q_cap;
  }
  p.q_count = s->qcount.p;
  const TransportFn<R> kt = est ? kt_est : pick(block > 512, lds_tables);
  if (!kt) return fail(LMC_ERR_UNSUPPORTED, "no transport kernel for this launch shape");
  int rc = raise_lds_limit(kt, lds);
  if (rc) return rc;
  void (*kr)(KParams<R>) = k_crossings<R, kAdj>;
  const int n_gp = 3 * std::min(p.n_layers, kMaxGradLayers);
  p.lds_grad = (kAdj && sizeof(double) * p.lds_grid * (1 + n_gp) <= 131072) ? 1 : 0;
  const size_t lds_vals = static_cast<size_t>(p.lds_grid) * (1 + (p.lds_grad ? n_gp : 0));
  const size_t lds_r = ((sizeof(double) * lds_vals + 15) & ~15ull) + sizeof(unsigned) * (p.n_sensors + 1);
  if (lds_r > g_max_lds_block) return fail(LMC_ERR_UNSUPPORTED, "too many sensors for the sensor kernel");
  if ((rc = raise_lds_limit(kr, lds_r))) return rc;
  // one resident round of sensor workgroups (a partial second round left the
  // tail of the grid-stride loop on a quarter of the machine)
  int kr_per_cu = 0;
  LMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&kr_per_cu, kr, 256, lds_r));
  const unsigned kr_grid = static_cast<unsigned>(g_num_cu * std::max(1, kr_per_cu));
  // Binned deposits (forward runs that may allocate, grids in HBM): per pass
  // over rec_sub records, count per tile, scan, write entries, sum per tile.
  void (*kr1)(KParams<R>) = nullptr;
  void (*kr2)(KParams<R>) = nullptr;
  int max_dep = 0;
  for (const PlaneDev& P : s->planes_host) {
    int d = 0;
    for (int k = 0; k < P.count; ++k) d += hbm_deposits(s->sensors_host[P.first + k], p.direct != 0);
    max_dep = std::max(max_dep, d);
  }
  const int64_t bin_tiles = (s->n_out + (int64_t(1) << kBinShift) - 1) >> kBinShift;
  const size_t lds_rb = ((sizeof(double) * lds_vals + 15) & ~15ull) +
                        sizeof(unsigned) * (((p.n_sensors + 1 + 3) & ~3) + static_cast<size_t>(bin_tiles));
  bool binned = !kAdj && !scratch && per_photon > 0 && max_dep > 0 && s->n_out < (int64_t(1) << 32) &&
                bin_tiles <= kMaxBinTiles && lds_rb <= g_max_lds_block;
  int bin_wgs = 0, items_cap = 0;
  int64_t rec_per_wg = 0, rec_sub = 0;
  if constexpr (!kAdj) {
    kr1 = k_crossings<R, kAdj, 1>;
    kr2 = k_crossings<R, kAdj, 2>;
  }
  if (binned) {
    if ((rc = raise_lds_limit(kr1, lds_rb)) || (rc = raise_lds_limit(kr2, lds_rb))) return rc;
    int per_cu = 0;
    LMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kr2, 256, lds_rb));
    bin_wgs = g_num_cu * std::max(1, per_cu);
    const int64_t rec_bound = chunk * per_photon + reserve;  // records one transport launch can queue
    int64_t e_cap = std::min<int64_t>(rec_bound * max_dep, kMaxBinEntries);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      e_cap = std::min<int64_t>(e_cap, static_cast<int64_t>(free_b / 8 / (sizeof(uint32_t) + sizeof(double))));
    rec_per_wg = e_cap / max_dep / bin_wgs;
    binned = rec_per_wg >= 1;
    if (binned) {
      rec_sub = rec_per_wg * bin_wgs;
      e_cap = rec_sub * max_dep;
      items_cap = static_cast<int>(bin_tiles + (e_cap + kBinChunk - 1) / kBinChunk + 1);
      if ((rc = ensure_buf(s, s->bin_keys, e_cap)) || (rc = ensure_buf(s, s->bin_vals, e_cap)) ||
          (rc = ensure_buf(s, s->bin_pos, static_cast<size_t>(bin_tiles) * bin_wgs)) ||
          (rc = ensure_buf(s, s->bin_tile, bin_tiles + 1)) || (rc = ensure_buf(s, s->bin_items, items_cap)) ||
          (rc = ensure_buf(s, s->bin_n_items, 2)))
        return rc;
      p.bin_keys = s->bin_keys.p;
      p.bin_vals = s->bin_vals.p;
      p.bin_pos = s->bin_pos.p;
      p.bin_tile = s->bin_tile.p;
      p.bin_items = s->bin_items.p;
      p.bin_n_items = s->bin_n_items.p;
      p.bin_flag = s->bin_n_items.p + 1;
      p.bin_cap = e_cap;
      p.bin_rec_per_wg = rec_per_wg;
      p.bin_tiles = static_cast<int32_t>(bin_tiles);
      p.bin_shift = kBinShift;
      p.bin_wgs = bin_wgs;
      p.bin_chunk = kBinChunk;
    }
  }
  // the sensor kernel(s) after one transport launch of up to n photons
  auto sensors = [&](uint64_t n) {
    if (!binned) {
      hipLaunchKernelGGL(kr, dim3(kr_grid), dim3(256), lds_r, s->stream, p);
      return;
    }
    const int64_t recs = static_cast<int64_t>(n) * per_photon + reserve;
    for (int64_t lo = 0; lo < recs; lo += rec_sub) {  // records past the queued count are skipped inside
      p.bin_rec_lo = lo;
      hipLaunchKernelGGL(kr1, dim3(bin_wgs), dim3(256), lds_rb, s->stream, p);
      hipLaunchKernelGGL(k_bin_scan_wgs, dim3(static_cast<unsigned>(bin_tiles)), dim3(256), 0, s->stream, p.bin_pos,
                         p.bin_tile, bin_wgs);
      hipLaunchKernelGGL(k_bin_scan_tiles, dim3(1), dim3(1024), 0, s->stream, p.bin_tile, p.bin_items,
                         p.bin_n_items, static_cast<int>(bin_tiles), kBinChunk, items_cap);
      hipLaunchKernelGGL(kr2, dim3(bin_wgs), dim3(256), lds_rb, s->stream, p);
      hipLaunchKernelGGL(k_bin_accumulate, dim3(static_cast<unsigned>(items_cap)), dim3(256),
                         sizeof(double) << kBinShift, s->stream, p.out, p.n_out, p.bin_keys, p.bin_vals, p.bin_items,
                         p.bin_n_items, kBinShift);
    }
  };
  // the launch sequence: [zero work counter + record count, transport, sensors] per launch
  struct Launch { uint64_t first, n; };
  std::vector<Launch> seq;
  for (uint64_t done = 0; done < total;) {
    uint64_t n = std::min<uint64_t>(static_cast<uint64_t>(chunk), total - done);
    if (!batch) {  // photon indices of one launch share their high word (a uniform register)
      const uint64_t first = r->photon_offset + done;
      const uint64_t to_wrap = (uint64_t(1) << 32) - (first & 0xFFFFFFFFull);
      if (n > to_wrap) n = to_wrap;
    }
    seq.push_back({done, n});
    done += n;
  }
  auto enqueue = [&](const Launch& L) -> int {
    p.n_photons = L.n;
    if (batch) {
      p.offset = L.first;
      p.idx_hi = 0;
    } else {
      p.offset = r->photon_offset + L.first;
      p.idx_hi = static_cast<uint32_t>(p.offset >> 32);
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->counter.p, 1);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->qcount.p, 1);
    if (p.q_cap > 0) {  // plane sentinel over the slots this launch can hand out
      const int64_t n_fill = std::min<int64_t>(p.q_cap, static_cast<int64_t>(L.n) * per_photon + reserve);
      hipLaunchKernelGGL(k_fill_i32, dim3(static_cast<unsigned>(std::min<int64_t>((n_fill + 255) / 256, 4096))),
                         dim3(256), 0, s->stream, p.qi + I_PLANE * p.q_cap, n_fill, -1);
    }
    LMC_HIP(hipGetLastError());
    return LMC_OK;
  };
  // stats and the overflow flags start at zero once per run
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->counter.p, 8);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->qcount.p, 2);
  if (binned)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s->stream, reinterpret_cast<int32_t*>(p.bin_flag), 1, 0);
  LMC_HIP(hipGetLastError());
  float ms_t = 0.f, ms_r = 0.f;
  if (use_graph) {
    std::vector<uint64_t> key(pre.begin(), pre.end());
    const uint64_t kv[] = {r->photon_offset, r->seed, static_cast<uint64_t>(r->max_events),
                           static_cast<uint64_t>(track), static_cast<uint64_t>(block), static_cast<uint64_t>(bpc),
                           static_cast<uint64_t>(p.claim), static_cast<uint64_t>(sizeof(R)), static_cast<uint64_t>(kXo),
                           static_cast<uint64_t>(kAdj), static_cast<uint64_t>(r->n_photons)};
    key.insert(key.end(), kv, kv + sizeof(kv) / sizeof(kv[0]));
    if (!s->graph || s->graph_key != key) {
      drop_graph(s);
      LMC_HIP(hipStreamSynchronize(s->stream));
      LMC_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      int crc = LMC_OK;
      for (const Launch& L : seq) {
        if ((crc = enqueue(L))) break;
        hipLaunchKernelGGL(kt, dim3(grid), dim3(block), lds, s->stream, p);
        if (p.q_cap > 0) sensors(L.n);
      }
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(s->stream, &g);
      if (crc) {
        if (g) (void)hipGraphDestroy(g);
        return crc;
      }
      if (ce != hipSuccess) return fail(LMC_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
      const hipError_t ie = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        s->graph = nullptr;
        return fail(LMC_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
      }
      s->graph_key = key;
    }
    LMC_HIP(hipEventRecord(s->ev0, s->stream));
    LMC_HIP(hipGraphLaunch(s->graph, s->stream));
    LMC_HIP(hipEventRecord(s->ev1, s->stream));
    LMC_HIP(hipEventSynchronize(s->ev1));
    LMC_HIP(hipEventElapsedTime(&ms_t, s->ev0, s->ev1));
    if (p.q_cap > 0) {
      unsigned long long qc[2];
      LMC_HIP(hipMemcpy(qc, s->qcount.p, sizeof(qc), hipMemcpyDeviceToHost));
      if (qc[1] != 0) return fail(LMC_ERR_DEVICE, "crossing record queue overflowed");
    }
  } else {
    for (const Launch& L : seq) {
      if ((rc = enqueue(L))) return rc;
      LMC_HIP(hipEventRecord(s->ev0, s->stream));
      hipLaunchKernelGGL(kt, dim3(grid), dim3(block), lds, s->stream, p);
      LMC_HIP(hipGetLastError());
      LMC_HIP(hipEventRecord(s->ev1, s->stream));
      if (rec_run && p.q_cap > 0 && (rc = prepare_records(s, p))) return rc;
      if (p.q_cap > 0) {
        sensors(L.n);
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
        if (qc[1] != 0) return fail(LMC_ERR_DEVICE, "crossing record queue overflowed");
      }
      if (rec_run && p.q_cap > 0 && (rc = gather_records(s))) return rc;
    }
  }
  if (binned) {
    uint32_t flag = 0;
    LMC_HIP(hipMemcpy(&flag, p.bin_flag, sizeof(flag), hipMemcpyDeviceToHost));
    if (flag) return fail(LMC_ERR_DEVICE, "binned sensor deposits: the counting and writing passes disagreed");
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
    st->transport_ms = ms_t;  // graph runs: the whole graph
    st->sensor_ms = ms_r;
    st->transport_launches = seq.size();
    st->block_size = block;
    st->blocks_per_cu = bpc;
    st->lds_tables = lds_tables ? 1 : 0;
    st->aux_mode = aux_mode;
  }
  return LMC_OK;
}

template <bool kAdj>
int run_dispatch(lmc_scene* s, const lmc_run_desc* r, const uint64_t* budgets, bool use_graph, lmc_stats* st,
                 unsigned char* scratch = nullptr) {
  if (!s || !r) return fail(LMC_ERR_INVALID, "null argument");
  if (r->max_events < 0) return fail(LMC_ERR_INVALID, "max_events must be >= 0");
  if (!budgets && r->n_photons == 0) {
    if (st) std::memset(st, 0, sizeof(*st));
    return LMC_OK;
  }
  const char* no_tables = "no-alloc run: this precision's tables are not on the device yet (FP32 tables are "
                          "uploaded by lmc_scene_create, FP64 ones by a first lmc_scene_run)";
  if (r->precision == LMC_PREC_FP32) {
    if (r->rng != LMC_RNG_PHILOX) return fail(LMC_ERR_UNSUPPORTED, "FP32 transport supports the Philox stream only");
    if (scratch && !s->t32.ready) return fail(LMC_ERR_INVALID, no_tables);
    int rc = build_tables<float>(s, s->t32);
    if (rc) return rc;
    return run_impl<float, false, kAdj>(s, s->t32, r, budgets, use_graph, st, scratch);
  }
  if (r->precision == LMC_PREC_FP64) {
    if (scratch && !s->t64.ready) return fail(LMC_ERR_INVALID, no_tables);
    int rc = build_tables<double>(s, s->t64);
    if (rc) return rc;
    if (r->rng == LMC_RNG_XOSHIRO) return run_impl<double, true, kAdj>(s, s->t64, r, budgets, use_graph, st, scratch);
    return run_impl<double, false, kAdj>(s, s->t64, r, budgets, use_graph, st, scratch);
  }
  return fail(LMC_ERR_INVALID, "unknown precision");
}

// Conservative bounds of a sensor's grid for the queueing test in the
// transport kernel: a crossing outside them cannot be binned by the sensor
// kernel (which applies the exact binning), so it is not queued.  Margins
// cover FP32 rounding of the compared quantities.
void queue_bounds(SensorDev& o) {
  const float inf = std::numeric_limits<float>::infinity();
  o.q_cmin = -inf;
  o.q_x0 = o.q_y0 = o.q_t0 = -inf;
  o.q_x1 = o.q_y1 = o.q_t1 = inf;
  auto cos_floor = [](double theta_max) -> float {
    if (!(theta_max < 3.0)) return -std::numeric_limits<float>::infinity();
    return static_cast<float>(std::cos(theta_max) - 1e-5);
  };
  switch (o.kind) {
    case LMC_SENSOR_FARFIELD_CBS:
      o.q_cmin = cos_floor(o.n0 * o.g[3]);
      if (o.g[5] > 0) {
        o.q_t0 = -1e-6f;
        o.q_t1 = static_cast<float>(o.g[2] * (1 + 1e-5) + 1e-6);
      }
      break;
    case LMC_SENSOR_FARFIELD_FLUENCE:
      o.q_cmin = cos_floor(o.n0 * o.g[2]);
      break;
    case LMC_SENSOR_PLANAR_FLUENCE:
    case LMC_SENSOR_PLANAR_FIELD: {
      const bool flu = o.kind == LMC_SENSOR_PLANAR_FLUENCE;
      const double wx = o.n0 * (flu ? o.g[3] : o.g[2]), wy = o.n1 * (flu ? o.g[4] : o.g[3]);
      const double mx = 1e-5 * (std::fabs(wx) + std::fabs(o.g[0])) + 1e-6;
      const double my = 1e-5 * (std::fabs(wy) + std::fabs(o.g[1])) + 1e-6;
      // the pixel index truncates toward zero: (-1, 0) still lands in pixel 0
      const double dx = flu ? o.g[3] : o.g[2], dy = flu ? o.g[4] : o.g[3];
      o.q_x0 = static_cast<float>(-0.5 * o.g[0] - std::fabs(dx) - mx);
      o.q_x1 = static_cast<float>(-0.5 * o.g[0] + wx + mx);
      o.q_y0 = static_cast<float>(-0.5 * o.g[1] - std::fabs(dy) - my);
      o.q_y1 = static_cast<float>(-0.5 * o.g[1] + wy + my);
      if (flu && o.g[5] > 0) {
        o.q_t0 = -1e-6f;
        o.q_t1 = static_cast<float>(o.g[2] * (1 + 1e-5) + 1e-6);
      }
      break;
    }
    default:
      break;
  }
}

// Append one validated scene to a set: absolute species, table, sensor and
// accumulator offsets; sensors grouped into planes with the reference's 1e-9
// key tolerance (detector.hpp:46-49), ascending z.
int add_scene(lmc_scene* s, const lmc_scene_desc* d, bool privatise) {
  int rc = validate(d);
  if (rc) return rc;
  const int ll = launch_layer_of(d, d->laser.pos[2]);
  if (ll < 0) return fail(LMC_ERR_INVALID, "laser position lies outside the sample");
  SceneInfo sc;
  sc.layer_base = static_cast<int>(s->layers.size());
  sc.n_layers = d->n_layers;
  sc.species_base = static_cast<int>(s->species.size());
  sc.n_species = d->n_species;
  sc.launch_layer = sc.layer_base + ll;
  sc.out_base = s->n_out;
  sc.n_out = d->n_out;
  sc.absorption = d->absorption;
  sc.k_wave = 2.0 * M_PI / d->laser.wavelength;
  sc.n_sensors = d->n_sensors;
  sc.laser = d->laser;
  const int64_t tbase = static_cast<int64_t>(s->tables.size());
  const int scene_idx = static_cast<int>(s->scenes.size());
  if (d->n_tables) s->tables.insert(s->tables.end(), d->tables, d->tables + d->n_tables);
  for (int i = 0; i < d->n_layers; ++i) {
    lmc_layer_desc L = d->layers[i];
    L.species_first += sc.species_base;
    if (L.species_count > 1) L.select_offset += tbase;
    s->layers.push_back(L);
    s->layer_scene.push_back(scene_idx);
  }
  for (int i = 0; i < d->n_species; ++i) {
    lmc_species_desc S = d->species[i];
    S.cdf_offset += tbase;
    S.s1_offset += tbase;
    S.s2_offset += tbase;
    s->species.push_back(S);
  }
  std::vector<double> zs;
  for (int i = 0; i < d->n_sensors; ++i) {
    const double z = d->sensors[i].z;
    bool found = false;
    for (double zz : zs)
      if (!(z < zz - 1e-9) && !(zz < z - 1e-9)) found = true;
    if (!found) zs.push_back(z);
  }
  std::sort(zs.begin(), zs.end());
  sc.plane_first = static_cast<int>(s->planes_host.size());
  sc.plane_count = static_cast<int>(zs.size());
  const double ztop_s = d->layers[d->n_layers - 1].z_max;
  for (size_t pl = 0; pl < zs.size(); ++pl) {
    PlaneDev P;
    std::memset(&P, 0, sizeof(P));
    P.z = zs[pl];
    P.zf = static_cast<float>(zs[pl]);
    P.first = static_cast<int32_t>(s->sensors_host.size());
    P.count = 0;
    P.scene = scene_idx;
    bool all_absorb = true;
    int n_rec_plane = 0;
    for (int i = 0; i < d->n_sensors; ++i) {
      const lmc_sensor_desc& sd = d->sensors[i];
      if ((sd.z < zs[pl] - 1e-9) || (zs[pl] < sd.z - 1e-9)) continue;
      SensorDev o;
      std::memset(&o, 0, sizeof(o));
      o.kind = sd.kind;
      o.absorb = sd.absorb;
      o.plane = static_cast<int32_t>(sc.plane_first + pl);
      o.id = i;
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
      o.out_off = sc.out_base + sd.out_offset;
      queue_bounds(o);
      if (sd.kind == LMC_SENSOR_STATISTICS) {  // histogram parameters; n0 = accumulator length
        o.aux = static_cast<int32_t>(s->aux_host.size());
        s->aux_host.insert(s->aux_host.end(), d->tables + sd.stat_offset, d->tables + sd.stat_offset + 27);
        o.n0 = static_cast<int32_t>(sd.out_len);
        o.n1 = o.n2 = 0;
      }
      o.est = sd.estimator ? 1 : 0;
      // forward runs deposit the light sensors' crossings in the transport
      // kernel (no queue record); gradient runs still queue them
      o.direct = (!o.est && (sd.kind == LMC_SENSOR_PLANAR_FLUENCE || sd.kind == LMC_SENSOR_FARFIELD_FLUENCE ||
                             sd.kind == LMC_SENSOR_REFLECTANCE))
                     ? 1
                     : 0;
      o.est_lds = -1;
      if (o.est) {
        // no crossings reach an estimator-mode sensor (detector.cpp:246-253)
        o.absorb = 0;
        o.q_cmin = 2.0f;
        if (sd.kind == LMC_SENSOR_FARFIELD_CBS) {
          // bin centres and theta-band solid angles (detector.cpp:884-907),
          // explicit phi columns (set_phi_slices, :909-943), theta cap (:1382)
          const double theta_max = sd.g[0], dth = sd.g[3], dph = sd.g[4];
          const double th_cap = sd.theta_pp_max > 0.0 ? std::min(sd.theta_pp_max, theta_max) : theta_max;
          o.est_imax = std::min(sd.n[0], static_cast<int>(th_cap / dth));
          o.aux = static_cast<int32_t>(s->aux_host.size());
          auto& eg = s->aux_host;
          for (int it = 0; it < sd.n[0]; ++it) eg.push_back(std::cos((it + 0.5) * dth));
          for (int it = 0; it < sd.n[0]; ++it) eg.push_back(std::sin((it + 0.5) * dth));
          for (int it = 0; it < sd.n[0]; ++it) eg.push_back(std::cos(it * dth) - std::cos((it + 1) * dth));
          for (int jp = 0; jp < sd.n[1]; ++jp) {
            const double ph = sd.n_phi_explicit > 0 ? d->tables[sd.phi_offset + jp] : (jp + 0.5) * dph;
            eg.push_back(std::cos(ph));
          }
          for (int jp = 0; jp < sd.n[1]; ++jp) {
            const double ph = sd.n_phi_explicit > 0 ? d->tables[sd.phi_offset + jp] : (jp + 0.5) * dph;
            eg.push_back(std::sin(ph));
          }
          // FP32 kernels read per bin the forced direction in the lab frame
          // (e1 a1 + e2 a2 - s_in cos th, the laser's frame) and dOmega dphi:
          // 16-byte records from the next multiple of 4
          while (eg.size() % 4) eg.push_back(0.0);
          const double* e1 = d->laser.m;
          const double* e2 = d->laser.n;
          const double* s_in = d->laser.dir;
          for (int it = 0; it < sd.n[0]; ++it) {
            const double cth = std::cos((it + 0.5) * dth), sth = std::sin((it + 0.5) * dth);
            const double dom = std::cos(it * dth) - std::cos((it + 1) * dth);
            for (int jp = 0; jp < sd.n[1]; ++jp) {
              const double ph = sd.n_phi_explicit > 0 ? d->tables[sd.phi_offset + jp] : (jp + 0.5) * dph;
              const double a1 = sth * std::cos(ph), a2 = sth * std::sin(ph);
              for (int j = 0; j < 3; ++j) eg.push_back(e1[j] * a1 + e2[j] * a2 - s_in[j] * cth);
              eg.push_back(dom * dph);
            }
          }
          // the transport kernel's LDS copy of the grid when it is small (<= 48 KiB in all)
          if ((s->est_lds_n + sd.out_len) * 8 <= 49152) {
            o.est_lds = s->est_lds_n;
            s->est_lds_n += static_cast<int>(sd.out_len);
          }
          s->has_est = true;
        }
      }
      // privatise small angular grids per workgroup of the sensor kernel (<= 48 KiB)
      const bool small = privatise && !o.est &&
                         (sd.kind == LMC_SENSOR_FARFIELD_CBS || sd.kind == LMC_SENSOR_FARFIELD_FLUENCE ||
                          sd.kind == LMC_SENSOR_REFLECTANCE || sd.kind == LMC_SENSOR_STATISTICS) &&
                         (s->lds_grid + sd.out_len) * 8 <= 49152;
      if (small) {
        o.lds = 1;
        o.lds_off = s->lds_grid;
        s->lds_grid += static_cast<int>(sd.out_len);
      }
      // binned deposits (forward runs): a grid far beyond the caches; then
      // the set's other HBM grids too (finish_set)
      o.bin = (bin_eligible(o) && sd.out_len * 8 >= kBinMinBytes) ? 1 : 0;
      s->sensors_host.push_back(o);
      s->sensor_scene.push_back(scene_idx);
      P.count++;
      if (sd.kind == LMC_SENSOR_FARFIELD_CBS && !o.est) P.flags |= kPlaneCbs;
      if (sd.kind == LMC_SENSOR_PHOTON_RECORD) {
        P.flags |= kPlaneRec;
        ++n_rec_plane;
      }
      if (!o.absorb) all_absorb = false;
    }
    // a plane on or outside the stack's faces is crossed only by a photon
    // leaving the sample; an interior plane whose sensors all absorb records
    // at most one crossing as well
    const bool boundary = zs[pl] <= 0.0 || (!std::isinf(ztop_s) && zs[pl] >= ztop_s);
    if (!boundary || zs[pl] == d->laser.pos[2]) s->flight_planes = true;
    if (boundary || all_absorb)
      sc.n_single_planes++;
    else
      sc.n_multi_planes++;
    s->planes_host.push_back(P);
    s->rec_per_crossing = std::max(s->rec_per_crossing, n_rec_plane);
  }
  s->n_out += d->n_out;
  s->scenes.push_back(sc);
  return LMC_OK;
}

int finish_set(lmc_scene* s) {
  // two plans: runs that queue every crossing (bin), and runs whose transport
  // kernel deposits the direct sensors (bin_d: those grids are no trigger)
  bool any_bin = false, any_bin_d = false;
  for (const SensorDev& o : s->sensors_host) {
    any_bin = any_bin || o.bin;
    any_bin_d = any_bin_d || (o.bin && !o.direct);
  }
  for (SensorDev& o : s->sensors_host) {
    o.bin_d = (any_bin_d && bin_eligible(o) && !o.direct) ? 1 : 0;
    if (any_bin) o.bin = bin_eligible(o) ? 1 : 0;
  }
  int rc = s->sensors_dev.upload(s->sensors_host.data(), s->sensors_host.size());
  if (!rc) rc = s->planes_dev.upload(s->planes_host.data(), s->planes_host.size());
  if (!rc) rc = s->out.alloc(std::max<int64_t>(1, s->n_out));
  if (!rc) rc = s->hits.alloc(std::max<size_t>(1, s->sensors_host.size()));
  if (!rc) rc = s->counter.alloc(8);
  if (!rc) rc = s->qcount.alloc(2);
  if (rc) return rc;
  if (!rc && s->rec_per_crossing > 0) rc = s->rec_count.alloc(2);
  if (rc) return rc;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess ||
      hipEventCreate(&s->ev2) != hipSuccess)
    return fail(LMC_ERR_DEVICE, "stream/event creation failed");
  return LMC_OK;
}

void destroy_set(lmc_scene* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_graph(s);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->ev2) (void)hipEventDestroy(s->ev2);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  s->ev0 = s->ev1 = s->ev2 = nullptr;
  s->stream = nullptr;
}

int zero_set(lmc_scene* s) {
  LMC_HIP(hipMemsetAsync(s->out.p, 0, sizeof(double) * s->out.n, s->stream));
  LMC_HIP(hipMemsetAsync(s->hits.p, 0, sizeof(unsigned long long) * s->hits.n, s->stream));
  if (s->grad.p) LMC_HIP(hipMemsetAsync(s->grad.p, 0, sizeof(double) * s->grad.n, s->stream));
  LMC_HIP(hipStreamSynchronize(s->stream));
  s->rec_host.clear();
  return LMC_OK;
}

// The sensor kernel deposits a time-resolved CBS or planar-fluence hit in its
// time bin t >= 1 only; the reference also adds it to the time-integrated bin
// 0 (detector.cpp:1223-1225, :829-840).  Bin 0 of a read-out copy of scene k's
// accumulators (`buf` holds n_out values from out_base on) becomes its own
// steady-state deposits plus the sum of bins 1..n_t-1.
void fold_time_bins(const lmc_scene* s, int k, double* buf) {
  const SceneInfo& sc = s->scenes[k];
  for (size_t j = 0; j < s->sensors_host.size(); ++j) {
    if (s->sensor_scene[j] != k) continue;
    const SensorDev& sd = s->sensors_host[j];
    const int nq = sd.kind == LMC_SENSOR_FARFIELD_CBS ? 8 : (sd.kind == LMC_SENSOR_PLANAR_FLUENCE ? 4 : 0);
    if (!nq || sd.n2 < 2) continue;
    const int64_t tplane = static_cast<int64_t>(sd.n0) * sd.n1, plane = tplane * sd.n2;
    for (int q = 0; q < nq; ++q) {
      double* g = buf + (sd.out_off - sc.out_base) + q * plane;
      for (int t = 1; t < sd.n2; ++t)
        for (int64_t i = 0; i < tplane; ++i) g[i] += g[t * tplane + i];
    }
  }
}

// Scene k's PhotonRecordSensor records: min(max_records, count) of them into
// `out` (may be null), the count into *n_records.
int copy_records(const lmc_scene* s, int k, double* out, int64_t max_records, int64_t* n_records) {
  const std::vector<double>* v = static_cast<size_t>(k) < s->rec_host.size() ? &s->rec_host[k] : nullptr;
  const int64_t n = v ? static_cast<int64_t>(v->size() / LMC_RECORD_FIELDS) : 0;
  if (n_records) *n_records = n;
  if (v && out && max_records > 0)
    std::memcpy(out, v->data(), sizeof(double) * LMC_RECORD_FIELDS * std::min(n, max_records));
  return LMC_OK;
}

// Add scene k's accumulators and hit counters into caller buffers.
int read_scene(lmc_scene* s, int k, double* out, int64_t n, uint64_t* hits) {
  const SceneInfo& sc = s->scenes[k];
  if (out) {
    if (n != sc.n_out) return fail(LMC_ERR_INVALID, "read size does not match the scene's accumulators");
    std::vector<double> tmp(sc.n_out);
    if (sc.n_out)
      LMC_HIP(hipMemcpy(tmp.data(), s->out.p + sc.out_base, sizeof(double) * sc.n_out, hipMemcpyDeviceToHost));
    fold_time_bins(s, k, tmp.data());
    for (int64_t i = 0; i < sc.n_out; ++i) out[i] += tmp[i];
  }
  if (hits) {
    std::vector<unsigned long long> h(s->hits.n);
    LMC_HIP(hipMemcpy(h.data(), s->hits.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
    for (size_t j = 0; j < s->sensors_host.size(); ++j)
      if (s->sensor_scene[j] == k) hits[s->sensors_host[j].id] += h[j];
  }
  return LMC_OK;
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
  lmc_scene* s = new lmc_scene();
  rc = add_scene(s, d, true);
  if (!rc) rc = finish_set(s);
  if (!rc) rc = zero_set(s);
  // the production tables now, so a no-alloc run can follow at once
  if (!rc) rc = build_tables<float>(s, s->t32);
  if (rc) {
    destroy_set(s);
    delete s;
    return rc;
  }
  *out = s;
  return LMC_OK;
}

void lmc_scene_destroy(lmc_scene* s) {
  if (!s) return;
  destroy_set(s);
  delete s;
}

int lmc_scene_zero(lmc_scene* s) {
  if (!s) return fail(LMC_ERR_INVALID, "null scene");
  return zero_set(s);
}

int lmc_scene_run(lmc_scene* s, const lmc_run_desc* r, lmc_stats* st) {
  return run_dispatch<false>(s, r, nullptr, false, st);
}

// lmc_scene_run without any device allocation: the crossing queue is the
// static `scratch` array.  Runs on one device are serialised by a mutex
// (they share that device's array); the array keeps the last run's records,
// and each launch rewrites the plane sentinels it can hand out before use.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int lmc_scene_run_noalloc(lmc_scene* s, const lmc_run_desc* r, lmc_stats* st) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  int dev = 0;
  LMC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(LMC_ERR_DEVICE, "device index out of range");
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(LMC_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  return run_dispatch<false>(s, r, nullptr, false, st, g_scratch_ptr[dev]);
}

int lmc_scene_read(lmc_scene* s, double* out, int64_t n, uint64_t* hits) {
  if (!s) return fail(LMC_ERR_INVALID, "null scene");
  return read_scene(s, 0, out, n, hits);
}

int lmc_scene_records(lmc_scene* s, double* out, int64_t max_records, int64_t* n_records) {
  if (!s) return fail(LMC_ERR_INVALID, "null scene");
  if (max_records < 0) return fail(LMC_ERR_INVALID, "max_records must be >= 0");
  return copy_records(s, 0, out, max_records, n_records);
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
  }
  LMC_HIP(hipMemset(s->grad.p, 0, sizeof(double) * need));
  int rc = run_dispatch<true>(s, r, nullptr, false, st);
  if (rc) return rc;
  if (grad) {
    std::vector<double> tmp(need);
    LMC_HIP(hipMemcpy(tmp.data(), s->grad.p, sizeof(double) * need, hipMemcpyDeviceToHost));
    for (int q = 0; q < 3 * nl; ++q) fold_time_bins(s, 0, tmp.data() + q * s->n_out);  // rows share the layout
    for (int64_t i = 0; i < need; ++i) grad[i] += tmp[i];
  }
  return LMC_OK;
}

int lmc_batch_create(int32_t n_scenes, const lmc_scene_desc* scenes, lmc_batch** out) {
  if (!out) return fail(LMC_ERR_INVALID, "null output pointer");
  *out = nullptr;
  if (n_scenes <= 0 || n_scenes > kMaxBatchScenes || !scenes)
    return fail(LMC_ERR_INVALID, "a batch holds 1..4096 scenes");
  int rc = ensure_device();
  if (rc) return rc;
  lmc_batch* b = new lmc_batch();
  for (int k = 0; k < n_scenes && !rc; ++k) rc = add_scene(&b->set, &scenes[k], false);
  if (!rc && b->set.sensors_host.size() > 8192) rc = fail(LMC_ERR_INVALID, "a batch holds at most 8192 sensors");
  if (!rc) rc = finish_set(&b->set);
  if (!rc) rc = zero_set(&b->set);
  if (rc) {
    destroy_set(&b->set);
    delete b;
    return rc;
  }
  *out = b;
  return LMC_OK;
}

void lmc_batch_destroy(lmc_batch* b) {
  if (!b) return;
  destroy_set(&b->set);
  delete b;
}

int lmc_batch_run(lmc_batch* b, const uint64_t* budgets, const lmc_run_desc* r, int32_t use_graph,
                  lmc_stats* st) {
  if (!b || !budgets || !r) return fail(LMC_ERR_INVALID, "null argument");
  return run_dispatch<false>(&b->set, r, budgets, use_graph != 0, st);
}

int lmc_batch_read(lmc_batch* b, int32_t scene, double* out, int64_t n, uint64_t* hits) {
  if (!b) return fail(LMC_ERR_INVALID, "null batch");
  if (scene < 0 || scene >= static_cast<int32_t>(b->set.scenes.size()))
    return fail(LMC_ERR_INVALID, "scene index out of range");
  return read_scene(&b->set, scene, out, n, hits);
}

int lmc_batch_records(lmc_batch* b, int32_t scene, double* out, int64_t max_records, int64_t* n_records) {
  if (!b) return fail(LMC_ERR_INVALID, "null batch");
  if (scene < 0 || scene >= static_cast<int32_t>(b->set.scenes.size()))
    return fail(LMC_ERR_INVALID, "scene index out of range");
  if (max_records < 0) return fail(LMC_ERR_INVALID, "max_records must be >= 0");
  return copy_records(&b->set, scene, out, max_records, n_records);
}

int lmc_batch_zero(lmc_batch* b) {
  if (!b) return fail(LMC_ERR_INVALID, "null batch");
  return zero_set(&b->set);
}

int lmc_probe_azimuth(int32_t n, int32_t precision, const double* abc, const uint32_t* words, double* out) {
  if (n <= 0 || !abc || !words || !out) return fail(LMC_ERR_INVALID, "bad probe arguments");
  if (precision != LMC_PREC_FP32 && precision != LMC_PREC_FP64) return fail(LMC_ERR_INVALID, "unknown precision");
  int rc = ensure_device();
  if (rc) return rc;
  double *d_abc = nullptr, *d_out = nullptr;
  uint32_t* d_w = nullptr;
  LMC_HIP(hipMalloc(&d_abc, sizeof(double) * 3 * n));
  LMC_HIP(hipMalloc(&d_out, sizeof(double) * 3 * n));
  LMC_HIP(hipMalloc(&d_w, sizeof(uint32_t) * 2 * n));
  LMC_HIP(hipMemcpy(d_abc, abc, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  LMC_HIP(hipMemcpy(d_w, words, sizeof(uint32_t) * 2 * n, hipMemcpyHostToDevice));
  if (precision == LMC_PREC_FP32)
    hipLaunchKernelGGL(k_azimuth_probe<float>, dim3((n + 255) / 256), dim3(256), 0, 0, n, d_abc, d_w, d_out);
  else
    hipLaunchKernelGGL(k_azimuth_probe<double>, dim3((n + 255) / 256), dim3(256), 0, 0, n, d_abc, d_w, d_out);
  LMC_HIP(hipGetLastError());
  LMC_HIP(hipMemcpy(out, d_out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  LMC_HIP(hipFree(d_abc));
  LMC_HIP(hipFree(d_out));
  LMC_HIP(hipFree(d_w));
  return LMC_OK;
}

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

int lmc_probe_reverse(int32_t n, int32_t precision, const double* frames, double* out) {
  if (n <= 0 || n > (1 << 22) || !frames || !out) return fail(LMC_ERR_INVALID, "bad probe arguments");
  if (precision != LMC_PREC_FP32 && precision != LMC_PREC_FP64) return fail(LMC_ERR_INVALID, "unknown precision");
  int rc = ensure_device();
  if (rc) return rc;
  double *din = nullptr, *dout = nullptr;
  LMC_HIP(hipMalloc(&din, sizeof(double) * 20 * n));
  LMC_HIP(hipMalloc(&dout, sizeof(double) * 4 * n));
  LMC_HIP(hipMemcpy(din, frames, sizeof(double) * 20 * n, hipMemcpyHostToDevice));
  if (precision == LMC_PREC_FP32)
    hipLaunchKernelGGL(k_reverse_probe<float>, dim3((n + 255) / 256), dim3(256), 0, 0, n, din, dout);
  else
    hipLaunchKernelGGL(k_reverse_probe<double>, dim3((n + 255) / 256), dim3(256), 0, 0, n, din, dout);
  LMC_HIP(hipGetLastError());
  LMC_HIP(hipMemcpy(out, dout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  LMC_HIP(hipFree(din));
  LMC_HIP(hipFree(dout));
  return LMC_OK;
}

int lmc_probe_phase(int32_t n, const float* x, float* out) {
  if (n <= 0 || n > (1 << 24) || !x || !out) return fail(LMC_ERR_INVALID, "bad probe arguments");
  int rc = ensure_device();
  if (rc) return rc;
  float *dx = nullptr, *dout = nullptr;
  LMC_HIP(hipMalloc(&dx, sizeof(float) * n));
  LMC_HIP(hipMalloc(&dout, sizeof(float) * 2 * n));
  LMC_HIP(hipMemcpy(dx, x, sizeof(float) * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_phase_probe, dim3((n + 255) / 256), dim3(256), 0, 0, n, dx, dout);
  LMC_HIP(hipGetLastError());
  LMC_HIP(hipMemcpy(out, dout, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
  LMC_HIP(hipFree(dx));
  LMC_HIP(hipFree(dout));
  return LMC_OK;
}

int lmc_probe_stream(uint32_t k0, uint32_t k1, uint64_t photon, int32_t n, uint32_t* out) {
  if (n <= 0 || n > (1 << 20) || !out) return fail(LMC_ERR_INVALID, "bad probe arguments");
  int rc = ensure_device();
  if (rc) return rc;
  uint32_t* dout = nullptr;
  LMC_HIP(hipMalloc(&dout, sizeof(uint32_t) * n));
  hipLaunchKernelGGL(k_stream_probe, dim3(1), dim3(64), 0, 0, k0, k1, photon, n, dout);
  LMC_HIP(hipGetLastError());
  LMC_HIP(hipMemcpy(out, dout, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
  LMC_HIP(hipFree(dout));
  return LMC_OK;
}

}  // extern "C"
