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

#include "lmc_kernels.h"

using namespace lmc;

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
  std::vector<int> plane_single;  // 1: at most one crossing record per photon
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
  p.lpos_x = L.pos[0];
  p.lpos_y = L.pos[1];
  p.lpos_z = static_cast<R>(L.pos[2]);
  for (int j = 0; j < 3; ++j) {
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

// Crossing records per transport launch.  A run is split into chunks of
// q_cap / (worst-case records per photon) photons so the queue never fills:
// one record per plane a photon can cross only when leaving the sample (or
// whose sensors all absorb), max_events + 2 for other interior planes.
constexpr int64_t kQueueRecords = int64_t(1) << 23;

template <class R> int ensure_queue(lmc_scene* s, int64_t cap) {
  cap = (cap + 1) & ~int64_t(1);
  const size_t rec = sizeof(R) * kRecF + sizeof(double) + sizeof(int32_t) * kRecI;
  if (s->queue.p && s->q_cap >= cap && s->q_rec_bytes == rec) return LMC_OK;
  int rc = s->queue.alloc(static_cast<size_t>(cap) * rec + 256);
  if (rc) return rc;
  s->q_cap = cap;
  s->q_rec_bytes = rec;
  return LMC_OK;
}

template <class R, bool kXo, bool kAdj>
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
  p.key0 = static_cast<uint32_t>(r->seed);
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
  size_t base_lds = 0;
  auto add = [&](size_t b) { base_lds += (b + 15) & ~static_cast<size_t>(15); };
  add(sizeof(LayerDev<R>) * p.n_layers);
  add(sizeof(SpeciesDev<R>) * p.n_species);
  add(sizeof(SensorDev) * p.n_sensors);
  add(sizeof(PlaneDev) * p.n_planes);
  add(sizeof(unsigned) * (kMaxSensors + 2));
  add(sizeof(LaunchConst<R>));
  const size_t tab_lds = ((sizeof(R) * T.cdf_total + 15) & ~15ull) +
                         ((sizeof(uint16_t) * T.guide_total + 4 + 15) & ~15ull) +
                         ((sizeof(STab<R>) * T.stab_total + 15) & ~15ull) + ((sizeof(R) * T.sel_total + 15) & ~15ull);
  // Launch shape: 16 resident waves per CU (4 per SIMD at <= 128 VGPRs).  The
  // tables go to LDS when two 512-thread workgroups per CU can hold them;
  // otherwise 256-thread workgroups read them through the caches.
  const bool fit512 = 2 * (base_lds + tab_lds) <= g_max_lds_cu && base_lds + tab_lds <= g_max_lds_block;
  int block = r->block_size > 0 ? r->block_size : (fit512 ? 512 : 256);
  if (block % 64 != 0 || block > 512) return fail(LMC_ERR_INVALID, "block_size must be a multiple of 64, <= 512");
  const int bpc = r->blocks_per_cu > 0 ? r->blocks_per_cu : std::max(1, 1024 / block);
  const int grid = g_num_cu * bpc;
  const bool lds_tables = (base_lds + tab_lds) * static_cast<size_t>(std::min(bpc, 1024 / block)) <= g_max_lds_cu &&
                          base_lds + tab_lds <= g_max_lds_block;
  const size_t lds = lds_tables ? base_lds + tab_lds : base_lds;
  if (lds > g_max_lds_block) return fail(LMC_ERR_UNSUPPORTED, "scene constants exceed LDS");
  const bool track = r->track_reverse_paths != 0;
  int64_t per_photon = 0;
  for (size_t i = 0; i < s->planes_host.size(); ++i)
    per_photon += s->plane_single[i] ? 1 : static_cast<int64_t>(r->max_events) + 2;
  int64_t chunk = static_cast<int64_t>(r->n_photons);
  if (per_photon > 0) {
    const int64_t cap = std::min<int64_t>(kQueueRecords, static_cast<int64_t>(r->n_photons) * per_photon);
    int rc = ensure_queue<R>(s, cap);
    if (rc) return rc;
    chunk = std::max<int64_t>(1, s->q_cap / per_photon);
    p.qf = reinterpret_cast<R*>(s->queue.p);
    p.qd = reinterpret_cast<double*>(s->queue.p + sizeof(R) * kRecF * s->q_cap);
    p.qi = reinterpret_cast<int32_t*>(s->queue.p + (sizeof(R) * kRecF + sizeof(double)) * s->q_cap);
    p.q_cap = s->q_cap;
  }
  p.q_count = s->qcount.p;
  void (*kt)(KParams<R>) = nullptr;
  if (track)
    kt = lds_tables ? k_transport<R, kXo, true, true, kAdj> : k_transport<R, kXo, true, false, kAdj>;
  else
    kt = lds_tables ? k_transport<R, kXo, false, true, kAdj> : k_transport<R, kXo, false, false, kAdj>;
  int rc = raise_lds_limit(kt, lds);
  if (rc) return rc;
  void (*kr)(KParams<R>) = k_crossings<R, kAdj>;
  const size_t lds_r = ((sizeof(double) * p.lds_grid + 15) & ~15ull) + sizeof(unsigned) * (kMaxSensors + 1);
  if ((rc = raise_lds_limit(kr, lds_r))) return rc;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s->stream, s->counter.p, 8);
  LMC_HIP(hipGetLastError());
  float ms_t = 0.f, ms_r = 0.f;
  uint64_t done = 0, launches = 0;
  while (done < r->n_photons) {
    ++launches;
    uint64_t n = std::min<uint64_t>(static_cast<uint64_t>(chunk), r->n_photons - done);
    // photon indices of one launch share their high word (a uniform register)
    const uint64_t first = r->photon_offset + done;
    const uint64_t to_wrap = (uint64_t(1) << 32) - (first & 0xFFFFFFFFull);
    if (n > to_wrap) n = to_wrap;
    p.n_photons = n;
    p.offset = first;
    p.idx_hi = static_cast<uint32_t>(first >> 32);
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
    st->sensor_ms = ms_r;
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
    return run_impl<float, false, kAdj>(s, s->t32, r, st);
  }
  if (r->precision == LMC_PREC_FP64) {
    int rc = build_tables<double>(s, s->t64);
    if (rc) return rc;
    if (r->rng == LMC_RNG_XOSHIRO) return run_impl<double, true, kAdj>(s, s->t64, r, st);
    return run_impl<double, false, kAdj>(s, s->t64, r, st);
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
    std::memset(&P, 0, sizeof(P));
    P.z = zs[pl];
    P.first = static_cast<int32_t>(s->sensors_host.size());
    P.count = 0;
    bool all_absorb = true;
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
      // privatise small angular grids per workgroup of the sensor kernel (<= 48 KiB)
      const bool small = (sd.kind == LMC_SENSOR_FARFIELD_CBS || sd.kind == LMC_SENSOR_FARFIELD_FLUENCE ||
                          sd.kind == LMC_SENSOR_REFLECTANCE) &&
                         (lds_grid + sd.out_len) * 8 <= 49152;
      if (small) {
        o.lds = 1;
        o.lds_off = lds_grid;
        lds_grid += static_cast<int>(sd.out_len);
      }
      s->sensors_host.push_back(o);
      P.count++;
      if (sd.kind == LMC_SENSOR_FARFIELD_CBS) P.has_cbs = 1;
      if (!sd.absorb) all_absorb = false;
    }
    // a plane on or outside the stack's faces is crossed only by a photon
    // leaving the sample; an interior plane whose sensors all absorb records
    // at most one crossing as well
    const double ztop_s = s->layers.back().z_max;
    const bool boundary = zs[pl] <= 0.0 || (!std::isinf(ztop_s) && zs[pl] >= ztop_s);
    s->plane_single.push_back(boundary || all_absorb ? 1 : 0);
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
