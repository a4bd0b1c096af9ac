/*
 * luminis_hip.h -- C ABI of the MI355X (gfx950) photon-transport library.
 *
 * This is the drop-in boundary for luminis-mc's transport hot path.  The
 * reference reaches that path through one pybind11 call,
 *   run_simulation_parallel(config)         python/python_module.cpp:664-673
 *                                           -> src/simulation.cpp:105-249
 * which takes a SimConfig holding raw pointers to a Sample, a Laser and a
 * SensorsGroup (include/luminis/core/simulation.hpp:67-90) and *adds* into
 * the sensors' accumulators (src/detector.cpp:283-297).  Here the same
 * objects are flattened by the host into plain-old-data descriptors
 * (lmc_scene_desc), uploaded once (lmc_scene_create), run any number of times
 * (lmc_scene_run adds into device-resident FP64 accumulators) and read back
 * by adding into caller-owned host buffers laid out exactly like the
 * reference's Matrix/CMatrix grids (include/luminis/math/vec.hpp:116-170).
 *
 * Plain C: status codes, no exceptions, no torch types, caller-owned buffers.
 * Every call is synchronous unless documented otherwise.  A scene object is
 * not re-entrant (like a SensorsGroup, src/simulation.cpp:105).
 */
#ifndef LUMINIS_HIP_H
#define LUMINIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LMC_ABI_VERSION 2

/* Status codes (reference: std::invalid_argument -> ValueError,
 * src/simulation.cpp:95-103; fatal exits become LMC_ERR_*). */
enum {
  LMC_OK = 0,
  LMC_ERR_INVALID = 1,     /* bad descriptor / argument (ValueError)      */
  LMC_ERR_DEVICE = 2,      /* HIP runtime failure / no device             */
  LMC_ERR_NOMEM = 3,       /* device allocation failed                    */
  LMC_ERR_UNSUPPORTED = 4  /* valid but not implemented on this path      */
};

/* ---- theta samplers (include/luminis/sample/phase.hpp; src/phase.cpp) ---- */
enum {
  LMC_PHASE_TABLE_THETA = 0, /* SamplingTable over theta: Mie, RayleighDebye(EMC)
                                (src/phase.cpp:132-137,162-167,262-268)        */
  LMC_PHASE_TABLE_COS = 1,   /* SamplingTable over cos(theta); theta = acos:
                                Rayleigh, Draine (src/phase.cpp:90-95,230-235) */
  LMC_PHASE_HG = 2,          /* Henyey-Greenstein closed form (src/phase.cpp:108-115) */
  LMC_PHASE_UNIFORM = 3      /* cos = 2u-1 (src/phase.cpp:78-83)               */
};

/* One scattering species: a ScatteringMedium plus its PhaseFunction
 * (include/luminis/core/medium.hpp:71-189).  All tables live in the
 * scene-wide `tables` array; offsets are in doubles. */
typedef struct lmc_species_desc {
  int32_t phase_kind;   /* LMC_PHASE_*                                        */
  int32_t cdf_n;        /* SamplingTable nodes (= nDiv+1), table kinds only    */
  int64_t cdf_offset;   /* normalised CDF, cdf_n doubles (src/table.cpp:6-48)  */
  double cdf_min;       /* node abscissa i = cdf_min + i*cdf_step              */
  double cdf_step;
  double hg_g;          /* LMC_PHASE_HG asymmetry g (must be != 0)             */
  int32_t s_n;          /* DataTable nodes on theta_i = i*s_dtheta             */
  int32_t reserved0;
  int64_t s1_offset;    /* S1(theta): s_n complex values, re/im interleaved    */
  int64_t s2_offset;    /* S2(theta): idem  (src/medium.cpp:99-121,203-236)   */
  double s_dtheta;
  double mu_s, mu_a;    /* species coefficients [1/mm]                         */
  double mfp;           /* species mean_free_path field [mm] (medium.cpp:123-126) */
  double sigma_s;       /* scattering cross-section (mixtures only)            */
} lmc_species_desc;

/* One z-slab of the Sample (include/luminis/core/sample.hpp:68-204).
 * species_count == 1: HomogeneousLayer; > 1: MixtureLayer whose normalised
 * selection CDF (species_count doubles) sits at tables[select_offset]. */
typedef struct lmc_layer_desc {
  double z_min, z_max;  /* [z_min, z_max); z_max may be +inf                  */
  double mu_s, mu_a, mu_t; /* aggregate coefficients (albedo split, interface
                              optical-depth rescale: simulation.cpp:336,373,580-587) */
  double mfp;           /* free-path mean of the layer (HomogeneousLayer:
                           medium mfp; MixtureLayer: 1/mu_s_total, sample.cpp:77) */
  double n;             /* refractive index used for the optical path (uniform
                           host index in the reference, sample.cpp:110-116) */
  int32_t species_first;
  int32_t species_count;
  int64_t select_offset;
} lmc_layer_desc;

/* Sensor kinds (include/luminis/core/detector.hpp:280-699). */
enum {
  LMC_SENSOR_FARFIELD_CBS = 1,   /* FarFieldCBSSensor, analog process_hit      */
  LMC_SENSOR_PLANAR_FLUENCE = 2, /* PlanarFluenceSensor                        */
  LMC_SENSOR_PLANAR_FIELD = 3,   /* PlanarFieldSensor                          */
  LMC_SENSOR_FARFIELD_FLUENCE = 4, /* extension: exit-direction Stokes histogram
                                      (no reference counterpart)             */
  LMC_SENSOR_REFLECTANCE = 5,    /* extension: sum of exit weights (the
                                    PhotonRecordSensor weight sum)          */
  LMC_SENSOR_STATISTICS = 6,     /* StatisticsSensor (detector.cpp:1571-1888) */
  LMC_SENSOR_PHOTON_RECORD = 7   /* PhotonRecordSensor (detector.cpp:427-485):
                                    one record per accepted crossing, read with
                                    lmc_scene_records; no accumulators      */
};

/* One PhotonRecordSensor record (PhotonRecord, photon.hpp:259-291), doubles:
 *  [0] index of the sensor in its scene's sensor list   [1] events
 *  [2] penetration_depth  [3] launch_time  [4] arrival_time (launch_time +
 *  opticalpath, velocity 1)  [5] opticalpath (geometric path at the hit)
 *  [6] weight  [7] k = 2 pi / wavelength
 *  [8..10] position_first_scattering  [11..13] position_last_scattering
 *  [14..16] position_detector  [17..19] direction  [20..22] m  [23..25] n
 *  [26..29] polarization_forward (m re, m im, n re, n im)
 *  [30..33] polarization_reverse: the reverse-path Jones vector of this hit
 *           when a FarFieldCBSSensor earlier in the plane's sensor list
 *           binned it with events >= 2 (detector.cpp:1281-1288), else 0. */
#define LMC_RECORD_FIELDS 34

/* Crossing-direction filter values (detector.hpp CrossingDirection). */
enum { LMC_DIR_FORWARD = 0, LMC_DIR_BACKWARD = 1, LMC_DIR_BOTH = 2 };

/* One sensor plane with its filters (Sensor base, detector.hpp:95-180,
 * check_conditions detector.cpp:374-417).  Grid geometry by kind:
 *  CBS:       g[0..5] = theta_max, phi_max, t_max, dtheta, dphi, dt;
 *             n[0..2] = N_theta, N_phi, N_t.  Output: 8 grids
 *             [S0c,S1c,S2c,S3c,S0i,S1i,S2i,S3i][t][theta][phi] doubles.
 *  FLUENCE:   g = len_x, len_y, len_t, dx, dy, dt; n = N_x, N_y, N_t.
 *             Output: 4 grids [S0..S3][t][x][y] doubles.
 *  FIELD:     g = len_x, len_y, dx, dy; n = N_x, N_y.  Output: Ex grid then
 *             Ey grid, [x][y] complex (re, im) doubles.
 *  FF_FLUENCE g = theta_max, phi_max, dtheta, dphi; n = N_theta, N_phi.
 *             Output: 4 grids [S0..S3][theta][phi].
 *  REFLECTANCE: output 1 double (sum of weights).
 *  STATISTICS: the histogram parameters are 27 doubles at tables[stat_offset]:
 *             [0..5] set flags of the events, theta, phi, depth, time and
 *             weight histograms; [6] max_events; [7..10] min_theta,
 *             max_theta, n_bins_theta, dtheta; [11..14] the same for phi;
 *             [15..17] max_depth, n_bins_depth, ddepth; [18..20] max_time,
 *             n_bins_time, dtime; [21..23] max_weight, n_bins_weight,
 *             dweight; [24..26] N_t, dt, t_max (set_time_resolution).
 *             Output: the set histograms in that order, counts as doubles:
 *             events [N_t][max_events], theta [N_t][n], phi [N_t][n],
 *             depth [N_t][n], time [n], weight [n]. */
typedef struct lmc_sensor_desc {
  int32_t kind;
  int32_t absorb;
  double z;
  int32_t f_dir_en, f_dir;
  int32_t f_ev_en, f_ev_min, f_ev_max;
  int32_t f_th_en;
  double cos_th_min, cos_th_max; /* cached cosine window (detector.cpp:336-345) */
  int32_t f_ph_en, reserved0;
  double ph_min, ph_max;
  int32_t f_pos_en, reserved1;
  double x_min, x_max, y_min, y_max;
  double g[8];
  int32_t n[4];
  int64_t out_offset;   /* offset (doubles) of this sensor's accumulators     */
  int64_t out_len;      /* number of doubles                                  */
  /* Estimator mode (Sensor::estimator_enabled, detector.cpp:314-331): the
   * sensor takes no crossings (detector.cpp:246-253); a FARFIELD_CBS sensor
   * instead runs process_estimation at every scattering vertex
   * (detector.cpp:1315-1566), every other kind records nothing
   * (detector.cpp:419-422, 571-575, 844-846). */
  int32_t estimator;
  int32_t n_phi_explicit; /* > 0: the FARFIELD_CBS phi columns are these
                             angles at tables[phi_offset] (set_phi_slices,
                             detector.cpp:909-943); 0: (j + 1/2) * d_phi    */
  int64_t phi_offset;
  double theta_pp_max;  /* estimator theta cap when > 0 (detector.cpp:1382)   */
  int64_t stat_offset;  /* STATISTICS: parameter block in tables (above)      */
} lmc_sensor_desc;

/* Laser (include/luminis/core/laser.hpp:45-90, src/laser.cpp:29-99). */
enum { LMC_SOURCE_POINT = 0, LMC_SOURCE_UNIFORM = 1, LMC_SOURCE_GAUSSIAN = 2 };
enum {
  LMC_TIME_DELTA = 0, LMC_TIME_GAUSSIAN = 1, LMC_TIME_TOPHAT = 2,
  LMC_TIME_EXPONENTIAL = 3, LMC_TIME_PULSE_TRAIN = 4, LMC_TIME_CW = 5
};
typedef struct lmc_laser_desc {
  double pos[3], dir[3], m[3], n[3];
  double pol[4];        /* Em.re, Em.im, En.re, En.im                         */
  double wavelength;    /* vacuum wavelength [mm]; k0 = 2*pi/wavelength        */
  double sigma;
  int32_t source;
  int32_t temporal;
  double pulse_duration, repetition_rate, time_offset; /* ps, Hz, ps        */
} lmc_laser_desc;

/* Absorption r-z(-t) grid (Absorption, src/absortion.cpp:13-130): at every
 * scattering vertex the absorbed weight w * mu_a / mu_t of the layer (1 when
 * mu_t == 0, simulation.cpp:577-592) is added to bin (floor(r / d_r),
 * floor(z / d_z)) of time slice 0 and, when d_t > 0, of slice
 * floor(t / d_t) + 1.  Output n_t * n_r * n_z doubles, [t][r][z], at
 * out_offset of the scene's accumulators; n_r = radius / d_r + 1, n_z =
 * depth / d_z + 1, n_t = ceil(t_max / d_t) + 1 (1 when d_t == 0). */
typedef struct lmc_absorption_desc {
  int32_t enabled;
  int32_t n_r, n_z, n_t;
  double radius, depth, d_r, d_z, d_t, t_max;
  int64_t out_offset;
} lmc_absorption_desc;

typedef struct lmc_scene_desc {
  int32_t n_layers, n_species, n_sensors, reserved0;
  const lmc_layer_desc* layers;
  const lmc_species_desc* species;
  const lmc_sensor_desc* sensors;
  const double* tables;
  int64_t n_tables;     /* doubles in `tables`                                */
  int64_t n_out;        /* total accumulator doubles (sensors, absorption grid) */
  lmc_laser_desc laser;
  double n_medium;      /* host index (Sample::refractive_index)              */
  lmc_absorption_desc absorption; /* SimConfig::absorption; single-scene runs */
} lmc_scene_desc;

/* Precision / RNG modes. */
enum { LMC_PREC_FP32 = 0, LMC_PREC_FP64 = 1 };
enum {
  LMC_RNG_PHILOX = 0,   /* per-photon xoshiro128** streams seeded by Philox4x32-10
                           of (seed, photon index); production path            */
  LMC_RNG_XOSHIRO = 1   /* xoshiro256++ re-seeded per photon with mix_seed, in
                           the reference's draw order (src/simulation.cpp:197);
                           FP64 only: trajectory-level comparison mode      */
};

/* SimConfig run controls (simulation.hpp:67-90) plus launch knobs. */
typedef struct lmc_run_desc {
  uint64_t n_photons;
  uint64_t photon_offset;  /* first photon index (chunked / resumed runs)     */
  uint64_t seed;
  int32_t max_events;      /* SimConfig::MAX_EVENTS                           */
  int32_t track_reverse_paths;
  int32_t precision;       /* LMC_PREC_*                                      */
  int32_t rng;             /* LMC_RNG_*                                       */
  int32_t block_size;      /* threads per workgroup, 0 = default              */
  int32_t blocks_per_cu;   /* persistent workgroups per CU, 0 = default       */
  int32_t claim;           /* photons claimed per wave refill; 0 = default:
                              64 single scene, 256 batch, 8 estimator runs   */
  int32_t reserved0;
} lmc_run_desc;

/* Exact device-side counters of the last run (feed the FLOP model). */
typedef struct lmc_stats {
  uint64_t photons;        /* photons launched                                */
  uint64_t events;         /* scattering events                               */
  uint64_t phi_trials;     /* azimuth draws: rejection trials in the xoshiro
                              check mode, one per event otherwise             */
  uint64_t crossings;      /* sensor-plane crossings processed                */
  uint64_t cbs_detections; /* photons deposited with a reverse path (events>=2) */
  double kernel_ms;        /* device time of all kernels of the run           */
  double transport_ms;     /* device time of the transport kernel launches    */
  double sensor_ms;        /* device time of the sensor (crossing) kernel     */
  uint64_t transport_launches;
  int32_t block_size;      /* launch shape the run used                      */
  int32_t blocks_per_cu;
  int32_t lds_tables;      /* 1: phase/amplitude tables staged in LDS         */
  int32_t aux_mode;        /* transport kernel variant: 0 = FP32 forward run
                              whose sensor planes lie on or outside the
                              stack's faces and off the launch point
                              (crossings recorded at exit only, no depth),
                              1 = general, 2 = with the Absorption grid      */
} lmc_stats;

/* Device facts used for launch shaping and roofline reporting. */
typedef struct lmc_device_info {
  int32_t n_cu;
  int32_t clock_khz;
  int64_t lds_per_block;
  int64_t lds_per_cu;
  int64_t global_mem;
  char name[64];
} lmc_device_info;
int lmc_get_device_info(lmc_device_info* info);

typedef struct lmc_scene lmc_scene;

int lmc_abi_version(void);
const char* lmc_last_error(void);
int lmc_device_count(void);

/* Upload a scene (tables converted to the run precision on device) and
 * allocate its zeroed FP64 accumulators.
 * Replaces: the per-thread SensorsGroup::clone (detector.cpp:299-307) and
 * the scene construction reached from python_module.cpp:579-614. */
int lmc_scene_create(const lmc_scene_desc* desc, lmc_scene** out);
void lmc_scene_destroy(lmc_scene* s);

/* Transport n_photons photons [photon_offset, photon_offset+n_photons) and
 * ADD into the scene's device accumulators.
 * Replaces: run_simulation_parallel(const SimConfig&) (simulation.cpp:105-249)
 * with run_photon (:255-617), record_hit (detector.cpp:178-270) and the
 * sensors' process_hit (:546-568, :780-841, :1231-1310). */
int lmc_scene_run(lmc_scene* s, const lmc_run_desc* run, lmc_stats* stats);

/* lmc_scene_run without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * crossing queue is a static 64 MiB device array (one per device, held until
 * the process ends) instead of a per-scene hipMalloc'd buffer, so runs split
 * into launches of ~2e5 photons (FP32, one sensor plane).  Returns
 * LMC_ERR_INVALID when one photon's worst-case crossing records do not fit
 * the array beside the per-wave reserve, or when the run's precision has no
 * device tables yet (FP32 tables come with lmc_scene_create, FP64 tables with
 * a first lmc_scene_run); LMC_ERR_UNSUPPORTED for scenes with a
 * PhotonRecordSensor.  Calls are serialised within the process (they share
 * the array); the array keeps the last call's records, and each launch clears
 * the slots it can use first. */
int lmc_scene_run_noalloc(lmc_scene* s, const lmc_run_desc* run, lmc_stats* stats);

/* Add the device accumulators into host buffers (n doubles at `out`,
 * layout of lmc_sensor_desc) and the per-sensor hit counters into `hits`
 * (n_sensors entries); either may be NULL.
 * Replaces: SensorsGroup::merge_from (detector.cpp:283-297). */
int lmc_scene_read(lmc_scene* s, double* out, int64_t n, uint64_t* hits);
/* Zero the device accumulators and hit counters (and drop the records). */
int lmc_scene_zero(lmc_scene* s);
/* PhotonRecordSensor records gathered since creation or the last zero, in
 * crossing-queue order (not photon order): copies min(max_records, count)
 * records of LMC_RECORD_FIELDS doubles into `out` (may be NULL to count
 * only) and sets *n_records to the count.  Forward and gradient runs alike.
 * Replaces: PhotonRecordSensor::recorded_photons (detector.hpp:280-298) and
 * its merge_from (detector.cpp:455-461). */
int lmc_scene_records(lmc_scene* s, double* out, int64_t max_records, int64_t* n_records);

/* ---- batched persistent launch over many independent scenes ----------- */
/* No counterpart in the reference (which runs one SimConfig per call): up to
 * 4096 scenes, each validated as by lmc_scene_create, share one persistent
 * transport launch.  Scene i's photon j draws the same random stream
 * (production or xoshiro256++ check mode) as photon j of lmc_scene_run on
 * that scene alone with the same seed,
 * so a batched scene reproduces its single-scene run (up to the summation
 * order of the FP64 accumulators).  Absorption grids and PhotonRecordSensor
 * records are kept per scene; the CBS estimator mode runs single scenes
 * only (LMC_ERR_UNSUPPORTED). */
typedef struct lmc_batch lmc_batch;
int lmc_batch_create(int32_t n_scenes, const lmc_scene_desc* scenes, lmc_batch** out);
void lmc_batch_destroy(lmc_batch* b);
/* Run budgets[i] photons of scene i (photon indices photon_offset + [0,
 * budgets[i]), below 2^32; run->n_photons is ignored) and add into the
 * per-scene device accumulators.  use_graph != 0: the launch sequence is
 * captured once into a hipGraph and replayed while budgets and run repeat
 * (a batch with PhotonRecordSensor records runs eagerly: its record list is
 * sized and gathered on the host around each sensor launch). */
int lmc_batch_run(lmc_batch* b, const uint64_t* budgets, const lmc_run_desc* run,
                  int32_t use_graph, lmc_stats* stats);
/* Add scene `scene`'s accumulators / hit counters into host buffers, as
 * lmc_scene_read. */
int lmc_batch_read(lmc_batch* b, int32_t scene, double* out, int64_t n, uint64_t* hits);
/* Scene `scene`'s PhotonRecordSensor records, as lmc_scene_records. */
int lmc_batch_records(lmc_batch* b, int32_t scene, double* out, int64_t max_records, int64_t* n_records);
int lmc_batch_zero(lmc_batch* b);

/* ---- adjoint (path-length / event-count score) ------------------------ */
/* Forward transport that additionally accumulates, for every accumulator
 * double d of the scene and every parameter p, score_p * contribution_d,
 * with per-layer parameters ordered [mu_s(l), mu_a(l), g(l)] for l in layers
 * (3*n_layers parameters; g rows are zero for non-HG layers).
 * `grad` receives (added) 3*n_layers*n_out doubles, [p][d]. The ordinary
 * accumulators are filled as by lmc_scene_run.  The derivative assumes each
 * layer's mean free path tracks 1/mu_t (the descriptor's mfp must equal
 * 1/mu_t for every layer). */
int lmc_scene_run_adjoint(lmc_scene* s, const lmc_run_desc* run, double* grad,
                          int64_t n_grad, lmc_stats* stats);

/* ---- host-side table builders (src/mie/, src/table.cpp, src/phase.cpp) --- */
/* Mie amplitudes at n_mu cosines for the refractive index m_re + i*m_im,
 * with MIEV0's sign conventions (src/mie/dmiev.cpp:49-62,
 * src/mie/MIEV0noP.f:160-161,472-485).  s1/s2: 2*n_mu doubles (re, im). */
int lmc_mie_amplitudes(double x, double m_re, double m_im, int32_t n_mu,
                       const double* mu, double* s1, double* s2);
/* Q_ext, Q_sca, g (src/mie/dmiev.cpp:65-81). */
int lmc_mie_info(double x, double m_re, double m_im, double* qext, double* qsca,
                 double* g);
/* Normalised SamplingTable CDF from pdf values at nDiv+1 nodes
 * (src/table.cpp:6-48). */
int lmc_sampling_cdf(int32_t n_div, double min_val, double max_val,
                     const double* pdf, double* cdf);

/* ---- known-answer probes (tests only; no counterpart in the reference) -- */
/* Philox4x32-10 on device: out[4*i..] = philox(ctr[4*i..], key). */
int lmc_probe_philox(int32_t n, const uint32_t* ctr, uint32_t k0, uint32_t k1,
                     uint32_t* out);

/* Production random stream on device: the first n (<= 2^20) outputs of the
 * xoshiro128** generator of photon `photon` under key (k0, k1) = (low, high
 * word of the seed), seeded by Philox4x32-10 of (photon low, photon high,
 * 0xFFFFFFFE, 0). */
int lmc_probe_stream(uint32_t k0, uint32_t k1, uint64_t photon, int32_t n, uint32_t* out);

/* Stage C of the CBS reverse field on device (reverse_field_suffix,
 * detector.cpp:1123-1161) for given frames: frames[20i..] = m1, s1 (the first
 * scatter's frame, n1 = s1 x m1), m_n, n_n, s_n (the exit frame), the prefix
 * field (xm.re, xm.im, xn.re, xn.im) and the launch frame's handedness
 * (+1: m x n = s, -1: m x n = -s); the amplitudes are the constants S1 = 0.6 +
 * 0.2i, S2 = 0.3 - 0.5i; out[4i..] = Erm.re, Erm.im, Ern.re, Ern.im. */
int lmc_probe_reverse(int32_t n, int32_t precision, const double* frames, double* out);

/* FP32 path phase on device: out[2i], out[2i+1] = sin x[i], cos x[i] as the
 * FP32 CBS reverse-path factor exp(i k0 (s_out + s_in).(r_n - r_1)) forms
 * them (detector.cpp:1231-1310), for phases of any size. */
int lmc_probe_phase(int32_t n, const float* x, float* out);

/* Azimuth sampler of the production path on device: for each i, the (cos phi,
 * sin phi, F(phi)) that the transport draws from the two 32-bit words
 * words[2i] (component and radius) and words[2i+1] (angle, half period) for
 * F(phi) = a cos^2 + b sin^2 + c sin cos, (a, b, c) = abc[3i..3i+2] (the
 * density of src/phase.cpp:27-49, sampled exactly as a uniform/cos^2 mixture
 * about the principal axis instead of by rejection). */
int lmc_probe_azimuth(int32_t n, int32_t precision, const double* abc, const uint32_t* words,
                      double* out);

#ifdef __cplusplus
}
#endif

#endif /* LUMINIS_HIP_H */
