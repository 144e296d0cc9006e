/*
 * gemray_hip.h -- C ABI of the MI355X (gfx950) spectral transport library.
 *
 * Drop-in for gemray's GPU render backend (crates/gemray/src/renderer/gpu_backend.rs,
 * renderer/gpu/frame.rs): facet planes + material + camera + environment in, an XYZ film
 * out.  Plain C types only; every pointer argument is borrowed for the duration of the
 * call and nothing is retained after it returns unless the entry point says so.
 *
 * Determinism: every (pixel, absolute sample) result is a pure function of its indices,
 * and each pixel's samples are summed in sample order, so results are bit-identical for
 * any chunking, batching or co-scheduling (gpu_backend.rs:45-53).
 *
 * Errors: every call returns a gr_status.  Any status other than GR_OK leaves the
 * caller's film untouched (frame.rs:194-228).  GR_DECLINE_* are the reference's
 * "decline, use the CPU tracer" outcomes; GR_ERR_* are caller or device errors.
 */
#ifndef GEMRAY_HIP_H
#define GEMRAY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GR_OK = 0,
    GR_DECLINE_NO_DEVICE = 1, /* no usable gfx950 device (GpuFrameError::Acquire)      */
    GR_DECLINE_ENV = 2,       /* environment kind not handled (UnsupportedEnvironment) */
    GR_DECLINE_MATERIAL = 3,  /* material/finish branch not handled (UnsupportedMaterial) */
    GR_DECLINE_BUDGET = 4,    /* the job does not fit the HBM budget                   */
    GR_ERR_INVALID = 5,       /* argument outside the limits (validate/mod.rs:23-192)  */
    GR_ERR_DEVICE = 6,        /* a HIP call failed                                     */
    GR_ERR_NOT_FOUND = 7,     /* unknown job id                                        */
    GR_PENDING = 8,           /* gr_job_poll: the job has not finished                 */
    GR_CANCELLED = 9          /* gr_job_poll: the job was cancelled                    */
} gr_status;

/* Limits (apps/gemray-worker/src/validate/mod.rs:23-57, renderer/buffers.rs:148). */
#define GR_MAX_PIXELS (7680u * 4320u)
#define GR_MAX_SAMPLES_PER_CALL 65536u
#define GR_MAX_BOUNCES 64u
#define GR_MAX_PLANES 1024u
#define GR_MAX_BANDS 8

typedef struct gr_ctx gr_ctx;

typedef struct {
    int32_t device;            /* HIP device ordinal                                    */
    uint64_t chunk_tuples;     /* (pixel, sample) tuples per launch; 0 = library default */
    uint64_t hbm_budget_bytes; /* admission budget for served jobs; 0 = 3/4 of free HBM */
} gr_options;

/* GpuCameraParams / Camera (renderer/buffers.rs:344-367, optics/raytracer/camera.rs:53-59). */
typedef struct {
    float origin[3];
    float forward[3];
    float right[3];
    float up[3];
    float fov_tan;
} gr_camera;

enum { GR_DISP_SELLMEIER1 = 0, GR_DISP_SELLMEIER3 = 1, GR_DISP_CAUCHY = 2 };

/* The estimator-visible fields of GemMaterial (optics/materials.rs:32-120), replacing the
 * 432-byte GpuGemMaterial encoding (renderer/buffers.rs:237-316).  Bands are Gaussian
 * absorption bands (center_nm, width_nm, peak) per eigenmode axis. */
typedef struct {
    int32_t disp_kind;        /* GR_DISP_*                                              */
    float disp[6];            /* S1: b1,c1 | S3: b0,b1,b2,c0,c1,c2 | Cauchy: a,b,c       */
    float birefringence_delta;
    int32_t crystal_cubic;    /* 1 if CrystalSystem::Cubic                               */
    float c_axis[3];
    int32_t n_o_bands, n_e_bands;
    float o_bands[GR_MAX_BANDS][3];
    float e_bands[GR_MAX_BANDS][3];
    float absorption_path_scale;
    /* Biaxial indicatrix (materials.rs:1508-1519): n_beta = n(lambda), n_alpha = n_beta -
     * biaxial_delta_beta_alpha, n_gamma = n_alpha + birefringence_delta, gamma axis =
     * c_axis.  NaN = None (uniaxial or isotropic).  Traced when the stone is anisotropic;
     * a biaxial value on an isotropic stone is declined (GR_DECLINE_MATERIAL).        */
    float biaxial_delta_beta_alpha;
    /* Henyey-Greenstein volume scattering (scattering.rs:132-330) when > 0: extinction
     * coefficient sigma_s, in the units of the absorption bands.                      */
    float scattering_sigma_s;
    /* Traced: edge-rounded shading normals (intersect.rs:196-229) when > 0.          */
    float edge_rounding_radius;
    /* AbsorptionTensor::beta_ray (optics/absorption.rs): the third band set of a
     * trichroic biaxial stone, read only when biaxial; has_beta_bands = 0 is None.    */
    int32_t has_beta_bands, n_beta_bands;
    float beta_bands[GR_MAX_BANDS][3];
    /* Henyey-Greenstein asymmetry g of the scattering phase function, |g| < 1.        */
    float scattering_g;
} gr_material;

/* FacetFinish (optics/raytracer/camera.rs:43-51): a frosted facet scatters diffusely
 * (scattering.rs:522-590) instead of reflecting and refracting specularly. */
enum { GR_FINISH_POLISHED = 0, GR_FINISH_FROSTED = 1 };

/* An EnvironmentMap resident in HBM (gr_env_map_upload): the reference borrows one map for
 * every render that shows it (EnvironmentSource::HdrMap(&EnvironmentMap),
 * environment.rs:188-196), so a map shared by many scenes and jobs is uploaded once. */
typedef struct gr_env_map gr_env_map;

/* EnvironmentSource (optics/raytracer/environment.rs:188-196). */
enum { GR_ENV_STUDIO = 0, GR_ENV_MAP = 1 };
typedef struct {
    int32_t kind;
    /* Studio: LightingPreset::params() (temp_k, spot_mult) plus exposure/yaw/pitch. */
    float temp_k, spot_mult, exposure, light_yaw, light_pitch;
    /* Map: an EnvironmentMap, map_w x map_h RGB f32 texels, row-major (a 1x1 map is the
     * uniform furnace environment), copied in by every call that uploads the scene ... */
    int32_t map_w, map_h;
    const float *map_rgb;
    /* ... or, when map is not NULL, that device-resident map (map_w, map_h and map_rgb are
     * then ignored): scenes and jobs hold a reference to it until they are freed. */
    const gr_env_map *map;
} gr_env;

/* GpuSceneRef (renderer/gpu_backend.rs:85-99). */
typedef struct {
    const gr_camera *camera;
    uint32_t width, height;
    const float *planes;       /* n_planes x {normal.x, normal.y, normal.z, d}           */
    uint32_t n_planes;
    const uint32_t *finishes;  /* NULL = all polished; else n_planes GR_FINISH_* values  */
    const gr_material *material;
    uint32_t max_bounces;
    const gr_env *env;
} gr_scene;

/* Per-class event counters of the last render (device-side, exact).  The first twelve are
 * the reference tracer's event classes; the last four count work the kernels skip or do once
 * per path: camera rays proved to miss by their pixel's certificate (two planes instead of
 * the full clip), exterior misses (one plane), paths whose first segment misses the stone
 * (no dispersion evaluation), and absorption-cache fills (the band sums, once per path). */
enum {
    GR_EV_SEGMENT = 0, GR_EV_MISS, GR_EV_HIT, GR_EV_ROTATE, GR_EV_ABSORB, GR_EV_TIR,
    GR_EV_REFLECT, GR_EV_REFRACT, GR_EV_ENTRY_ANISO, GR_EV_WALKOFF, GR_EV_RR, GR_EV_SCATTER,
    GR_EV_CERT_MISS, GR_EV_EXTERIOR_MISS, GR_EV_FIRST_MISS, GR_EV_ABSORPTION_FILL, GR_EV_COUNT
};
typedef struct {
    uint64_t events[GR_EV_COUNT]; /* events[GR_EV_SEGMENT] = 8-lambda path segments     */
    uint64_t paths;
    uint32_t launches;
    double trace_kernel_ms;       /* sum of trace-kernel durations (HIP events)          */
    double total_ms;              /* wall time of the call on its stream                 */
    uint32_t n_planes;
    double finish_kernel_ms;      /* wavefront mode: sum of finish-kernel durations      */
    uint32_t hero_tables;         /* tables the trace read: 1 spectral, 2 material         */
    uint32_t reserved;
} gr_stats;

/* Trace kernels.  Both evaluate identical per-tuple results; they differ in scheduling.
 * GR_TRACE_PLAIN (default): one lane per (pixel, sample) tuple, as transport_main
 * dispatches (frame.rs:437-452).  GR_TRACE_WAVEFRONT: persistent wavefronts that refill
 * finished lanes with the next tuples of their range, plus a separate finish pass. */
enum { GR_TRACE_PLAIN = 0, GR_TRACE_WAVEFRONT = 1 };

/* GpuBackend::acquire / GpuFrameRenderer::new (gpu_backend.rs:114-126, frame.rs:574-594).
 * Returns GR_DECLINE_NO_DEVICE (and *out = NULL) when no gfx950 device is usable. */
gr_status gr_open(gr_ctx **out, const gr_options *opts);
/* Drop of GpuBackend. */
void gr_close(gr_ctx *ctx);
/* GpuBackend::adapter_label (gpu_backend.rs:146-153).  Writes a NUL-terminated label;
 * returns the label length. */
int gr_adapter_label(gr_ctx *ctx, char *buf, size_t len);
/* GpuFrameRenderer::set_chunk_budget_bytes (frame.rs:602), in tuples. */
gr_status gr_set_chunk_tuples(gr_ctx *ctx, uint64_t tuples);
/* Selects the trace kernel (GR_TRACE_PLAIN / GR_TRACE_WAVEFRONT).  No reference
 * counterpart: the reference has one dispatch shape (frame.rs:437-452). */
gr_status gr_set_trace_mode(gr_ctx *ctx, int mode);

/* Plain trace mode: paths still running after split_bounce segments are parked and
 * continued by further, fully populated launches (default 6; 0 = never).  Results are
 * bit-identical for every value.  No reference counterpart. */
gr_status gr_set_split_bounce(gr_ctx *ctx, uint32_t split_bounce);

/* Hero tables: per hero wavelength, the comb's colour-matching values and map-spectrum
 * lobes (one table per context, built at gr_open) and an absorbing material's dispersion
 * indices and absorption spectra (one per scene), evaluated by the kernels' own functions,
 * so results are bit-identical in every mode.  mode 0: every path evaluates them; 1
 * (default): traces read the tables, and a scene's material table is built by its first
 * trace of at least 2^26 tuples (a smaller one would not repay it); 2: as 1, with the
 * material table built by any trace of an absorbing scene.  No reference counterpart (the
 * reference evaluates them per path, spectral_transport.wgsl:672-1470). */
gr_status gr_set_hero_tables(gr_ctx *ctx, int mode);

/* GpuBackend::try_accumulate / GpuFrameRenderer::accumulate (gpu_backend.rs:160-193,
 * frame.rs:632-793): ADDS each pixel's sum over samples [sample_offset,
 * sample_offset + spp) into accum_xyz[3 * pixel + c].  accum_pixels must equal
 * width * height (the reference asserts this, frame.rs:673-677).  Thread-safe: calls on
 * one context are serialized. */
gr_status gr_accumulate(gr_ctx *ctx, const gr_scene *scene, uint32_t sample_offset,
                        uint32_t spp, float *accum_xyz, size_t accum_pixels);

/* The per-(pixel, sample) estimator for a pixel range, for the parity tiers: the dispatch
 * of transport_main with per-tuple outputs (spectral_transport.wgsl:672-1470,
 * frame.rs:437-452).  xyz_out receives pixel_count * spp * 3 floats, tuple (p, s) at
 * (p - pixel_begin) * spp + s.  info_out, when non-NULL, receives 4 u32 per tuple:
 * segments, termination (0 escaped, 2 Russian roulette, 3 bounce cap), an FNV-1a hash of
 * the per-bounce (facet, event, eigenmode) sequence, and an FNV-1a hash of the per-bounce
 * channel fates (channels past their critical angle, channels with zero path pdf). */
gr_status gr_trace_samples(gr_ctx *ctx, const gr_scene *scene, uint32_t pixel_begin,
                           uint32_t pixel_count, uint32_t sample_offset, uint32_t spp,
                           float *xyz_out, uint32_t *info_out);

/* Camera::new (optics/raytracer/camera.rs:63-93): orbit camera at (yaw, pitch), distance
 * from the origin, vertical field of view in degrees. */
gr_status gr_camera_new(float yaw, float pitch, float distance, float fov_deg, gr_camera *out);

/* Primary-hit guide buffers for the denoiser (generate_guide_buffers,
 * apps/diagram-gui/src/bridge/guide_pass.rs:37-110): for each pixel (x, y) of a width x height
 * frame, the camera ray Camera::generate_ray(x, y, width, height, 0, 0) against the planes
 * (intersect_polyhedron, intersect.rs:9-78): depth[p] = hit distance, normal[3p..3p+2] = the
 * hit facet's plane normal, facet[p] = its index; a miss gives 1e6, (0, 0, 0) and -1.  All
 * three outputs are caller-owned host arrays of width * height entries; a zero-sized frame
 * writes nothing. */
gr_status gr_guide_buffers(gr_ctx *ctx, const gr_camera *camera, uint32_t width, uint32_t height, const float *planes,
                           uint32_t n_planes, float *depth, float *normal, int32_t *facet);

/* Per-function probe (no reference counterpart; for the per-function ULP tests): evaluates
 * one device function of the transport kernel on n items, 9 input floats and 3 output
 * floats per item, and copies the outputs to out (3 * n floats).
 *   GR_PROBE_NEFF             in (n_o, n_e, cos theta)          out (n_eff)
 *   GR_PROBE_POYNTING         in (w.xyz, c.xyz, n_o, n_e)        out (S.xyz)
 *   GR_PROBE_TIR_RETARDATION  in (n1, cos_i, sin_i)              out (cos delta, sin delta)
 *   GR_PROBE_FRAME_ROTATION   in (prev.xyz, cur.xyz, axis.xyz)   out (cos 2psi, sin 2psi)
 *   GR_PROBE_FIELD_QFORM      in (Q, U, S.xyz, P.xyz, ao)        out (e.A.e with ae = 1 - ao)
 *   GR_PROBE_EXP              in (x)                             out (e^x as the lobes evaluate it)
 *   GR_PROBE_CMF              in (lambda_nm)                     out (xbar, ybar, zbar)
 *   GR_PROBE_ACOS             in (x)                             out (acos x as the metrics evaluate it)
 * The biaxial indicatrix (optics/birefringence.rs:435-836), in (n_alpha, n_beta, n_gamma,
 * w.xyz, gamma_axis.xyz) -- the principal frame is BiaxialIndicatrix::from_gamma_axis's:
 *   GR_PROBE_BX_WAVE_INDICES  out (n_slow, n_fast)             wave_indices(w)
 *   GR_PROBE_BX_EIGEN_SLOW/_FAST      out (D.xyz)              eigen_polarizations(w).0 / .1
 *   GR_PROBE_BX_POYNTING_SLOW/_FAST   out (S.xyz)              mode_poynting_dir(w, slow / fast)
 * The integer sampling streams (optics/raytracer/sampling.rs:12-271), u32 values passed and
 * returned as their bit patterns in the float arrays:
 *   GR_PROBE_SAMPLE_DRAWS     in (pixel, sample_num)             out (jx, jy, hero) of sample_draws,
 *                             the radical inverses (bases 3, 5) taken from the host's sample table
 *   GR_PROBE_SEED             in (pixel, sample_num, x)          out (path seed, hash_u32(x),
 *                                                                 x with its bits reversed)
 *   GR_PROBE_STREAM_DRAW      in (seed, bounce, salt0..2)        out (the three per-bounce draws)
 *   GR_PROBE_RGB_SPECTRAL     in (r, g, b, lambda_nm)            out (rgb_to_spectral_radiance,
 *                                                                 renderer/env_map_spectrum.rs:60-82)
 *   GR_PROBE_DOP              in (I, Q, U, V)                    out (degree_of_polarization as the
 *                                                                 absorption step evaluates it,
 *                                                                 polarization.rs:38-48)
 *   GR_PROBE_FRESNEL_COMPANION in (n1, n2, cos_i, cos_t)         out (r_s, r_p, the transmission
 *                                                                 Mueller element m00) as a
 *                                                                 companion channel evaluates them
 *                                                                 (refraction.rs:1049-1056,
 *                                                                 polarization.rs:119-137) */
enum { GR_PROBE_NEFF = 0, GR_PROBE_POYNTING = 1, GR_PROBE_TIR_RETARDATION = 2, GR_PROBE_FRAME_ROTATION = 3,
       GR_PROBE_FIELD_QFORM = 4, GR_PROBE_EXP = 5, GR_PROBE_CMF = 6, GR_PROBE_ACOS = 7,
       GR_PROBE_BX_WAVE_INDICES = 8, GR_PROBE_BX_EIGEN_SLOW = 9, GR_PROBE_BX_EIGEN_FAST = 10,
       GR_PROBE_BX_POYNTING_SLOW = 11, GR_PROBE_BX_POYNTING_FAST = 12, GR_PROBE_SAMPLE_DRAWS = 13,
       GR_PROBE_SEED = 14, GR_PROBE_STREAM_DRAW = 15, GR_PROBE_RGB_SPECTRAL = 16, GR_PROBE_DOP = 17,
       GR_PROBE_FRESNEL_COMPANION = 18, GR_PROBE_COUNT = 19 };
gr_status gr_probe(gr_ctx *ctx, int fn, const float *in, size_t n, float *out);

/* Uploads a w x h RGB f32 lat-long map (EnvironmentMap::new, renderer/env_map.rs) to HBM;
 * *out holds one reference, which gr_env_map_release drops.  The device copy is freed when
 * the last reference (the caller's, a device scene's or a job's) is dropped. */
gr_status gr_env_map_upload(gr_ctx *ctx, const float *rgb, int32_t w, int32_t h, gr_env_map **out);
void gr_env_map_release(gr_ctx *ctx, gr_env_map *map);

/* Statistics of the last gr_accumulate / gr_trace_samples / gr_render call. */
gr_status gr_last_stats(gr_ctx *ctx, gr_stats *out);

/* ---- Device-resident scenes and films (the served / batched path) ------------------
 * No reference counterpart as separate calls: GpuFrameRenderer re-uploads every input
 * per chunk (frame.rs:375-426) and reads the film back per chunk (:799-814).  These keep
 * scene inputs and the film in HBM so a caller can accumulate progressively
 * (sample_offset resumes, gpu_backend.rs:45-53) and read the film once. */
typedef struct gr_scene_dev gr_scene_dev;
typedef struct gr_film gr_film;

gr_status gr_scene_upload(gr_ctx *ctx, const gr_scene *scene, gr_scene_dev **out);
void gr_scene_free(gr_ctx *ctx, gr_scene_dev *scene);
gr_status gr_film_create(gr_ctx *ctx, uint32_t width, uint32_t height, gr_film **out);
gr_status gr_film_clear(gr_ctx *ctx, gr_film *film);
void gr_film_free(gr_ctx *ctx, gr_film *film);
/* Copies the film (width * height * 3 floats) to host memory. */
gr_status gr_film_download(gr_ctx *ctx, gr_film *film, float *xyz_out, size_t n_floats);
/* Adds samples [sample_offset, sample_offset + spp) of every pixel into film. */
gr_status gr_render(gr_ctx *ctx, const gr_scene_dev *scene, uint32_t sample_offset,
                    uint32_t spp, gr_film *film);
/* gr_render without any allocation: no counterpart in the reference.  For callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The per-tuple staging, continuation records, sample
 * table and counters live in one fixed 56 MB device array per GPU (its address taken once
 * per device at gr_open), so every call traces in launches of at most 2^19 tuples and
 * calls no hipMalloc, hipFree or allocating library routine; spp above
 * GR_MAX_SAMPLES_PER_CALL is rejected (GR_ERR_INVALID).  Results are bit-identical to
 * gr_render's.  Calls on one context are serialized; calls that overlap on one GPU from
 * different contexts share the array and race. */
gr_status gr_render_noalloc(gr_ctx *ctx, const gr_scene_dev *scene, uint32_t sample_offset,
                            uint32_t spp, gr_film *film);

/* ---- Autotuning (no reference counterpart: the reference fixes its workgroup size,
 * spectral_transport.wgsl:672, and its chunk budget, frame.rs:192) -----------------------
 * A configuration of the plain trace path.  Every configuration gives bit-identical films;
 * they differ only in speed. */
typedef struct {
    uint32_t waves_per_eu;  /* occupancy variant of the trace kernels: 2, 3 or 4 waves per SIMD */
    uint32_t split_bounce;  /* paths running after this many segments are parked (0: never)    */
    uint32_t cont_div;      /* continuation records for 1 / cont_div of a launch's tuples     */
    uint32_t trace_mode;    /* GR_TRACE_PLAIN or GR_TRACE_WAVEFRONT                            */
    uint64_t chunk_tuples;  /* tuples per launch                                               */
    double probe_ms;        /* gr_autotune: the chosen configuration's probe time              */
    uint32_t probes;        /* gr_autotune: configurations timed                               */
    uint32_t park_lanes;    /* from 3 segments on, a wave with fewer live lanes than this parks
                               them early (0: only at split_bounce); at most 64               */
} gr_tune_choice;
/* Times probe renders of the scene (its middle rows at spp samples, about 2^25 tuples) over
 * the configurations by coordinate descent (occupancy, split bounce, early parking,
 * continuation capacity, launch size, trace mode; a change must win by more than 1 %) and caches the fastest under
 * the scene's class (material and environment kernels), plane-count bucket and launch-size
 * bucket; writes it to *out. */
gr_status gr_autotune(gr_ctx *ctx, const gr_scene_dev *scene, uint32_t spp, gr_tune_choice *out);
/* With autotuning on, gr_render (and so gr_accumulate) traces with the configuration cached
 * for the scene's class and size, tuning it on first use; off (the default), with the
 * context's settings (gr_set_chunk_tuples, gr_set_trace_mode, gr_set_split_bounce). */
gr_status gr_set_autotune(gr_ctx *ctx, int enabled);
/* gr_render with a given configuration (GR_ERR_INVALID outside the ranges above). */
gr_status gr_render_tuned(gr_ctx *ctx, const gr_scene_dev *scene, uint32_t sample_offset, uint32_t spp,
                          gr_film *film, const gr_tune_choice *choice);


/* ---- Gemological metrics ------------------------------------------------------------
 * evaluate_gem_optical_metrics (crates/gemray/src/color/metrics.rs:1082-1248): brilliance,
 * fire, scintillation, windowing and extinction of a cut stone seen from a camera pose
 * under the studio rig's key/fill/ring lights.  Only the material's dispersion is read
 * (the d, F and C line indices).  One call evaluates n_poses poses of one stone on the
 * device; every result is bit-identical to the reference's for the same inputs.  The
 * reference's angular profiles (evaluate_angular_profile_at_azimuth, metrics.rs:1262-1290:
 * cam_pitch = 0, 5, ..., 90 degrees at one azimuth) are one call of 19 poses.
 * n_planes = 0 gives the reference's placeholder values (85, 25, 75, 5, 5). */
typedef struct {
    float cam_yaw, cam_pitch;      /* radians                                             */
    float light_yaw, light_pitch;  /* radians (StudioRig::new)                            */
} gr_metrics_pose;

typedef struct {
    float brilliance_pct, fire_index, scintillation_pct, windowing_pct, extinction_pct;
} gr_gem_metrics;

gr_status gr_optical_metrics(gr_ctx *ctx, const float *planes, uint32_t n_planes, const gr_material *material,
                             const gr_metrics_pose *poses, uint32_t n_poses, gr_gem_metrics *out);
/* gr_optical_metrics for many stones in one launch (the batched jobs' metrics, reduced on the
 * device together): every pose of every stone is evaluated by one items launch and one
 * reduction; out receives the stones' results in order, n_poses of each. */
typedef struct {
    const float *planes;
    uint32_t n_planes;
    const gr_material *material;
    const gr_metrics_pose *poses;
    uint32_t n_poses;
} gr_metrics_stone;
gr_status gr_optical_metrics_batch(gr_ctx *ctx, const gr_metrics_stone *stones, uint32_t n_stones,
                                   gr_gem_metrics *out);

/* ---- Batched and served rendering: the device-resident job loop ---------------------
 * The reference serves one RenderRequest{request_id, scene, first_sample, samples} at a
 * time through GpuFrameRenderer::accumulate (apps/gemray-worker render path;
 * messages/render.rs:42-70, frame.rs:632-793).  Here requests become jobs: a job is
 * admitted when its HBM footprint (film, planes, environment map, sample table) fits the
 * context's budget (gr_options.hbm_budget_bytes; 0 = 3/4 of the HBM free when the loop
 * starts), its inputs then stay in HBM, and a scheduler thread traces all admitted jobs
 * together: every tick gives each job a piece of its frame, traces the pieces of each
 * material class in one launch, and adds them into the jobs' device films.  Jobs
 * submitted while others run join at the next tick.  Submission never blocks on the GPU.
 *
 * Determinism: a job's film is bit-identical to gr_accumulate of the same scene and sample
 * range into a zeroed film, alone or in any mix of jobs. */
enum { GR_JOB_QUEUED = 0, GR_JOB_RUNNING = 1, GR_JOB_DONE = 2, GR_JOB_CANCELLED = 3, GR_JOB_FAILED = 4 };

typedef struct {
    const gr_scene *scene;  /* copied during gr_job_submit                               */
    uint32_t sample_offset; /* first absolute sample (RenderRequest.first_sample)        */
    uint32_t spp;           /* samples per pixel (RenderRequest.samples)                 */
    /* Progressive delivery (the worker's tracer, stream_emit/tracer.rs:64-117): with
     * subbatch_ms > 0 the job traces its range in contiguous sample sub-batches, each
     * over the whole frame, sized toward subbatch_ms of tracing each (next_batch_size,
     * stream_emit/sizing.rs:12-30: at most x4 growth per step, never 0), starting from
     * subbatch_samples (0 = 1) -- or always subbatch_samples when subbatch_fixed is set.
     * Every finished sub-batch is folded into a pending delta (gr_job_take_delta) and
     * into the running total that gr_job_film returns.  subbatch_ms = 0: one batch.    */
    uint32_t subbatch_ms;
    uint32_t subbatch_samples;
    uint32_t subbatch_fixed;
} gr_job_desc;

typedef struct {
    uint32_t state;                /* GR_JOB_*                                            */
    uint32_t width, height, sample_offset, spp;
    uint32_t samples_done;         /* samples of every pixel folded into the film so far  */
    uint32_t subbatches;           /* sub-batches finished (1 for a one-batch job)       */
    uint64_t footprint_bytes;      /* HBM held from admission to completion               */
    uint64_t paths;                /* paths traced so far                                 */
    uint64_t events[GR_EV_COUNT];  /* per-class counters (when done)                      */
    double queue_ms;               /* submit -> admission                                 */
    double run_ms;                 /* admission -> completion                             */
    double latency_ms;             /* submit -> completion                                */
} gr_job_stats;

typedef struct {
    uint64_t submitted, admitted, completed, cancelled, failed, declined;
    uint64_t queued, running;      /* now                                                 */
    uint64_t paths, segments;      /* of completed jobs                                   */
    uint64_t ticks, launches;
    double busy_ms;                /* summed tick durations (HIP events; two overlap)     */
    double wall_ms;                /* first admission -> last completion                  */
    double latency_p50_ms, latency_p99_ms, latency_mean_ms, latency_max_ms;
    uint64_t hbm_in_use, hbm_budget, tick_tuples;
    uint32_t policy, reserved;     /* GR_SERVE_*                                           */
    double slack;
    double rate_tuples_per_ms;     /* service-rate estimate behind the deadlines          */
} gr_serve_metrics;

/* Queues a job; *job_id identifies it until gr_job_release.  GR_DECLINE_* / GR_ERR_INVALID
 * as gr_accumulate; GR_DECLINE_BUDGET when the job alone exceeds the HBM budget. */
gr_status gr_job_submit(gr_ctx *ctx, const gr_job_desc *desc, uint64_t *job_id);
/* GR_OK (done), GR_PENDING (queued or running), GR_CANCELLED, GR_ERR_NOT_FOUND, or the
 * error that failed the job. */
gr_status gr_job_poll(gr_ctx *ctx, uint64_t job_id);
/* Blocks until the job settles or timeout_ms passes (< 0: no limit); returns as poll. */
gr_status gr_job_wait(gr_ctx *ctx, uint64_t job_id, int64_t timeout_ms);
/* Requests cancellation (the reference's cancel epochs, tracer.rs:64-117); the job stops
 * at the next tick boundary and frees its HBM. */
gr_status gr_job_cancel(gr_ctx *ctx, uint64_t job_id);
/* ADDS a finished job's film into accum_xyz (accum_pixels = width * height), like
 * gr_accumulate.  GR_PENDING if the job has not finished. */
gr_status gr_job_film(gr_ctx *ctx, uint64_t job_id, float *accum_xyz, size_t accum_pixels);
gr_status gr_job_info(gr_ctx *ctx, uint64_t job_id, gr_job_stats *out);
/* PendingDelta::take (stream_emit/emitter.rs:46-95): the sum, per pixel, of the samples
 * of every sub-batch finished since the last take, as one contiguous range
 * [*first_sample, *first_sample + *samples); written (not added) into delta_xyz
 * (n_pixels = width * height).  *samples = 0 and delta_xyz untouched when nothing is
 * pending.  Works in any state, so sub-batches finished before a cancel can be drained. */
gr_status gr_job_take_delta(gr_ctx *ctx, uint64_t job_id, uint32_t *first_sample, uint32_t *samples,
                            float *delta_xyz, size_t n_pixels);
/* Forgets a settled job (GR_PENDING if it has not settled). */
gr_status gr_job_release(gr_ctx *ctx, uint64_t job_id);
/* Loop metrics since the last reset: throughput, p50/p99 submit-to-completion latency. */
gr_status gr_metrics(gr_ctx *ctx, gr_serve_metrics *out);
gr_status gr_metrics_reset(gr_ctx *ctx);
/* Tuples traced per tick (default 2^24; at least GR_MAX_SAMPLES_PER_CALL). */
gr_status gr_serve_set_tick_tuples(gr_ctx *ctx, uint64_t tuples);
/* How a tick's tuples are shared among running jobs (no counterpart in the reference, which
 * serves one request at a time):
 *   GR_SERVE_DEADLINE (default, slack 0.25): in order of deadline = submit time + slack x the
 *     job's estimated service time (its tuples over the measured service rate).  Slack 0 is
 *     first-come-first-served (the best worst-case latency); larger slacks let short jobs
 *     overtake long ones (a lower median).  Deadlines are fixed at submission.
 *   GR_SERVE_EQUAL_SHARE: every running job gets an equal share of every tick (slack unused).
 * Films do not depend on the policy. */
enum { GR_SERVE_EQUAL_SHARE = 0, GR_SERVE_DEADLINE = 1 };
gr_status gr_serve_set_policy(gr_ctx *ctx, int policy, double slack);

#ifdef __cplusplus
}
#endif
#endif
