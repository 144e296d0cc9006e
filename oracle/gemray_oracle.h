/*
 * gemray_oracle.h -- C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a CPU restatement of gemray's spectral
 * tracer (crates/gemray/src/optics/raytracer/transport.rs:420-940 and the physics it
 * calls).  It exists to check the HIP product path and to time a CPU baseline.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library (libgemray_hip.so) never links or calls it.
 *
 * Every struct here is plain data; every pointer is borrowed for the duration of the
 * call.
 */
#ifndef GEMRAY_ORACLE_H
#define GEMRAY_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_BANDS 8

/* DispersionModel (crates/gemray/src/optics/dispersion.rs:3-7). */
enum { OR_DISP_SELLMEIER1 = 0, OR_DISP_SELLMEIER3 = 1, OR_DISP_CAUCHY = 2 };

/* The subset of GemMaterial (optics/materials.rs:32-120) the non-scattering estimator
 * reads (laid out as gr_material of include/gemray_hip.h). */
typedef struct {
    int32_t disp_kind;
    float disp[6];          /* S1: b1,c1 | S3: b0,b1,b2,c0,c1,c2 | Cauchy: a,b,c */
    float birefringence_delta;
    int32_t crystal_cubic;  /* CrystalSystem::Cubic */
    float c_axis[3];
    int32_t n_o_bands, n_e_bands;
    float o_bands[OR_MAX_BANDS][3]; /* center_nm, width_nm, peak */
    float e_bands[OR_MAX_BANDS][3];
    float absorption_path_scale;
    float biaxial_delta_beta_alpha;  /* NaN = None; else a biaxial indicatrix            */
    float scattering_sigma_s;        /* > 0: Henyey-Greenstein volume scattering         */
    float edge_rounding_radius;      /* > 0: edge-rounded shading normals                */
    int32_t has_beta_bands;          /* AbsorptionTensor::beta_ray is Some               */
    int32_t n_beta_bands;
    float beta_bands[OR_MAX_BANDS][3];
    float scattering_g;              /* Henyey-Greenstein asymmetry                     */
} or_material;

/* EnvironmentSource (optics/raytracer/environment.rs:188-196).  kind 0 = Studio,
 * kind 1 = HdrMap (an EnvironmentMap of map_w x map_h RGB texels, row-major). */
typedef struct {
    int32_t kind;
    float temp_k, spot_mult, exposure, light_yaw, light_pitch;
    int32_t map_w, map_h;
    const float *map_rgb;
} or_env;

/* Camera (optics/raytracer/camera.rs:53-59). */
typedef struct {
    float origin[3], forward[3], right[3], up[3], fov_tan;
} or_camera;

/* Path termination, transport.rs:373-382. */
enum { OR_TERM_ESCAPED = 0, OR_TERM_SCATTER = 1, OR_TERM_RR = 2, OR_TERM_CAP = 3 };

/* Per-class event counters used by the segment op-count model. */
enum {
    OR_EV_SEGMENT = 0,   /* bounce-loop iterations entered                 */
    OR_EV_MISS,          /* escapes (environment sampled)                   */
    OR_EV_HIT,           /* facet hits                                      */
    OR_EV_ROTATE,        /* Stokes frame rotations applied                  */
    OR_EV_ABSORB,        /* interior segments (absorption evaluated)        */
    OR_EV_TIR,           /* hero TIR events                                 */
    OR_EV_REFLECT,       /* hero partial reflections                        */
    OR_EV_REFRACT,       /* hero refractions                                */
    OR_EV_ENTRY_ANISO,   /* air->crystal entries of an anisotropic stone    */
    OR_EV_WALKOFF,       /* extraordinary Poynting evaluations (hero)       */
    OR_EV_RR,            /* Russian-roulette evaluations                    */
    OR_EV_SCATTER,       /* free-path draws of a scattering stone           */
    OR_EV_COUNT
};

typedef struct {
    uint32_t segments;     /* bounce+1 on escape / RR, max_bounces on HitCap  */
    uint32_t termination;  /* OR_TERM_*                                        */
    uint32_t facet_hash;   /* FNV-1a over (facet index, event) per bounce      */
    uint32_t fate_hash;    /* FNV-1a over per-bounce channel-fate masks        */
} or_path_info;

uint32_t or_hash_u32(uint32_t x);
float or_stream_draw(uint32_t seed, uint32_t bounce, uint32_t stream);
float or_low_discrepancy_base2(uint32_t n);
float or_cranley_patterson_rotate(float x, float offset);
float or_rgb_to_spectral_radiance(const float rgb[3], float lambda_nm);
void or_rng_records(size_t n, const uint32_t *pixel, const uint32_t *sample, const uint32_t *bounce,
                    const uint32_t *salts, int n_salts, uint32_t *out);
void or_pixel_rotations(uint32_t pixel, float out[3]);
void or_sample_draws(uint32_t pixel, uint32_t sample_num, uint32_t *seed, float out[3]);
float or_radical_inverse(uint32_t n, uint32_t base);
void or_wrapped_hero_wavelengths(float hero_rand, float out[8]);
void or_exp_f32x8(const float in[8], float out[8]);
void or_cie_1931_cmf(float lambda_nm, float out[3]);
void or_cie_1931_cmf_x8(const float lambdas[8], float out[24]);
float or_dispersion_evaluate(const or_material *m, float lambda_nm);
float or_blackbody(float lambda_nm, float temp_k);
void or_white_balance(float temp_k, float out[3]);
float or_sample_studio_environment(const float dir[3], float lambda_nm, float temp_k,
                                   float spot_mult, float exposure, float light_yaw,
                                   float light_pitch);
float or_env_map_radiance(const or_env *env, const float dir[3], float lambda_nm);
void or_analytic_furnace_target(float l0, float out[3]);
int or_standard_round_brilliant(float *planes_out /* 4 * 74 */);
void or_plane_new(const float n[3], float d, float out[4]);
void or_camera_new(float yaw, float pitch, float distance, float fov_deg, or_camera *out);
void or_generate_ray(const or_camera *cam, float sx, float sy, float w, float h, float jx,
                     float jy, float origin[3], float dir[3]);
/* Unit-level probes used by the tests. */
void or_shading_normal_near_edge(const float *planes, int n_planes, const float hit_point[3], int hit_facet,
                                 const float hit_normal[3], float rounding_radius, float out[3]);
int or_intersect(const float *planes, int n_planes, const float origin[3], const float dir[3],
                 float *t, float normal[3], int *facet);
int or_intersect_scalar(const float *planes, int n_planes, const float origin[3], const float dir[3],
                 float *t, float normal[3], int *facet);
float or_effective_extraordinary_index(float n_o, float n_e, float theta);
float or_walk_off_angle(float n_o, float n_e, float theta);
float or_neff_from_cos(float n_o, float n_e, float c);
void or_frame_rotation_cs(const float prev[3], const float cur[3], const float axis[3], float out[2]);
float or_field_qform(float q, float u, const float s_l[3], const float p_l[3], float ao, float ae);
float or_degree_of_polarization(float i, float q, float u, float v);
void or_fresnel_companion(float n1, float n2, float cos_i, float cos_t, float out[3]);
void or_extraordinary_poynting_dir(const float w[3], const float c[3], float n_o, float n_e,
                                   float out[3]);
float or_tir_phase_delta(float n1, float cos_i, float sin_i);
float or_signed_frame_rotation_psi(const float prev[3], const float cur[3], const float axis[3]);
/* BiaxialIndicatrix probes (optics/birefringence.rs:349-870).  An indicatrix is 12 floats:
 * n_alpha, n_beta, n_gamma, then the alpha, beta and gamma principal axes. */
int or_material_biaxial_indicatrix(const or_material *m, float lambda_nm, float out[12]);
void or_biaxial_from_gamma_axis(float n_alpha, float n_beta, float n_gamma, const float g[3], float out[12]);
void or_biaxial_wave_indices(const float ind[12], const float w[3], float out[2] /* slow, fast */);
void or_biaxial_eigen_polarizations(const float ind[12], const float w[3], float out[6] /* d_slow, d_fast */);
void or_biaxial_d_to_e(const float ind[12], const float d[3], float out[3]);
void or_biaxial_mode_poynting_dir(const float ind[12], const float w[3], int want_slow, float out[3]);
float or_biaxial_resolve_entry_mode(const float ind[12], const float incident[3], const float normal[3],
                                    float cos_i, float n_seed, int want_slow, float wave[3]);
void or_poynting_direction(const float k[3], const float e[3], float out[3]);
void or_uniaxial_eigen_polarizations(const float w[3], const float c[3], float out[6] /* o, e */);
float or_effective_pleochroic_alpha(const float alpha[3], int trichroic, const float axis[3], const float e[3],
                                    const float ea[3], const float eb[3], float dop);
/* generate_guide_buffers (apps/diagram-gui/src/bridge/guide_pass.rs:37-110): the denoiser's
 * primary-hit depth, normal (3 floats) and facet (-1 on a miss) per pixel. */
void or_guide_buffers(const or_camera *cam, uint32_t width, uint32_t height, const float *planes, int n_planes,
                      float *depth, float *normal, int32_t *facet);
/* Henyey-Greenstein scattering probes (optics/raytracer/scattering.rs:132-232, 428-447;
 * raytracer/absorption.rs:174-214).  Stokes vectors travel as 8 x (I, Q, U, V). */
void or_sample_hg_batch(int n, const float *u1, const float *u2, float g, const float forward[3], float *out);
void or_channel_absorption_alphas(const or_material *m, const float lambdas[8], const float s_axis[3],
                                  const float k_hat[3], const float stokes[32], float out[8]);
int or_maybe_scatter_or_extinguish(const float alphas[8], float sigma_s, float g, const float dir[3], float hit_t,
                                   float path_scale, uint32_t seed, uint32_t bounce, float stokes[32], float pdf[8],
                                   float *t_model, float new_dir[3]);

/* trace_spectral_ray_with_finish_instrumented (transport.rs:476-537) with
 * enable_internal_mode_coupling = true.  finishes: per-plane FacetFinish (0 polished,
 * 1 frosted), or NULL for all polished.  events may be NULL; when given it receives
 * OR_EV_COUNT counters (added into). */
void or_trace_spectral_ray(const float origin[3], const float dir[3], const float *planes,
                           int n_planes, const uint32_t *finishes, const or_material *m, uint32_t max_bounces,
                           const or_env *env, uint32_t rng_seed, float hero_rand,
                           float out_xyz[3], or_path_info *info, uint64_t *events);

/* render_core::trace_into (apps/gemray-worker/src/render_core/mod.rs:220-317): each pixel
 * sums its samples [first_sample, first_sample+spp) in order and ADDS the sum into
 * accum[pixel].  pixel_begin/pixel_end restrict the rows of work to a pixel range (for a
 * bounded CPU-baseline sample).  per_sample_xyz/per_sample_info, when non-NULL, receive
 * one entry per (pixel - pixel_begin) * spp + s.  Returns total segments traced. */
uint64_t or_render(const or_camera *cam, uint32_t width, uint32_t height, const float *planes,
                   int n_planes, const uint32_t *finishes, const or_material *m, uint32_t max_bounces,
                   const or_env *env,
                   uint32_t first_sample, uint32_t spp, uint32_t pixel_begin,
                   uint32_t pixel_end, int threads, float *accum, float *per_sample_xyz,
                   or_path_info *per_sample_info, uint64_t *events);

/* StandardGemCuts::emerald_cut (geometry/cuts.rs:147-218): writes 34 planes, returns 34. */
int or_emerald_cut(float *p);
/* evaluate_gem_optical_metrics (crates/gemray/src/color/metrics.rs:1082-1248): out =
 * {brilliance_pct, fire_index, scintillation_pct, windowing_pct, extinction_pct}. */
enum { OR_MX_NONE = 0, OR_MX_BLOCKED = 1, OR_MX_WINDOWED = 2, OR_MX_EXTINCT = 3, OR_MX_RETURNED = 4,
       OR_MX_RETURNED_FIRE = 5 };
#define OR_MX_ITEMS (5 * 18 * 18)
void or_gem_optical_metrics(const float *planes, int n_planes, const or_material *m, float cam_yaw,
                            float cam_pitch, float light_yaw, float light_pitch, float out[5]);
/* Its per-ray records (OR_MX_ITEMS each; see the definition) and the ordered reduction. */
void or_metrics_items(const float *planes, int n_planes, const or_material *m, float cam_yaw, float cam_pitch,
                      float light_yaw, float light_pitch, int32_t *cls, float *angle, float *weight,
                      uint8_t *temporal);
void or_metrics_reduce(const int32_t *cls, const float *angle, const float *weight, const uint8_t *temporal,
                       float out[5]);

#ifdef __cplusplus
}
#endif
#endif
