// transport.h -- the by-value scene block the transport kernel consumes, and the
// launchers the C ABI calls.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gemray_hip.h"

namespace gr {

// Everything one trace launch needs besides the facet planes and the env-map texels.
// Passed as a kernel argument (it lands in the kernarg segment and is read with scalar
// loads).  Host-derived constants are computed in host C++ with the same operation order
// as the reference's host/CPU code: the studio rig (optics/studio_rig.rs:50-80), the
// white-balance LMS scale (optics/raytracer/color.rs:281-315), the 560 nm blackbody
// denominator (environment.rs:16-18) and the absorption frame
// (optics/birefringence.rs:126-131, 202-209).
struct DevScene {
    float cam_o[3], cam_f[3], cam_r[3], cam_u[3];
    float fov_tan;
    uint32_t width, height, n_planes, max_bounces;
    // material
    int32_t disp_kind;
    float disp[6];
    float birefringence_delta;
    int32_t anisotropic;      // crystal not cubic and |delta n| > 1e-4
    int32_t has_absorption;   // any absorption band present
    float c_axis[3];
    float abs_frame[3][3];    // a1, a2, normalized c
    int32_t n_o_bands, n_e_bands;
    float o_bands[GR_MAX_BANDS][3];
    float e_bands[GR_MAX_BANDS][3];
    float absorption_path_scale;
    // Biaxial indicatrix (materials.rs:1508-1519): its principal frame is abs_frame.
    int32_t biaxial;              // anisotropic with biaxial_delta_beta_alpha set
    float delta_beta_alpha;       // n_beta - n_alpha
    int32_t has_beta_bands;       // a trichroic tensor (biaxial with beta_ray Some)
    int32_t n_beta_bands;
    float beta_bands[GR_MAX_BANDS][3];
    float edge_rounding_radius;   // > 0: edge-rounded shading normals
    float sigma_s, scattering_g;  // Henyey-Greenstein volume scattering when sigma_s > 0
    uint32_t n_frosted;           // facets with FacetFinish::Frosted (0: finishes unused)
    const uint32_t *finishes;     // device, n_planes FacetFinish values, or null
    // environment
    int32_t env_kind;         // GR_ENV_STUDIO / GR_ENV_MAP
    float temp_k_clamped, spot_mult, exposure, denom_560;
    float key[3], fill[3], ring[16][3];
    float wb[3];
    int32_t map_w, map_h;
    const float *map_rgb;     // device pointer
    // Hero tables (device, or null: the kernels then evaluate per path): the colour-matching
    // values and map-spectrum lobes of every comb channel (kSpecFloats per entry, shared by
    // every scene), and this material's dispersion indices and absorption spectra
    // (mat_floats per entry; absorbing scenes only).
    const float *spec_tab;
    const float *mat_tab;
    // Bounds over the facet planes for the interior clip's first-pass screen
    // (clip_scan_screened, whose kernels pad with copies of plane 0): W >= max |d_i| and
    // B >= max |n_x| + |n_y| + |n_z|.
    float clip_w_max, clip_b_max;
};

// Hero tables.  Every per-path quantity that depends on the path only through its comb
// wavelengths -- the colour-matching values and map-spectrum lobes of the finish, and the
// material's dispersion indices and absorption spectra -- is a function of the hero offset
// hero_off = lambda_hero - 380 that init_path forms as fma(hero, 400, 380) - 380.  That
// difference is exact (Sterbenz), so hero offsets lie on lambda_hero's float grid: multiples
// of 2^-15 in [0, 132) (lambda_hero in [380, 512)) and of 2^-14 in [132, 400], kHeroEntries
// values in all.  Entry e of a hero table holds the eight channels' values at offset
// hero_key(e), evaluated by the same device functions the kernels evaluate per path, so a
// lookup returns the same bits, and one path's lookups touch one contiguous entry.
constexpr uint32_t kHeroLow = 132u << 15;                      // entries below offset 132
constexpr uint32_t kHeroEntries = kHeroLow + (268u << 14) + 1u;  // through offset 400
constexpr int kSpecFloats = 48;  // per channel k: CMF x, y, z at 3k; map lobes r, g, b at 24 + 3k
// Material table: per channel k, the dispersion index at k, the ordinary and extraordinary
// absorption sums at 8 + k and 16 + k, and (biaxial stones) the beta-axis sum at 24 + k.
inline constexpr int mat_floats(bool biaxial) { return biaxial ? 32 : 24; }

// Floats per continuation record (the parked state of one path), and the number of
// continuation launches after the plain trace launch (paths are parked again at each
// multiple of the split bounce until the last level).
constexpr int kContFields = 94;
constexpr int kContLevels = 3;
// Iterations of a launch before a wave may park early (park_lanes); measured with the
// single-job workload, where 28 % of the paths end after exactly three segments.
constexpr uint32_t kParkMin = 3;

// Arguments of one trace launch over tuples [0, n_tuples): tuple t is pixel
// pixel_begin + t / spp, sample sample_offset + t % spp.
struct TraceArgs {
    const float4 *planes;        // device, n_planes
    const float2 *ri35;          // device, spp entries: radical inverses (base 3, base 5)
    const uint32_t *cert;        // device, per-pixel camera-ray miss certificates of the frame
    uint32_t pixel_begin, sample_offset, spp, n_tuples;
    float *xyz;                  // device, 3 floats per tuple
    uint32_t *info;              // device, 4 u32 per tuple, or null
    unsigned long long *stats;   // device, GR_EV_COUNT counters (added into)
    bool wavefront;              // persistent wavefront kernel + finish pass
    float *rec;                  // device escape records (wavefront mode), 20 x rec_cap
    uint32_t rec_cap;
    hipEvent_t trace_done;       // recorded after the trace kernel (before the finish kernel or the
                                 // continuation levels), or null
    // Plain mode: paths running after split_bounce segments are parked in cont (two halves
    // of kContFields x cont_cap floats, kContLevels + 1 counters at cont_count) and
    // continued by further launches; 0 = no split.
    uint32_t split_bounce;
    // Early parking (park_due): from park_min iterations on, a wave with fewer than
    // park_lanes lanes still tracing parks them (0: only at split_bounce).
    uint32_t park_min, park_lanes;
    float *cont;
    uint32_t cont_cap;
    uint32_t *cont_count;
    // Occupancy variant of the plain trace kernels (2, 3 or 4 waves per SIMD; 0 = default).
    uint32_t waves_per_eu;
};

// One piece of one job in a batched launch: pixels [pixel_begin, pixel_begin + n_pixels)
// x samples [sample_offset, sample_offset + spp) of the job's frame.  Tuple t of the
// slice is pixel pixel_begin + t / spp, sample sample_offset + t % spp, exactly as in a
// plain launch.  Slices of one material class are traced by one launch; the trace kernel
// finds its slice by binary search of trace_block_begin over the launch's slices, the
// film reduce by reduce_block_begin over all slices of the tick.
struct BatchSlice {
    const DevScene *scene;       // device
    const float4 *planes;        // device, scene->n_planes
    const float2 *ri35;          // device, radical inverses of samples [sample_offset, +spp)
    const uint32_t *cert;        // device, the job's per-pixel miss certificates
    float *xyz;                  // device staging, 3 floats per tuple of this slice
    float *film;                 // device, the job's film (3 floats per pixel)
    unsigned long long *stats;   // device, the job's GR_EV_COUNT counters
    uint32_t pixel_begin, n_pixels, sample_offset, spp;
    uint32_t trace_block_begin;  // first workgroup of this slice in its class launch
    uint32_t reduce_block_begin; // first workgroup of this slice in the film reduce
    // Path-length split (see trace_kernel): this slice's continuation records, in two
    // halves of kContFields x cont_cap floats, and its kContLevels + 1 counters.
    uint32_t split_bounce;       // 0 = no split
    uint32_t park_lanes;         // early parking below this many live lanes (0 = none)
    uint32_t cont_cap;
    uint32_t cont_block_begin;   // first workgroup of this slice in its class's continuation launches
    float *cont[2];
    uint32_t *cont_count;
};

// One job's frame in a batched certificate launch: workgroups [block_begin, block_begin +
// ceil(n_pixels / kBlock)) fill cert[0, n_pixels) from the scene block and planes.
struct CertSlice {
    const DevScene *scene;  // device
    const float4 *planes;   // device, scene->n_planes
    uint32_t *cert;         // device, n_pixels
    uint32_t n_pixels;
    uint32_t block_begin;
};

// Whether a scene needs the kernels built for the rare features (kFrost): some facet
// frosted, or a scattering material.
inline bool rare_features(const DevScene &sc) { return sc.n_frosted != 0 || sc.sigma_s > 0.0f; }

// Material/environment class of a scene: bit 0 anisotropic, bit 1 absorbing, bit 2 HDR map,
// bit 3 rare features (frosted facets, volume scattering), bit 4 biaxial (only with bit 0).
inline int scene_class(const DevScene &sc) {
    return (sc.anisotropic ? 1 : 0) | (sc.has_absorption ? 2 : 0) | (sc.env_kind == GR_ENV_MAP ? 4 : 0) |
           (rare_features(sc) ? 8 : 0) | (sc.biaxial ? 16 : 0);
}

// The launchers of one scene class (scene_class bits), defined in the per-class translation
// units (trace_launch.h) and dispatched to by gr_launch_trace / gr_launch_trace_batch.
template <int C>
hipError_t launch_trace_class(const DevScene &sc, const TraceArgs &a, hipStream_t stream);
template <int C>
hipError_t launch_batch_class(const BatchSlice *slices, uint32_t n_slices, uint32_t n_blocks, uint32_t cont_blocks,
                              uint32_t max_planes, hipStream_t stream);

}  // namespace gr

hipError_t gr_launch_trace(const gr::DevScene &sc, const gr::TraceArgs &args, hipStream_t stream);
// Evaluates device function fn (GR_PROBE_*) on n items (9 floats in, 3 out each).
hipError_t gr_launch_probe(int fn, const float *in, float *out, uint32_t n, hipStream_t stream);
// Fills the primary-hit guide buffers of the frame (depth, normal xyz, facet per pixel).
hipError_t gr_launch_guide(const gr::DevScene &sc, const float4 *planes, float *depth, float *normal, int32_t *facet,
                           hipStream_t stream);
// Fills cert[0, width * height) with the frame's camera-ray miss certificates.
hipError_t gr_launch_cert(const gr::DevScene &sc, const float4 *planes, uint32_t *cert, hipStream_t stream);
// Fills the certificates of slices [0, n_slices) (block_begin ascending, the first 0) in one
// launch of n_blocks workgroups; bit-identical to one gr_launch_cert per slice.
hipError_t gr_launch_cert_batch(const gr::CertSlice *slices, uint32_t n_slices, uint32_t n_blocks, hipStream_t stream);
// Traces slices [0, n_slices) of one class (all with scene_class == cls) in one launch of
// n_blocks workgroups, then, when the slices split paths, kContLevels continuation
// launches of cont_blocks workgroups.
hipError_t gr_launch_trace_batch(const gr::BatchSlice *slices, uint32_t n_slices, uint32_t n_blocks,
                                 uint32_t cont_blocks, uint32_t max_planes, int cls, hipStream_t stream);
// Adds each slice's per-pixel ordered sample sums into its job's film.
hipError_t gr_launch_film_reduce_batch(const gr::BatchSlice *slices, uint32_t n_slices, uint32_t n_blocks,
                                       hipStream_t stream);
// Fills the spectral hero table (kHeroEntries x kSpecFloats floats at tab).
hipError_t gr_launch_spec_table(float *tab, hipStream_t stream);
// Fills scene sc's material hero table (kHeroEntries x mat_floats(sc.biaxial) floats at tab).
hipError_t gr_launch_mat_table(const gr::DevScene &sc, float *tab, hipStream_t stream);
// Fills ri35[0, n) with the radical inverses (bases 3, 5) of samples offset + i.
hipError_t gr_launch_ri35(uint32_t offset, uint32_t n, float2 *ri35, hipStream_t stream);
hipError_t gr_launch_film_reduce(const float *xyz, uint32_t n_pixels, uint32_t spp,
                                 uint32_t pixel_begin, float *film, hipStream_t stream);
