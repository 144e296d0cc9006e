// runtime.hip -- the C ABI of include/gemray_hip.h: context, validation and decline rules,
// host-side scene constants, chunked launches, film accumulation and statistics.
//
// Host-side arithmetic here (studio rig, white balance, absorption frame) follows the
// reference's CPU code operation for operation: the reference computes these on the host
// too (frame.rs:691 compute_illuminant_white_balance) or once per ray with values that do
// not depend on the ray (optics/studio_rig.rs:50-80).  The library is built with
// -ffp-contract=off so the host compiler does not fuse them either.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/gemray_hip.h"
#include "autotune.h"
#include "runtime_internal.h"
#include "transport.h"

namespace {

constexpr float kPiF = 3.14159265358979323846f;
constexpr uint64_t kDefaultChunkTuples = 1ull << 25;
constexpr uint64_t kMaxChunkTuples = 1ull << 30;

// ---- host restatements of reference host-side maths --------------------------------

struct H3 {
    float x, y, z;
};
inline float hdot(H3 a, H3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
inline H3 hmul(H3 a, float s) { return H3{a.x * s, a.y * s, a.z * s}; }
inline H3 hsub(H3 a, H3 b) { return H3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline H3 hcross(H3 a, H3 b) {
    return H3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline H3 hnormalize(H3 a) { return hmul(a, 1.0f / std::sqrt(hdot(a, a))); }
inline H3 hnormalize_or_zero(H3 a) {
    float rcp = 1.0f / std::sqrt(hdot(a, a));
    if (std::isfinite(rcp) && rcp > 0.0f) return hmul(a, rcp);
    return H3{0.0f, 0.0f, 0.0f};
}
inline float hclamp(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
// x^n rounded once (the reference's powi, as pinned by its golden vectors).
inline float hpowi(float x, int n) { return (float)std::pow((double)x, (double)n); }

// blackbody_spectrum (optics/raytracer/environment.rs:12-21).
float h_blackbody(float lambda_nm, float temp_k) {
    float t_k = std::fmax(temp_k, 1000.0f);
    const float h_c_k = 14388000.0f;
    float exp_val = std::exp(std::fmin(h_c_k / (lambda_nm * t_k), 80.0f));
    float exp_560 = std::exp(std::fmin(h_c_k / (560.0f * t_k), 80.0f));
    float denom = std::fmax(exp_val - 1.0f, 1e-6f);
    float denom_560 = std::fmax(exp_560 - 1.0f, 1e-6f);
    return hclamp(hpowi(560.0f / lambda_nm, 5) * (denom_560 / denom), 0.01f, 20.0f);
}
inline float h_lobe(float x, float mu, float lo, float hi) {
    float sigma = x < mu ? lo : hi;
    float t = (x - mu) / sigma;
    return std::exp(-0.5f * t * t);
}
// cie_1931_cmf, scalar form (color/cie1931.rs:29-42).
H3 h_cmf(float l) {
    float x = std::fmaf(0.065f, -h_lobe(l, 501.1f, 20.4f, 26.2f),
                        std::fmaf(1.056f, h_lobe(l, 599.8f, 37.9f, 31.0f), 0.362f * h_lobe(l, 442.0f, 16.0f, 26.7f)));
    float y = std::fmaf(0.286f, h_lobe(l, 530.9f, 16.3f, 31.1f), 0.821f * h_lobe(l, 568.8f, 46.9f, 40.5f));
    float z = std::fmaf(0.681f, h_lobe(l, 459.0f, 26.0f, 13.8f), 1.217f * h_lobe(l, 437.0f, 11.8f, 36.0f));
    return H3{std::fmax(x, 0.0f), std::fmax(y, 0.0f), std::fmax(z, 0.0f)};
}
inline H3 h_xyz_to_lms(H3 v) {
    return H3{std::fmaf(0.8951f, v.x, std::fmaf(0.2664f, v.y, -0.1614f * v.z)),
              std::fmaf(-0.7502f, v.x, std::fmaf(1.7135f, v.y, 0.0367f * v.z)),
              std::fmaf(0.0389f, v.x, std::fmaf(-0.0685f, v.y, 1.0296f * v.z))};
}
// compute_illuminant_white_balance (optics/raytracer/color.rs:281-315).
H3 h_white_balance(float temp_k) {
    H3 w{0.0f, 0.0f, 0.0f};
    for (int step = 0; step <= 400; ++step) {
        float lambda = 380.0f + (float)step;
        H3 c = h_cmf(lambda);
        float b = h_blackbody(lambda, temp_k);
        w = H3{w.x + c.x * b, w.y + c.y * b, w.z + c.z * b};
    }
    const float dx = 0.3127f, dy = 0.3290f;
    float ty = std::fmax(w.y, 1e-6f);
    H3 target{(dx / dy) * ty, ty, ((1.0f - dx - dy) / dy) * ty};
    H3 src = h_xyz_to_lms(w), tgt = h_xyz_to_lms(target);
    return H3{src.x > 1e-6f ? tgt.x / src.x : 1.0f, src.y > 1e-6f ? tgt.y / src.y : 1.0f,
              src.z > 1e-6f ? tgt.z / src.z : 1.0f};
}

// Per-temperature white-balance cache (the reference caches one per preset,
// color.rs:329-350).
struct WbCache {
    std::mutex mu;
    float temps[16];
    H3 vals[16];
    int n = 0;
    H3 get(float t) {
        std::lock_guard<std::mutex> g(mu);
        for (int i = 0; i < n; ++i)
            if (temps[i] == t) return vals[i];
        H3 v = h_white_balance(t);
        if (n < 16) {
            temps[n] = t;
            vals[n] = v;
            ++n;
        }
        return v;
    }
};
WbCache g_wb;

bool finite3(const float *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

double dispersion_f64(const gr_material *m, double lambda_nm) {
    double l2 = (lambda_nm * 1e-3) * (lambda_nm * 1e-3);
    if (m->disp_kind == GR_DISP_SELLMEIER1)
        return std::sqrt(std::fmax(1.0 + m->disp[0] * l2 / (l2 - m->disp[1]), 1.0));
    if (m->disp_kind == GR_DISP_SELLMEIER3) {
        double n2 = 1.0;
        for (int i = 0; i < 3; ++i) n2 += m->disp[i] * l2 / (l2 - m->disp[3 + i]);
        return std::sqrt(std::fmax(n2, 1.0));
    }
    return m->disp[0] + m->disp[1] / l2 + m->disp[2] / (l2 * l2);
}

}  // namespace

namespace gr {

gr_status validate_scene(const gr_scene *s) {
    if (!s || !s->camera || !s->material || !s->env || !s->planes) return GR_ERR_INVALID;
    if (s->width == 0 || s->height == 0) return GR_ERR_INVALID;
    if ((uint64_t)s->width * s->height > GR_MAX_PIXELS) return GR_ERR_INVALID;
    if (s->max_bounces == 0 || s->max_bounces > GR_MAX_BOUNCES) return GR_ERR_INVALID;
    if (s->n_planes == 0 || s->n_planes > GR_MAX_PLANES) return GR_ERR_INVALID;
    for (uint32_t i = 0; i < s->n_planes; ++i) {
        const float *p = s->planes + 4 * i;
        if (!finite3(p) || !std::isfinite(p[3])) return GR_ERR_INVALID;
        double len = std::sqrt((double)p[0] * p[0] + (double)p[1] * p[1] + (double)p[2] * p[2]);
        if (len <= 1e-6) return GR_ERR_INVALID;
    }
    const gr_camera *c = s->camera;
    if (!finite3(c->origin) || !finite3(c->forward) || !finite3(c->right) || !finite3(c->up) ||
        !std::isfinite(c->fov_tan))
        return GR_ERR_INVALID;
    const gr_material *m = s->material;
    if (m->disp_kind < 0 || m->disp_kind > 2) return GR_ERR_INVALID;
    if (!finite3(m->c_axis) || !std::isfinite(m->birefringence_delta)) return GR_ERR_INVALID;
    if (!std::isfinite(m->absorption_path_scale) || m->absorption_path_scale <= 0.0f) return GR_ERR_INVALID;
    if (m->n_o_bands < 0 || m->n_o_bands > GR_MAX_BANDS || m->n_e_bands < 0 || m->n_e_bands > GR_MAX_BANDS)
        return GR_ERR_INVALID;
    for (double l : {380.0, 589.3, 780.0}) {
        double n = dispersion_f64(m, l);
        if (!std::isfinite(n) || n < 1.0 || n > 6.0) return GR_ERR_INVALID;
    }
    const gr_env *e = s->env;
    if (e->kind == GR_ENV_STUDIO) {
        if (!std::isfinite(e->temp_k) || !std::isfinite(e->spot_mult) || !std::isfinite(e->exposure) ||
            !std::isfinite(e->light_yaw) || !std::isfinite(e->light_pitch))
            return GR_ERR_INVALID;
    } else if (e->kind == GR_ENV_MAP) {
        if (e->map) {
            if (!e->map->d || e->map->w <= 0 || e->map->h <= 0) return GR_ERR_INVALID;
        } else {
            if (e->map_w <= 0 || e->map_h <= 0 || !e->map_rgb) return GR_ERR_INVALID;
            if ((uint64_t)e->map_w * (uint64_t)e->map_h > (1ull << 28)) return GR_ERR_INVALID;
        }
    } else {
        return GR_DECLINE_ENV;
    }
    if (!std::isnan(m->biaxial_delta_beta_alpha)) {
        // A biaxial indicatrix: finite, with physical principal indices over the visible.
        if (!std::isfinite(m->biaxial_delta_beta_alpha)) return GR_ERR_INVALID;
        for (double l : {380.0, 589.3, 780.0}) {
            double nb = dispersion_f64(m, l), na = nb - m->biaxial_delta_beta_alpha,
                   ng = na + m->birefringence_delta;
            if (!(na >= 1.0 && na <= 6.0 && ng >= 1.0 && ng <= 6.0)) return GR_ERR_INVALID;
        }
        if (m->has_beta_bands != 0 && m->has_beta_bands != 1) return GR_ERR_INVALID;
        if (m->has_beta_bands && (m->n_beta_bands < 0 || m->n_beta_bands > GR_MAX_BANDS)) return GR_ERR_INVALID;
        // Branch not built here: the biaxial arms of an isotropic stone (a cubic crystal,
        // or |delta n| <= 1e-4).  Decline so the caller falls back to its CPU tracer.
        if (m->crystal_cubic || !(std::fabs(m->birefringence_delta) > 1e-4f)) return GR_DECLINE_MATERIAL;
    }
    // Henyey-Greenstein scattering: a finite extinction (<= 0 is off) and |g| <= 1.
    if (!std::isfinite(m->scattering_sigma_s)) return GR_ERR_INVALID;
    if (m->scattering_sigma_s > 0.0f && !(std::fabs(m->scattering_g) <= 1.0f)) return GR_ERR_INVALID;
    if (!std::isfinite(m->edge_rounding_radius)) return GR_ERR_INVALID;
    if (s->finishes)
        for (uint32_t i = 0; i < s->n_planes; ++i)
            if (s->finishes[i] > GR_FINISH_FROSTED) return GR_ERR_INVALID;
    return GR_OK;
}

void fill_dev_scene(const gr_scene *s, DevScene *d) {
    std::memset(d, 0, sizeof(*d));
    const gr_camera *c = s->camera;
    for (int i = 0; i < 3; ++i) {
        d->cam_o[i] = c->origin[i];
        d->cam_f[i] = c->forward[i];
        d->cam_r[i] = c->right[i];
        d->cam_u[i] = c->up[i];
    }
    d->fov_tan = c->fov_tan;
    d->width = s->width;
    d->height = s->height;
    d->n_planes = s->n_planes;
    // Bounds for the interior clip's first-pass screen (clip_scan_screened), rounded up: W
    // over |d|, B over |n_x| + |n_y| + |n_z| of the planes (the staged padding copies plane 0).
    double w_max = 1e-20, b_max = 0.0;
    for (uint32_t i = 0; i < s->n_planes; ++i) {
        const float *p = s->planes + 4 * (size_t)i;
        w_max = std::max(w_max, (double)std::fabs(p[3]));
        b_max = std::max(b_max, (double)std::fabs(p[0]) + (double)std::fabs(p[1]) + (double)std::fabs(p[2]));
    }
    d->clip_w_max = std::nextafter((float)w_max, INFINITY);
    d->clip_b_max = std::nextafter((float)b_max, INFINITY);
    d->max_bounces = s->max_bounces;
    const gr_material *m = s->material;
    d->disp_kind = m->disp_kind;
    for (int i = 0; i < 6; ++i) d->disp[i] = m->disp[i];
    d->birefringence_delta = m->birefringence_delta;
    d->anisotropic = (!m->crystal_cubic && std::fabs(m->birefringence_delta) > 1e-4f) ? 1 : 0;
    d->has_absorption = (m->n_o_bands > 0 || m->n_e_bands > 0) ? 1 : 0;
    for (int i = 0; i < 3; ++i) d->c_axis[i] = m->c_axis[i];
    // AbsorptionTensor3::uniaxial frame (birefringence.rs:126-131, 202-209).
    H3 cn = hnormalize_or_zero(H3{m->c_axis[0], m->c_axis[1], m->c_axis[2]});
    H3 a = std::fabs(cn.x) > 0.9f ? H3{0.0f, 1.0f, 0.0f} : H3{1.0f, 0.0f, 0.0f};
    H3 a1 = hnormalize_or_zero(hsub(a, hmul(cn, hdot(cn, a))));
    H3 a2 = hcross(cn, a1);
    H3 fr[3] = {a1, a2, cn};
    for (int j = 0; j < 3; ++j) {
        d->abs_frame[j][0] = fr[j].x;
        d->abs_frame[j][1] = fr[j].y;
        d->abs_frame[j][2] = fr[j].z;
    }
    d->n_o_bands = m->n_o_bands;
    d->n_e_bands = m->n_e_bands;
    std::memcpy(d->o_bands, m->o_bands, sizeof(d->o_bands));
    std::memcpy(d->e_bands, m->e_bands, sizeof(d->e_bands));
    d->absorption_path_scale = m->absorption_path_scale;
    d->biaxial = (d->anisotropic && !std::isnan(m->biaxial_delta_beta_alpha)) ? 1 : 0;
    if (d->biaxial) {
        d->delta_beta_alpha = m->biaxial_delta_beta_alpha;
        d->has_beta_bands = m->has_beta_bands ? 1 : 0;
        d->n_beta_bands = m->has_beta_bands ? m->n_beta_bands : 0;
        std::memcpy(d->beta_bands, m->beta_bands, sizeof(d->beta_bands));
        if (d->n_beta_bands > 0) d->has_absorption = 1;
    }
    d->edge_rounding_radius = m->edge_rounding_radius;
    d->sigma_s = m->scattering_sigma_s > 0.0f ? m->scattering_sigma_s : 0.0f;
    d->scattering_g = m->scattering_g;
    d->n_frosted = 0;
    if (s->finishes)
        for (uint32_t i = 0; i < s->n_planes; ++i) d->n_frosted += s->finishes[i] == GR_FINISH_FROSTED ? 1u : 0u;
    d->finishes = nullptr;  // device copy set by the uploader
    const gr_env *e = s->env;
    d->env_kind = e->kind;
    if (e->kind == GR_ENV_STUDIO) {
        d->temp_k_clamped = std::fmax(e->temp_k, 1000.0f);
        d->spot_mult = e->spot_mult;
        d->exposure = e->exposure;
        float exp_560 = std::exp(std::fmin(14388000.0f / (560.0f * d->temp_k_clamped), 80.0f));
        d->denom_560 = std::fmax(exp_560 - 1.0f, 1e-6f);
        // StudioRig::new (optics/studio_rig.rs:50-80).
        float cos_lp = std::cos(e->light_pitch), sin_lp = std::sin(e->light_pitch);
        float cos_ly = std::cos(e->light_yaw), sin_ly = std::sin(e->light_yaw);
        H3 key = hnormalize(H3{cos_lp * sin_ly, sin_lp, cos_lp * cos_ly});
        float fill_yaw = std::fmaf(kPiF, 0.78f, e->light_yaw);
        float fill_pitch = hclamp(e->light_pitch * 0.65f, 0.15f, 1.2f);
        H3 fill = hnormalize(H3{std::cos(fill_pitch) * std::sin(fill_yaw), std::sin(fill_pitch),
                                std::cos(fill_pitch) * std::cos(fill_yaw)});
        d->key[0] = key.x; d->key[1] = key.y; d->key[2] = key.z;
        d->fill[0] = fill.x; d->fill[1] = fill.y; d->fill[2] = fill.z;
        const float step = kPiF * 2.0f / 16.0f;
        for (int i = 0; i < 16; ++i) {
            float ang = std::fmaf((float)i, step, e->light_yaw);
            H3 r = hnormalize(H3{std::cos(ang) * 0.75f, sin_lp * 0.8f, std::sin(ang) * 0.75f});
            d->ring[i][0] = r.x; d->ring[i][1] = r.y; d->ring[i][2] = r.z;
        }
        H3 wb = g_wb.get(e->temp_k);
        d->wb[0] = wb.x; d->wb[1] = wb.y; d->wb[2] = wb.z;
    } else {
        // HdrMap: no white balance (environment.rs:238-243), LMS round trip with scale 1.
        d->wb[0] = d->wb[1] = d->wb[2] = 1.0f;
        d->map_w = e->map ? e->map->w : e->map_w;
        d->map_h = e->map ? e->map->h : e->map_h;
        if (e->map) d->map_rgb = e->map->d;
    }
}

// radical_inverse_base (optics/raytracer/sampling.rs:170-181), the same operations as the
// device's, evaluated once per call for its sample range.
void env_map_retain(const gr_env_map *m) { const_cast<gr_env_map *>(m)->refs.fetch_add(1); }
void env_map_release(const gr_env_map *m, hipStream_t stream) {
    gr_env_map *mm = const_cast<gr_env_map *>(m);
    if (mm->refs.fetch_sub(1) == 1) {
        // No work that reads the texels is outstanding: the scenes and jobs holding
        // references have completed their work before dropping them.  A stream-ordered free
        // does not wait for the device (the job loop's other tick keeps running).
        (void)hipFreeAsync(mm->d, stream);
        delete mm;
    }
}

float h_radical_inverse(uint32_t n, uint32_t base) {
    uint32_t remaining = n;
    float val = 0.0f;
    float inv_base = 1.0f / (float)base;
    while (remaining > 0) {
        uint32_t digit = remaining % base;
        val = std::fmaf((float)digit, inv_base, val);
        inv_base /= (float)base;
        remaining /= base;
    }
    return val;
}

}  // namespace gr

using gr::fill_dev_scene;
using gr::h_radical_inverse;
using gr::validate_scene;

struct gr_scene_dev {
    gr::DevScene sc;
    float4 *planes = nullptr;
    uint32_t *finishes = nullptr;  // only when some facet is frosted
    float *map = nullptr;             // the scene's own copy of its map, or
    const gr_env_map *env_map = nullptr;  // the device-resident map it references
    uint32_t *cert = nullptr;  // per-pixel camera-ray miss certificates
    // The material's hero table, built by the first trace large enough to repay it (an
    // absorbing scene only); the scene is otherwise read-only to traces.
    mutable float *mat_tab = nullptr;
};

struct gr_film {
    float *d = nullptr;
    uint32_t width = 0, height = 0;
};

namespace {

#define GR_CHECK(call)                                     \
    do {                                                   \
        hipError_t err_ = (call);                          \
        if (err_ != hipSuccess) {                          \
            std::fprintf(stderr, "gemray_hip: %s failed: %s\n", #call, hipGetErrorString(err_)); \
            return GR_ERR_DEVICE;                          \
        }                                                  \
    } while (0)

gr_status ensure_xyz(gr_ctx *ctx, uint64_t tuples) {
    if (tuples <= ctx->xyz_cap) return GR_OK;
    if (ctx->xyz) GR_CHECK(hipFree(ctx->xyz));
    ctx->xyz = nullptr;
    ctx->xyz_cap = 0;
    GR_CHECK(hipMalloc(&ctx->xyz, tuples * 3 * sizeof(float)));
    ctx->xyz_cap = tuples;
    return GR_OK;
}
gr_status ensure_info(gr_ctx *ctx, uint64_t tuples) {
    if (tuples <= ctx->info_cap) return GR_OK;
    if (ctx->info) GR_CHECK(hipFree(ctx->info));
    ctx->info = nullptr;
    ctx->info_cap = 0;
    GR_CHECK(hipMalloc(&ctx->info, tuples * 4 * sizeof(uint32_t)));
    ctx->info_cap = tuples;
    return GR_OK;
}

gr_status ensure_rec(gr_ctx *ctx, uint64_t tuples) {
    if (tuples <= ctx->rec_cap) return GR_OK;
    if (ctx->rec) GR_CHECK(hipFree(ctx->rec));
    ctx->rec = nullptr;
    ctx->rec_cap = 0;
    GR_CHECK(hipMalloc(&ctx->rec, tuples * 20 * sizeof(float)));
    ctx->rec_cap = tuples;
    return GR_OK;
}

// Two halves of continuation records for 1 / div of a launch's tuples each (paths beyond
// that carry on in the launch that would have parked them), and the counters of two lanes.
gr_status ensure_cont(gr_ctx *ctx, uint64_t tuples, uint32_t div) {
    const uint64_t cap = (tuples + div - 1) / div;
    if (cap <= ctx->cont_cap) return GR_OK;
    if (ctx->cont) GR_CHECK(hipFree(ctx->cont));
    ctx->cont = nullptr;
    ctx->cont_cap = 0;
    if (!ctx->cont_count) GR_CHECK(hipMalloc(&ctx->cont_count, 2 * (gr::kContLevels + 1) * sizeof(uint32_t)));
    GR_CHECK(hipMalloc(&ctx->cont, 2 * cap * gr::kContFields * sizeof(float)));
    ctx->cont_cap = (uint32_t)cap;
    return GR_OK;
}

gr_status ensure_ri35(gr_ctx *ctx, uint32_t sample_offset, uint32_t spp) {
    if (ctx->ri35 && ctx->ri35_offset == sample_offset && ctx->ri35_n == spp) return GR_OK;
    if (spp > ctx->ri35_cap) {
        if (ctx->ri35) GR_CHECK(hipFree(ctx->ri35));
        ctx->ri35 = nullptr;
        ctx->ri35_cap = 0;
        GR_CHECK(hipMalloc(&ctx->ri35, (size_t)spp * sizeof(float2)));
        ctx->ri35_cap = spp;
    }
    ctx->ri35_host.resize(spp);
    for (uint32_t i = 0; i < spp; ++i) {
        uint32_t s = sample_offset + i;
        ctx->ri35_host[i] = make_float2(h_radical_inverse(s, 3), h_radical_inverse(s, 5));
    }
    GR_CHECK(hipMemcpy(ctx->ri35, ctx->ri35_host.data(), (size_t)spp * sizeof(float2), hipMemcpyHostToDevice));
    ctx->ri35_offset = sample_offset;
    ctx->ri35_n = spp;
    return GR_OK;
}

gr_status upload_scene(gr_ctx *ctx, const gr_scene *s, gr_scene_dev *out) {
    fill_dev_scene(s, &out->sc);
    GR_CHECK(hipMalloc(&out->planes, s->n_planes * sizeof(float4)));
    GR_CHECK(hipMemcpyAsync(out->planes, s->planes, s->n_planes * sizeof(float4), hipMemcpyHostToDevice,
                            ctx->stream));
    if (out->sc.n_frosted) {
        GR_CHECK(hipMalloc(&out->finishes, s->n_planes * sizeof(uint32_t)));
        GR_CHECK(hipMemcpyAsync(out->finishes, s->finishes, s->n_planes * sizeof(uint32_t), hipMemcpyHostToDevice,
                                ctx->stream));
        out->sc.finishes = out->finishes;
    }
    if (s->env->kind == GR_ENV_MAP && s->env->map) {
        gr::env_map_retain(s->env->map);
        out->env_map = s->env->map;
        out->sc.map_rgb = s->env->map->d;
    } else if (s->env->kind == GR_ENV_MAP) {
        size_t bytes = (size_t)s->env->map_w * s->env->map_h * 3 * sizeof(float);
        GR_CHECK(hipMalloc(&out->map, bytes));
        GR_CHECK(hipMemcpyAsync(out->map, s->env->map_rgb, bytes, hipMemcpyHostToDevice, ctx->stream));
        out->sc.map_rgb = out->map;
    }
    GR_CHECK(hipMalloc(&out->cert, (size_t)s->width * s->height * sizeof(uint32_t)));
    GR_CHECK(gr_launch_cert(out->sc, out->planes, out->cert, ctx->stream));
    GR_CHECK(hipStreamSynchronize(ctx->stream));
    return GR_OK;
}

void free_scene(gr_scene_dev *s) {
    if (!s) return;
    if (s->planes) (void)hipFree(s->planes);
    if (s->finishes) (void)hipFree(s->finishes);
    if (s->map) (void)hipFree(s->map);
    if (s->env_map) gr::env_map_release(s->env_map, nullptr);
    if (s->cert) (void)hipFree(s->cert);
    if (s->mat_tab) (void)hipFree(s->mat_tab);
    s->mat_tab = nullptr;
    s->planes = nullptr;
    s->finishes = nullptr;
    s->map = nullptr;
    s->env_map = nullptr;
    s->cert = nullptr;
}

// Builds scene s's material hero table, stream-ordered before the traces that read it, when
// the scene absorbs, the context reads hero tables and a trace of `tuples` tuples repays it.
// Not having one (no memory) is not an error: the kernels then evaluate per path.
gr_status ensure_mat_table(gr_ctx *ctx, const gr_scene_dev *s, uint64_t tuples) {
    if (s->mat_tab || !gr::wants_mat_table(s->sc, tuples, ctx->hero_tables.load())) return GR_OK;
    float *tab = nullptr;
    const size_t bytes = (size_t)gr::kHeroEntries * gr::mat_floats(s->sc.biaxial != 0) * sizeof(float);
    if (hipMalloc(&tab, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return GR_OK;
    }
    if (gr_launch_mat_table(s->sc, tab, ctx->stream) != hipSuccess) {
        (void)hipFree(tab);
        return GR_ERR_DEVICE;
    }
    s->mat_tab = tab;
    return GR_OK;
}

// The scene block a trace launch takes: the scene's own, with the hero tables the context
// reads.
gr::DevScene launch_scene(gr_ctx *ctx, const gr_scene_dev *s) {
    gr::DevScene sc = s->sc;
    const bool tables = ctx->hero_tables.load() != 0;
    sc.spec_tab = tables ? ctx->spec_tab : nullptr;
    sc.mat_tab = tables ? s->mat_tab : nullptr;
    return sc;
}

// The device buffers one traced call uses besides the scene and the film, per lane: per-tuple
// XYZ staging, per-tuple info, escape records (wavefront mode, rec_cap per field),
// continuation records (two halves of kContFields x cont_cap floats) and their counters;
// and, shared, the call's sample table (radical inverses of its samples) and the event
// counters.  With two lanes, consecutive chunks alternate between them on two streams, so a
// chunk's continuation levels -- a few long paths in a mostly idle device -- overlap the
// next chunk's trace; each lane reuses its buffers only after its previous chunk, in stream
// order.  Chunks write disjoint pixels of the film and add into the counters atomically.
struct Workspace {
    struct Lane {
        float *xyz = nullptr;
        uint32_t *info = nullptr;
        float *rec = nullptr;
        float *cont = nullptr;
        uint32_t *cont_count = nullptr;
    } lane[2];
    uint32_t n_lanes = 1;
    uint64_t xyz_cap = 0;
    uint64_t rec_cap = 0;
    uint32_t cont_cap = 0;
    const float2 *ri35 = nullptr;
    unsigned long long *stats = nullptr;
};

// Traces pixels [pixel_begin, pixel_begin + n_pixels) x samples [sample_offset, +spp) in
// launches of at most chunk_pixels pixels, with the buffers of w.  With film != NULL each
// pixel's ordered sum is added into film; with xyz_host/info_host != NULL per-tuple outputs
// are copied out.  Chunks are queued ahead: the host waits for a chunk's timing events only
// when their slot of the ring comes round again, or at the end.
gr_status trace_chunks(gr_ctx *ctx, const gr_scene_dev *s, uint32_t pixel_begin, uint32_t n_pixels,
                       uint32_t sample_offset, uint32_t spp, float *film, float *xyz_host, uint32_t *info_host,
                       const Workspace &w, uint64_t chunk_pixels, bool wavefront, uint32_t split_bounce,
                       uint32_t park_lanes, uint32_t waves_per_eu) {
    GR_CHECK(hipMemsetAsync(w.stats, 0, GR_EV_COUNT * sizeof(unsigned long long), ctx->stream));
    GR_CHECK(hipEventRecord(ctx->ev_begin, ctx->stream));
    const uint32_t n_lanes = n_pixels > chunk_pixels ? w.n_lanes : 1u;
    const hipStream_t lane_stream[2] = {ctx->stream, ctx->stream2};
    if (n_lanes > 1) GR_CHECK(hipStreamWaitEvent(ctx->stream2, ctx->ev_begin, 0));
    const gr::DevScene sc = launch_scene(ctx, s);
    double kernel_ms = 0.0, finish_ms = 0.0;
    uint32_t queued = 0;  // chunks whose timing events have not been read
    auto read_slot = [&](uint32_t k) -> gr_status {
        const gr_ctx::ChunkEvents &e = ctx->chunk_ev[k % gr_ctx::kChunkRing];
        GR_CHECK(hipEventSynchronize(e.k2));
        float ms = 0.0f, fms = 0.0f;
        GR_CHECK(hipEventElapsedTime(&ms, e.k0, e.k1));
        GR_CHECK(hipEventElapsedTime(&fms, e.k1, e.k2));
        kernel_ms += ms;
        finish_ms += fms;
        return GR_OK;
    };
    uint32_t chunk = 0;
    hipEvent_t prev_trace_done = nullptr;
    for (uint64_t done = 0; done < n_pixels; done += chunk_pixels, ++chunk) {
        uint32_t np = (uint32_t)((n_pixels - done) < chunk_pixels ? (n_pixels - done) : chunk_pixels);
        uint32_t first = pixel_begin + (uint32_t)done;
        uint32_t n_tuples = np * spp;
        if (chunk >= gr_ctx::kChunkRing) {
            gr_status st = read_slot(chunk - gr_ctx::kChunkRing);
            if (st != GR_OK) return st;
            --queued;
        }
        const gr_ctx::ChunkEvents &e = ctx->chunk_ev[chunk % gr_ctx::kChunkRing];
        const Workspace::Lane &L = w.lane[chunk % n_lanes];
        const hipStream_t stream = lane_stream[chunk % n_lanes];
        gr::TraceArgs args{};
        args.planes = s->planes;
        args.ri35 = w.ri35;
        args.cert = s->cert;
        args.pixel_begin = first;
        args.sample_offset = sample_offset;
        args.spp = spp;
        args.n_tuples = n_tuples;
        args.xyz = L.xyz;
        args.info = info_host ? L.info : nullptr;
        args.stats = w.stats;
        args.wavefront = wavefront;
        args.rec = L.rec;
        args.rec_cap = (uint32_t)w.rec_cap;
        // The trace kernels of consecutive chunks do not overlap (each starts when the previous
        // one has ended); a chunk's continuation levels or finish kernel overlap the next
        // chunk's trace kernel.
        const hipEvent_t trace_done = wavefront ? e.k1 : e.kt;
        args.trace_done = trace_done;
        args.split_bounce = wavefront ? 0u : split_bounce;
        args.park_min = gr::kParkMin;
        args.park_lanes = park_lanes;
        args.cont = L.cont;
        args.cont_cap = w.cont_cap;
        args.cont_count = L.cont_count;
        args.waves_per_eu = waves_per_eu;
        if (n_lanes > 1 && prev_trace_done) GR_CHECK(hipStreamWaitEvent(stream, prev_trace_done, 0));
        prev_trace_done = trace_done;
        GR_CHECK(hipEventRecord(e.k0, stream));
        GR_CHECK(gr_launch_trace(sc, args, stream));
        if (!wavefront) GR_CHECK(hipEventRecord(e.k1, stream));
        GR_CHECK(hipEventRecord(e.k2, stream));
        if (film) GR_CHECK(gr_launch_film_reduce(L.xyz, np, spp, first, film, stream));
        if (xyz_host)
            GR_CHECK(hipMemcpyAsync(xyz_host + 3 * done * spp, L.xyz, (size_t)n_tuples * 3 * sizeof(float),
                                    hipMemcpyDeviceToHost, stream));
        if (info_host)
            GR_CHECK(hipMemcpyAsync(info_host + 4 * done * spp, L.info, (size_t)n_tuples * 4 * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, stream));
        ++queued;
        ctx->last.launches++;
    }
    for (uint32_t k = chunk - queued; k < chunk; ++k) {
        gr_status st = read_slot(k);
        if (st != GR_OK) return st;
    }
    if (n_lanes > 1) {
        GR_CHECK(hipEventRecord(ctx->ev_join, ctx->stream2));
        GR_CHECK(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    }
    GR_CHECK(hipEventRecord(ctx->ev_end, ctx->stream));
    GR_CHECK(hipStreamSynchronize(ctx->stream));
    float total = 0.0f;
    GR_CHECK(hipEventElapsedTime(&total, ctx->ev_begin, ctx->ev_end));
    unsigned long long ev[GR_EV_COUNT];
    GR_CHECK(hipMemcpy(ev, w.stats, sizeof(ev), hipMemcpyDeviceToHost));
    for (int i = 0; i < GR_EV_COUNT; ++i) ctx->last.events[i] = ev[i];
    ctx->last.paths = (uint64_t)n_pixels * spp;
    ctx->last.hero_tables = (sc.spec_tab ? 1u : 0u) | (sc.mat_tab ? 2u : 0u);
    ctx->last.trace_kernel_ms = kernel_ms;
    ctx->last.finish_kernel_ms = finish_ms;
    ctx->last.total_ms = total;
    return GR_OK;
}

// Traces with the context's own buffers, grown on demand (gr_render, gr_accumulate,
// gr_trace_samples): launches of at most chunk_tuples tuples.
// The context's own settings as a configuration (occupancy variant 0 = each class's default).
gr::TuneConfig context_config(gr_ctx *ctx) {
    gr::TuneConfig c;
    c.waves_per_eu = 0;
    c.split_bounce = ctx->split_bounce.load(std::memory_order_relaxed);
    c.park_lanes = ctx->park_lanes.load(std::memory_order_relaxed);
    c.cont_div = 4;
    c.trace_mode = (uint32_t)ctx->trace_mode;
    c.chunk_tuples = ctx->chunk_tuples;
    return c;
}

gr_status run_trace(gr_ctx *ctx, const gr_scene_dev *s, uint32_t pixel_begin, uint32_t n_pixels,
                    uint32_t sample_offset, uint32_t spp, float *film, float *xyz_host,
                    uint32_t *info_host, const gr::TuneConfig *cfg = nullptr) {
    std::memset(&ctx->last, 0, sizeof(ctx->last));
    ctx->last.n_planes = s->sc.n_planes;
    if (spp == 0 || n_pixels == 0) return GR_OK;
    const gr::TuneConfig c = cfg ? *cfg : context_config(ctx);
    uint64_t chunk_pixels = c.chunk_tuples / spp;
    if (chunk_pixels == 0) chunk_pixels = 1;
    if (chunk_pixels > n_pixels) chunk_pixels = n_pixels;
    uint64_t cap = chunk_pixels * spp;
    const bool wavefront = c.trace_mode == GR_TRACE_WAVEFRONT;
    const uint32_t split_bounce = c.split_bounce;
    // Two lanes (Workspace) when the call has more than one chunk.
    const uint32_t n_lanes = n_pixels > chunk_pixels ? 2u : 1u;
    gr_status st = ensure_xyz(ctx, n_lanes * cap);
    if (st != GR_OK) return st;
    if (info_host) {
        st = ensure_info(ctx, n_lanes * cap);
        if (st != GR_OK) return st;
    }
    if (wavefront) {
        st = ensure_rec(ctx, n_lanes * cap);
        if (st != GR_OK) return st;
    } else if (split_bounce) {
        st = ensure_cont(ctx, n_lanes * cap, c.cont_div);
        if (st != GR_OK) return st;
    }
    st = ensure_ri35(ctx, sample_offset, spp);
    if (st != GR_OK) return st;
    st = ensure_mat_table(ctx, s, (uint64_t)n_pixels * spp);
    if (st != GR_OK) return st;
    Workspace w;
    w.n_lanes = n_lanes;
    w.xyz_cap = cap;
    w.rec_cap = cap;
    w.cont_cap = (uint32_t)std::min<uint64_t>(ctx->cont_cap / n_lanes, (cap + c.cont_div - 1) / c.cont_div);
    for (uint32_t l = 0; l < n_lanes; ++l) {
        Workspace::Lane &L = w.lane[l];
        L.xyz = ctx->xyz + (size_t)l * cap * 3;
        L.info = ctx->info ? ctx->info + (size_t)l * cap * 4 : nullptr;
        L.rec = ctx->rec ? ctx->rec + (size_t)l * cap * 20 : nullptr;
        L.cont = ctx->cont ? ctx->cont + (size_t)l * 2 * w.cont_cap * gr::kContFields : nullptr;
        L.cont_count = ctx->cont_count ? ctx->cont_count + (size_t)l * (gr::kContLevels + 1) : nullptr;
    }
    w.ri35 = ctx->ri35;
    w.stats = ctx->stats;
    return trace_chunks(ctx, s, pixel_begin, n_pixels, sample_offset, spp, film, xyz_host, info_host, w,
                        chunk_pixels, wavefront, split_bounce, c.park_lanes, c.waves_per_eu);
}

// Probe launches of the autotuner: about this many tuples (at most the call's own) of the
// frame's middle rows at the call's samples per pixel -- the densest band of a centred stone,
// traced in the waves (samples of a pixel) the call will trace, so a configuration whose
// continuation records run out in the frame's stone-filled chunks shows it in the probes.
constexpr uint64_t kTuneProbeTuples = 1ull << 25;

gr::TuneConfig from_choice(const gr_tune_choice &c) {
    gr::TuneConfig t;
    t.waves_per_eu = c.waves_per_eu;
    t.split_bounce = c.split_bounce;
    t.park_lanes = c.park_lanes;
    t.cont_div = c.cont_div;
    t.trace_mode = c.trace_mode;
    t.chunk_tuples = c.chunk_tuples;
    return t;
}
void to_choice(const gr::TuneConfig &t, double ms, int probes, gr_tune_choice *c) {
    std::memset(c, 0, sizeof(*c));
    c->waves_per_eu = t.waves_per_eu;
    c->split_bounce = t.split_bounce;
    c->park_lanes = t.park_lanes;
    c->cont_div = t.cont_div;
    c->trace_mode = t.trace_mode;
    c->chunk_tuples = t.chunk_tuples;
    c->probe_ms = ms;
    c->probes = (uint32_t)probes;
}
bool valid_choice(const gr_tune_choice &c) {
    return c.waves_per_eu >= 2 && c.waves_per_eu <= 4 && c.split_bounce <= GR_MAX_BOUNCES && c.park_lanes <= 64 &&
           c.cont_div >= 1 &&
           c.cont_div <= 64 && (c.trace_mode == GR_TRACE_PLAIN || c.trace_mode == GR_TRACE_WAVEFRONT) &&
           c.chunk_tuples >= 1 && c.chunk_tuples <= kMaxChunkTuples;
}

// Tunes the configuration for renders of scene at spp samples per pixel (coordinate descent
// over probe renders into a scratch film, autotune.h) and caches it under the scene's
// class and size.  ctx->mu is held.
gr_status tune(gr_ctx *ctx, const gr_scene_dev *s, uint32_t spp, gr::TuneConfig *out, double *ms, int *probes) {
    const uint32_t n_pixels = s->sc.width * s->sc.height;
    const uint32_t probe_spp = std::max<uint32_t>(spp, 1);
    const uint32_t probe_pixels = (uint32_t)std::min<uint64_t>(n_pixels, std::max<uint64_t>(1, kTuneProbeTuples / probe_spp));
    const uint32_t probe_begin = (n_pixels - probe_pixels) / 2;
    // The probes trace with the hero tables the tuned renders will read.
    gr_status tst = ensure_mat_table(ctx, s, (uint64_t)n_pixels * spp);
    if (tst != GR_OK) return tst;
    float *film = nullptr;
    GR_CHECK(hipMalloc(&film, (size_t)n_pixels * 3 * sizeof(float)));
    gr::TuneConfig start;
    start.chunk_tuples = ctx->chunk_tuples;
    start.split_bounce = ctx->split_bounce.load(std::memory_order_relaxed);
    start.park_lanes = ctx->park_lanes.load(std::memory_order_relaxed);
    gr_status st = GR_OK;
    auto time = [&](const gr::TuneConfig &c) {
        const gr_status r = run_trace(ctx, s, probe_begin, probe_pixels, 0, probe_spp, film, nullptr, nullptr, &c);
        if (r != GR_OK) st = r;
        return r == GR_OK ? ctx->last.total_ms : 1e30;
    };
    (void)time(start);  // first-launch costs outside the comparison
    const gr::TuneSpace space;
    *out = gr::coordinate_descent(start, space, 0.01, 2, time, ms, probes);
    (void)hipFree(film);
    if (st != GR_OK) return st;
    ctx->tuned[gr::tune_key(gr::scene_class(s->sc), s->sc.n_planes, (uint64_t)n_pixels * spp)] = *out;
    return GR_OK;
}

// gr_render_noalloc's workspace, one fixed device array per GPU (56 MB): XYZ staging for a
// launch of kNoAllocTuples tuples, continuation records for an eighth of them (a path that
// finds the records full carries on in its launch, so results do not depend on the
// capacity), a sample table for GR_MAX_SAMPLES_PER_CALL samples, the event counters and the
// continuation counters, each section 256-byte aligned.
constexpr uint64_t kNoAllocTuples = 1ull << 19;
constexpr uint32_t kNoAllocContCap = (uint32_t)(kNoAllocTuples / 8);
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kScratchXyz = 0;
constexpr size_t kScratchCont = kScratchXyz + align256(kNoAllocTuples * 3 * sizeof(float));
constexpr size_t kScratchRi35 = kScratchCont + align256(2ull * kNoAllocContCap * gr::kContFields * sizeof(float));
constexpr size_t kScratchStats = kScratchRi35 + align256((size_t)GR_MAX_SAMPLES_PER_CALL * sizeof(float2));
constexpr size_t kScratchCounts = kScratchStats + align256(GR_EV_COUNT * sizeof(unsigned long long));
constexpr size_t kScratchBytes = kScratchCounts + align256((gr::kContLevels + 1) * sizeof(uint32_t));

bool create_chunk_events(gr_ctx *ctx) {
    for (gr_ctx::ChunkEvents &c : ctx->chunk_ev)
        if (hipEventCreate(&c.k0) != hipSuccess || hipEventCreate(&c.k1) != hipSuccess ||
            hipEventCreate(&c.k2) != hipSuccess || hipEventCreateWithFlags(&c.kt, hipEventDisableTiming) != hipSuccess)
            return false;
    return true;
}

}  // namespace

// The host-side name of this array is a stand-in, not its device address: host code uses the
// address hipGetSymbolAddress returns for the current device (gr_ctx::scratch) only.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

extern "C" {

gr_status gr_open(gr_ctx **out, const gr_options *opts) {
    if (!out) return GR_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return GR_DECLINE_NO_DEVICE;
    int dev = opts ? opts->device : 0;
    if (dev < 0 || dev >= count) return GR_DECLINE_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return GR_DECLINE_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GR_DECLINE_NO_DEVICE;
    gr_ctx *ctx = new (std::nothrow) gr_ctx();
    if (!ctx) return GR_ERR_DEVICE;
    ctx->device = dev;
    ctx->chunk_tuples = kDefaultChunkTuples;
    if (opts) ctx->hbm_budget = opts->hbm_budget_bytes;
    if (opts && opts->chunk_tuples) ctx->chunk_tuples = opts->chunk_tuples < kMaxChunkTuples ? opts->chunk_tuples : kMaxChunkTuples;
    // Some drivers report no marketing name: the label then names the device by ordinal and
    // compute-unit count, so it is never empty (gpu_backend.rs:146-153 shows the adapter).
    if (prop.name[0] != '\0')
        std::snprintf(ctx->label, sizeof(ctx->label), "%s (%s, HIP)", prop.name, prop.gcnArchName);
    else
        std::snprintf(ctx->label, sizeof(ctx->label), "HIP device %d, %d CUs (%s, HIP)", dev,
                      prop.multiProcessorCount, prop.gcnArchName);
    if (hipSetDevice(dev) != hipSuccess ||
        hipGetSymbolAddress((void **)&ctx->scratch, HIP_SYMBOL(scratch)) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev_begin) != hipSuccess || hipEventCreate(&ctx->ev_end) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
        !create_chunk_events(ctx) ||
        hipMalloc(&ctx->stats, GR_EV_COUNT * sizeof(unsigned long long)) != hipSuccess) {
        gr_close(ctx);
        return GR_ERR_DEVICE;
    }
    // The spectral hero table (1.7 GB): without the memory for it the kernels evaluate the
    // finish per path instead.
    if (hipMalloc(&ctx->spec_tab, (size_t)gr::kHeroEntries * gr::kSpecFloats * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        ctx->spec_tab = nullptr;
    } else if (gr_launch_spec_table(ctx->spec_tab, ctx->stream) != hipSuccess ||
               hipStreamSynchronize(ctx->stream) != hipSuccess) {
        gr_close(ctx);
        return GR_ERR_DEVICE;
    }
    *out = ctx;
    return GR_OK;
}

void gr_close(gr_ctx *ctx) {
    if (!ctx) return;
    gr::serve_shutdown(ctx);
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    if (ctx->xyz) (void)hipFree(ctx->xyz);
    if (ctx->info) (void)hipFree(ctx->info);
    if (ctx->rec) (void)hipFree(ctx->rec);
    if (ctx->ri35) (void)hipFree(ctx->ri35);
    if (ctx->cont) (void)hipFree(ctx->cont);
    if (ctx->cont_count) (void)hipFree(ctx->cont_count);
    if (ctx->mx_buf) (void)hipFree(ctx->mx_buf);
    if (ctx->stats) (void)hipFree(ctx->stats);
    if (ctx->spec_tab) (void)hipFree(ctx->spec_tab);
    for (hipEvent_t e : {ctx->ev_begin, ctx->ev_end, ctx->ev_join})
        if (e) (void)hipEventDestroy(e);
    for (gr_ctx::ChunkEvents &c : ctx->chunk_ev)
        for (hipEvent_t e : {c.k0, c.k1, c.k2, c.kt})
            if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
}

int gr_adapter_label(gr_ctx *ctx, char *buf, size_t len) {
    if (!ctx) return -1;
    int n = (int)std::strlen(ctx->label);
    if (buf && len) std::snprintf(buf, len, "%s", ctx->label);
    return n;
}

gr_status gr_set_trace_mode(gr_ctx *ctx, int mode) {
    if (!ctx || (mode != GR_TRACE_PLAIN && mode != GR_TRACE_WAVEFRONT)) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->trace_mode = mode;
    return GR_OK;
}

gr_status gr_set_split_bounce(gr_ctx *ctx, uint32_t split_bounce) {
    if (!ctx || split_bounce > GR_MAX_BOUNCES) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->split_bounce = split_bounce;
    return GR_OK;
}

gr_status gr_set_hero_tables(gr_ctx *ctx, int mode) {
    if (!ctx || mode < 0 || mode > 2) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->hero_tables.store(mode);
    return GR_OK;
}

gr_status gr_set_chunk_tuples(gr_ctx *ctx, uint64_t tuples) {
    if (!ctx || tuples == 0) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->chunk_tuples = tuples < kMaxChunkTuples ? tuples : kMaxChunkTuples;
    return GR_OK;
}

gr_status gr_camera_new(float yaw, float pitch, float distance, float fov_deg, gr_camera *out) {
    if (!out || !std::isfinite(yaw) || !std::isfinite(pitch) || !std::isfinite(distance) ||
        !(distance > 0.0f) || !std::isfinite(fov_deg))
        return GR_ERR_INVALID;
    float cos_p = std::cos(pitch), sin_p = std::sin(pitch);
    float cos_y = std::cos(yaw), sin_y = std::sin(yaw);
    H3 origin{distance * cos_p * sin_y, distance * sin_p, distance * cos_p * cos_y};
    H3 forward = hnormalize(H3{-origin.x, -origin.y, -origin.z});
    H3 world_up = std::fabs(cos_p) < 1e-4f ? H3{0.0f, 0.0f, -1.0f} : H3{0.0f, 1.0f, 0.0f};
    H3 right = hnormalize(hcross(forward, world_up));
    H3 up = hnormalize(hcross(right, forward));
    const float rads_per_deg = kPiF / 180.0f;
    float fov_tan = std::tan(fov_deg * rads_per_deg * 0.5f);
    H3 v[4] = {origin, forward, right, up};
    float *dst[4] = {out->origin, out->forward, out->right, out->up};
    for (int i = 0; i < 4; ++i) {
        dst[i][0] = v[i].x;
        dst[i][1] = v[i].y;
        dst[i][2] = v[i].z;
    }
    out->fov_tan = fov_tan;
    return GR_OK;
}

gr_status gr_probe(gr_ctx *ctx, int fn, const float *in, size_t n, float *out) {
    if (!ctx || !in || !out || fn < 0 || fn >= GR_PROBE_COUNT || n == 0 || n > (1u << 24)) return GR_ERR_INVALID;
    // The sample-draw probe takes its radical inverses from the host's sample table, exactly
    // as a render does (ensure_ri35).
    std::vector<float> staged;
    if (fn == GR_PROBE_SAMPLE_DRAWS) {
        staged.assign(in, in + n * 9);
        for (size_t i = 0; i < n; ++i) {
            uint32_t s;
            std::memcpy(&s, &staged[9 * i + 1], sizeof(s));
            staged[9 * i + 2] = h_radical_inverse(s, 3);
            staged[9 * i + 3] = h_radical_inverse(s, 5);
        }
        in = staged.data();
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    GR_CHECK(hipSetDevice(ctx->device));
    float *d_in = nullptr, *d_out = nullptr;
    gr_status st = GR_OK;
    if (hipMalloc(&d_in, n * 9 * sizeof(float)) != hipSuccess || hipMalloc(&d_out, n * 3 * sizeof(float)) != hipSuccess ||
        hipMemcpyAsync(d_in, in, n * 9 * sizeof(float), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        gr_launch_probe(fn, d_in, d_out, (uint32_t)n, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(out, d_out, n * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = GR_ERR_DEVICE;
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    return st;
}

gr_status gr_guide_buffers(gr_ctx *ctx, const gr_camera *camera, uint32_t width, uint32_t height, const float *planes,
                           uint32_t n_planes, float *depth, float *normal, int32_t *facet) {
    if (!ctx || !camera) return GR_ERR_INVALID;
    if (width == 0 || height == 0) return GR_OK;
    if (!depth || !normal || !facet || (n_planes > 0 && !planes)) return GR_ERR_INVALID;
    if ((uint64_t)width * height > GR_MAX_PIXELS || n_planes > GR_MAX_PLANES) return GR_ERR_INVALID;
    if (!finite3(camera->origin) || !finite3(camera->forward) || !finite3(camera->right) || !finite3(camera->up) ||
        !std::isfinite(camera->fov_tan))
        return GR_ERR_INVALID;
    for (uint32_t i = 0; i < n_planes; ++i)
        if (!finite3(planes + 4 * i) || !std::isfinite(planes[4 * i + 3])) return GR_ERR_INVALID;
    gr::DevScene sc;
    std::memset(&sc, 0, sizeof(sc));
    for (int i = 0; i < 3; ++i) {
        sc.cam_o[i] = camera->origin[i];
        sc.cam_f[i] = camera->forward[i];
        sc.cam_r[i] = camera->right[i];
        sc.cam_u[i] = camera->up[i];
    }
    sc.fov_tan = camera->fov_tan;
    sc.width = width;
    sc.height = height;
    sc.n_planes = n_planes;
    const size_t n = (size_t)width * height;
    std::lock_guard<std::mutex> g(ctx->mu);
    GR_CHECK(hipSetDevice(ctx->device));
    float4 *d_planes = nullptr;
    float *d_depth = nullptr, *d_normal = nullptr;
    int32_t *d_facet = nullptr;
    gr_status st = GR_OK;
    if (hipMalloc(&d_planes, (n_planes ? n_planes : 1) * sizeof(float4)) != hipSuccess ||
        hipMalloc(&d_depth, n * sizeof(float)) != hipSuccess || hipMalloc(&d_normal, 3 * n * sizeof(float)) != hipSuccess ||
        hipMalloc(&d_facet, n * sizeof(int32_t)) != hipSuccess ||
        (n_planes && hipMemcpyAsync(d_planes, planes, n_planes * sizeof(float4), hipMemcpyHostToDevice, ctx->stream) !=
                         hipSuccess) ||
        gr_launch_guide(sc, d_planes, d_depth, d_normal, d_facet, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(depth, d_depth, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(normal, d_normal, 3 * n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(facet, d_facet, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = GR_ERR_DEVICE;
    if (d_planes) (void)hipFree(d_planes);
    if (d_depth) (void)hipFree(d_depth);
    if (d_normal) (void)hipFree(d_normal);
    if (d_facet) (void)hipFree(d_facet);
    return st;
}

gr_status gr_env_map_upload(gr_ctx *ctx, const float *rgb, int32_t w, int32_t h, gr_env_map **out) {
    if (!ctx || !rgb || !out || w <= 0 || h <= 0 || (uint64_t)w * (uint64_t)h > (1ull << 28)) return GR_ERR_INVALID;
    *out = nullptr;
    const size_t n = (size_t)w * h * 3;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(rgb[i])) return GR_ERR_INVALID;
    gr_env_map *m = new (std::nothrow) gr_env_map();
    if (!m) return GR_ERR_DEVICE;
    m->w = w;
    m->h = h;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess ||
        hipMallocAsync((void **)&m->d, n * sizeof(float), ctx->stream) != hipSuccess ||
        hipMemcpyAsync(m->d, rgb, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        if (m->d) (void)hipFreeAsync(m->d, ctx->stream);
        delete m;
        return GR_ERR_DEVICE;
    }
    *out = m;
    return GR_OK;
}

void gr_env_map_release(gr_ctx *ctx, gr_env_map *map) {
    if (!ctx || !map) return;
    (void)hipSetDevice(ctx->device);
    gr::env_map_release(map, ctx->stream);
}

gr_status gr_last_stats(gr_ctx *ctx, gr_stats *out) {
    if (!ctx || !out) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    *out = ctx->last;
    return GR_OK;
}

gr_status gr_scene_upload(gr_ctx *ctx, const gr_scene *scene, gr_scene_dev **out) {
    if (!ctx || !out) return GR_ERR_INVALID;
    *out = nullptr;
    gr_status st = validate_scene(scene);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return GR_ERR_DEVICE;
    gr_scene_dev *s = new (std::nothrow) gr_scene_dev();
    if (!s) return GR_ERR_DEVICE;
    st = upload_scene(ctx, scene, s);
    if (st != GR_OK) {
        free_scene(s);
        delete s;
        return st;
    }
    *out = s;
    return GR_OK;
}

void gr_scene_free(gr_ctx *ctx, gr_scene_dev *scene) {
    if (!ctx || !scene) return;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_scene(scene);
    delete scene;
}

gr_status gr_film_create(gr_ctx *ctx, uint32_t width, uint32_t height, gr_film **out) {
    if (!ctx || !out || width == 0 || height == 0 || (uint64_t)width * height > GR_MAX_PIXELS) return GR_ERR_INVALID;
    *out = nullptr;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return GR_ERR_DEVICE;
    gr_film *f = new (std::nothrow) gr_film();
    if (!f) return GR_ERR_DEVICE;
    f->width = width;
    f->height = height;
    size_t bytes = (size_t)width * height * 3 * sizeof(float);
    if (hipMalloc(&f->d, bytes) != hipSuccess || hipMemsetAsync(f->d, 0, bytes, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        if (f->d) (void)hipFree(f->d);
        delete f;
        return GR_ERR_DEVICE;
    }
    *out = f;
    return GR_OK;
}

gr_status gr_film_clear(gr_ctx *ctx, gr_film *film) {
    if (!ctx || !film) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    GR_CHECK(hipSetDevice(ctx->device));
    GR_CHECK(hipMemsetAsync(film->d, 0, (size_t)film->width * film->height * 3 * sizeof(float), ctx->stream));
    GR_CHECK(hipStreamSynchronize(ctx->stream));
    return GR_OK;
}

void gr_film_free(gr_ctx *ctx, gr_film *film) {
    if (!ctx || !film) return;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (film->d) (void)hipFree(film->d);
    delete film;
}

gr_status gr_film_download(gr_ctx *ctx, gr_film *film, float *xyz_out, size_t n_floats) {
    if (!ctx || !film || !xyz_out) return GR_ERR_INVALID;
    size_t need = (size_t)film->width * film->height * 3;
    if (n_floats != need) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    GR_CHECK(hipSetDevice(ctx->device));
    GR_CHECK(hipMemcpyAsync(xyz_out, film->d, need * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    GR_CHECK(hipStreamSynchronize(ctx->stream));
    return GR_OK;
}

gr_status gr_render(gr_ctx *ctx, const gr_scene_dev *scene, uint32_t sample_offset, uint32_t spp,
                    gr_film *film) {
    if (!ctx || !scene || !film) return GR_ERR_INVALID;
    if (film->width != scene->sc.width || film->height != scene->sc.height) return GR_ERR_INVALID;
    if (spp > GR_MAX_SAMPLES_PER_CALL) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    GR_CHECK(hipSetDevice(ctx->device));
    const uint32_t n_pixels = film->width * film->height;
    if (ctx->autotune && spp > 0) {
        const uint64_t key = gr::tune_key(gr::scene_class(scene->sc), scene->sc.n_planes, (uint64_t)n_pixels * spp);
        auto it = ctx->tuned.find(key);
        gr::TuneConfig cfg;
        if (it != ctx->tuned.end()) {
            cfg = it->second;
        } else {
            double ms = 0.0;
            int probes = 0;
            gr_status st = tune(ctx, scene, spp, &cfg, &ms, &probes);
            if (st != GR_OK) return st;
        }
        return run_trace(ctx, scene, 0, n_pixels, sample_offset, spp, film->d, nullptr, nullptr, &cfg);
    }
    return run_trace(ctx, scene, 0, n_pixels, sample_offset, spp, film->d, nullptr, nullptr);
}

gr_status gr_autotune(gr_ctx *ctx, const gr_scene_dev *scene, uint32_t spp, gr_tune_choice *out) {
    if (!ctx || !scene || !out || spp == 0 || spp > GR_MAX_SAMPLES_PER_CALL) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    GR_CHECK(hipSetDevice(ctx->device));
    gr::TuneConfig cfg;
    double ms = 0.0;
    int probes = 0;
    gr_status st = tune(ctx, scene, spp, &cfg, &ms, &probes);
    if (st != GR_OK) return st;
    to_choice(cfg, ms, probes, out);
    return GR_OK;
}

gr_status gr_set_autotune(gr_ctx *ctx, int enabled) {
    if (!ctx) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->autotune = enabled != 0;
    return GR_OK;
}

gr_status gr_render_tuned(gr_ctx *ctx, const gr_scene_dev *scene, uint32_t sample_offset, uint32_t spp, gr_film *film,
                          const gr_tune_choice *choice) {
    if (!ctx || !scene || !film || !choice || !valid_choice(*choice)) return GR_ERR_INVALID;
    if (film->width != scene->sc.width || film->height != scene->sc.height) return GR_ERR_INVALID;
    if (spp > GR_MAX_SAMPLES_PER_CALL) return GR_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    GR_CHECK(hipSetDevice(ctx->device));
    const gr::TuneConfig cfg = from_choice(*choice);
    return run_trace(ctx, scene, 0, film->width * film->height, sample_offset, spp, film->d, nullptr, nullptr, &cfg);
}

gr_status gr_render_noalloc(gr_ctx *ctx, const gr_scene_dev *scene, uint32_t sample_offset, uint32_t spp,
                            gr_film *film) {
    if (!ctx || !scene || !film) return GR_ERR_INVALID;
    if (film->width != scene->sc.width || film->height != scene->sc.height) return GR_ERR_INVALID;
    if (spp > GR_MAX_SAMPLES_PER_CALL) {
        std::fprintf(stderr, "gemray_hip: gr_render_noalloc: %u samples per call exceed its sample table (%u)\n",
                     spp, GR_MAX_SAMPLES_PER_CALL);
        return GR_ERR_INVALID;
    }
    if (!ctx->scratch) return GR_ERR_DEVICE;
    std::lock_guard<std::mutex> g(ctx->mu);
    GR_CHECK(hipSetDevice(ctx->device));
    std::memset(&ctx->last, 0, sizeof(ctx->last));
    ctx->last.n_planes = scene->sc.n_planes;
    const uint32_t n_pixels = film->width * film->height;
    if (spp == 0) return GR_OK;
    uint64_t chunk_pixels = kNoAllocTuples / spp;
    if (chunk_pixels > n_pixels) chunk_pixels = n_pixels;
    Workspace w;  // one lane: chunks run in order on the context's stream
    w.lane[0].xyz = reinterpret_cast<float *>(ctx->scratch + kScratchXyz);
    w.xyz_cap = kNoAllocTuples;
    w.lane[0].cont = reinterpret_cast<float *>(ctx->scratch + kScratchCont);
    w.cont_cap = kNoAllocContCap;
    w.lane[0].cont_count = reinterpret_cast<uint32_t *>(ctx->scratch + kScratchCounts);
    float2 *ri35 = reinterpret_cast<float2 *>(ctx->scratch + kScratchRi35);
    w.ri35 = ri35;
    w.stats = reinterpret_cast<unsigned long long *>(ctx->scratch + kScratchStats);
    GR_CHECK(gr_launch_ri35(sample_offset, spp, ri35, ctx->stream));
    return trace_chunks(ctx, scene, 0, n_pixels, sample_offset, spp, film->d, nullptr, nullptr, w, chunk_pixels,
                        false, ctx->split_bounce.load(std::memory_order_relaxed),
                        ctx->park_lanes.load(std::memory_order_relaxed), 0u);
}

gr_status gr_accumulate(gr_ctx *ctx, const gr_scene *scene, uint32_t sample_offset, uint32_t spp,
                        float *accum_xyz, size_t accum_pixels) {
    if (!ctx || !scene || !accum_xyz) return GR_ERR_INVALID;
    gr_status st = validate_scene(scene);
    if (st != GR_OK) return st;
    size_t n_pixels = (size_t)scene->width * scene->height;
    if (accum_pixels != n_pixels) return GR_ERR_INVALID;
    if (spp > GR_MAX_SAMPLES_PER_CALL) return GR_ERR_INVALID;
    if (spp == 0) return GR_OK;
    gr_scene_dev *s = nullptr;
    st = gr_scene_upload(ctx, scene, &s);
    if (st != GR_OK) return st;
    gr_film *f = nullptr;
    st = gr_film_create(ctx, scene->width, scene->height, &f);
    if (st == GR_OK) st = gr_render(ctx, s, sample_offset, spp, f);
    float *host = nullptr;
    if (st == GR_OK) {
        host = new (std::nothrow) float[n_pixels * 3];
        if (!host) st = GR_ERR_DEVICE;
    }
    if (st == GR_OK) st = gr_film_download(ctx, f, host, n_pixels * 3);
    if (st == GR_OK)
        for (size_t i = 0; i < n_pixels * 3; ++i) accum_xyz[i] += host[i];
    delete[] host;
    gr_film_free(ctx, f);
    gr_scene_free(ctx, s);
    return st;
}

gr_status gr_trace_samples(gr_ctx *ctx, const gr_scene *scene, uint32_t pixel_begin, uint32_t pixel_count,
                           uint32_t sample_offset, uint32_t spp, float *xyz_out, uint32_t *info_out) {
    if (!ctx || !scene || !xyz_out) return GR_ERR_INVALID;
    gr_status st = validate_scene(scene);
    if (st != GR_OK) return st;
    uint64_t n_pixels = (uint64_t)scene->width * scene->height;
    if ((uint64_t)pixel_begin + pixel_count > n_pixels) return GR_ERR_INVALID;
    if (spp > GR_MAX_SAMPLES_PER_CALL) return GR_ERR_INVALID;
    gr_scene_dev *s = nullptr;
    st = gr_scene_upload(ctx, scene, &s);
    if (st != GR_OK) return st;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (hipSetDevice(ctx->device) != hipSuccess) st = GR_ERR_DEVICE;
        else st = run_trace(ctx, s, pixel_begin, pixel_count, sample_offset, spp, nullptr, xyz_out, info_out);
    }
    gr_scene_free(ctx, s);
    return st;
}

}  // extern "C"
