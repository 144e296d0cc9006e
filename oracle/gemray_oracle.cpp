/*
 * gemray_oracle.cpp -- CPU restatement of gemray's spectral transport estimator.
 *
 * TEST INFRASTRUCTURE ONLY (see gemray_oracle.h).  It is the checker for the HIP
 * product path and the "port" CPU baseline; nothing in the product links it.
 *
 * It follows the reference Rust source operation for operation, because the
 * reference's own golden vectors are f32 bit patterns:
 *   - glam 0.33 scalar Vec3 arithmetic: left-associative unfused dot, the glam cross
 *     order, normalize = v * (1 / sqrt(dot(v, v))), normalize_or_zero guarding a
 *     non-finite or zero reciprocal;
 *   - glam's SSE2 Mat4 * Vec4: ((c0*x + c1*y) + c2*z) + c3*w, zero entries included;
 *   - f32::mul_add -> fmaf; nothing else is fused (build with -ffp-contract=off);
 *   - f32::midpoint -> ((double)a + (double)b) / 2 rounded to f32 (Rust std on x86_64);
 *   - powi(n) -> LLVM's constant-exponent expansion: multiply the squares x^(2^i) of the
 *     set bits of n in increasing i;
 *   - exp_f32x8 -> the Cephes polynomial of crates/gemray/src/simd/exp_poly.rs:23-40;
 *   - every other transcendental -> the platform libm (sinf, cosf, tanf, acosf, atanf,
 *     atan2f, expf, floorf, fmodf).
 *
 * Scope: polished and frosted facets, edge-rounded shading normals, isotropic, uniaxial
 * and biaxial materials, Henyey-Greenstein volume scattering.
 */
#include "gemray_oracle.h"

#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#if defined(__AVX2__)
#include <immintrin.h>
#endif
#include <thread>
#include <vector>

namespace {

// Transcendentals.  The reference calls the platform's f32 libm; GEMRAY_ORACLE_CR=1
// evaluates each function in f64 and rounds once, which is the correctly rounded f32
// result except in vanishingly rare double-rounding cases.
#ifndef GEMRAY_ORACLE_CR
#define GEMRAY_ORACLE_CR 0
#endif
#if GEMRAY_ORACLE_CR
inline float tf_sin(float x) { return (float)std::sin((double)x); }
inline float tf_cos(float x) { return (float)std::cos((double)x); }
inline float tf_tan(float x) { return (float)std::tan((double)x); }
inline float tf_acos(float x) { return (float)std::acos((double)x); }
inline float tf_atan(float x) { return (float)std::atan((double)x); }
inline float tf_atan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline float tf_exp(float x) { return (float)std::exp((double)x); }
inline float tf_log(float x) { return (float)std::log((double)x); }
#else
inline float tf_sin(float x) { return std::sin(x); }
inline float tf_cos(float x) { return std::cos(x); }
inline float tf_tan(float x) { return std::tan(x); }
inline float tf_acos(float x) { return std::acos(x); }
inline float tf_atan(float x) { return std::atan(x); }
inline float tf_atan2(float y, float x) { return std::atan2(y, x); }
inline float tf_exp(float x) { return std::exp(x); }
inline float tf_log(float x) { return std::log(x); }
#endif

constexpr int NUM_CHANNELS = 8;
constexpr float PI_F = 3.14159265358979323846f;
constexpr float FLT_EPSILON_ = 1.1920928955078125e-7f;  // f32::EPSILON

// ---------------------------------------------------------------------------------
// Scalar helpers with Rust f32 semantics.
// ---------------------------------------------------------------------------------

// f32::clamp (core::f32::clamp): NaN passes through.
inline float clampf(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
// f32::max / f32::min (IEEE maxNum/minNum).
inline float maxf(float a, float b) { return std::fmax(a, b); }
inline float minf(float a, float b) { return std::fmin(a, b); }
inline float fma_(float a, float b, float c) { return std::fmaf(a, b, c); }
// f32::midpoint on x86_64: computed in f64.
inline float midpoint(float a, float b) { return (float)(((double)a + (double)b) / 2.0); }
// f32::rem_euclid.
inline float rem_euclid(float x, float rhs) {
    float r = std::fmod(x, rhs);
    return r < 0.0f ? r + std::fabs(rhs) : r;
}
// powi(x, n), n > 0 constant.  The reference's golden vectors pin this to the correctly
// rounded power (they differ from a repeated-squaring product by up to 3 ULP at n = 28
// and agree with x^n evaluated in f64 and rounded once), so it is computed that way.
inline float powi(float x, unsigned n) { return (float)std::pow((double)x, (double)n); }
// f32::to_radians.
inline float to_radians(float deg) {
    const float rads_per_deg = PI_F / 180.0f;
    return deg * rads_per_deg;
}

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline float dot(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline float length_squared(V3 a) { return dot(a, a); }
inline float length_recip(V3 a) { return 1.0f / std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return mul(a, length_recip(a)); }
inline V3 normalize_or_zero(V3 a) {
    float rcp = length_recip(a);
    if (std::isfinite(rcp) && rcp > 0.0f) return mul(a, rcp);
    return v3(0.0f, 0.0f, 0.0f);
}
inline V3 load3(const float *p) { return v3(p[0], p[1], p[2]); }

// ---------------------------------------------------------------------------------
// Stokes / Mueller (optics/polarization.rs).
// ---------------------------------------------------------------------------------

struct Stokes {
    float i, q, u, v;
};
// Mat4 stored column-major as the from_cols_array argument.
struct M4 {
    float c[16];
};
inline Stokes apply(const Stokes &s, const M4 &m) {
    // glam SSE2 Mat4::mul_vec4: ((col0*x + col1*y) + col2*z) + col3*w per lane.
    float r[4];
    for (int row = 0; row < 4; ++row) {
        float acc = m.c[0 + row] * s.i;
        acc = acc + m.c[4 + row] * s.q;
        acc = acc + m.c[8 + row] * s.u;
        acc = acc + m.c[12 + row] * s.v;
        r[row] = acc;
    }
    return Stokes{r[0], r[1], r[2], r[3]};
}
inline Stokes scale(const Stokes &s, float k) { return Stokes{s.i * k, s.q * k, s.u * k, s.v * k}; }
inline float intensity(const Stokes &s) { return maxf(s.i, 0.0f); }
// polarization.rs:38-48.
inline float degree_of_polarization(const Stokes &s) {
    if (s.i <= 1e-7f) return 0.0f;
    return clampf(std::sqrt(fma_(s.v, s.v, fma_(s.u, s.u, s.q * s.q))) / s.i, 0.0f, 1.0f);
}
// polarization.rs:95-101.
inline M4 frame_rotation(float psi) {
    float c2 = tf_cos(2.0f * psi);
    float s2 = tf_sin(2.0f * psi);
    return M4{{1, 0, 0, 0, 0, c2, -s2, 0, 0, s2, c2, 0, 0, 0, 0, 1}};
}
// polarization.rs:105-115.
inline M4 fresnel_reflection(float r_s, float r_p) {
    float rs2 = r_s * r_s;
    float rp2 = r_p * r_p;
    float a = midpoint(rs2, rp2);
    float b = 0.5f * (rs2 - rp2);
    float c = r_s * r_p;
    return M4{{a, b, 0, 0, b, a, 0, 0, 0, 0, c, 0, 0, 0, 0, c}};
}
// polarization.rs:119-137.
inline M4 fresnel_transmission(float n1, float n2, float cos_i, float cos_t, float t_s,
                               float t_p) {
    float factor = (n2 * cos_t) / maxf(n1 * cos_i, 1e-6f);
    float ts2 = t_s * t_s * factor;
    float tp2 = t_p * t_p * factor;
    float a = midpoint(ts2, tp2);
    float b = 0.5f * (ts2 - tp2);
    float c = t_s * t_p * factor;
    return M4{{a, b, 0, 0, b, a, 0, 0, 0, 0, c, 0, 0, 0, 0, c}};
}
// polarization.rs:141-147.
inline M4 tir_retardation(float delta) {
    float cd = tf_cos(delta);
    float sd = tf_sin(delta);
    return M4{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, cd, -sd, 0, 0, sd, cd}};
}
// polarization.rs:197-200.
inline V3 arbitrary_perpendicular(V3 n) {
    V3 a = std::fabs(n.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
    return normalize_or_zero(sub(a, mul(n, dot(n, a))));
}
// polarization.rs:172-192.
inline V3 electric_field_direction(const Stokes &st, V3 s_axis, V3 prop) {
    V3 k = normalize_or_zero(prop);
    V3 s_raw = sub(s_axis, mul(k, dot(k, s_axis)));
    V3 s_hat = length_squared(s_raw) > 1e-8f ? normalize(s_raw) : arbitrary_perpendicular(k);
    V3 p_hat = cross(k, s_hat);
    float psi = 0.5f * tf_atan2(st.u, st.q);
    V3 e = add(mul(s_hat, tf_cos(psi)), mul(p_hat, tf_sin(psi)));
    return length_squared(e) > 1e-8f ? normalize(e) : s_hat;
}

// ---------------------------------------------------------------------------------
// exp_f32x8 (simd/exp_poly.rs:8-40) and the CIE 1931 fit (color/cie1931.rs).
// ---------------------------------------------------------------------------------

inline float exp_lane(float x) {
    const float LOG2E = 1.44269504088896340736f;
    uint32_t c1bits = 0x3F318000u;
    float C1;
    std::memcpy(&C1, &c1bits, 4);
    const float C2 = -2.1219444e-4f;
    const float P0 = 1.9875691e-4f, P1 = 1.3981999e-3f, P2 = 8.333452e-3f,
                P3 = 4.1665796e-2f, P4 = 1.6666665e-1f, P5 = 5.0000003e-1f;
    x = clampf(x, -87.33654f, 88.02875f);
    float n = std::floor(fma_(x, LOG2E, 0.5f));
    x = fma_(n, -C1, x);
    x = fma_(n, -C2, x);
    float z = x * x;
    float p = P0;
    p = fma_(p, x, P1);
    p = fma_(p, x, P2);
    p = fma_(p, x, P3);
    p = fma_(p, x, P4);
    p = fma_(p, x, P5);
    p = fma_(p, z, x) + 1.0f;
    uint32_t bits = (uint32_t)(((int32_t)n + 127) << 23);
    float pow2n;
    std::memcpy(&pow2n, &bits, 4);
    return p * pow2n;
}

inline float cmf_g(float x, float mu, float lo, float hi) {
    float sigma = x < mu ? lo : hi;
    float t = (x - mu) / sigma;
    return tf_exp(-0.5f * t * t);
}

void cmf_scalar(float l, float out[3]) {
    float x = fma_(0.065f, -cmf_g(l, 501.1f, 20.4f, 26.2f),
                   fma_(1.056f, cmf_g(l, 599.8f, 37.9f, 31.0f), 0.362f * cmf_g(l, 442.0f, 16.0f, 26.7f)));
    float y = fma_(0.286f, cmf_g(l, 530.9f, 16.3f, 31.1f), 0.821f * cmf_g(l, 568.8f, 46.9f, 40.5f));
    float z = fma_(0.681f, cmf_g(l, 459.0f, 26.0f, 13.8f), 1.217f * cmf_g(l, 437.0f, 11.8f, 36.0f));
    out[0] = maxf(x, 0.0f);
    out[1] = maxf(y, 0.0f);
    out[2] = maxf(z, 0.0f);
}

inline void g_x8(const float l[8], float mu, float lo, float hi, float out[8]) {
    for (int k = 0; k < 8; ++k) {
        float sigma = l[k] < mu ? lo : hi;
        float t = (l[k] - mu) / sigma;
        out[k] = exp_lane(-0.5f * t * t);
    }
}

void cmf_x8(const float l[8], float out[8][3]) {
    float x0[8], x1[8], x2[8], y0[8], y1[8], z0[8], z1[8];
    g_x8(l, 501.1f, 20.4f, 26.2f, x0);
    g_x8(l, 599.8f, 37.9f, 31.0f, x1);
    g_x8(l, 442.0f, 16.0f, 26.7f, x2);
    g_x8(l, 530.9f, 16.3f, 31.1f, y0);
    g_x8(l, 568.8f, 46.9f, 40.5f, y1);
    g_x8(l, 459.0f, 26.0f, 13.8f, z0);
    g_x8(l, 437.0f, 11.8f, 36.0f, z1);
    for (int k = 0; k < 8; ++k) {
        float x = fma_(0.065f, -x0[k], fma_(1.056f, x1[k], 0.362f * x2[k]));
        float y = fma_(0.286f, y0[k], 0.821f * y1[k]);
        float z = fma_(0.681f, z0[k], 1.217f * z1[k]);
        out[k][0] = maxf(x, 0.0f);
        out[k][1] = maxf(y, 0.0f);
        out[k][2] = maxf(z, 0.0f);
    }
}

// ---------------------------------------------------------------------------------
// Sampling (optics/raytracer/sampling.rs).
// ---------------------------------------------------------------------------------

constexpr uint32_t FRESNEL_BRANCH_STREAM = 0x9E3779B1u;
constexpr uint32_t FROSTED_DIR_U_STREAM = 0x27D4EB2Fu;  // sampling.rs:28-29
constexpr uint32_t FROSTED_DIR_V_STREAM = 0x165667B1u;
constexpr uint32_t DISTANCE_SAMPLE_STREAM = 0xA24BAED4u;  // sampling.rs:30-41
constexpr uint32_t PHASE_DIR_U_STREAM = 0x9FB21C65u;
constexpr uint32_t PHASE_DIR_V_STREAM = 0x1CE4E5B9u;
constexpr uint32_t RUSSIAN_ROULETTE_STREAM = 0x517CC1B7u;
constexpr uint32_t BIREFRINGENT_SPLIT_STREAM = 0x2545F491u;
constexpr uint32_t MODE_COUPLING_STREAM = 0xCC9E2D51u;
constexpr uint32_t PIXEL_JITTER_X_ROTATION_STREAM = 0xA511E9B3u;
constexpr uint32_t PIXEL_JITTER_Y_ROTATION_STREAM = 0x63D81B23u;
constexpr uint32_t HERO_WAVELENGTH_ROTATION_STREAM = 0x1B873593u;

inline uint32_t hash_u32(uint32_t x) {
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}
inline float unit_draw(uint32_t h) { return (float)h / 4294967295.0f; }
inline float stream_draw(uint32_t seed, uint32_t bounce, uint32_t stream) {
    return unit_draw(hash_u32(seed ^ hash_u32(bounce ^ stream)));
}
inline uint32_t reverse_bits(uint32_t v) {
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
    v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
    return (v >> 16) | (v << 16);
}
inline float low_discrepancy_base2(uint32_t n) { return (float)reverse_bits(n) / 4294967296.0f; }
inline float radical_inverse_base(uint32_t n, uint32_t base) {
    uint32_t remaining = n;
    float val = 0.0f;
    float inv_base = 1.0f / (float)base;
    while (remaining > 0) {
        uint32_t digit = remaining % base;
        val = fma_((float)digit, inv_base, val);
        inv_base /= (float)base;
        remaining /= base;
    }
    return val;
}
inline float cranley_patterson_rotate(float x, float offset) {
    float sum = x + offset;
    return sum - std::floor(sum);
}

struct Draws {
    uint32_t seed;
    float jx, jy, hero;
};
inline void pixel_rotations(uint32_t pixel, float rot[3]) {
    rot[0] = low_discrepancy_base2(hash_u32(pixel ^ PIXEL_JITTER_X_ROTATION_STREAM));
    rot[1] = low_discrepancy_base2(hash_u32(pixel ^ PIXEL_JITTER_Y_ROTATION_STREAM));
    rot[2] = low_discrepancy_base2(hash_u32(pixel ^ HERO_WAVELENGTH_ROTATION_STREAM));
}
inline Draws sample_draws(uint32_t pixel, uint32_t sample_num, const float rot[3]) {
    Draws d;
    d.seed = hash_u32((pixel * 0x9e3779b9u) ^ (sample_num * 0x85ebca6bu));
    d.jx = cranley_patterson_rotate(low_discrepancy_base2(sample_num), rot[0]) - 0.5f;
    d.jy = cranley_patterson_rotate(radical_inverse_base(sample_num, 3), rot[1]) - 0.5f;
    d.hero = cranley_patterson_rotate(radical_inverse_base(sample_num, 5), rot[2]);
    return d;
}

// transport.rs:45-53.
inline void wrapped_hero_wavelengths(float hero_rand, float out[8]) {
    const float span = 780.0f - 380.0f;
    const float channel_width = span / 8.0f;
    float lambda_hero = fma_(hero_rand, span, 380.0f);
    for (int k = 0; k < 8; ++k) {
        float offset = fma_((float)k, channel_width, lambda_hero - 380.0f);
        out[k] = 380.0f + rem_euclid(offset, span);
    }
}

// ---------------------------------------------------------------------------------
// Material physics (optics/dispersion.rs, optics/absorption.rs, birefringence.rs).
// ---------------------------------------------------------------------------------

float dispersion_evaluate(const or_material *m, float lambda_nm) {
    float lambda_um = lambda_nm * 1e-3f;
    float l2 = lambda_um * lambda_um;
    if (m->disp_kind == OR_DISP_SELLMEIER1) {
        float n2 = 1.0f + (m->disp[0] * l2) / (l2 - m->disp[1]);
        return std::sqrt(maxf(n2, 1.0f));
    }
    if (m->disp_kind == OR_DISP_SELLMEIER3) {
        float n2 = 1.0f;
        n2 += (m->disp[0] * l2) / (l2 - m->disp[3]);
        n2 += (m->disp[1] * l2) / (l2 - m->disp[4]);
        n2 += (m->disp[2] * l2) / (l2 - m->disp[5]);
        return std::sqrt(maxf(n2, 1.0f));
    }
    float l4 = l2 * l2;
    return m->disp[0] + (m->disp[1] / l2) + (m->disp[2] / l4);
}

// optics/absorption.rs:68-71 summed as raytracer/absorption.rs:28-30 (fold from -0.0).
float spectral_absorption(const float (*bands)[3], int n, float lambda_nm) {
    float sum = -0.0f;
    for (int i = 0; i < n; ++i) {
        float t = (lambda_nm - bands[i][0]) / bands[i][1];
        sum = sum + bands[i][2] * tf_exp(-0.5f * t * t);
    }
    return sum;
}

// birefringence.rs:27-38.
float effective_extraordinary_index(float n_o, float n_e, float theta) {
    if (std::fabs(n_o - n_e) < 1e-5f) return n_o;
    float sin_t = tf_sin(theta);
    float cos_t = tf_cos(theta);
    float denom_sq = fma_(n_e * cos_t, n_e * cos_t, powi(n_o * sin_t, 2));
    if (denom_sq <= 1e-8f) return n_o;
    return (n_o * n_e) / std::sqrt(denom_sq);
}
// birefringence.rs:42-55.
float walk_off_angle(float n_o, float n_e, float theta) {
    if (std::fabs(n_o - n_e) < 1e-5f) return 0.0f;
    float n_o2 = n_o * n_o;
    float n_e2 = n_e * n_e;
    float sin_t = tf_sin(theta);
    float cos_t = tf_cos(theta);
    float tan_rho = ((n_o2 - n_e2) * sin_t * cos_t) /
                    maxf(fma_(n_e2 * cos_t, cos_t, n_o2 * sin_t * sin_t), 1e-6f);
    return tf_atan(tan_rho);
}
// birefringence.rs:74-91.
V3 extraordinary_poynting_dir(V3 w, V3 c, float n_o, float n_e) {
    float cos_theta = clampf(dot(w, c), -1.0f, 1.0f);
    float theta = tf_acos(std::fabs(cos_theta));
    float delta = walk_off_angle(n_o, n_e, theta);
    if (std::fabs(delta) < 1e-5f) return w;
    V3 c_proj = normalize_or_zero(sub(c, mul(w, cos_theta)));
    if (length_squared(c_proj) < 1e-6f) return w;
    float sign = cos_theta >= 0.0f ? 1.0f : -1.0f;
    return normalize(sub(mul(w, tf_cos(delta)), mul(mul(c_proj, sign), tf_sin(delta))));
}
// birefringence.rs:126-131.
inline void stable_orthonormal_basis(V3 n, V3 *t_out, V3 *b_out) {
    V3 a = std::fabs(n.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
    V3 t = normalize_or_zero(sub(a, mul(n, dot(n, a))));
    *t_out = t;
    *b_out = cross(n, t);
}
// birefringence.rs:101-118.
V3 ordinary_eigen_polarization(V3 w, V3 c) {
    V3 cr = cross(w, c);
    if (length_squared(cr) > 1e-8f) return normalize(cr);
    V3 t, b;
    stable_orthonormal_basis(normalize_or_zero(w), &t, &b);
    return t;
}
V3 extraordinary_eigen_polarization(V3 w, V3 c) {
    V3 o = ordinary_eigen_polarization(w, c);
    return normalize_or_zero(cross(w, o));
}

// AbsorptionTensor3 (birefringence.rs:185-255).
struct Tensor3 {
    V3 alpha;
    V3 a1, a2, a3;  // columns of `axes`
};
inline Tensor3 tensor_uniaxial(float alpha_o, float alpha_e, V3 c_axis) {
    V3 c = normalize_or_zero(c_axis);
    Tensor3 t;
    stable_orthonormal_basis(c, &t.a1, &t.a2);
    t.a3 = c;
    t.alpha = v3(alpha_o, alpha_o, alpha_e);
    return t;
}
// AbsorptionTensor3::biaxial (birefringence.rs:238-245): (alpha, beta, gamma) on the
// same frame as the uniaxial tensor and the biaxial indicatrix.
inline Tensor3 tensor_biaxial(float alpha, float beta, float gamma, V3 gamma_axis) {
    V3 g = normalize_or_zero(gamma_axis);
    Tensor3 t;
    stable_orthonormal_basis(g, &t.a1, &t.a2);
    t.a3 = g;
    t.alpha = v3(alpha, beta, gamma);
    return t;
}
inline float quadratic_form(const Tensor3 &t, V3 e) {
    // axes.transpose() * e: glam Mat3 * Vec3 (scalar) = x_axis*e.x + y_axis*e.y + z_axis*e.z
    // with the transposed columns; lane j is a_j.x*e.x + a_j.y*e.y + a_j.z*e.z.
    V3 local = v3(dot(t.a1, e), dot(t.a2, e), dot(t.a3, e));
    return dot(t.alpha, mulv(local, local));
}
// birefringence.rs:272-286.
inline float effective_pleochroic_alpha(const Tensor3 &t, V3 e_hat, V3 ea, V3 eb, float dop) {
    float p = clampf(dop, 0.0f, 1.0f);
    float alpha_pol = quadratic_form(t, e_hat);
    float alpha_unpol = midpoint(quadratic_form(t, ea), quadratic_form(t, eb));
    return fma_(p, alpha_pol - alpha_unpol, alpha_unpol);
}

// ---------------------------------------------------------------------------------
// Biaxial indicatrix (optics/birefringence.rs:139-170, 349-870).
// ---------------------------------------------------------------------------------

// birefringence.rs:139-150.
inline V3 canonicalize_eigenvector_sign(V3 v) {
    float ax = std::fabs(v.x), ay = std::fabs(v.y), az = std::fabs(v.z);
    float largest = (ax >= ay && ax >= az) ? v.x : (ay >= az ? v.y : v.z);
    return largest < 0.0f ? neg(v) : v;
}
// birefringence.rs:157-170.
inline V3 cross_fma(V3 a, V3 b) {
    return v3(fma_(a.y, b.z, -(a.z * b.y)), fma_(a.z, b.x, -(a.x * b.z)), fma_(a.x, b.y, -(a.y * b.x)));
}
inline float dot_fma(V3 a, V3 b) { return fma_(a.x, b.x, fma_(a.y, b.y, a.z * b.z)); }

struct Biaxial {
    float n_alpha, n_beta, n_gamma;
    V3 ax, ay, az;  // columns of `axes`: the alpha, beta and gamma principal directions
};
// axes.transpose() * v (glam Mat3 * Vec3: lane j is dot(column j, v)).
inline V3 to_local(const Biaxial &b, V3 v) { return v3(dot(b.ax, v), dot(b.ay, v), dot(b.az, v)); }
// axes * v: ((ax * v.x) + (ay * v.y)) + (az * v.z) per lane.
inline V3 to_world(const Biaxial &b, V3 v) { return add(add(mul(b.ax, v.x), mul(b.ay, v.y)), mul(b.az, v.z)); }

// BiaxialIndicatrix::from_gamma_axis (birefringence.rs:386-391).
Biaxial biaxial_from_gamma_axis(float n_alpha, float n_beta, float n_gamma, V3 gamma_axis) {
    Biaxial b;
    b.n_alpha = n_alpha;
    b.n_beta = n_beta;
    b.n_gamma = n_gamma;
    b.az = normalize_or_zero(gamma_axis);
    stable_orthonormal_basis(b.az, &b.ax, &b.ay);
    return b;
}
// GemMaterial::biaxial_indicatrix (materials.rs:1508-1519); false for None.
bool material_biaxial_indicatrix(const or_material *m, float lambda_nm, Biaxial *out) {
    if (std::isnan(m->biaxial_delta_beta_alpha)) return false;
    float n_beta = dispersion_evaluate(m, lambda_nm);
    float n_alpha = n_beta - m->biaxial_delta_beta_alpha;
    float n_gamma = n_alpha + m->birefringence_delta;
    *out = biaxial_from_gamma_axis(n_alpha, n_beta, n_gamma, load3(m->c_axis));
    return true;
}
// birefringence.rs:394-400.
inline void b_coeffs(const Biaxial &b, float out[3]) {
    out[0] = 1.0f / (b.n_alpha * b.n_alpha);
    out[1] = 1.0f / (b.n_beta * b.n_beta);
    out[2] = 1.0f / (b.n_gamma * b.n_gamma);
}
// birefringence.rs:519-522.
inline bool indices_are_degenerate(float a, float b) {
    float scale = maxf(maxf(std::fabs(a), std::fabs(b)), 1.0f);
    return std::fabs(a - b) <= std::sqrt(FLT_EPSILON_) * scale;
}
// birefringence.rs:495-499.
inline void uniaxial_wave_indices(V3 k_hat, V3 c_axis, float n_o, float n_e, float *n_slow, float *n_fast) {
    float theta = tf_acos(std::fabs(clampf(dot(k_hat, c_axis), -1.0f, 1.0f)));
    float n_eff = effective_extraordinary_index(n_o, n_e, theta);
    *n_slow = maxf(n_o, n_eff);
    *n_fast = minf(n_o, n_eff);
}
// BiaxialIndicatrix::wave_indices (birefringence.rs:435-487): (n_slow, n_fast).
void wave_indices(const Biaxial &b, V3 wave_normal, float *n_slow, float *n_fast) {
    V3 k = normalize_or_zero(wave_normal);
    bool ab = indices_are_degenerate(b.n_alpha, b.n_beta);
    bool bg = indices_are_degenerate(b.n_beta, b.n_gamma);
    if (ab && bg) {
        float n = (b.n_alpha + b.n_beta + b.n_gamma) / 3.0f;
        *n_slow = n;
        *n_fast = n;
        return;
    }
    if (ab) {
        uniaxial_wave_indices(k, b.az, midpoint(b.n_alpha, b.n_beta), b.n_gamma, n_slow, n_fast);
        return;
    }
    if (bg) {
        uniaxial_wave_indices(k, b.ax, midpoint(b.n_beta, b.n_gamma), b.n_alpha, n_slow, n_fast);
        return;
    }
    V3 l = to_local(b, k);
    float a2 = l.x * l.x, b2 = l.y * l.y, g2 = l.z * l.z;
    float bc[3];
    b_coeffs(b, bc);
    float big_b = fma_(a2, bc[1] + bc[2], fma_(b2, bc[0] + bc[2], g2 * (bc[0] + bc[1])));
    float big_c = fma_(a2, bc[1] * bc[2], fma_(b2, bc[0] * bc[2], g2 * bc[0] * bc[1]));
    float disc = std::sqrt(maxf(fma_(big_b, big_b, -4.0f * big_c), 0.0f));
    float x_lo = 0.5f * (big_b - disc);
    float x_hi = midpoint(big_b, disc);
    *n_slow = 1.0f / std::sqrt(maxf(x_lo, 1e-12f));
    *n_fast = 1.0f / std::sqrt(maxf(x_hi, 1e-12f));
}
// birefringence.rs:728-756.
float precise_root_near(V3 l, const float bc[3], float x) {
    float a = l.x * l.x, bb = l.y * l.y, cc = l.z * l.z;
    float big_b = fma_(a, bc[1] + bc[2], fma_(bb, bc[0] + bc[2], cc * (bc[0] + bc[1])));
    float xdiff = bc[0] - bc[1];
    float ydiff = bc[2] - bc[0];
    float a_plus_c = a + cc;
    float a_plus_bb = a + bb;
    float two_a_minus_bc = 2.0f * fma_(bb, -cc, a);
    float disc_sq = fma_(a_plus_c * a_plus_c, xdiff * xdiff,
                         fma_(a_plus_bb * a_plus_bb, ydiff * ydiff, two_a_minus_bc * xdiff * ydiff));
    float disc = std::sqrt(maxf(disc_sq, 0.0f));
    float x_lo = 0.5f * (big_b - disc);
    float x_hi = midpoint(big_b, disc);
    return std::fabs(x - x_lo) <= std::fabs(x - x_hi) ? x_lo : x_hi;
}
// birefringence.rs:659-700.
V3 eigenvector_world(const Biaxial &b, V3 l, const float bc[3], float x_in, V3 k_hat) {
    float x = precise_root_near(l, bc, x_in);
    float s = fma_(bc[0], l.x * l.x, fma_(bc[1], l.y * l.y, bc[2] * l.z * l.z));
    float m00 = fma_(l.x * l.x, fma_(2.0f, -bc[0], s), bc[0] - x);
    float m11 = fma_(l.y * l.y, fma_(2.0f, -bc[1], s), bc[1] - x);
    float m22 = fma_(l.z * l.z, fma_(2.0f, -bc[2], s), bc[2] - x);
    float m01 = l.x * l.y * (s - bc[0] - bc[1]);
    float m02 = l.x * l.z * (s - bc[0] - bc[2]);
    float m12 = l.y * l.z * (s - bc[1] - bc[2]);
    V3 row0 = v3(m00, m01, m02), row1 = v3(m01, m11, m12), row2 = v3(m02, m12, m22);
    V3 c01 = cross_fma(row0, row1);
    V3 c02 = cross_fma(row0, row2);
    V3 c12 = cross_fma(row1, row2);
    float sign02 = dot_fma(c01, c02) < 0.0f ? -1.0f : 1.0f;
    float sign12 = dot_fma(c01, c12) < 0.0f ? -1.0f : 1.0f;
    V3 v_local = add(add(c01, mul(c02, sign02)), mul(c12, sign12));
    V3 v_world = to_world(b, v_local);
    if (length_squared(v_world) > 1e-12f) return normalize(canonicalize_eigenvector_sign(v_world));
    V3 t, bb;
    stable_orthonormal_basis(k_hat, &t, &bb);
    return t;
}
// BiaxialIndicatrix::eigen_polarizations (birefringence.rs:542-553): (d_slow, d_fast).
void eigen_polarizations(const Biaxial &b, V3 wave_normal, V3 *d_slow, V3 *d_fast) {
    V3 k = normalize_or_zero(wave_normal);
    V3 l = to_local(b, k);
    float bc[3];
    b_coeffs(b, bc);
    float n_slow, n_fast;
    wave_indices(b, wave_normal, &n_slow, &n_fast);
    float x_slow = 1.0f / (n_slow * n_slow);
    float x_fast = 1.0f / (n_fast * n_fast);
    *d_slow = eigenvector_world(b, l, bc, x_slow, k);
    *d_fast = eigenvector_world(b, l, bc, x_fast, k);
}
// birefringence.rs:763-771.
V3 d_to_e_direction(const Biaxial &b, V3 d_hat) {
    V3 d = to_local(b, d_hat);
    V3 e = v3(d.x / (b.n_alpha * b.n_alpha), d.y / (b.n_beta * b.n_beta), d.z / (b.n_gamma * b.n_gamma));
    return normalize_or_zero(to_world(b, e));
}
// birefringence.rs:855-865.
V3 poynting_direction(V3 wave_normal, V3 e_field_hat) {
    V3 k_hat = normalize_or_zero(wave_normal);
    V3 e_hat = normalize_or_zero(e_field_hat);
    V3 perp = sub(k_hat, mul(e_hat, dot(e_hat, k_hat)));
    float len2 = length_squared(perp);
    if (len2 > 1e-10f) {
        float len = std::sqrt(len2);
        return v3(perp.x / len, perp.y / len, perp.z / len);
    }
    return k_hat;
}
// birefringence.rs:831-836.
V3 mode_poynting_dir(const Biaxial &b, V3 wave_normal, bool want_slow) {
    V3 d_slow, d_fast;
    eigen_polarizations(b, wave_normal, &d_slow, &d_fast);
    V3 e_hat = d_to_e_direction(b, want_slow ? d_slow : d_fast);
    return poynting_direction(wave_normal, e_hat);
}
// birefringence.rs:795-818: (n, wave_normal) of one mode at an air->crystal entry.
float resolve_entry_mode(const Biaxial &b, V3 incident, V3 normal, float cos_i, float n_seed, bool want_slow,
                         V3 *wave_out) {
    float n_guess = n_seed;
    V3 wave_dir = incident;
    for (int it = 0; it < 2; ++it) {
        float eta = 1.0f / n_guess;
        float sin2 = eta * eta * fma_(cos_i, -cos_i, 1.0f);
        if (sin2 > 1.0f) break;
        float cos_t = std::sqrt(maxf(1.0f - sin2, 0.0f));
        wave_dir = normalize(add(mul(incident, eta), mul(normal, fma_(eta, cos_i, -cos_t))));
        float ns, nf;
        wave_indices(b, wave_dir, &ns, &nf);
        n_guess = want_slow ? ns : nf;
    }
    *wave_out = wave_dir;
    return n_guess;
}

// ---------------------------------------------------------------------------------
// Environment (optics/raytracer/environment.rs, optics/studio_rig.rs,
// renderer/env_map.rs, renderer/env_map_spectrum.rs) and colour (raytracer/color.rs).
// ---------------------------------------------------------------------------------

float blackbody_spectrum(float lambda_nm, float temp_k) {
    float t_k = maxf(temp_k, 1000.0f);
    const float h_c_k = 14388000.0f;
    float exp_val = tf_exp(minf(h_c_k / (lambda_nm * t_k), 80.0f));
    float exp_560 = tf_exp(minf(h_c_k / (560.0f * t_k), 80.0f));
    float denom = maxf(exp_val - 1.0f, 1e-6f);
    float denom_560 = maxf(exp_560 - 1.0f, 1e-6f);
    float ratio = denom_560 / denom;
    return clampf(powi(560.0f / lambda_nm, 5) * ratio, 0.01f, 20.0f);
}

struct Rig {
    V3 key, fill, ring[16];
};
Rig studio_rig(float light_yaw, float light_pitch) {
    Rig r;
    float cos_lp = tf_cos(light_pitch), sin_lp = tf_sin(light_pitch);
    float cos_ly = tf_cos(light_yaw), sin_ly = tf_sin(light_yaw);
    r.key = normalize(v3(cos_lp * sin_ly, sin_lp, cos_lp * cos_ly));
    float fill_yaw = fma_(PI_F, 0.78f, light_yaw);
    float fill_pitch = clampf(light_pitch * 0.65f, 0.15f, 1.2f);
    r.fill = normalize(v3(tf_cos(fill_pitch) * tf_sin(fill_yaw), tf_sin(fill_pitch),
                          tf_cos(fill_pitch) * tf_cos(fill_yaw)));
    const float step = PI_F * 2.0f / 16.0f;
    for (int i = 0; i < 16; ++i) {
        float angle = fma_((float)i, step, light_yaw);
        r.ring[i] = normalize(v3(tf_cos(angle) * 0.75f, sin_lp * 0.8f, tf_sin(angle) * 0.75f));
    }
    return r;
}

float studio_with_rig(V3 dir, float lambda_nm, float temp_k, float spot_mult, float exposure,
                      const Rig &rig) {
    V3 d = normalize(dir);
    float spec = blackbody_spectrum(lambda_nm, temp_k);
    float bg_val = maxf(fma_(0.012f, fma_(d.y, 0.5f, 0.5f), 0.015f), 0.005f) * exposure;
    float radiance = bg_val * spec;
    float key_dot = maxf(dot(d, rig.key), 0.0f);
    if (key_dot > 0.0f) {
        float softbox = powi(key_dot, 28) * 12.0f * spot_mult * exposure;
        radiance = fma_(softbox, spec, radiance);
    }
    float fill_dot = maxf(dot(d, rig.fill), 0.0f);
    if (fill_dot > 0.0f) {
        float fill = powi(fill_dot, 18) * 4.5f * exposure;
        radiance = fma_(fill, spec, radiance);
    }
    for (int i = 0; i < 16; ++i) {
        float ring_dot = maxf(dot(d, rig.ring[i]), 0.0f);
        if (ring_dot > 0.96f) {
            float spark = (ring_dot - 0.96f) / 0.04f;
            float inten = powi(spark, 6) * 22.0f * spot_mult * exposure;
            radiance = fma_(inten, spec, radiance);
        }
    }
    return radiance;
}

inline float asymmetric_gaussian(float x, float mu, float lo, float hi) {
    float sigma = x < mu ? lo : hi;
    float t = (x - mu) / sigma;
    return tf_exp(-0.5f * t * t);
}
float rgb_to_spectral(const float rgb[3], float lambda_nm) {
    float r = maxf(rgb[0], 0.0f), g = maxf(rgb[1], 0.0f), b = maxf(rgb[2], 0.0f);
    return fma_(r, asymmetric_gaussian(lambda_nm, 615.0f, 45.0f, 65.0f),
                fma_(g, asymmetric_gaussian(lambda_nm, 545.0f, 45.0f, 45.0f),
                     b * asymmetric_gaussian(lambda_nm, 465.0f, 40.0f, 45.0f)));
}

void map_radiance_rgb(const or_env *env, V3 dir, float out[3]) {
    // direction_to_uv (env_map.rs:231-238) is evaluated even for a 1x1 map.
    V3 d = normalize_or_zero(dir);
    float theta = tf_acos(clampf(d.y, -1.0f, 1.0f));
    float phi = tf_atan2(d.x, d.z);
    float v = theta / PI_F;
    float u = rem_euclid(phi / (2.0f * PI_F), 1.0f);
    const int w = env->map_w, h = env->map_h;
    const float *px = env->map_rgb;
    if (w == 1 && h == 1) {
        out[0] = px[0];
        out[1] = px[1];
        out[2] = px[2];
        return;
    }
    float fx = fma_(rem_euclid(u, 1.0f), (float)w, -0.5f);
    float fy = fma_(clampf(v, 0.0f, 1.0f), (float)h, -0.5f);
    float x0 = std::floor(fx), y0 = std::floor(fy);
    float tx = fx - x0, ty = fy - y0;
    auto wrap_x = [w](int64_t x) -> int64_t {
        int64_t r = x % w;
        return r < 0 ? r + w : r;
    };
    auto clamp_y = [h](int64_t y) -> int64_t { return y < 0 ? 0 : (y > h - 1 ? h - 1 : y); };
    int64_t x0i = wrap_x((int64_t)x0), x1i = wrap_x((int64_t)x0 + 1);
    int64_t y0i = clamp_y((int64_t)y0), y1i = clamp_y((int64_t)y0 + 1);
    const float *p00 = px + 3 * (y0i * w + x0i);
    const float *p10 = px + 3 * (y0i * w + x1i);
    const float *p01 = px + 3 * (y1i * w + x0i);
    const float *p11 = px + 3 * (y1i * w + x1i);
    for (int c = 0; c < 3; ++c) {
        float top = fma_(p10[c], tx, p00[c] * (1.0f - tx));
        float bottom = fma_(p11[c], tx, p01[c] * (1.0f - tx));
        out[c] = fma_(bottom, ty, top * (1.0f - ty));
    }
}

// color.rs:186-251.
const float BRADFORD_XYZ_TO_LMS[3][3] = {
    {0.8951f, 0.2664f, -0.1614f}, {-0.7502f, 1.7135f, 0.0367f}, {0.0389f, -0.0685f, 1.0296f}};
const float BRADFORD_LMS_TO_XYZ[3][3] = {{0.986993f, -0.147054f, 0.159963f},
                                         {0.432305f, 0.518360f, 0.049291f},
                                         {-0.008529f, 0.040043f, 0.968487f}};
inline V3 mat3_fma(const float m[3][3], V3 v) {
    return v3(fma_(m[0][0], v.x, fma_(m[0][1], v.y, m[0][2] * v.z)),
              fma_(m[1][0], v.x, fma_(m[1][1], v.y, m[1][2] * v.z)),
              fma_(m[2][0], v.x, fma_(m[2][1], v.y, m[2][2] * v.z)));
}
inline V3 von_kries(V3 xyz, V3 lms_scale) {
    return mat3_fma(BRADFORD_LMS_TO_XYZ, mulv(mat3_fma(BRADFORD_XYZ_TO_LMS, xyz), lms_scale));
}
V3 compute_white_balance(float temp_k) {
    V3 xyz_w = v3(0, 0, 0);
    for (int step = 0; step <= 400; ++step) {
        float lambda = 380.0f + (float)step;
        float c[3];
        cmf_scalar(lambda, c);
        xyz_w = add(xyz_w, mul(load3(c), blackbody_spectrum(lambda, temp_k)));
    }
    const float D65_X = 0.3127f, D65_Y = 0.3290f;
    float target_y = maxf(xyz_w.y, 1e-6f);
    V3 xyz_target = v3((D65_X / D65_Y) * target_y, target_y,
                       ((1.0f - D65_X - D65_Y) / D65_Y) * target_y);
    V3 src = mat3_fma(BRADFORD_XYZ_TO_LMS, xyz_w);
    V3 tgt = mat3_fma(BRADFORD_XYZ_TO_LMS, xyz_target);
    return v3(src.x > 1e-6f ? tgt.x / src.x : 1.0f, src.y > 1e-6f ? tgt.y / src.y : 1.0f,
              src.z > 1e-6f ? tgt.z / src.z : 1.0f);
}

struct WbCache {
    std::mutex mu;
    std::vector<std::pair<float, V3>> entries;
    V3 get(float temp_k) {
        std::lock_guard<std::mutex> g(mu);
        for (auto &e : entries)
            if (e.first == temp_k) return e.second;
        V3 v = compute_white_balance(temp_k);
        entries.push_back({temp_k, v});
        return v;
    }
};
WbCache g_wb;

// ---------------------------------------------------------------------------------
// Geometry: slab scan (simd/slab.rs:94-152) + hit selection (intersect.rs:93-132).
// ---------------------------------------------------------------------------------

struct Hit {
    bool hit;
    float t;
    V3 normal;
    int facet;
};

// shading_normal_near_edge (optics/raytracer/intersect.rs:196-229): blends the flat normal
// toward the bisector with the nearest other facet's normal within rounding_radius of
// their shared edge (smoothstep ease); radius <= 0 returns the flat normal unchanged.
V3 shading_normal_near_edge(const float *planes, int n, V3 hit_point, int hit_facet, V3 hit_normal,
                            float rounding_radius) {
    if (rounding_radius <= 0.0f) return hit_normal;
    float nearest_dist = INFINITY;
    V3 nearest_normal = hit_normal;
    for (int i = 0; i < n; ++i) {
        if (i == hit_facet) continue;
        V3 nrm = load3(planes + 4 * i);
        float dist = -(planes[4 * i + 3] + dot(nrm, hit_point));
        if (dist < nearest_dist) {
            nearest_dist = dist;
            nearest_normal = nrm;
        }
    }
    if (nearest_dist >= rounding_radius) return hit_normal;
    float t = clampf(1.0f - nearest_dist / rounding_radius, 0.0f, 1.0f);
    float smooth_t = t * t * fma_(2.0f, -t, 3.0f);
    V3 bisector = normalize_or_zero(add(hit_normal, nearest_normal));
    return normalize_or_zero(add(mul(hit_normal, 1.0f - smooth_t), mul(bisector, smooth_t)));
}
// One plane of the scan (slab.rs:113-128): false when a near-parallel plane has the origin
// outside it (the intersection is empty whatever the other planes give).
struct SlabState {
    float t_near = -1e30f, t_far = 1e30f;
    int near_idx = -1, far_idx = -1;
    bool step(int i, float denom, float side, float t) {
        if (std::fabs(denom) > 1e-7f) {
            if (denom < 0.0f) {
                if (t > t_near) {
                    t_near = t;
                    near_idx = i;
                }
            } else if (t < t_far) {
                t_far = t;
                far_idx = i;
            }
            return true;
        }
        return !(side > 0.0f);
    }
};

Hit select_hit(const float *planes, const SlabState &st) {
    const float t_near = st.t_near, t_far = st.t_far;
    const int near_idx = st.near_idx, far_idx = st.far_idx;
    if (t_near > t_far) return Hit{false, 0, v3(0, 0, 0), 0};
    if (t_near > 1e-4f)
        return Hit{true, t_near, near_idx >= 0 ? load3(planes + 4 * near_idx) : v3(0, 0, 0),
                   near_idx >= 0 ? near_idx : 0};
    if (t_far > 1e-4f)
        return Hit{true, t_far, far_idx >= 0 ? load3(planes + 4 * far_idx) : v3(0, 0, 0),
                   far_idx >= 0 ? far_idx : 0};
    return Hit{false, 0, v3(0, 0, 0), 0};
}

// The scan one plane at a time (slab_scan_scalar, slab.rs:140-152).
Hit intersect_planes_scalar(const float *planes, int n, V3 o, V3 d) {
    SlabState st;
    for (int i = 0; i < n; ++i) {
        V3 nrm = load3(planes + 4 * i);
        float denom = dot(nrm, d);
        float side = planes[4 * i + 3] + dot(nrm, o);
        float t = (-side) / denom;
        if (!st.step(i, denom, side, t)) return Hit{false, 0, v3(0, 0, 0), 0};
    }
    return select_hit(planes, st);
}

#if defined(__AVX2__)
// The scan eight planes at a time, the way the reference's CPU tracer runs it on AVX2
// (slab_scan_avx2, simd/slab.rs:186-270): every lane forms denom, side and t with the
// scalar scan's operations in the same order, so the values are bit-identical; a block
// whose entering maximum cannot beat t_near and whose exiting minimum cannot beat t_far
// changes nothing and is skipped, any other block is replayed plane by plane in order; a
// near-parallel plane with the origin outside empties the intersection wherever it lies.
// The planes are the interleaved (n, d) records, transposed eight at a time.
Hit intersect_planes(const float *planes, int n, V3 o, V3 d) {
    const __m256 dx = _mm256_set1_ps(d.x), dy = _mm256_set1_ps(d.y), dz = _mm256_set1_ps(d.z);
    const __m256 ox = _mm256_set1_ps(o.x), oy = _mm256_set1_ps(o.y), oz = _mm256_set1_ps(o.z);
    const __m256 thr = _mm256_set1_ps(1e-7f), zero = _mm256_setzero_ps();
    const __m256 abs_mask = _mm256_castsi256_ps(_mm256_set1_epi32(0x7FFFFFFF));
    const __m256 neg_inf = _mm256_set1_ps(-INFINITY), pos_inf = _mm256_set1_ps(INFINITY);
    SlabState st;
    alignas(32) float denom_l[8], side_l[8], t_l[8];
    int i = 0;
    for (; i + 8 <= n; i += 8) {
        const float *p = planes + 4 * i;
        __m128 a0 = _mm_loadu_ps(p), a1 = _mm_loadu_ps(p + 4), a2 = _mm_loadu_ps(p + 8), a3 = _mm_loadu_ps(p + 12);
        __m128 b0 = _mm_loadu_ps(p + 16), b1 = _mm_loadu_ps(p + 20), b2 = _mm_loadu_ps(p + 24),
               b3 = _mm_loadu_ps(p + 28);
        _MM_TRANSPOSE4_PS(a0, a1, a2, a3);
        _MM_TRANSPOSE4_PS(b0, b1, b2, b3);
        const __m256 nx = _mm256_set_m128(b0, a0), ny = _mm256_set_m128(b1, a1), nz = _mm256_set_m128(b2, a2),
                     dd = _mm256_set_m128(b3, a3);
        const __m256 denom =
            _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(nx, dx), _mm256_mul_ps(ny, dy)), _mm256_mul_ps(nz, dz));
        const __m256 dot_o =
            _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(nx, ox), _mm256_mul_ps(ny, oy)), _mm256_mul_ps(nz, oz));
        const __m256 side = _mm256_add_ps(dd, dot_o);
        const __m256 t = _mm256_div_ps(_mm256_sub_ps(zero, side), denom);
        const __m256 ad = _mm256_and_ps(denom, abs_mask);
        const __m256 valid = _mm256_cmp_ps(ad, thr, _CMP_GT_OQ);
        const __m256 par_out = _mm256_and_ps(_mm256_cmp_ps(ad, thr, _CMP_LE_OQ), _mm256_cmp_ps(side, zero, _CMP_GT_OQ));
        if (_mm256_movemask_ps(par_out)) {
            _mm256_zeroupper();
            return Hit{false, 0, v3(0, 0, 0), 0};
        }
        const __m256 entering = _mm256_and_ps(valid, _mm256_cmp_ps(denom, zero, _CMP_LT_OQ));
        const __m256 exiting = _mm256_andnot_ps(entering, valid);
        const __m256 can_near = _mm256_cmp_ps(_mm256_blendv_ps(neg_inf, t, entering), _mm256_set1_ps(st.t_near),
                                              _CMP_GT_OQ);
        const __m256 can_far = _mm256_cmp_ps(_mm256_blendv_ps(pos_inf, t, exiting), _mm256_set1_ps(st.t_far),
                                             _CMP_LT_OQ);
        if (_mm256_movemask_ps(_mm256_or_ps(can_near, can_far))) {
            _mm256_store_ps(denom_l, denom);
            _mm256_store_ps(side_l, side);
            _mm256_store_ps(t_l, t);
            for (int k = 0; k < 8; ++k)
                if (!st.step(i + k, denom_l[k], side_l[k], t_l[k])) {
                    _mm256_zeroupper();
                    return Hit{false, 0, v3(0, 0, 0), 0};
                }
        }
    }
    // Clean upper register halves before the scalar tail and the tracer's libm calls: legacy
    // SSE code after dirty 256-bit state runs several times slower.
    _mm256_zeroupper();
    for (; i < n; ++i) {
        V3 nrm = load3(planes + 4 * i);
        float denom = dot(nrm, d);
        float side = planes[4 * i + 3] + dot(nrm, o);
        float t = (-side) / denom;
        if (!st.step(i, denom, side, t)) return Hit{false, 0, v3(0, 0, 0), 0};
    }
    return select_hit(planes, st);
}
#else
Hit intersect_planes(const float *planes, int n, V3 o, V3 d) { return intersect_planes_scalar(planes, n, o, d); }
#endif

// ---------------------------------------------------------------------------------
// Per-ray context (refraction.rs:58-174) and per-bounce geometry (:189-547).
// ---------------------------------------------------------------------------------

struct RayCtx {
    const or_material *m;
    float lambdas[8];
    V3 c_axis;
    bool is_anisotropic;
    bool is_biaxial;        // hero_indicatrix.is_some()
    float n_o_ch[8];
    Biaxial biaxial_ch[8];  // material.biaxial_indicatrix(lambdas[k]) when is_biaxial
    Tensor3 tensor_ch[8];
};

void build_ray_ctx(const or_material *m, const float lambdas[8], RayCtx *ctx) {
    ctx->m = m;
    for (int k = 0; k < 8; ++k) ctx->lambdas[k] = lambdas[k];
    ctx->c_axis = load3(m->c_axis);
    ctx->is_anisotropic = !m->crystal_cubic && std::fabs(m->birefringence_delta) > 1e-4f;
    for (int k = 0; k < 8; ++k) {
        // per_channel_uniaxial_indices(ctx, 0.0): the n_eff half is discarded.
        ctx->n_o_ch[k] = dispersion_evaluate(m, lambdas[k]);
    }
    // build_ray_wavelength_cache (refraction.rs:128-177): per-channel indicatrices, and a
    // trichroic tensor when the stone is biaxial and carries a beta band set.
    ctx->is_biaxial = false;
    for (int k = 0; k < 8; ++k) {
        bool bx = material_biaxial_indicatrix(m, lambdas[k], &ctx->biaxial_ch[k]);
        if (k == 0) ctx->is_biaxial = bx;
    }
    for (int k = 0; k < 8; ++k) {
        float ao = spectral_absorption(m->o_bands, m->n_o_bands, lambdas[k]);
        float ae = spectral_absorption(m->e_bands, m->n_e_bands, lambdas[k]);
        if (ctx->is_biaxial && m->has_beta_bands) {
            float ab = spectral_absorption(m->beta_bands, m->n_beta_bands, lambdas[k]);
            ctx->tensor_ch[k] = tensor_biaxial(ao, ab, ae, ctx->c_axis);
        } else {
            ctx->tensor_ch[k] = tensor_uniaxial(ao, ae, ctx->c_axis);
        }
    }
}

struct Geo {
    float cos_i, sin_i;
    float n_o_hero, n_e_hero;
    float n_o_ch[8];
    float n_biax_a_ch[8], n_biax_b_ch[8];  // biaxial fast (A) / slow (B) mode indices
    float n_biax_a_hero, n_biax_b_hero;
    float n1, n2, sin2_t;
    float n1_ch[8], n2_ch[8], sin2_t_ch[8];
};

// refraction.rs:235-273.
float theta_c_for_bounce(const RayCtx &ctx, V3 normal, V3 k_hat, float cos_i, bool inside,
                         float n_o_hero_seed) {
    if (!inside && ctx.is_anisotropic && !ctx.is_biaxial) {
        float n_e_hero_seed = n_o_hero_seed + ctx.m->birefringence_delta;
        float n_guess = n_o_hero_seed;
        float theta = 0.0f;
        for (int it = 0; it < 2; ++it) {
            float eta_guess = 1.0f / n_guess;
            float sin2_t_guess = eta_guess * eta_guess * fma_(cos_i, -cos_i, 1.0f);
            if (sin2_t_guess > 1.0f) break;
            float cos_t_guess = std::sqrt(maxf(1.0f - sin2_t_guess, 0.0f));
            V3 wave = normalize(add(mul(k_hat, eta_guess),
                                    mul(normal, fma_(eta_guess, cos_i, -cos_t_guess))));
            float cos_theta_wave = std::fabs(clampf(dot(wave, ctx.c_axis), -1.0f, 1.0f));
            theta = tf_acos(cos_theta_wave);
            n_guess = effective_extraordinary_index(n_o_hero_seed, n_e_hero_seed, theta);
        }
        return theta;
    }
    return tf_acos(std::fabs(clampf(dot(k_hat, ctx.c_axis), -1.0f, 1.0f)));
}

Geo compute_geo(const RayCtx &ctx, V3 normal, V3 k_hat, bool inside, bool is_extra) {
    Geo g;
    g.cos_i = clampf(dot(neg(k_hat), normal), 0.0f, 1.0f);
    g.sin_i = std::sqrt(maxf(fma_(g.cos_i, -g.cos_i, 1.0f), 0.0f));
    float theta_c = theta_c_for_bounce(ctx, normal, k_hat, g.cos_i, inside, ctx.n_o_ch[0]);
    float n_eff_ch[8];
    for (int k = 0; k < 8; ++k) {
        float n_o_k = ctx.n_o_ch[k];
        float n_e_k = n_o_k + ctx.m->birefringence_delta;
        n_eff_ch[k] =
            ctx.is_anisotropic ? effective_extraordinary_index(n_o_k, n_e_k, theta_c) : n_o_k;
        g.n_o_ch[k] = n_o_k;
    }
    g.n_o_hero = ctx.n_o_ch[0];
    g.n_e_hero = g.n_o_hero + ctx.m->birefringence_delta;
    // hero_biaxial_wave_dirs + per_channel_biaxial_indices (refraction.rs:345-400): the
    // hero's indicatrix resolves one wave normal per mode (inside the stone, the tracked
    // wave normal), and every channel's own indicatrix gives its mode index there.
    for (int k = 0; k < 8; ++k) g.n_biax_a_ch[k] = g.n_biax_b_ch[k] = 0.0f;
    if (ctx.is_biaxial) {
        V3 dir_a = k_hat, dir_b = k_hat;
        if (!inside) {
            resolve_entry_mode(ctx.biaxial_ch[0], k_hat, normal, g.cos_i, ctx.n_o_ch[0], false, &dir_a);
            resolve_entry_mode(ctx.biaxial_ch[0], k_hat, normal, g.cos_i, ctx.n_o_ch[0], true, &dir_b);
        }
        for (int k = 0; k < 8; ++k) {
            float ns, nf;
            wave_indices(ctx.biaxial_ch[k], dir_a, &ns, &nf);
            g.n_biax_a_ch[k] = nf;
            wave_indices(ctx.biaxial_ch[k], dir_b, &ns, &nf);
            g.n_biax_b_ch[k] = ns;
        }
    }
    g.n_biax_a_hero = g.n_biax_a_ch[0];
    g.n_biax_b_hero = g.n_biax_b_ch[0];
    const float *n_mode_a = ctx.is_biaxial ? g.n_biax_a_ch : ctx.n_o_ch;
    const float *n_mode_b = ctx.is_biaxial ? g.n_biax_b_ch : n_eff_ch;
    const float *n_medium = (ctx.is_anisotropic && inside && !is_extra) ? n_mode_a : n_mode_b;
    float n_medium_hero = n_medium[0];
    g.n1 = inside ? n_medium_hero : 1.0f;
    g.n2 = inside ? 1.0f : n_medium_hero;
    float eta = g.n1 / g.n2;
    g.sin2_t = eta * eta * fma_(g.cos_i, -g.cos_i, 1.0f);
    for (int k = 0; k < 8; ++k) {
        float n1k = inside ? n_medium[k] : 1.0f;
        float n2k = inside ? 1.0f : n_medium[k];
        float etak = n1k / n2k;
        g.n1_ch[k] = n1k;
        g.n2_ch[k] = n2k;
        g.sin2_t_ch[k] = etak * etak * fma_(g.cos_i, -g.cos_i, 1.0f);
    }
    return g;
}

inline float tir_phase_delta(float n1k, float cos_i, float sin_i) {
    float tan_half = (cos_i * std::sqrt(maxf(fma_(n1k * n1k * sin_i, sin_i, -1.0f), 0.0f))) /
                     maxf(n1k * sin_i * sin_i, 1e-6f);
    return 2.0f * tf_atan(tan_half);
}

inline V3 reflect(V3 k, V3 n) { return sub(k, mul(n, 2.0f * dot(k, n))); }

inline V3 poynting_dir_for_mode(const RayCtx &ctx, const Geo &g, V3 k, bool inside, bool is_extra,
                                uint64_t *ev) {
    if (!inside || !ctx.is_anisotropic) return k;
    if (ctx.is_biaxial) {
        // Both biaxial modes walk off (refraction.rs:574-576).
        if (ev) ev[OR_EV_WALKOFF]++;
        return mode_poynting_dir(ctx.biaxial_ch[0], k, is_extra);
    }
    if (is_extra) {
        if (ev) ev[OR_EV_WALKOFF]++;
        return extraordinary_poynting_dir(k, ctx.c_axis, g.n_o_hero, g.n_e_hero);
    }
    return k;
}

inline void fresnel_rs_rp(float n1k, float n2k, float cos_i, float cos_t, float *r_s, float *r_p) {
    *r_s = fma_(n2k, -cos_t, n1k * cos_i) / fma_(n2k, cos_t, n1k * cos_i);
    *r_p = fma_(n1k, -cos_t, n2k * cos_i) / fma_(n1k, cos_t, n2k * cos_i);
}

// frosted_orthonormal_basis and cosine_weighted_hemisphere
// (optics/raytracer/scattering.rs:339-355).
V3 cosine_weighted_hemisphere(float u1, float u2, V3 n) {
    float r = std::sqrt(u1);
    float theta = 2.0f * PI_F * u2;
    float sin_t = tf_sin(theta), cos_t = tf_cos(theta);
    V3 a = std::fabs(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    V3 t = normalize_or_zero(sub(a, mul(n, dot(n, a))));
    V3 b = cross(n, t);
    return normalize_or_zero(add(add(mul(t, r * cos_t), mul(b, r * sin_t)), mul(n, std::sqrt(maxf(1.0f - u1, 0.0f)))));
}

// sample_henyey_greenstein_direction (scattering.rs:428-447): HG-distributed cosine about
// `forward` (isotropic for |g| < 1e-3), azimuth 2 pi u2, on the frosted basis of `forward`.
V3 sample_henyey_greenstein_direction(float u1, float u2, float g, V3 forward) {
    float cos_theta;
    if (std::fabs(g) < 1e-3f) {
        cos_theta = fma_(2.0f, -u1, 1.0f);
    } else {
        float one_minus_g2 = fma_(g, -g, 1.0f);
        float denom = fma_(2.0f * g, u1, 1.0f - g);
        float sq = one_minus_g2 / denom;
        cos_theta = fma_(sq, -sq, fma_(g, g, 1.0f)) / (2.0f * g);
    }
    cos_theta = clampf(cos_theta, -1.0f, 1.0f);
    float sin_theta = std::sqrt(maxf(fma_(cos_theta, -cos_theta, 1.0f), 0.0f));
    float phi = 2.0f * PI_F * u2;
    float sin_p = tf_sin(phi), cos_p = tf_cos(phi);
    V3 a = std::fabs(forward.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    V3 t = normalize_or_zero(sub(a, mul(forward, dot(forward, a))));
    V3 b = cross(forward, t);
    return normalize_or_zero(
        add(add(mul(t, sin_theta * cos_p), mul(b, sin_theta * sin_p)), mul(forward, cos_theta)));
}

// refraction.rs:599-637.
V3 apply_tir_bounce(const Geo &g, V3 k_hat, V3 normal, Stokes *st, float *pdf) {
    for (int k = 0; k < 8; ++k) {
        float n1k = g.n1_ch[k], n2k = g.n2_ch[k];
        if (g.sin2_t_ch[k] > 1.0f) {
            st[k] = apply(st[k], tir_retardation(tir_phase_delta(n1k, g.cos_i, g.sin_i)));
        } else {
            float cos_t_k = std::sqrt(maxf(1.0f - g.sin2_t_ch[k], 0.0f));
            float r_s, r_p;
            fresnel_rs_rp(n1k, n2k, g.cos_i, cos_t_k, &r_s, &r_p);
            st[k] = apply(st[k], fresnel_reflection(r_s, r_p));
            float r_unpol_k = clampf(0.5f * fma_(r_p, r_p, r_s * r_s), 1e-4f, 1.0f - 1e-4f);
            pdf[k] *= r_unpol_k;
        }
    }
    return reflect(k_hat, normal);
}

// refraction.rs:650-690.
V3 apply_partial_reflect_bounce(const Geo &g, float r_unpol, V3 k_hat, V3 normal, Stokes *st,
                                float *pdf) {
    for (int k = 0; k < 8; ++k) {
        float n1k = g.n1_ch[k], n2k = g.n2_ch[k];
        M4 m;
        if (g.sin2_t_ch[k] > 1.0f) {
            m = tir_retardation(tir_phase_delta(n1k, g.cos_i, g.sin_i));
        } else {
            float cos_t_k = std::sqrt(maxf(1.0f - g.sin2_t_ch[k], 0.0f));
            float r_s, r_p;
            fresnel_rs_rp(n1k, n2k, g.cos_i, cos_t_k, &r_s, &r_p);
            float r_unpol_k = clampf(0.5f * fma_(r_p, r_p, r_s * r_s), 1e-4f, 1.0f - 1e-4f);
            pdf[k] *= r_unpol_k;
            m = fresnel_reflection(r_s, r_p);
        }
        st[k] = scale(apply(st[k], m), 1.0f / r_unpol);
    }
    return reflect(k_hat, normal);
}

struct RefractOut {
    V3 new_k, new_s;
    bool has_update;
    bool update;
};

// refraction.rs:748-1011.
RefractOut apply_refract_bounce(const RayCtx &ctx, const Geo &g, float r_unpol, V3 k_hat,
                                V3 normal, bool inside, bool is_extra, uint32_t seed,
                                uint32_t bounce, Stokes *st, float *pdf, uint64_t *ev) {
    const float DIRECTION_MATCH_COS_TOL = 1.0f - 1e-6f;
    V3 c_axis = ctx.c_axis;
    bool entering_aniso = !inside && ctx.is_anisotropic;
    bool use_extra = entering_aniso ? (stream_draw(seed, bounce, BIREFRINGENT_SPLIT_STREAM) < 0.5f)
                                    : is_extra;
    if (ev && entering_aniso) ev[OR_EV_ENTRY_ANISO]++;
    const bool entering_biaxial = entering_aniso && ctx.is_biaxial;
    float n2_hero_dir;
    if (entering_biaxial)
        n2_hero_dir = use_extra ? g.n_biax_b_hero : g.n_biax_a_hero;
    else
        n2_hero_dir = (entering_aniso && !use_extra) ? g.n_o_hero : g.n2;
    float eta_dir = g.n1 / n2_hero_dir;
    float sin2_t_dir = minf(eta_dir * eta_dir * fma_(g.cos_i, -g.cos_i, 1.0f), 1.0f);
    float cos_t_dir = std::sqrt(maxf(1.0f - sin2_t_dir, 0.0f));
    V3 refr_wave_dir =
        normalize(add(mul(k_hat, eta_dir), mul(normal, fma_(eta_dir, g.cos_i, -cos_t_dir))));
    V3 s = refr_wave_dir;
    if (entering_biaxial) {
        // Both biaxial modes walk off at entry (refraction.rs:803-818).
        if (ev) ev[OR_EV_WALKOFF]++;
        s = mode_poynting_dir(ctx.biaxial_ch[0], refr_wave_dir, use_extra);
    } else if (entering_aniso && use_extra) {
        if (ev) ev[OR_EV_WALKOFF]++;
        s = extraordinary_poynting_dir(refr_wave_dir, c_axis, g.n_o_hero, g.n_e_hero);
    }
    V3 final_refr_dir = s;

    for (int k = 0; k < 8; ++k) {
        float n1k = g.n1_ch[k];
        float n2k = (entering_aniso && !use_extra) ? (ctx.is_biaxial ? g.n_biax_a_ch[k] : g.n_o_ch[k])
                                                   : g.n2_ch[k];
        float q = n1k / n2k;
        float sin2_t_k = powi(q, 2) * fma_(g.cos_i, -g.cos_i, 1.0f);
        if (sin2_t_k > 1.0f) {
            st[k] = scale(st[k], 0.0f);
            pdf[k] = 0.0f;
            continue;
        }
        float cos_t_k = std::sqrt(maxf(1.0f - sin2_t_k, 0.0f));
        float eta_dir_k = n1k / n2k;
        V3 refr_k = normalize(
            add(mul(k_hat, eta_dir_k), mul(normal, fma_(eta_dir_k, g.cos_i, -cos_t_k))));
        V3 final_k = refr_k;
        if (entering_biaxial) {
            final_k = mode_poynting_dir(ctx.biaxial_ch[k], refr_k, use_extra);
        } else if (entering_aniso && use_extra) {
            float n_e_k = g.n_o_ch[k] + ctx.m->birefringence_delta;
            final_k = extraordinary_poynting_dir(refr_k, c_axis, g.n_o_ch[k], n_e_k);
        }
        bool matches = dot(final_k, final_refr_dir) >= DIRECTION_MATCH_COS_TOL;
        if (matches) {
            float t_s = (2.0f * n1k * g.cos_i) / fma_(n2k, cos_t_k, n1k * g.cos_i);
            float t_p = (2.0f * n1k * g.cos_i) / fma_(n1k, cos_t_k, n2k * g.cos_i);
            M4 tm = fresnel_transmission(n1k, n2k, g.cos_i, cos_t_k, t_s, t_p);
            st[k] = scale(apply(st[k], tm), 1.0f / (1.0f - r_unpol));
            float r_s, r_p;
            fresnel_rs_rp(n1k, n2k, g.cos_i, cos_t_k, &r_s, &r_p);
            float r_unpol_k = clampf(0.5f * fma_(r_p, r_p, r_s * r_s), 1e-4f, 1.0f - 1e-4f);
            pdf[k] *= 1.0f - r_unpol_k;
        } else {
            st[k] = scale(st[k], 0.0f);
            pdf[k] = 0.0f;
        }
    }
    RefractOut out;
    out.new_k = refr_wave_dir;
    out.new_s = final_refr_dir;
    out.has_update = entering_aniso;
    out.update = use_extra;
    return out;
}

// channel_absorption_alphas (raytracer/absorption.rs:174-214): each channel's pleochroic
// coefficient for its own Stokes vector, s axis s_axis and wave normal k_hat.
void channel_absorption_alphas(const RayCtx &ctx, V3 s_axis, V3 k_hat, const Stokes *st, float alphas[8]) {
    V3 ea, eb;
    if (ctx.is_biaxial) {
        eigen_polarizations(ctx.biaxial_ch[0], k_hat, &ea, &eb);
    } else {
        ea = ordinary_eigen_polarization(k_hat, ctx.c_axis);
        eb = extraordinary_eigen_polarization(k_hat, ctx.c_axis);
    }
    for (int k = 0; k < 8; ++k) {
        V3 e_hat = electric_field_direction(st[k], s_axis, k_hat);
        alphas[k] = effective_pleochroic_alpha(ctx.tensor_ch[k], e_hat, ea, eb, degree_of_polarization(st[k]));
    }
}

// maybe_scatter_or_extinguish (scattering.rs:132-232): an exponential free path on the
// hero's extinction alphas[0] + sigma_s (absorption-length units; hit_t is in model units,
// scaled by path_scale).  Before the boundary: every channel is depolarized and weighted
// tr_k sigma_s / (sigma_t,hero (1 - u)), its pdf multiplied by sigma_t,k tr_k, and the
// new direction is HG-sampled; returns true with the free path in model units.  Else every
// channel takes survive_k / survive_hero and pdf *= survive_k; returns false.
bool maybe_scatter_or_extinguish(const float alphas[8], float sigma_s, float g, V3 ray_dir, float hit_t,
                                 float path_scale, uint32_t seed, uint32_t bounce, Stokes *st, float *pdf,
                                 float *t_model, V3 *new_dir) {
    float sigma_t_hero = alphas[0] + sigma_s;
    float dist_rand = stream_draw(seed, bounce, DISTANCE_SAMPLE_STREAM);
    float one_minus_u = maxf(1.0f - dist_rand, 1e-7f);
    float t_free = -(tf_log(one_minus_u)) / sigma_t_hero;
    float hit_t_scaled = hit_t * path_scale;
    if (t_free < hit_t_scaled) {
        float pdf_hero = sigma_t_hero * one_minus_u;
        for (int k = 0; k < 8; ++k) {
            float sigma_t_k = alphas[k] + sigma_s;
            float tr_k = exp_lane(-(alphas[k] + sigma_s) * t_free);
            float weight = tr_k * sigma_s / pdf_hero;
            st[k] = Stokes{intensity(st[k]) * weight, 0.0f, 0.0f, 0.0f};
            pdf[k] *= sigma_t_k * tr_k;
        }
        float u1 = stream_draw(seed, bounce, PHASE_DIR_U_STREAM);
        float u2 = stream_draw(seed, bounce, PHASE_DIR_V_STREAM);
        *new_dir = sample_henyey_greenstein_direction(u1, u2, g, ray_dir);
        *t_model = t_free / path_scale;
        return true;
    }
    float survive[8];
    for (int k = 0; k < 8; ++k) survive[k] = exp_lane(-(alphas[k] + sigma_s) * hit_t_scaled);
    float survive_hero = maxf(survive[0], 1e-30f);
    for (int k = 0; k < 8; ++k) {
        st[k] = scale(st[k], survive[k] / survive_hero);
        pdf[k] *= survive[k];
    }
    return false;
}

inline void fnv(uint32_t &h, uint32_t v) {
    h ^= v;
    h *= 16777619u;
}

}  // namespace

// ---------------------------------------------------------------------------------
// The estimator: trace_spectral_ray_inner (transport.rs:617-872).
// ---------------------------------------------------------------------------------

extern "C" void or_trace_spectral_ray(const float origin[3], const float dir[3],
                                      const float *planes, int n_planes, const uint32_t *finishes,
                                      const or_material *m, uint32_t max_bounces, const or_env *env,
                                      uint32_t rng_seed, float hero_rand, float out_xyz[3],
                                      or_path_info *info, uint64_t *ev) {
    float lambdas[8];
    wrapped_hero_wavelengths(hero_rand, lambdas);
    Stokes stokes[8];
    float radiance[8], path_pdf[8];
    for (int k = 0; k < 8; ++k) {
        stokes[k] = Stokes{1.0f, 0.0f, 0.0f, 0.0f};
        radiance[k] = 0.0f;
        path_pdf[k] = 1.0f;
    }
    V3 ray_o = load3(origin), ray_d = load3(dir);
    V3 current_k = ray_d;
    bool inside = false;
    bool have_prev = false;
    V3 prev_plane_normal = v3(0, 0, 0);
    bool is_extra = false;
    RayCtx ctx;
    build_ray_ctx(m, lambdas, &ctx);

    uint32_t term_bounce = max_bounces, term_reason = OR_TERM_CAP;
    uint32_t fh = 2166136261u;
    uint32_t fate_h = 2166136261u;

    for (uint32_t bounce = 0; bounce < max_bounces; ++bounce) {
        if (ev) ev[OR_EV_SEGMENT]++;
        Hit hit = intersect_planes(planes, n_planes, ray_o, ray_d);
        if (!hit.hit) {
            // accumulate_miss_radiance (transport.rs:67-90).
            if (ev) ev[OR_EV_MISS]++;
            if (env->kind == 0) {
                Rig rig = studio_rig(env->light_yaw, env->light_pitch);
                for (int k = 0; k < 8; ++k) {
                    float e = studio_with_rig(ray_d, lambdas[k], env->temp_k, env->spot_mult,
                                              env->exposure, rig);
                    radiance[k] = fma_(intensity(stokes[k]), e, radiance[k]);
                }
            } else {
                for (int k = 0; k < 8; ++k) {
                    float rgb[3];
                    map_radiance_rgb(env, ray_d, rgb);
                    float e = rgb_to_spectral(rgb, lambdas[k]);
                    radiance[k] = fma_(intensity(stokes[k]), e, radiance[k]);
                }
            }
            fnv(fh, 0xFFFFu);
            term_bounce = bounce;
            term_reason = OR_TERM_ESCAPED;
            break;
        }
        if (ev) ev[OR_EV_HIT]++;
        // try_scatter_step (scattering.rs:291-336) -> maybe_scatter_or_extinguish
        // (:132-232), inside a scattering stone before its facet hit is processed: an
        // exponential free path on the hero's extinction sigma_a + sigma_s; a scatter before
        // the facet depolarizes, reweights every channel, turns the ray by the HG phase
        // function and skips the facet (transport.rs:805-848); otherwise every channel
        // takes its extinction over the segment relative to the hero's.
        if (inside && !(m->scattering_sigma_s <= 0.0f)) {
            if (ev) ev[OR_EV_SCATTER]++;
            float alphas[8];
            channel_absorption_alphas(ctx, have_prev ? prev_plane_normal : v3(0, 0, 0), current_k, stokes, alphas);
            float t_free_model;
            V3 new_dir;
            if (maybe_scatter_or_extinguish(alphas, m->scattering_sigma_s, m->scattering_g, ray_d, hit.t,
                                            m->absorption_path_scale, rng_seed, bounce, stokes, path_pdf,
                                            &t_free_model, &new_dir)) {
                ray_o = add(ray_o, mul(ray_d, t_free_model));
                ray_d = new_dir;
                current_k = new_dir;
                fnv(fh, 0xFFFEu);
                have_prev = false;
                prev_plane_normal = v3(0, 0, 0);
                if (bounce > 4) {
                    if (ev) ev[OR_EV_RR]++;
                    float max_i = 0.0f;
                    for (int k = 0; k < 8; ++k) max_i = maxf(max_i, intensity(stokes[k]));
                    float q = clampf(max_i, 0.05f, 1.0f);
                    if (stream_draw(rng_seed, bounce, RUSSIAN_ROULETTE_STREAM) > q) {
                        term_bounce = bounce;
                        term_reason = OR_TERM_SCATTER;
                        break;
                    }
                    for (int k = 0; k < 8; ++k) stokes[k] = scale(stokes[k], 1.0f / q);
                }
                continue;
            }
        }
        V3 hit_point = add(ray_o, mul(ray_d, hit.t));
        // Facet edge rounding (transport.rs:850-857).
        V3 normal = shading_normal_near_edge(planes, n_planes, hit_point, hit.facet, hit.normal,
                                             m->edge_rounding_radius);

        // rotate_stokes_to_plane_of_incidence (raytracer/absorption.rs:71-89).
        V3 cpn = normalize_or_zero(cross(current_k, normal));
        if (have_prev && length_squared(cpn) > 1e-6f && length_squared(prev_plane_normal) > 1e-6f) {
            float cos_psi = clampf(dot(prev_plane_normal, cpn), -1.0f, 1.0f);
            float sin_psi = dot(cross(prev_plane_normal, cpn), normalize_or_zero(current_k));
            float psi = tf_atan2(sin_psi, cos_psi);
            M4 rot = frame_rotation(psi);
            for (int k = 0; k < 8; ++k) stokes[k] = apply(stokes[k], rot);
            if (ev) ev[OR_EV_ROTATE]++;
        }
        prev_plane_normal = cpn;
        have_prev = true;

        // apply_interior_segment (transport.rs:538-557) -> apply_absorption
        // (raytracer/absorption.rs:115-141).
        // (A scattering stone's extinction was taken by the scatter step above.)
        if (inside) {
            normal = neg(normal);
        }
        if (inside && m->scattering_sigma_s <= 0.0f) {
            if (ev) ev[OR_EV_ABSORB]++;
            float alphas[8];
            channel_absorption_alphas(ctx, cpn, current_k, stokes, alphas);
            float spl = hit.t * m->absorption_path_scale;
            for (int k = 0; k < 8; ++k) {
                float trans = exp_lane(-alphas[k] * spl);
                stokes[k] = scale(stokes[k], trans);
            }
        }

        Geo g = compute_geo(ctx, normal, current_k, inside, is_extra);

        // dispatch_bounce (transport.rs:128-193): FacetFinish::Frosted (1) takes
        // apply_frosted_bounce (scattering.rs:522-590), Polished (0 / no finishes) the
        // specular TIR / partial Fresnel dispatch.
        V3 new_k, new_s;
        bool new_inside;
        bool has_update = false, update = false;
        uint32_t event;
        if (finishes && finishes[hit.facet] == 1u) {
            float u1 = stream_draw(rng_seed, bounce, FROSTED_DIR_U_STREAM);
            float u2 = stream_draw(rng_seed, bounce, FROSTED_DIR_V_STREAM);
            new_inside = inside;
            if (g.sin2_t > 1.0f) {
                if (ev) ev[OR_EV_TIR]++;
                event = 4;
                new_k = cosine_weighted_hemisphere(u1, u2, normal);
                for (int k = 0; k < 8; ++k) stokes[k] = Stokes{intensity(stokes[k]), 0.0f, 0.0f, 0.0f};
            } else {
                float cos_t = std::sqrt(maxf(1.0f - g.sin2_t, 0.0f));
                float r_s, r_p;
                fresnel_rs_rp(g.n1, g.n2, g.cos_i, cos_t, &r_s, &r_p);
                float r_unpol = clampf(0.5f * fma_(r_p, r_p, r_s * r_s), 1e-4f, 1.0f - 1e-4f);
                if (stream_draw(rng_seed, bounce, FRESNEL_BRANCH_STREAM) < r_unpol) {
                    if (ev) ev[OR_EV_REFLECT]++;
                    event = 5;
                    new_k = cosine_weighted_hemisphere(u1, u2, normal);
                    for (int k = 0; k < 8; ++k) {
                        stokes[k] = Stokes{intensity(stokes[k]) / r_unpol, 0.0f, 0.0f, 0.0f};
                        path_pdf[k] *= r_unpol;
                    }
                } else {
                    if (ev) ev[OR_EV_REFRACT]++;
                    event = 6;
                    new_k = cosine_weighted_hemisphere(u1, u2, neg(normal));
                    bool entering = !inside && ctx.is_anisotropic;
                    bool use_extra = is_extra;
                    if (entering) {
                        if (ev) ev[OR_EV_ENTRY_ANISO]++;
                        use_extra = stream_draw(rng_seed, bounce, BIREFRINGENT_SPLIT_STREAM) < 0.5f;
                    }
                    float t_unpol = 1.0f - r_unpol;
                    for (int k = 0; k < 8; ++k) {
                        stokes[k] = Stokes{intensity(stokes[k]) / t_unpol, 0.0f, 0.0f, 0.0f};
                        path_pdf[k] *= t_unpol;
                    }
                    new_inside = !inside;
                    has_update = entering;
                    update = use_extra;
                }
            }
            new_s = new_k;
        } else if (g.sin2_t > 1.0f) {
            if (ev) ev[OR_EV_TIR]++;
            event = 1;
            new_k = apply_tir_bounce(g, current_k, normal, stokes, path_pdf);
            new_s = poynting_dir_for_mode(ctx, g, new_k, inside, is_extra, ev);
            new_inside = inside;
        } else {
            // apply_partial_fresnel_bounce (refraction.rs:1036-1086).
            float cos_t = std::sqrt(1.0f - g.sin2_t);
            float r_s, r_p;
            fresnel_rs_rp(g.n1, g.n2, g.cos_i, cos_t, &r_s, &r_p);
            float r_unpol = clampf(0.5f * fma_(r_p, r_p, r_s * r_s), 1e-4f, 1.0f - 1e-4f);
            float rng_bounce = stream_draw(rng_seed, bounce, FRESNEL_BRANCH_STREAM);
            if (rng_bounce < r_unpol) {
                if (ev) ev[OR_EV_REFLECT]++;
                event = 2;
                new_k = apply_partial_reflect_bounce(g, r_unpol, current_k, normal, stokes,
                                                     path_pdf);
                new_s = poynting_dir_for_mode(ctx, g, new_k, inside, is_extra, ev);
                new_inside = inside;
            } else {
                if (ev) ev[OR_EV_REFRACT]++;
                event = 3;
                RefractOut ro = apply_refract_bounce(ctx, g, r_unpol, current_k, normal, inside,
                                                     is_extra, rng_seed, bounce, stokes,
                                                     path_pdf, ev);
                new_k = ro.new_k;
                new_s = ro.new_s;
                new_inside = !inside;
                has_update = ro.has_update;
                update = ro.update;
            }
        }
        bool was_internal_reflection = inside && !has_update && new_inside == inside;
        ray_d = new_s;
        ray_o = add(hit_point, mul(new_s, 1e-4f));
        current_k = new_k;
        inside = new_inside;
        if (has_update) is_extra = update;
        if (ctx.is_anisotropic && was_internal_reflection)
            is_extra = stream_draw(rng_seed, bounce, MODE_COUPLING_STREAM) < 0.5f;
        fnv(fh, (uint32_t)hit.facet);
        fnv(fh, event | (is_extra ? 16u : 0u));
        // Per-channel fates of this bounce: channels past their own critical angle, and
        // channels whose path pdf is zero (chromatically terminated).
        uint32_t fate = 0;
        for (int k = 0; k < 8; ++k) {
            if (g.sin2_t_ch[k] > 1.0f) fate |= 1u << k;
            if (path_pdf[k] == 0.0f) fate |= 1u << (8 + k);
        }
        fnv(fate_h, fate);

        // apply_russian_roulette (transport.rs:204-220).
        if (bounce > 4) {
            if (ev) ev[OR_EV_RR]++;
            float max_i = 0.0f;
            for (int k = 0; k < 8; ++k) max_i = maxf(max_i, intensity(stokes[k]));
            float q = clampf(max_i, 0.05f, 1.0f);
            float rr = stream_draw(rng_seed, bounce, RUSSIAN_ROULETTE_STREAM);
            if (rr > q) {
                term_bounce = bounce;
                term_reason = OR_TERM_RR;
                break;
            }
            for (int k = 0; k < 8; ++k) stokes[k] = scale(stokes[k], 1.0f / q);
        }
    }

    // integrate_channels_to_xyz (color.rs:121-164).
    float sum_pdf = -0.0f;
    for (int k = 0; k < 8; ++k) sum_pdf = sum_pdf + path_pdf[k];
    float mis = sum_pdf <= 1e-12f ? 1.0f : (8.0f * path_pdf[0]) / sum_pdf;
    const float norm_factor = (400.0f / 8.0f) / 106.856f;
    float cmfs[8][3];
    cmf_x8(lambdas, cmfs);
    V3 xyz = v3(0, 0, 0);
    for (int k = 0; k < 8; ++k) {
        float wr = radiance[k] * mis;
        xyz = add(xyz, mul(load3(cmfs[k]), wr * norm_factor));
    }
    V3 wb = env->kind == 0 ? g_wb.get(env->temp_k) : v3(1.0f, 1.0f, 1.0f);
    V3 outv = von_kries(xyz, wb);
    out_xyz[0] = outv.x;
    out_xyz[1] = outv.y;
    out_xyz[2] = outv.z;
    if (info) {
        info->segments = term_reason == OR_TERM_CAP ? max_bounces : term_bounce + 1;
        info->termination = term_reason;
        info->facet_hash = fh;
        info->fate_hash = fate_h;
    }
}

// ---------------------------------------------------------------------------------
// Thin C wrappers used by the tests.
// ---------------------------------------------------------------------------------

extern "C" uint32_t or_hash_u32(uint32_t x) { return hash_u32(x); }
// The per-bounce draw of one stream (sampling.rs:45-65), the base-2 van der Corput value
// (:139-147) and the Cranley-Patterson rotation (:189-197), for the integer-exact probes.
extern "C" float or_stream_draw(uint32_t seed, uint32_t bounce, uint32_t stream) {
    return stream_draw(seed, bounce, stream);
}
extern "C" float or_low_discrepancy_base2(uint32_t n) { return low_discrepancy_base2(n); }
extern "C" float or_cranley_patterson_rotate(float x, float offset) { return cranley_patterson_rotate(x, offset); }
// The RNG record of renderer/gpu/rng_check.rs:1-35 for n (pixel, sample, bounce) tuples, as
// 32-bit patterns, out[(8 + n_salts) * i + k]: the path seed, the three pixel-rotation hashes
// hash_u32(pixel ^ ROTATION_STREAM), the sample's reversed bits, the stratified draws
// (jx, jy, hero) of sample_draws, then stream_draw(seed, bounce, salt) for each salt.
extern "C" void or_rng_records(size_t n, const uint32_t *pixel, const uint32_t *sample, const uint32_t *bounce,
                               const uint32_t *salts, int n_salts, uint32_t *out) {
    auto bits = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, sizeof(u));
        return u;
    };
    for (size_t i = 0; i < n; ++i) {
        uint32_t *o = out + (size_t)(8 + n_salts) * i;
        float rot[3];
        pixel_rotations(pixel[i], rot);
        const Draws d = sample_draws(pixel[i], sample[i], rot);
        o[0] = d.seed;
        o[1] = hash_u32(pixel[i] ^ PIXEL_JITTER_X_ROTATION_STREAM);
        o[2] = hash_u32(pixel[i] ^ PIXEL_JITTER_Y_ROTATION_STREAM);
        o[3] = hash_u32(pixel[i] ^ HERO_WAVELENGTH_ROTATION_STREAM);
        o[4] = reverse_bits(sample[i]);
        o[5] = bits(d.jx);
        o[6] = bits(d.jy);
        o[7] = bits(d.hero);
        for (int k = 0; k < n_salts; ++k) o[8 + k] = bits(stream_draw(d.seed, bounce[i], salts[k]));
    }
}
// rgb_to_spectral_radiance (renderer/env_map_spectrum.rs:60-82).
extern "C" float or_rgb_to_spectral_radiance(const float rgb[3], float lambda_nm) {
    return rgb_to_spectral(rgb, lambda_nm);
}
extern "C" void or_pixel_rotations(uint32_t pixel, float out[3]) { pixel_rotations(pixel, out); }
extern "C" void or_sample_draws(uint32_t pixel, uint32_t sample_num, uint32_t *seed,
                                float out[3]) {
    float rot[3];
    pixel_rotations(pixel, rot);
    Draws d = sample_draws(pixel, sample_num, rot);
    *seed = d.seed;
    out[0] = d.jx;
    out[1] = d.jy;
    out[2] = d.hero;
}
extern "C" float or_radical_inverse(uint32_t n, uint32_t base) {
    return radical_inverse_base(n, base);
}
extern "C" void or_wrapped_hero_wavelengths(float hero_rand, float out[8]) {
    wrapped_hero_wavelengths(hero_rand, out);
}
extern "C" void or_exp_f32x8(const float in[8], float out[8]) {
    for (int k = 0; k < 8; ++k) out[k] = exp_lane(in[k]);
}
extern "C" void or_cie_1931_cmf(float lambda_nm, float out[3]) { cmf_scalar(lambda_nm, out); }
extern "C" void or_cie_1931_cmf_x8(const float lambdas[8], float out[24]) {
    float tmp[8][3];
    cmf_x8(lambdas, tmp);
    std::memcpy(out, tmp, sizeof(tmp));
}
extern "C" float or_dispersion_evaluate(const or_material *m, float lambda_nm) {
    return dispersion_evaluate(m, lambda_nm);
}
extern "C" float or_blackbody(float lambda_nm, float temp_k) {
    return blackbody_spectrum(lambda_nm, temp_k);
}
extern "C" void or_white_balance(float temp_k, float out[3]) {
    V3 v = compute_white_balance(temp_k);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}
extern "C" float or_sample_studio_environment(const float dir[3], float lambda_nm, float temp_k,
                                              float spot_mult, float exposure, float light_yaw,
                                              float light_pitch) {
    Rig rig = studio_rig(light_yaw, light_pitch);
    return studio_with_rig(load3(dir), lambda_nm, temp_k, spot_mult, exposure, rig);
}
extern "C" float or_env_map_radiance(const or_env *env, const float dir[3], float lambda_nm) {
    float rgb[3];
    map_radiance_rgb(env, load3(dir), rgb);
    return rgb_to_spectral(rgb, lambda_nm);
}
// renderer/gpu/estimator_check/furnace.rs:294-302.
extern "C" void or_analytic_furnace_target(float l0, float out[3]) {
    V3 sum = v3(0, 0, 0);
    float rgb[3] = {l0, l0, l0};
    for (int step = 0; step <= 400; ++step) {
        float lambda = 380.0f + (float)step;
        float spec = rgb_to_spectral(rgb, lambda);
        float c[3];
        cmf_scalar(lambda, c);
        sum = add(sum, mul(load3(c), spec));
    }
    // Vec3 / f32: each lane divided.
    out[0] = sum.x / 106.856f;
    out[1] = sum.y / 106.856f;
    out[2] = sum.z / 106.856f;
}
// geometry/plane.rs:154-161.
extern "C" void or_plane_new(const float n[3], float d, float out[4]) {
    V3 u = normalize(load3(n));
    out[0] = u.x;
    out[1] = u.y;
    out[2] = u.z;
    out[3] = d;
}
// geometry/cuts.rs:147-218 with its helpers :599-768 (34 planes).
extern "C" int or_emerald_cut(float *p) {
    int count = 0;
    auto push = [&](V3 n, float d) {
        float nn[3] = {n.x, n.y, n.z};
        or_plane_new(nn, d, p + 4 * count);
        ++count;
    };
    const float table_y = 0.30f, half_band = 0.03f, crown_crease_y = 0.15f, pav_crease = 0.35f;
    const float r_z = 0.80f, r_x = 1.10f, r_diag = 1.00f, keel_y = -0.86f;
    const float diag = 0.70710678118654752440f;
    auto d_girdle = [](float a, float r, float hb) { return -fma_(tf_sin(a), r, tf_cos(a) * hb); };
    auto d_crease = [](float ga, float ca, float r, float cm, float hb) {
        float z_c = r - tf_cos(ga) * (cm - hb) / tf_sin(ga);
        return -fma_(tf_sin(ca), z_c, tf_cos(ca) * cm);
    };
    auto corners = [&](float a, float ysign) {
        float d = -fma_(tf_cos(a), half_band, tf_sin(a) * r_diag);
        float y = ysign * tf_cos(a);
        float sd = tf_sin(a) * diag;
        push(v3(sd, y, sd), d);
        push(v3(-sd, y, sd), d);
        push(v3(sd, y, -sd), d);
        push(v3(-sd, y, -sd), d);
    };
    auto tiers = [&](float a, float ysign, float dz, float dx) {
        push(v3(0.0f, ysign * tf_cos(a), tf_sin(a)), dz);
        push(v3(0.0f, ysign * tf_cos(a), -tf_sin(a)), dz);
        push(v3(tf_sin(a), ysign * tf_cos(a), 0.0f), dx);
        push(v3(-tf_sin(a), ysign * tf_cos(a), 0.0f), dx);
    };
    push(v3(0.0f, 1.0f, 0.0f), -table_y);
    const float a1 = to_radians(35.0f), a2 = to_radians(45.0f);
    tiers(a1, 1.0f, d_crease(a2, a1, r_z, crown_crease_y, half_band), d_crease(a2, a1, r_x, crown_crease_y, half_band));
    tiers(a2, 1.0f, d_girdle(a2, r_z, half_band), d_girdle(a2, r_x, half_band));
    corners(to_radians(40.0f), 1.0f);
    push(v3(0.0f, 0.0f, 1.0f), -r_z);
    push(v3(0.0f, 0.0f, -1.0f), -r_z);
    push(v3(1.0f, 0.0f, 0.0f), -r_x);
    push(v3(-1.0f, 0.0f, 0.0f), -r_x);
    push(v3(diag, 0.0f, diag), -r_diag);
    push(v3(-diag, 0.0f, diag), -r_diag);
    push(v3(diag, 0.0f, -diag), -r_diag);
    push(v3(-diag, 0.0f, -diag), -r_diag);
    const float p1 = to_radians(53.0f), p2 = to_radians(43.0f);
    tiers(p1, -1.0f, d_girdle(p1, r_z, half_band), d_girdle(p1, r_x, half_band));
    tiers(p2, -1.0f, d_crease(p1, p2, r_z, pav_crease, half_band), d_crease(p1, p2, r_x, pav_crease, half_band));
    corners(to_radians(48.0f), -1.0f);
    push(v3(0.0f, -1.0f, 0.0f), keel_y);
    return count;
}
// geometry/cuts.rs:45-129 (74 planes).
extern "C" int or_standard_round_brilliant(float *p) {
    int count = 0;
    const float gear = 96.0f;
    auto push = [&](V3 n, float d) {
        float nn[3] = {n.x, n.y, n.z};
        or_plane_new(nn, d, p + 4 * count);
        ++count;
    };
    push(v3(0.0f, 1.0f, 0.0f), -0.32f);
    float star = to_radians(15.0f);
    const float star_g[8] = {6, 18, 30, 42, 54, 66, 78, 90};
    for (float g : star_g) {
        float phi = 2.0f * PI_F * g / gear;
        push(v3(tf_sin(star) * tf_cos(phi), tf_cos(star), tf_sin(star) * tf_sin(phi)),
             -0.45f);
    }
    float crown = to_radians(34.5f);
    const float main_g[8] = {0, 12, 24, 36, 48, 60, 72, 84};
    for (float g : main_g) {
        float phi = 2.0f * PI_F * g / gear;
        push(v3(tf_sin(crown) * tf_cos(phi), tf_cos(crown), tf_sin(crown) * tf_sin(phi)),
             -0.59f);
    }
    float ug = to_radians(41.0f);
    const float girdle_g[16] = {95, 1, 11, 13, 23, 25, 35, 37, 47, 49, 59, 61, 71, 73, 83, 85};
    for (float g : girdle_g) {
        float phi = 2.0f * PI_F * g / gear;
        push(v3(tf_sin(ug) * tf_cos(phi), tf_cos(ug), tf_sin(ug) * tf_sin(phi)), -0.67f);
    }
    for (int i = 0; i < 16; ++i) {
        float phi = 2.0f * PI_F * (float)i / 16.0f;
        push(v3(tf_cos(phi), 0.0f, tf_sin(phi)), -1.0f);
    }
    float pav = to_radians(41.0f);
    for (float g : main_g) {
        float phi = 2.0f * PI_F * g / gear;
        push(v3(tf_sin(pav) * tf_cos(phi), -tf_cos(pav), tf_sin(pav) * tf_sin(phi)),
             -0.67f);
    }
    float lg = to_radians(42.5f);
    for (float g : girdle_g) {
        float phi = 2.0f * PI_F * g / gear;
        push(v3(tf_sin(lg) * tf_cos(phi), -tf_cos(lg), tf_sin(lg) * tf_sin(phi)), -0.68f);
    }
    push(v3(0.0f, -1.0f, 0.0f), -0.88f);
    return count;
}
// optics/raytracer/camera.rs:63-93.
extern "C" void or_camera_new(float yaw, float pitch, float distance, float fov_deg,
                              or_camera *out) {
    float cos_p = tf_cos(pitch), sin_p = tf_sin(pitch);
    float cos_y = tf_cos(yaw), sin_y = tf_sin(yaw);
    V3 origin = v3(distance * cos_p * sin_y, distance * sin_p, distance * cos_p * cos_y);
    V3 forward = normalize(neg(origin));
    V3 world_up = std::fabs(cos_p) < 1e-4f ? v3(0.0f, 0.0f, -1.0f) : v3(0.0f, 1.0f, 0.0f);
    V3 right = normalize(cross(forward, world_up));
    V3 up = normalize(cross(right, forward));
    float fov_tan = tf_tan(to_radians(fov_deg) * 0.5f);
    out->origin[0] = origin.x; out->origin[1] = origin.y; out->origin[2] = origin.z;
    out->forward[0] = forward.x; out->forward[1] = forward.y; out->forward[2] = forward.z;
    out->right[0] = right.x; out->right[1] = right.y; out->right[2] = right.z;
    out->up[0] = up.x; out->up[1] = up.y; out->up[2] = up.z;
    out->fov_tan = fov_tan;
}
// optics/raytracer/camera.rs:96-114.
extern "C" void or_generate_ray(const or_camera *cam, float sx, float sy, float w, float h,
                                float jx, float jy, float origin[3], float dir[3]) {
    float aspect = w / h;
    float u = ((sx + jx) / w - 0.5f) * 2.0f * aspect * cam->fov_tan;
    float v = (0.5f - (sy + jy) / h) * 2.0f * cam->fov_tan;
    V3 d = normalize(add(add(load3(cam->forward), mul(load3(cam->right), u)),
                         mul(load3(cam->up), v)));
    origin[0] = cam->origin[0];
    origin[1] = cam->origin[1];
    origin[2] = cam->origin[2];
    dir[0] = d.x;
    dir[1] = d.y;
    dir[2] = d.z;
}
extern "C" void or_shading_normal_near_edge(const float *planes, int n_planes, const float hit_point[3],
                                            int hit_facet, const float hit_normal[3], float rounding_radius,
                                            float out[3]) {
    V3 r = shading_normal_near_edge(planes, n_planes, load3(hit_point), hit_facet, load3(hit_normal),
                                    rounding_radius);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
extern "C" int or_intersect(const float *planes, int n_planes, const float origin[3],
                            const float dir[3], float *t, float normal[3], int *facet) {
    Hit h = intersect_planes(planes, n_planes, load3(origin), load3(dir));
    *t = h.t;
    normal[0] = h.normal.x;
    normal[1] = h.normal.y;
    normal[2] = h.normal.z;
    *facet = h.facet;
    return h.hit ? 1 : 0;
}
// The one-plane-at-a-time scan, for the test that the vector scan matches it bit for bit.
extern "C" int or_intersect_scalar(const float *planes, int n_planes, const float origin[3], const float dir[3],
                                   float *t, float normal[3], int *facet) {
    Hit h = intersect_planes_scalar(planes, n_planes, load3(origin), load3(dir));
    *t = h.t;
    normal[0] = h.normal.x;
    normal[1] = h.normal.y;
    normal[2] = h.normal.z;
    *facet = h.facet;
    return h.hit ? 1 : 0;
}
extern "C" float or_effective_extraordinary_index(float n_o, float n_e, float theta) {
    return effective_extraordinary_index(n_o, n_e, theta);
}
// The theta_c_for_bounce call site (refraction.rs:248-269): theta = acos(|cos|), then
// effective_extraordinary_index.
extern "C" float or_neff_from_cos(float n_o, float n_e, float c) {
    return effective_extraordinary_index(n_o, n_e, tf_acos(std::fabs(clampf(c, -1.0f, 1.0f))));
}
// rotate_stokes_to_plane_of_incidence's matrix (raytracer/absorption.rs:71-89,
// polarization.rs:95-101): frame_rotation(signed_frame_rotation_psi(prev, cur, axis)) as
// (cos 2psi, sin 2psi), its Q<-Q and Q<-U entries.
extern "C" float or_signed_frame_rotation_psi(const float prev[3], const float cur[3], const float axis[3]);
extern "C" void or_frame_rotation_cs(const float prev[3], const float cur[3], const float axis[3], float out[2]) {
    M4 m = frame_rotation(or_signed_frame_rotation_psi(prev, cur, axis));
    out[0] = m.c[5];
    out[1] = m.c[9];
}
// The polarization quadratic form on the reference's angle route (polarization.rs:58-60,
// 172-192; birefringence.rs:252-270): psi = atan2(U, Q) / 2, e = cos psi s + sin psi p,
// here with s and p given by their components in the tensor's principal frame, where
// A = diag(ao, ao, ae).
extern "C" float or_field_qform(float q, float u, const float s_l[3], const float p_l[3], float ao, float ae) {
    float psi = 0.5f * tf_atan2(u, q);
    V3 e = add(mul(load3(s_l), tf_cos(psi)), mul(load3(p_l), tf_sin(psi)));
    return dot(v3(ao, ao, ae), mulv(e, e));
}
// The refraction channel's amplitudes and transmission Mueller element m00 on the
// reference's operations (refraction.rs:1049-1056 and the t_s, t_p of :748-1011,
// polarization.rs:119-137).
extern "C" void or_fresnel_companion(float n1, float n2, float cos_i, float cos_t, float out[3]) {
    fresnel_rs_rp(n1, n2, cos_i, cos_t, &out[0], &out[1]);
    const float t_s = (2.0f * n1 * cos_i) / fma_(n2, cos_t, n1 * cos_i);
    const float t_p = (2.0f * n1 * cos_i) / fma_(n1, cos_t, n2 * cos_i);
    out[2] = fresnel_transmission(n1, n2, cos_i, cos_t, t_s, t_p).c[0];
}
extern "C" float or_degree_of_polarization(float i, float q, float u, float v) {
    return degree_of_polarization(Stokes{i, q, u, v});
}
extern "C" float or_walk_off_angle(float n_o, float n_e, float theta) {
    return walk_off_angle(n_o, n_e, theta);
}
extern "C" void or_extraordinary_poynting_dir(const float w[3], const float c[3], float n_o,
                                              float n_e, float out[3]) {
    V3 r = extraordinary_poynting_dir(load3(w), load3(c), n_o, n_e);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
extern "C" float or_tir_phase_delta(float n1, float cos_i, float sin_i) {
    return tir_phase_delta(n1, cos_i, sin_i);
}
// Biaxial indicatrix probes: an indicatrix travels as 12 floats (n_alpha, n_beta,
// n_gamma, then the alpha, beta and gamma principal axes).
namespace {
Biaxial load_biaxial(const float p[12]) {
    Biaxial b;
    b.n_alpha = p[0];
    b.n_beta = p[1];
    b.n_gamma = p[2];
    b.ax = load3(p + 3);
    b.ay = load3(p + 6);
    b.az = load3(p + 9);
    return b;
}
void store_biaxial(const Biaxial &b, float p[12]) {
    p[0] = b.n_alpha;
    p[1] = b.n_beta;
    p[2] = b.n_gamma;
    const V3 v[3] = {b.ax, b.ay, b.az};
    for (int j = 0; j < 3; ++j) {
        p[3 + 3 * j] = v[j].x;
        p[4 + 3 * j] = v[j].y;
        p[5 + 3 * j] = v[j].z;
    }
}
void store3(V3 v, float *p) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
}  // namespace
extern "C" int or_material_biaxial_indicatrix(const or_material *m, float lambda_nm, float out[12]) {
    Biaxial b;
    if (!material_biaxial_indicatrix(m, lambda_nm, &b)) return 0;
    store_biaxial(b, out);
    return 1;
}
extern "C" void or_biaxial_from_gamma_axis(float n_alpha, float n_beta, float n_gamma, const float g[3],
                                           float out[12]) {
    store_biaxial(biaxial_from_gamma_axis(n_alpha, n_beta, n_gamma, load3(g)), out);
}
extern "C" void or_biaxial_wave_indices(const float ind[12], const float w[3], float out[2]) {
    wave_indices(load_biaxial(ind), load3(w), &out[0], &out[1]);
}
extern "C" void or_biaxial_eigen_polarizations(const float ind[12], const float w[3], float out[6]) {
    V3 a, b;
    eigen_polarizations(load_biaxial(ind), load3(w), &a, &b);
    store3(a, out);
    store3(b, out + 3);
}
extern "C" void or_biaxial_d_to_e(const float ind[12], const float d[3], float out[3]) {
    store3(d_to_e_direction(load_biaxial(ind), load3(d)), out);
}
extern "C" void or_biaxial_mode_poynting_dir(const float ind[12], const float w[3], int want_slow, float out[3]) {
    store3(mode_poynting_dir(load_biaxial(ind), load3(w), want_slow != 0), out);
}
extern "C" float or_biaxial_resolve_entry_mode(const float ind[12], const float incident[3], const float normal[3],
                                               float cos_i, float n_seed, int want_slow, float wave[3]) {
    V3 w;
    float n = resolve_entry_mode(load_biaxial(ind), load3(incident), load3(normal), cos_i, n_seed, want_slow != 0, &w);
    store3(w, wave);
    return n;
}
extern "C" void or_poynting_direction(const float k[3], const float e[3], float out[3]) {
    store3(poynting_direction(load3(k), load3(e)), out);
}
extern "C" void or_uniaxial_eigen_polarizations(const float w[3], const float c[3], float out[6]) {
    store3(ordinary_eigen_polarization(load3(w), load3(c)), out);
    store3(extraordinary_eigen_polarization(load3(w), load3(c)), out + 3);
}
// generate_guide_buffers (apps/diagram-gui/src/bridge/guide_pass.rs:37-110): per pixel the
// ray Camera::generate_ray(x, y, w, h, 0, 0) against the planes (intersect_polyhedron,
// intersect.rs:9-78): the hit distance, normal and facet, or 1e6, zero and -1 on a miss.
extern "C" void or_guide_buffers(const or_camera *cam, uint32_t width, uint32_t height, const float *planes,
                                 int n_planes, float *depth, float *normal, int32_t *facet) {
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            const size_t p = (size_t)y * width + x;
            float o[3], d[3];
            or_generate_ray(cam, (float)x, (float)y, (float)width, (float)height, 0.0f, 0.0f, o, d);
            Hit h = intersect_planes(planes, n_planes, load3(o), load3(d));
            depth[p] = h.hit ? h.t : 1.0e6f;
            normal[3 * p + 0] = h.hit ? h.normal.x : 0.0f;
            normal[3 * p + 1] = h.hit ? h.normal.y : 0.0f;
            normal[3 * p + 2] = h.hit ? h.normal.z : 0.0f;
            facet[p] = h.hit ? h.facet : -1;
        }
    }
}

// Scattering probes (scattering.rs:132-232, 428-447; raytracer/absorption.rs:174-214).
extern "C" void or_sample_hg_batch(int n, const float *u1, const float *u2, float g, const float forward[3],
                                   float *out) {
    for (int i = 0; i < n; ++i) {
        V3 d = sample_henyey_greenstein_direction(u1[i], u2[i], g, load3(forward));
        out[3 * i] = d.x;
        out[3 * i + 1] = d.y;
        out[3 * i + 2] = d.z;
    }
}
extern "C" void or_channel_absorption_alphas(const or_material *m, const float lambdas[8], const float s_axis[3],
                                             const float k_hat[3], const float stokes[32], float out[8]) {
    RayCtx ctx;
    build_ray_ctx(m, lambdas, &ctx);
    Stokes st[8];
    for (int k = 0; k < 8; ++k) st[k] = Stokes{stokes[4 * k], stokes[4 * k + 1], stokes[4 * k + 2], stokes[4 * k + 3]};
    channel_absorption_alphas(ctx, load3(s_axis), load3(k_hat), st, out);
}
// stokes (32) and pdf (8) are updated in place; returns 1 on a scatter (t_model and dir set).
extern "C" int or_maybe_scatter_or_extinguish(const float alphas[8], float sigma_s, float g, const float dir[3],
                                              float hit_t, float path_scale, uint32_t seed, uint32_t bounce,
                                              float stokes[32], float pdf[8], float *t_model, float new_dir[3]) {
    Stokes st[8];
    for (int k = 0; k < 8; ++k) st[k] = Stokes{stokes[4 * k], stokes[4 * k + 1], stokes[4 * k + 2], stokes[4 * k + 3]};
    V3 nd = v3(0, 0, 0);
    float t = 0.0f;
    bool sc = maybe_scatter_or_extinguish(alphas, sigma_s, g, load3(dir), hit_t, path_scale, seed, bounce, st, pdf,
                                          &t, &nd);
    for (int k = 0; k < 8; ++k) {
        stokes[4 * k] = st[k].i;
        stokes[4 * k + 1] = st[k].q;
        stokes[4 * k + 2] = st[k].u;
        stokes[4 * k + 3] = st[k].v;
    }
    *t_model = t;
    store3(nd, new_dir);
    return sc ? 1 : 0;
}
// effective_pleochroic_alpha (birefringence.rs:272-286) on AbsorptionTensor3::uniaxial
// (alpha[0], alpha[2]) when trichroic == 0, else ::biaxial (alpha[0], alpha[1], alpha[2]).
extern "C" float or_effective_pleochroic_alpha(const float alpha[3], int trichroic, const float axis[3],
                                               const float e[3], const float ea[3], const float eb[3], float dop) {
    Tensor3 t = trichroic ? tensor_biaxial(alpha[0], alpha[1], alpha[2], load3(axis))
                          : tensor_uniaxial(alpha[0], alpha[2], load3(axis));
    return effective_pleochroic_alpha(t, load3(e), load3(ea), load3(eb), dop);
}
// raytracer/absorption.rs:44-54.
extern "C" float or_signed_frame_rotation_psi(const float prev[3], const float cur[3],
                                              const float axis[3]) {
    V3 p = load3(prev), c = load3(cur);
    float cos_psi = clampf(dot(p, c), -1.0f, 1.0f);
    float sin_psi = dot(cross(p, c), normalize_or_zero(load3(axis)));
    return tf_atan2(sin_psi, cos_psi);
}

// render_core::trace_into: rows claimed from an atomic counter, every pixel sums its own
// samples in order and adds the sum into accum.
extern "C" uint64_t or_render(const or_camera *cam, uint32_t width, uint32_t height,
                              const float *planes, int n_planes, const uint32_t *finishes,
                              const or_material *m,
                              uint32_t max_bounces, const or_env *env, uint32_t first_sample,
                              uint32_t spp, uint32_t pixel_begin, uint32_t pixel_end,
                              int threads, float *accum, float *per_sample_xyz,
                              or_path_info *per_sample_info, uint64_t *events) {
    if (pixel_end > width * height) pixel_end = width * height;
    if (pixel_begin >= pixel_end) return 0;
    if (env->kind == 0) (void)g_wb.get(env->temp_k);
    if (threads < 1) threads = 1;
    const uint32_t block = 64;
    std::atomic<uint32_t> next{pixel_begin};
    std::atomic<uint64_t> total_segments{0};
    std::mutex ev_mu;
    auto worker = [&]() {
        uint64_t local_ev[OR_EV_COUNT] = {0};
        uint64_t segs = 0;
        for (;;) {
            uint32_t b0 = next.fetch_add(block);
            if (b0 >= pixel_end) break;
            uint32_t b1 = b0 + block < pixel_end ? b0 + block : pixel_end;
            for (uint32_t pixel = b0; pixel < b1; ++pixel) {
                uint32_t x = pixel % width, y = pixel / width;
                float rot[3];
                pixel_rotations(pixel, rot);
                V3 sum = v3(0, 0, 0);
                for (uint32_t s = 0; s < spp; ++s) {
                    uint32_t sample_num = first_sample + s;
                    Draws d = sample_draws(pixel, sample_num, rot);
                    float o[3], dir[3], xyz[3];
                    or_generate_ray(cam, (float)x, (float)y, (float)width, (float)height, d.jx,
                                    d.jy, o, dir);
                    or_path_info info;
                    or_trace_spectral_ray(o, dir, planes, n_planes, finishes, m, max_bounces, env, d.seed,
                                          d.hero, xyz, &info, events ? local_ev : nullptr);
                    segs += info.segments;
                    sum = add(sum, load3(xyz));
                    size_t slot = (size_t)(pixel - pixel_begin) * spp + s;
                    if (per_sample_xyz) {
                        per_sample_xyz[3 * slot + 0] = xyz[0];
                        per_sample_xyz[3 * slot + 1] = xyz[1];
                        per_sample_xyz[3 * slot + 2] = xyz[2];
                    }
                    if (per_sample_info) per_sample_info[slot] = info;
                }
                if (accum) {
                    accum[3 * pixel + 0] += sum.x;
                    accum[3 * pixel + 1] += sum.y;
                    accum[3 * pixel + 2] += sum.z;
                }
            }
        }
        total_segments += segs;
        if (events) {
            std::lock_guard<std::mutex> g(ev_mu);
            for (int i = 0; i < OR_EV_COUNT; ++i) events[i] += local_ev[i];
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; ++i) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    return total_segments.load();
}

// ---------------------------------------------------------------------------------
// Gemological metrics (crates/gemray/src/color/metrics.rs:10-1248).
// ---------------------------------------------------------------------------------
namespace {

// camera_view_basis (metrics.rs:15-30).
void view_basis(float cam_yaw, float cam_pitch, V3 *fwd, V3 *right, V3 *up) {
    float cos_cp = tf_cos(cam_pitch), sin_cp = tf_sin(cam_pitch);
    float cos_cy = tf_cos(cam_yaw), sin_cy = tf_sin(cam_yaw);
    V3 f = normalize(v3(-cos_cp * sin_cy, -sin_cp, -cos_cp * cos_cy));
    V3 world_up = std::fabs(cos_cp) < 1e-4f ? v3(0.0f, 0.0f, -1.0f) : v3(0.0f, 1.0f, 0.0f);
    V3 r = normalize(cross(f, world_up));
    *up = normalize(cross(r, f));
    *fwd = f;
    *right = r;
}

// fresnel_transmittance (metrics.rs:117-131).
float fresnel_transmittance(float n1, float n2, float cos_i, float cos_t) {
    float denom_s = fma_(n2, cos_t, n1 * cos_i);
    float denom_p = fma_(n2, cos_i, n1 * cos_t);
    if (std::fabs(denom_s) < 1e-6f || std::fabs(denom_p) < 1e-6f) return 0.0f;
    float rs = fma_(n2, -cos_t, n1 * cos_i) / denom_s;
    rs = rs * rs;
    float rp = fma_(n2, -cos_i, n1 * cos_t) / denom_p;
    rp = rp * rp;
    float r = midpoint(rs, rp);
    return clampf(1.0f - r, 0.0f, 1.0f);
}

enum { FATE_BLOCKED = 0, FATE_LEAKED = 1, FATE_UP = 2, FATE_ABSORBED = 3 };
struct Fate {
    int kind;
    V3 dir;
    float transmittance, exit_cos;
    int facet;
    uint32_t bounces;
};

// trace_wavelength (metrics.rs:138-222).
Fate trace_wavelength(V3 entry_point, V3 incoming, V3 n_entry, float cos_i, const float *planes, int n,
                      float index) {
    Fate f{FATE_ABSORBED, v3(0, 0, 0), 0.0f, 0.0f, -1, 0};
    float sin2_t = (1.0f / (index * index)) * fma_(cos_i, -cos_i, 1.0f);
    if (sin2_t > 1.0f) {
        f.kind = FATE_BLOCKED;
        return f;
    }
    float cos_t = std::sqrt(1.0f - sin2_t);
    float inv = 1.0f / index;
    V3 dir = normalize(add(mul(incoming, inv), mul(n_entry, fma_(inv, cos_i, -cos_t))));
    V3 hit_point = add(entry_point, mul(dir, 1e-4f));
    float sin_crit = 1.0f / index;
    float entry_t = fresnel_transmittance(1.0f, index, cos_i, cos_t);
    for (uint32_t bounce = 0; bounce < 10; ++bounce) {
        Hit h = intersect_planes(planes, n, hit_point, dir);
        if (!h.hit) break;
        V3 next_point = add(hit_point, mul(dir, h.t));
        V3 n_out = h.normal;
        float cos_theta = clampf(dot(dir, n_out), 0.0f, 1.0f);
        float sin_theta = std::sqrt(maxf(fma_(cos_theta, -cos_theta, 1.0f), 0.0f));
        if (sin_theta < sin_crit) {
            float sin2_out = (index * index) * fma_(cos_theta, -cos_theta, 1.0f);
            if (sin2_out <= 1.0f) {
                float cos_out = std::sqrt(1.0f - sin2_out);
                V3 out_dir = normalize(add(mul(dir, index), mul(n_out, fma_(index, -cos_theta, cos_out))));
                if (n_out.y < -0.05f) {
                    f.kind = FATE_LEAKED;
                } else if (out_dir.y > 0.05f) {
                    f.kind = FATE_UP;
                    f.dir = out_dir;
                    f.transmittance = entry_t * fresnel_transmittance(index, 1.0f, cos_theta, cos_out);
                    f.exit_cos = cos_out;
                    f.facet = h.facet;
                    f.bounces = bounce;
                }
                return f;
            }
        }
        dir = normalize(sub(dir, mul(n_out, 2.0f * cos_theta)));
        hit_point = add(next_point, mul(dir, 1e-4f));
    }
    return f;
}

// ray_is_visibly_returned (metrics.rs:233-251).
bool visibly_returned(V3 exit_dir, V3 fwd, V3 key, V3 fill, float sin_lp) {
    bool head_shadow = dot(neg(exit_dir), fwd) > 0.96f;
    float key_dot = maxf(dot(exit_dir, key), 0.0f);
    float fill_dot = maxf(dot(exit_dir, fill), 0.0f);
    bool ring = std::fabs(fma_(sin_lp, -0.8f, exit_dir.y)) < 0.35f;
    bool lit = key_dot > 0.70f || fill_dot > 0.75f || (ring && exit_dir.y > 0.2f);
    return !head_shadow && lit;
}

// f32::acos of the fire angle, correctly rounded (the f64 arc cosine rounded once).  The
// reference calls the platform's acosf.  glibc 2.35's is not correctly rounded (it differs
// by one ulp on 0.25 % of [-1, 1], none of them in [0.9999, 1], where nearly all fire angles
// lie); the reference's golden metrics and every value its tests report hold with either,
// and the correctly rounded value is the platform-independent one (newer glibc releases
// round acosf correctly too).
inline float acos_cr(float x) { return (float)std::acos((double)x); }

constexpr int kGrid = 18;
constexpr float kAperture[5][2] = {{0.0f, 0.0f}, {0.08f, 0.0f}, {-0.08f, 0.0f}, {0.0f, 0.08f}, {0.0f, -0.08f}};
constexpr float kYawOffsetsDeg[5] = {-3.0f, -1.5f, 0.0f, 1.5f, 3.0f};

inline float grid_coord(int i) { return fma_(((float)i + 0.5f) / (float)kGrid, 2.0f, -1.0f); }

}  // namespace

// Per-ray records of evaluate_gem_optical_metrics: for cell c = ix * 18 + iz, aperture
// sample k: cls[5c + k] (OR_MX_*), and for a returned ray with a Fire sample its
// (angle_deg, weight); temporal[5c + k] = 1 when the cell's ray at yaw offset k is visibly
// returned (classify_aperture_sample metrics.rs:519-771, cell_returned_at_yaw_offset
// :309-334).  Cells outside the perimeter get OR_MX_NONE and temporal 0.
extern "C" void or_metrics_items(const float *planes, int n_planes, const or_material *m, float cam_yaw,
                                 float cam_pitch, float light_yaw, float light_pitch, int32_t *cls,
                                 float *angle, float *weight, uint8_t *temporal) {
    const float nd = maxf(dispersion_evaluate(m, 589.3f), 1.1f);
    const float n_f = maxf(dispersion_evaluate(m, 486.1f), 1.001f);
    const float n_c = maxf(dispersion_evaluate(m, 656.3f), 1.001f);
    V3 fwd, right, up;
    view_basis(cam_yaw, cam_pitch, &fwd, &right, &up);
    Rig rig = studio_rig(light_yaw, light_pitch);
    const float sin_lp = tf_sin(light_pitch);
    V3 tf[5], tr[5], tu[5];
    for (int k = 0; k < 5; ++k) view_basis(cam_yaw + to_radians(kYawOffsetsDeg[k]), cam_pitch, &tf[k], &tr[k], &tu[k]);
    for (int ix = 0; ix < kGrid; ++ix) {
        for (int iz = 0; iz < kGrid; ++iz) {
            const int c = ix * kGrid + iz;
            const float u = grid_coord(ix), v = grid_coord(iz);
            for (int k = 0; k < 5; ++k) {
                cls[5 * c + k] = OR_MX_NONE;
                angle[5 * c + k] = 0.0f;
                weight[5 * c + k] = 0.0f;
                temporal[5 * c + k] = 0;
            }
            if (fma_(v, v, u * u) > 0.70f) continue;
            for (int k = 0; k < 5; ++k) {
                V3 dir = normalize(add(add(fwd, mul(right, kAperture[k][0])), mul(up, kAperture[k][1])));
                V3 org = add(add(mul(neg(fwd), 2.5f), mul(right, u * 0.95f)), mul(up, v * 0.95f));
                Hit h = intersect_planes(planes, n_planes, org, dir);
                if (!h.hit) continue;
                V3 entry = add(org, mul(dir, h.t));
                float cos_i = clampf(dot(neg(dir), h.normal), 0.0f, 1.0f);
                Fate fd = trace_wavelength(entry, dir, h.normal, cos_i, planes, n_planes, nd);
                int out = OR_MX_EXTINCT;
                if (fd.kind == FATE_BLOCKED) {
                    out = OR_MX_BLOCKED;
                } else if (fd.kind == FATE_LEAKED) {
                    out = OR_MX_WINDOWED;
                } else if (fd.kind == FATE_UP && visibly_returned(fd.dir, fwd, rig.key, rig.fill, sin_lp)) {
                    out = OR_MX_RETURNED;
                    Fate ff = trace_wavelength(entry, dir, h.normal, cos_i, planes, n_planes, n_f);
                    Fate fc = trace_wavelength(entry, dir, h.normal, cos_i, planes, n_planes, n_c);
                    if (ff.kind == FATE_UP && fc.kind == FATE_UP && ff.facet == fc.facet &&
                        ff.bounces == fc.bounces) {
                        float cos_sep = clampf(dot(ff.dir, fc.dir), -1.0f, 1.0f);
                        angle[5 * c + k] = acos_cr(cos_sep) * 57.2957795130823208767981548141051703f;
                        float energy = fd.transmittance * ff.transmittance * fc.transmittance;
                        float radiance = maxf(ff.exit_cos, 0.0f) * maxf(fc.exit_cos, 0.0f);
                        weight[5 * c + k] = energy * radiance;
                        out = OR_MX_RETURNED_FIRE;
                    }
                }
                cls[5 * c + k] = out;
            }
            for (int k = 0; k < 5; ++k) {
                V3 org = add(add(mul(neg(tf[k]), 2.5f), mul(tr[k], u * 0.95f)), mul(tu[k], v * 0.95f));
                Hit h = intersect_planes(planes, n_planes, org, tf[k]);
                if (!h.hit) continue;
                V3 entry = add(org, mul(tf[k], h.t));
                float cos_i = clampf(dot(neg(tf[k]), h.normal), 0.0f, 1.0f);
                Fate fd = trace_wavelength(entry, tf[k], h.normal, cos_i, planes, n_planes, nd);
                temporal[5 * c + k] = fd.kind == FATE_UP && visibly_returned(fd.dir, tf[k], rig.key, rig.fill, sin_lp);
            }
        }
    }
}

// The accumulation and final metrics of evaluate_gem_optical_metrics (metrics.rs:1082-1248,
// spatial/temporal/combined scintillation :366-450) over records in or_metrics_items'
// layout, in the reference's (ix, iz, sample) order.
extern "C" void or_metrics_reduce(const int32_t *cls, const float *angle, const float *weight,
                                  const uint8_t *temporal, float out[5]) {
    uint32_t total = 0, windowed = 0, extinct = 0, returned = 0, cell_count = 0;
    float fire_sum = 0.0f, cfs = 0.0f, cfs2 = 0.0f, tvs = 0.0f;
    for (int c = 0; c < kGrid * kGrid; ++c) {
        uint32_t cell_total = 0, cell_returned = 0;
        for (int k = 0; k < 5; ++k) {
            const int r = cls[5 * c + k];
            if (r == OR_MX_NONE) continue;
            ++total;
            ++cell_total;
            if (r == OR_MX_WINDOWED) ++windowed;
            else if (r == OR_MX_EXTINCT) ++extinct;
            else if (r == OR_MX_RETURNED || r == OR_MX_RETURNED_FIRE) {
                ++returned;
                ++cell_returned;
                if (r == OR_MX_RETURNED_FIRE) fire_sum = fma_(angle[5 * c + k], weight[5 * c + k], fire_sum);
            }
        }
        if (cell_total > 0) {
            float frac = (float)cell_returned / (float)cell_total;
            cfs += frac;
            cfs2 = fma_(frac, frac, cfs2);
            ++cell_count;
            uint32_t tr = 0;
            for (int k = 0; k < 5; ++k) tr += temporal[5 * c + k] ? 1u : 0u;
            float p = (float)tr / 5.0f;
            tvs += fma_(p, -p, p);
        }
    }
    const float n_total = (float)(total > 1u ? total : 1u);
    const float windowing = clampf((float)windowed / n_total * 100.0f, 0.0f, 100.0f);
    const float extinction = clampf((float)extinct / n_total * 100.0f, 0.0f, 100.0f);
    const float brilliance = clampf((float)returned / n_total * 100.0f, 0.0f, 100.0f);
    const float fire = maxf(fire_sum / n_total * 275.0f, 0.1f);
    float spatial = 0.0f;
    if (cell_count > 0) {
        float mean = cfs / (float)cell_count;
        if (mean > 1e-4f) {
            float var = maxf(fma_(mean, -mean, cfs2 / (float)cell_count), 0.0f);
            float cv = std::sqrt(var) / mean;
            spatial = clampf(cv / (1.0f + cv) * 100.0f, 0.0f, 100.0f);
        }
    }
    float temporal_pct = 0.0f;
    if (cell_count > 0) temporal_pct = clampf((tvs / (float)cell_count) / 0.24f * 100.0f, 0.0f, 100.0f);
    const float scint = clampf(fma_(0.4f, temporal_pct, 0.6f * spatial), 0.0f, 100.0f);
    out[0] = brilliance;
    out[1] = fire;
    out[2] = scint;
    out[3] = windowing;
    out[4] = extinction;
}

// evaluate_gem_optical_metrics (metrics.rs:1082-1248): out = brilliance_pct, fire_index,
// scintillation_pct, windowing_pct, extinction_pct.
extern "C" void or_gem_optical_metrics(const float *planes, int n_planes, const or_material *m, float cam_yaw,
                                       float cam_pitch, float light_yaw, float light_pitch, float out[5]) {
    if (n_planes == 0) {
        out[0] = 85.0f;
        out[1] = 25.0f;
        out[2] = 75.0f;
        out[3] = 5.0f;
        out[4] = 5.0f;
        return;
    }
    const int n = 5 * kGrid * kGrid;
    std::vector<int32_t> cls(n);
    std::vector<float> angle(n), weight(n);
    std::vector<uint8_t> temporal(n);
    or_metrics_items(planes, n_planes, m, cam_yaw, cam_pitch, light_yaw, light_pitch, cls.data(), angle.data(),
                     weight.data(), temporal.data());
    or_metrics_reduce(cls.data(), angle.data(), weight.data(), temporal.data(), out);
}
