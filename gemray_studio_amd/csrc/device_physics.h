// device_physics.h -- gfx950 device functions of the spectral transport estimator.
//
// The arithmetic follows gemray's CPU tracer (crates/gemray/src/optics/raytracer/*.rs and
// optics/*.rs): every `mul_add` of the reference is an explicit fmaf here and nothing
// else is fused (the library is compiled with -ffp-contract=off), dot products are
// left-associative, cross products use glam's operand order, and normalize is
// v * (1 / sqrt(dot(v, v))).  Division and sqrt are correctly rounded (HIP's default for
// f32).  Mueller products skip the structural zeros of each matrix: that changes at most
// the sign of a zero Stokes component, which no output depends on.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gr {

constexpr float kPi = 3.14159265358979323846f;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ float len2(V3 a) { return dot(a, a); }
__device__ __forceinline__ V3 normalize(V3 a) { return mul(a, 1.0f / sqrtf(dot(a, a))); }
// f32::acos correctly rounded: the f64 arc cosine rounded once to f32.
__device__ __forceinline__ float acos_cr(float x) { return (float)acos((double)x); }
__device__ __forceinline__ V3 normalize_or_zero(V3 a) {
    float rcp = 1.0f / sqrtf(dot(a, a));
    if (isfinite(rcp) && rcp > 0.0f) return mul(a, rcp);
    return v3(0.0f, 0.0f, 0.0f);
}

// cosine_weighted_hemisphere with frosted_orthonormal_basis
// (optics/raytracer/scattering.rs:339-355): a cosine-distributed direction about n.  Out
// of line: only frosted facets call it, and its sincos would otherwise add to the
// register pressure of every bounce.
__device__ __noinline__ V3 cosine_weighted_hemisphere(float u1, float u2, V3 n) {
    const float r = sqrtf(u1);
    const float theta = (2.0f * 3.14159265358979323846f) * u2;
    float sin_t, cos_t;
    sincosf(theta, &sin_t, &cos_t);
    const V3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    const V3 t = normalize_or_zero(sub(a, mul(n, dot(n, a))));
    const V3 b = cross(n, t);
    return normalize_or_zero(add(add(mul(t, r * cos_t), mul(b, r * sin_t)), mul(n, sqrtf(fmaxf(1.0f - u1, 0.0f)))));
}

// f32::clamp keeps NaN; fmaxf/fminf are IEEE maxNum/minNum like f32::max/min.
__device__ __forceinline__ float clampf(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
__device__ __forceinline__ float rem_euclid(float x, float rhs) {
    float r = fmodf(x, rhs);
    return r < 0.0f ? r + fabsf(rhs) : r;
}
// x^n, n a small positive constant, rounded once: the product of the set-bit squares
// is formed in f64 (at most 5 roundings of 2^-53 each), so the single f32 rounding at
// the end is the correctly rounded power except within ~1e-15 of a rounding boundary.
template <unsigned N>
__device__ __forceinline__ float powi_cr(float x) {
    double sq = (double)x;
    double res = 1.0;
    unsigned n = N;
    while (n) {
        if (n & 1u) res *= sq;
        n >>= 1;
        if (n) sq *= sq;
    }
    return (float)res;
}

// ---- Stokes vectors and the Mueller matrices of optics/polarization.rs:95-147 --------

struct Stokes {
    float i, q, u, v;
};
__device__ __forceinline__ Stokes scale(Stokes s, float k) {
    return Stokes{s.i * k, s.q * k, s.u * k, s.v * k};
}
__device__ __forceinline__ float intensity(const Stokes &s) { return fmaxf(s.i, 0.0f); }

// The hardware square root, reciprocal square root and reciprocal (v_sqrt_f32, v_rsq_f32,
// v_rcp_f32: within 1 ulp each) serve the quantities that only scale an absorption
// coefficient -- never a direction, a clip t or a branch of the path, which keep the
// correctly rounded sqrtf and division.  sqrt_hw scales an argument below 2^-64 by 2^64
// first (exactly, and the root back by 2^-32), so the instruction never sees a denormal
// or the small range where its error grows.
__device__ __forceinline__ float sqrt_hw(float x) {  // x >= 0
    const bool tiny = x < 0x1p-64f;
    return __builtin_amdgcn_sqrtf(tiny ? x * 0x1p64f : x) * (tiny ? 0x1p-32f : 1.0f);
}
// a / b as a times the hardware reciprocal (within 2 ulp), for the same intensity-only uses.
__device__ __forceinline__ float div_hw(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// 1 / sqrt(x) for finite x > 0, scaled like sqrt_hw.
__device__ __forceinline__ float rsqrt_hw(float x) {
    const bool tiny = x < 0x1p-64f;
    return __builtin_amdgcn_rsqf(tiny ? x * 0x1p64f : x) * (tiny ? 0x1p32f : 1.0f);
}
// normalize and normalize_or_zero for vectors that only orient a polarization frame or an
// absorption axis: the same results to within 2 ulp per component.  normalize_or_zero_hw
// returns the zero vector exactly where normalize_or_zero does (|a|^2 zero, infinite or
// NaN).
__device__ __forceinline__ V3 normalize_hw(V3 a) { return mul(a, rsqrt_hw(dot(a, a))); }
__device__ __forceinline__ V3 normalize_or_zero_hw(V3 a) {
    const float d = dot(a, a);
    if (d > 0.0f && d < __builtin_huge_valf()) return mul(a, rsqrt_hw(d));
    return v3(0.0f, 0.0f, 0.0f);
}
// (cos 2psi, sin 2psi) = (Q, U) / |(Q, U)| with the hardware reciprocal square root, on
// (Q, U) scaled by the power of two that brings the larger magnitude into [0.5, 1) (exact),
// so neither the squares nor the instruction meet an underflow or a denormal.  False when
// Q = U = 0 (or NaN): psi = 0.
__device__ __forceinline__ bool unit_qu(float q, float u, float &c2, float &s2) {
    const float m = __uint_as_float(max(__float_as_uint(q) & 0x7FFFFFFFu, __float_as_uint(u) & 0x7FFFFFFFu));
    if (!(m > 0.0f)) return false;
    const int e = __builtin_amdgcn_frexp_expf(m);
    const float qs = __builtin_amdgcn_ldexpf(q, -e), us = __builtin_amdgcn_ldexpf(u, -e);
    const float inv_r = __builtin_amdgcn_rsqf(fmaf(qs, qs, us * us));
    c2 = qs * inv_r;
    s2 = us * inv_r;
    return true;
}

// degree_of_polarization (polarization.rs:38-48): sqrt(Q^2 + U^2 + V^2) / I, clamped to
// [0, 1], with the hardware square root and reciprocal (it only weights the absorption
// coefficient; within 4 ulp of the reference's correctly rounded route).
__device__ __forceinline__ float degree_of_polarization(const Stokes &s) {
    if (s.i <= 1e-7f) return 0.0f;
    return clampf(sqrt_hw(fmaf(s.v, s.v, fmaf(s.u, s.u, s.q * s.q))) * __builtin_amdgcn_rcpf(s.i), 0.0f, 1.0f);
}
// f32::midpoint on x86_64 is computed in f64; for finite normal operands that equals
// (a + b) * 0.5 in f32 (halving commutes with rounding).
__device__ __forceinline__ float midpoint(float a, float b) { return (a + b) * 0.5f; }

// Mueller for a diagonal-block matrix [[a,b,0,0],[b,a,0,0],[0,0,c,0],[0,0,0,c]].
__device__ __forceinline__ Stokes apply_abc(Stokes s, float a, float b, float c) {
    return Stokes{a * s.i + b * s.q, b * s.i + a * s.q, c * s.u, c * s.v};
}
// frame_rotation(psi) with c2 = cos 2psi, s2 = sin 2psi.
__device__ __forceinline__ Stokes apply_rotation(Stokes s, float c2, float s2) {
    return Stokes{s.i, c2 * s.q + s2 * s.u, (-s2) * s.q + c2 * s.u, s.v};
}
// tir_retardation(delta) with cd = cos delta, sd = sin delta.
__device__ __forceinline__ Stokes apply_retardation(Stokes s, float cd, float sd) {
    return Stokes{s.i, s.q, cd * s.u + sd * s.v, (-sd) * s.u + cd * s.v};
}
__device__ __forceinline__ Stokes apply_fresnel_reflection(Stokes s, float r_s, float r_p) {
    float rs2 = r_s * r_s;
    float rp2 = r_p * r_p;
    return apply_abc(s, midpoint(rs2, rp2), 0.5f * (rs2 - rp2), r_s * r_p);
}
// The transmission factor (n2 cos_t) / (n1 cos_i) on the hardware reciprocal (div_hw).
__device__ __forceinline__ Stokes apply_fresnel_transmission(Stokes s, float n1, float n2,
                                                             float cos_i, float cos_t,
                                                             float t_s, float t_p) {
    float factor = div_hw(n2 * cos_t, fmaxf(n1 * cos_i, 1e-6f));
    float ts2 = t_s * t_s * factor;
    float tp2 = t_p * t_p * factor;
    return apply_abc(s, midpoint(ts2, tp2), 0.5f * (ts2 - tp2), t_s * t_p * factor);
}

// e^x on the hardware exponential (HIP's __expf: v_exp_f32 of x log2 e).  The reference's
// CPU tracer evaluates its Gaussian lobes and the Beer-Lambert factor with a Cephes
// polynomial (exp_f32x8, simd/exp_poly.rs:8-40) and its GPU shader with the native exp;
// this is the latter, within 2.4e-7 + |x| 1.2e-7 relative of e^x (the argument's rounding
// dominates for large |x|, where the lobes are negligible).
__device__ __forceinline__ float exp_lobe(float x) { return __expf(x); }

// ---- CIE 1931 analytic fit, batched form (color/cie1931.rs:48-98) ---------------------

// x / c for a divisor fixed at compile time: q0 = x * rc with rc = RN(1/c), the exact
// remainder r = x - q0 c (one fma), then q0 + r rc (one fma).  For the Gaussian lobes'
// constant widths this equals the correctly rounded x / c for every comb wavelength in
// [380, 780] nm (checked exhaustively on the host: tests/test_div_const.py), at three
// instructions instead of the division sequence.
__device__ __forceinline__ float div_by_const(float x, float c, float rc) {
    const float q0 = x * rc;
    const float r = fmaf(-q0, c, x);
    return fmaf(r, rc, q0);
}

__device__ __forceinline__ float cmf_lobe(float l, float mu, float lo, float hi) {
    const bool below = l < mu;
    float t = div_by_const(l - mu, below ? lo : hi, below ? 1.0f / lo : 1.0f / hi);
    return exp_lobe(-0.5f * t * t);
}
__device__ __forceinline__ V3 cmf_x8_channel(float l) {
    float x0 = cmf_lobe(l, 501.1f, 20.4f, 26.2f);
    float x1 = cmf_lobe(l, 599.8f, 37.9f, 31.0f);
    float x2 = cmf_lobe(l, 442.0f, 16.0f, 26.7f);
    float y0 = cmf_lobe(l, 530.9f, 16.3f, 31.1f);
    float y1 = cmf_lobe(l, 568.8f, 46.9f, 40.5f);
    float z0 = cmf_lobe(l, 459.0f, 26.0f, 13.8f);
    float z1 = cmf_lobe(l, 437.0f, 11.8f, 36.0f);
    float x = fmaf(0.065f, -x0, fmaf(1.056f, x1, 0.362f * x2));
    float y = fmaf(0.286f, y0, 0.821f * y1);
    float z = fmaf(0.681f, z0, 1.217f * z1);
    return v3(fmaxf(x, 0.0f), fmaxf(y, 0.0f), fmaxf(z, 0.0f));
}

// ---- Sampling streams (optics/raytracer/sampling.rs) -----------------------------------

constexpr uint32_t FRESNEL_BRANCH_STREAM = 0x9E3779B1u;
constexpr uint32_t FROSTED_DIR_U_STREAM = 0x27D4EB2Fu;  // sampling.rs:28-29
constexpr uint32_t FROSTED_DIR_V_STREAM = 0x165667B1u;
constexpr uint32_t RUSSIAN_ROULETTE_STREAM = 0x517CC1B7u;
constexpr uint32_t BIREFRINGENT_SPLIT_STREAM = 0x2545F491u;
constexpr uint32_t MODE_COUPLING_STREAM = 0xCC9E2D51u;
constexpr uint32_t PIXEL_JITTER_X_ROTATION_STREAM = 0xA511E9B3u;
constexpr uint32_t PIXEL_JITTER_Y_ROTATION_STREAM = 0x63D81B23u;
constexpr uint32_t HERO_WAVELENGTH_ROTATION_STREAM = 0x1B873593u;
constexpr uint32_t DISTANCE_SAMPLE_STREAM = 0xA24BAED4u;  // sampling.rs:30-41
constexpr uint32_t PHASE_DIR_U_STREAM = 0x9FB21C65u;
constexpr uint32_t PHASE_DIR_V_STREAM = 0x1CE4E5B9u;

// sample_henyey_greenstein_direction (optics/raytracer/scattering.rs:428-447): the
// Henyey-Greenstein cosine about `forward` by inversion of its CDF (isotropic for
// |g| < 1e-3), azimuth 2 pi u2, on frosted_orthonormal_basis(forward).  Out of line, like
// cosine_weighted_hemisphere: only scattering stones call it.
__device__ __noinline__ V3 sample_henyey_greenstein_direction(float u1, float u2, float g, V3 forward) {
    float cos_theta;
    if (fabsf(g) < 1e-3f) {
        cos_theta = fmaf(2.0f, -u1, 1.0f);
    } else {
        const float one_minus_g2 = fmaf(g, -g, 1.0f);
        const float denom = fmaf(2.0f * g, u1, 1.0f - g);
        const float sq = one_minus_g2 / denom;
        cos_theta = fmaf(sq, -sq, fmaf(g, g, 1.0f)) / (2.0f * g);
    }
    cos_theta = clampf(cos_theta, -1.0f, 1.0f);
    const float sin_theta = sqrtf(fmaxf(fmaf(cos_theta, -cos_theta, 1.0f), 0.0f));
    const float phi = (2.0f * 3.14159265358979323846f) * u2;
    float sin_p, cos_p;
    sincosf(phi, &sin_p, &cos_p);
    const V3 a = fabsf(forward.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    const V3 t = normalize_or_zero(sub(a, mul(forward, dot(forward, a))));
    const V3 b = cross(forward, t);
    return normalize_or_zero(
        add(add(mul(t, sin_theta * cos_p), mul(b, sin_theta * sin_p)), mul(forward, cos_theta)));
}

__host__ __device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ float unit_draw(uint32_t h) { return (float)h / 4294967295.0f; }
__device__ __forceinline__ float stream_draw(uint32_t seed, uint32_t bounce, uint32_t stream) {
    return unit_draw(hash_u32(seed ^ hash_u32(bounce ^ stream)));
}
__device__ __forceinline__ float low_discrepancy_base2(uint32_t n) {
    return (float)__builtin_bitreverse32(n) / 4294967296.0f;
}
__device__ __forceinline__ float radical_inverse(uint32_t n, uint32_t base) {
    uint32_t remaining = n;
    float val = 0.0f;
    float inv_base = 1.0f / (float)base;
    while (remaining > 0) {
        uint32_t digit = remaining % base;
        val = fmaf((float)digit, inv_base, val);
        inv_base /= (float)base;
        remaining /= base;
    }
    return val;
}
__device__ __forceinline__ float cranley_patterson(float x, float offset) {
    float sum = x + offset;
    return sum - floorf(sum);
}

// ---- Uniaxial birefringence (optics/birefringence.rs:27-131) ---------------------------

// effective_extraordinary_index (birefringence.rs:27-40) with sin and cos of theta supplied
// by the caller (shared across channels).
__device__ __forceinline__ float effective_extraordinary_index_sc(float n_o, float n_e,
                                                                  float sin_t, float cos_t) {
    if (fabsf(n_o - n_e) < 1e-5f) return n_o;
    float os = n_o * sin_t;
    float denom_sq = fmaf(n_e * cos_t, n_e * cos_t, os * os);
    if (denom_sq <= 1e-8f) return n_o;
    return (n_o * n_e) / sqrtf(denom_sq);
}
// effective_extraordinary_index(n_o, n_e, acos(|cos theta|)) (refraction.rs:248-269,
// birefringence.rs:27-40) from the cosine: sin theta = sqrt(1 - cos^2).
__device__ __forceinline__ void theta_sc_from_cos(float cos_theta, float *sin_t, float *cos_t) {
    const float ct = fabsf(clampf(cos_theta, -1.0f, 1.0f));
    *cos_t = ct;
    *sin_t = sqrtf(fmaxf(fmaf(ct, -ct, 1.0f), 0.0f));
}
__device__ __forceinline__ float neff_from_cos(float n_o, float n_e, float cos_theta) {
    float st, ct;
    theta_sc_from_cos(cos_theta, &st, &ct);
    return effective_extraordinary_index_sc(n_o, n_e, st, ct);
}

// MuellerMatrix::frame_rotation(signed_frame_rotation_psi(prev, cur, axis))
// (raytracer/absorption.rs:44-54, polarization.rs:95-103) needs cos 2psi and sin 2psi
// only: the double-angle forms of the (unnormalized) pair (cos psi, sin psi) replace
// atan2 + sincos.
__device__ __forceinline__ void frame_rotation_cs(V3 prev, V3 cur, V3 axis, float *c2, float *s2) {
    const float cos_psi = clampf(dot(prev, cur), -1.0f, 1.0f);
    const float sin_psi = dot(cross(prev, cur), normalize_or_zero(axis));
    const float cc = cos_psi * cos_psi, ss = sin_psi * sin_psi;
    const float r2 = cc + ss;
    *c2 = 1.0f;
    *s2 = 0.0f;
    if (r2 > 0.0f) {
        const float inv = 1.0f / r2;
        *c2 = (cc - ss) * inv;
        *s2 = (2.0f * cos_psi * sin_psi) * inv;
    }
}

// The polarization quadratic form e . A . e of pleochroic absorption
// (raytracer/absorption.rs:115-216, birefringence.rs:252-270, polarization.rs:172-192) for
// the field e = cos psi s + sin psi p, psi = atan2(U, Q) / 2, in the absorption tensor's
// principal frame where A = diag(ao, ao, ae).  With S_i, P_i the principal-frame
// components of s and p, and cos^2 psi = (1 + cos 2psi) / 2, sin^2 psi = (1 - cos 2psi) / 2,
// 2 cos psi sin psi = sin 2psi:
//   e.A.e = sum_i w_i (S_i^2 + P_i^2) / 2 + cos 2psi sum_i w_i (S_i^2 - P_i^2) / 2
//           + sin 2psi sum_i w_i S_i P_i,   w = (ao, ao, ae),
// and cos 2psi = Q / r, sin 2psi = U / r with r = sqrt(Q^2 + U^2) (psi = 0 when r = 0;
// unit_qu).
// The six channel-invariant sums are formed once per interior segment.
struct QfInv {
    float h_xy, h_z, d_xy, d_z, x_xy, x_z;
};
__device__ __forceinline__ QfInv qf_invariants(V3 s_l, V3 p_l) {
    QfInv v;
    const float sx = s_l.x * s_l.x, sy = s_l.y * s_l.y, sz = s_l.z * s_l.z;
    const float px = p_l.x * p_l.x, py = p_l.y * p_l.y, pz = p_l.z * p_l.z;
    v.h_xy = ((sx + px) + (sy + py)) * 0.5f;
    v.h_z = (sz + pz) * 0.5f;
    v.d_xy = ((sx - px) + (sy - py)) * 0.5f;
    v.d_z = (sz - pz) * 0.5f;
    v.x_xy = fmaf(s_l.x, p_l.x, s_l.y * p_l.y);
    v.x_z = s_l.z * p_l.z;
    return v;
}
__device__ __forceinline__ float qf_channel(const QfInv &v, float q, float u, float ao, float ae) {
    const float h = fmaf(ao, v.h_xy, ae * v.h_z);
    float c2, s2;
    if (!unit_qu(q, u, c2, s2)) return h + fmaf(ao, v.d_xy, ae * v.d_z);
    const float d = fmaf(ao, v.d_xy, ae * v.d_z);
    const float x = fmaf(ao, v.x_xy, ae * v.x_z);
    return fmaf(c2, d, fmaf(s2, x, h));
}
// The same form for a trichroic tensor, w = (ao, ab, ae): the a1 and a2 sums kept apart.
struct QfInv3 {
    float h_x, h_y, h_z, d_x, d_y, d_z, x_x, x_y, x_z;
};
__device__ __forceinline__ QfInv3 qf_invariants3(V3 s_l, V3 p_l) {
    QfInv3 v;
    const float sx = s_l.x * s_l.x, sy = s_l.y * s_l.y, sz = s_l.z * s_l.z;
    const float px = p_l.x * p_l.x, py = p_l.y * p_l.y, pz = p_l.z * p_l.z;
    v.h_x = (sx + px) * 0.5f;
    v.h_y = (sy + py) * 0.5f;
    v.h_z = (sz + pz) * 0.5f;
    v.d_x = (sx - px) * 0.5f;
    v.d_y = (sy - py) * 0.5f;
    v.d_z = (sz - pz) * 0.5f;
    v.x_x = s_l.x * p_l.x;
    v.x_y = s_l.y * p_l.y;
    v.x_z = s_l.z * p_l.z;
    return v;
}
__device__ __forceinline__ float qf_channel3(const QfInv3 &v, float q, float u, float ao, float ab, float ae) {
    const float h = fmaf(ao, v.h_x, fmaf(ab, v.h_y, ae * v.h_z));
    const float d = fmaf(ao, v.d_x, fmaf(ab, v.d_y, ae * v.d_z));
    float c2, s2;
    if (!unit_qu(q, u, c2, s2)) return h + d;
    const float x = fmaf(ao, v.x_x, fmaf(ab, v.x_y, ae * v.x_z));
    return fmaf(c2, d, fmaf(s2, x, h));
}

// extraordinary_poynting_dir (birefringence.rs:74-100) without the angle round trips.  The
// reference takes theta = acos(|cos theta|), the walk-off delta = atan(tan rho) from sin and
// cos of theta (walk_off_angle, :42-55) and then sin and cos of delta; here
// sin theta = sqrt(1 - cos^2) and cos delta = 1 / sqrt(1 + tan^2 rho),
// sin delta = tan rho cos delta.  |delta| < 1e-5 is tested as |tan rho| < 1e-5 (atan is
// monotonic and atan(x) = x to single precision there).
__device__ __forceinline__ V3 extraordinary_poynting_dir(V3 w, V3 c, float n_o, float n_e) {
    const float cos_theta = clampf(dot(w, c), -1.0f, 1.0f);
    if (fabsf(n_o - n_e) < 1e-5f) return w;
    const float ct = fabsf(cos_theta);
    const float st = sqrtf(fmaxf(fmaf(ct, -ct, 1.0f), 0.0f));
    const float n_o2 = n_o * n_o;
    const float n_e2 = n_e * n_e;
    const float tan_rho = ((n_o2 - n_e2) * st * ct) / fmaxf(fmaf(n_e2 * ct, ct, n_o2 * st * st), 1e-6f);
    if (fabsf(tan_rho) < 1e-5f) return w;
    V3 c_proj = normalize_or_zero(sub(c, mul(w, cos_theta)));
    if (len2(c_proj) < 1e-6f) return w;
    const float sign = cos_theta >= 0.0f ? 1.0f : -1.0f;
    const float cd = 1.0f / sqrtf(fmaf(tan_rho, tan_rho, 1.0f));
    const float sd = tan_rho * cd;
    return normalize(sub(mul(w, cd), mul(mul(c_proj, sign), sd)));
}
__device__ __forceinline__ void stable_orthonormal_basis(V3 n, V3 *t_out, V3 *b_out) {
    V3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    V3 t = normalize_or_zero(sub(a, mul(n, dot(n, a))));
    *t_out = t;
    *b_out = cross(n, t);
}
// ordinary_eigen_polarization and extraordinary_eigen_polarization (birefringence.rs:101-118)
// as the absorption step uses them: the ordinary mode along w x c (a stable perpendicular of
// w when w is along the axis), the extraordinary one along w x e_o.  They only orient the
// absorption tensor's frame, so the normalizations are the hardware's (normalize_hw).
__device__ __forceinline__ void uniaxial_eigen_polarizations_hw(V3 w, V3 c, V3 *eo, V3 *ee) {
    const V3 cr = cross(w, c);
    V3 o;
    if (len2(cr) > 1e-8f) {
        o = normalize_hw(cr);
    } else {
        V3 b;
        stable_orthonormal_basis(normalize_or_zero(w), &o, &b);
    }
    *eo = o;
    *ee = normalize_or_zero_hw(cross(w, o));
}

// polarization.rs:172-200.
__device__ __forceinline__ V3 arbitrary_perpendicular(V3 n) {
    V3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    return normalize_or_zero(sub(a, mul(n, dot(n, a))));
}

// refraction.rs:45-50: delta = 2 atan(x) with x = cos_i sqrt(A) / B the tangent of the half
// retardation, A = n1^2 sin_i^2 - 1, B = n1 sin_i^2.  The retardation matrix
// (polarization.rs:141-147) needs only cos delta and sin delta, the rational functions
// (1 - x^2) / (1 + x^2) and 2x / (1 + x^2) -- with numerator and denominator times B^2,
// (B^2 - cos_i^2 A) / (B^2 + cos_i^2 A) and 2 cos_i sqrt(A) B / (B^2 + cos_i^2 A): one
// reciprocal and one square root.  They only rotate a channel's Stokes vector, so both are
// the hardware's (v_rcp_f32, sqrt_hw).
__device__ __forceinline__ void tir_retardation_cs(float n1k, float cos_i, float sin_i,
                                                   float *cos_d, float *sin_d) {
    const float a = fmaxf(fmaf(n1k * n1k * sin_i, sin_i, -1.0f), 0.0f);
    const float b = fmaxf(n1k * sin_i * sin_i, 1e-6f);
    const float ca = (cos_i * cos_i) * a;
    const float b2 = b * b;
    const float inv = __builtin_amdgcn_rcpf(b2 + ca);
    *cos_d = (b2 - ca) * inv;
    *sin_d = ((2.0f * cos_i) * sqrt_hw(a)) * (b * inv);
}

__device__ __forceinline__ void fresnel_rs_rp(float n1k, float n2k, float cos_i, float cos_t,
                                              float *r_s, float *r_p) {
    *r_s = fmaf(n2k, -cos_t, n1k * cos_i) / fmaf(n2k, cos_t, n1k * cos_i);
    *r_p = fmaf(n1k, -cos_t, n2k * cos_i) / fmaf(n1k, cos_t, n2k * cos_i);
}
// The same amplitudes on the hardware reciprocal, for the channels whose amplitudes only
// scale their own Stokes vector and sampling weight (every channel but the hero's, whose
// r_unpol decides the reflect/refract draw).
__device__ __forceinline__ void fresnel_rs_rp_hw(float n1k, float n2k, float cos_i, float cos_t,
                                                 float *r_s, float *r_p) {
    *r_s = div_hw(fmaf(n2k, -cos_t, n1k * cos_i), fmaf(n2k, cos_t, n1k * cos_i));
    *r_p = div_hw(fmaf(n1k, -cos_t, n2k * cos_i), fmaf(n1k, cos_t, n2k * cos_i));
}

}  // namespace gr
