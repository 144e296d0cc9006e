// biaxial.h -- the biaxial indicatrix on the device (optics/birefringence.rs:139-170,
// 349-870) for the trace kernels built for biaxial stones.
//
// A biaxial material's indicatrix at wavelength lambda (materials.rs:1508-1519) has
// n_beta = n(lambda) (the dispersion curve), n_alpha = n_beta - biaxial_delta_beta_alpha,
// n_gamma = n_alpha + birefringence_delta, and principal axes (alpha, beta, gamma) =
// (a1, a2, c): c the normalized c axis and (a1, a2) its stable orthonormal completion --
// the scene's absorption frame, which is the same construction on the same input
// (birefringence.rs:386-391, 238-245).  Every channel shares the frame; only the three
// indices change with the channel's wavelength.
//
// Operation order follows the reference (explicit fmaf for each mul_add, nothing else
// fused).  The one departure: the uniaxial short-cut for a degenerate index pair takes
// the sine and cosine of theta from its cosine, as the uniaxial kernels do, instead of
// acos followed by sin and cos.
#pragma once

#include <hip/hip_runtime.h>

#include "device_physics.h"

namespace gr {

// The principal frame: columns alpha, beta, gamma.
struct BxFrame {
    V3 ax, ay, az;
};
// axes^T v: lane j is dot(axis j, v).
__device__ __forceinline__ V3 bx_local(const BxFrame &f, V3 v) {
    return v3(dot(f.ax, v), dot(f.ay, v), dot(f.az, v));
}
// axes v: ((ax v.x) + (ay v.y)) + (az v.z) per lane.
__device__ __forceinline__ V3 bx_world(const BxFrame &f, V3 v) {
    return add(add(mul(f.ax, v.x), mul(f.ay, v.y)), mul(f.az, v.z));
}

// One channel's principal indices.
struct Bx {
    float na, nb, ng;
};
__device__ __forceinline__ Bx bx_of(float n_beta, float delta_beta_alpha, float birefringence_delta) {
    Bx b;
    b.nb = n_beta;
    b.na = n_beta - delta_beta_alpha;
    b.ng = b.na + birefringence_delta;
    return b;
}

// indices_are_degenerate (birefringence.rs:519-522): |a - b| <= sqrt(f32::EPSILON) * scale.
__device__ __forceinline__ bool bx_degenerate(float a, float b) {
    const float scale = fmaxf(fmaxf(fabsf(a), fabsf(b)), 1.0f);
    return fabsf(a - b) <= 3.45266983e-4f * scale;
}

// wave_indices (birefringence.rs:435-487) for the wave normal whose principal-frame
// direction cosines are l (= axes^T normalize_or_zero(w)): (n_slow, n_fast).
__device__ __forceinline__ void bx_wave_indices(const Bx &b, V3 l, float *n_slow, float *n_fast) {
    const bool ab = bx_degenerate(b.na, b.nb), bg = bx_degenerate(b.nb, b.ng);
    if (ab && bg) {
        const float n = (b.na + b.nb + b.ng) / 3.0f;
        *n_slow = n;
        *n_fast = n;
        return;
    }
    if (ab || bg) {
        // uniaxial_wave_indices (:495-499): optic axis gamma (ab) or alpha (bg).
        const float n_o = ab ? midpoint(b.na, b.nb) : midpoint(b.nb, b.ng);
        const float n_e = ab ? b.ng : b.na;
        const float n_eff = neff_from_cos(n_o, n_e, ab ? l.z : l.x);
        *n_slow = fmaxf(n_o, n_eff);
        *n_fast = fminf(n_o, n_eff);
        return;
    }
    const float a2 = l.x * l.x, b2 = l.y * l.y, g2 = l.z * l.z;
    const float b1 = 1.0f / (b.na * b.na), b2c = 1.0f / (b.nb * b.nb), b3 = 1.0f / (b.ng * b.ng);
    const float big_b = fmaf(a2, b2c + b3, fmaf(b2, b1 + b3, g2 * (b1 + b2c)));
    const float big_c = fmaf(a2, b2c * b3, fmaf(b2, b1 * b3, g2 * b1 * b2c));
    const float disc = sqrtf(fmaxf(fmaf(big_b, big_b, -4.0f * big_c), 0.0f));
    const float x_lo = 0.5f * (big_b - disc);
    const float x_hi = midpoint(big_b, disc);
    *n_slow = 1.0f / sqrtf(fmaxf(x_lo, 1e-12f));
    *n_fast = 1.0f / sqrtf(fmaxf(x_hi, 1e-12f));
}

// precise_root_near (birefringence.rs:728-756): the root of the same quadratic nearest x,
// its discriminant from differences of the principal 1/n^2.
__device__ __forceinline__ float bx_precise_root_near(V3 l, float p, float q, float r, float x) {
    const float a = l.x * l.x, bb = l.y * l.y, cc = l.z * l.z;
    const float big_b = fmaf(a, q + r, fmaf(bb, p + r, cc * (p + q)));
    const float xdiff = p - q;
    const float ydiff = r - p;
    const float a_plus_c = a + cc;
    const float a_plus_bb = a + bb;
    const float two_a_minus_bc = 2.0f * fmaf(bb, -cc, a);
    const float disc_sq = fmaf(a_plus_c * a_plus_c, xdiff * xdiff,
                               fmaf(a_plus_bb * a_plus_bb, ydiff * ydiff, two_a_minus_bc * xdiff * ydiff));
    const float disc = sqrtf(fmaxf(disc_sq, 0.0f));
    const float x_lo = 0.5f * (big_b - disc);
    const float x_hi = midpoint(big_b, disc);
    return fabsf(x - x_lo) <= fabsf(x - x_hi) ? x_lo : x_hi;
}

__device__ __forceinline__ V3 bx_cross_fma(V3 a, V3 b) {
    return v3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ float bx_dot_fma(V3 a, V3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// eigenvector_world (birefringence.rs:659-700): the D eigenvector of root x of the mode at
// k_hat (l its principal-frame cosines): the null vector of Gamma - x I, Gamma = P B P, as
// the sign-aligned sum of its row-pair cross products, largest component made positive.
__device__ __forceinline__ V3 bx_eigenvector(const BxFrame &f, V3 l, float p, float q, float r, float x_in,
                                             V3 k_hat) {
    const float x = bx_precise_root_near(l, p, q, r, x_in);
    const float s = fmaf(p, l.x * l.x, fmaf(q, l.y * l.y, r * l.z * l.z));
    const float m00 = fmaf(l.x * l.x, fmaf(2.0f, -p, s), p - x);
    const float m11 = fmaf(l.y * l.y, fmaf(2.0f, -q, s), q - x);
    const float m22 = fmaf(l.z * l.z, fmaf(2.0f, -r, s), r - x);
    const float m01 = l.x * l.y * (s - p - q);
    const float m02 = l.x * l.z * (s - p - r);
    const float m12 = l.y * l.z * (s - q - r);
    const V3 row0 = v3(m00, m01, m02), row1 = v3(m01, m11, m12), row2 = v3(m02, m12, m22);
    const V3 c01 = bx_cross_fma(row0, row1);
    const V3 c02 = bx_cross_fma(row0, row2);
    const V3 c12 = bx_cross_fma(row1, row2);
    const float sign02 = bx_dot_fma(c01, c02) < 0.0f ? -1.0f : 1.0f;
    const float sign12 = bx_dot_fma(c01, c12) < 0.0f ? -1.0f : 1.0f;
    const V3 v_local = add(add(c01, mul(c02, sign02)), mul(c12, sign12));
    V3 v = bx_world(f, v_local);
    if (len2(v) > 1e-12f) {
        // canonicalize_eigenvector_sign (:139-150), then normalize.
        const float ax = fabsf(v.x), ay = fabsf(v.y), az = fabsf(v.z);
        const float largest = (ax >= ay && ax >= az) ? v.x : (ay >= az ? v.y : v.z);
        if (largest < 0.0f) v = neg(v);
        return normalize(v);
    }
    V3 t, b;
    stable_orthonormal_basis(k_hat, &t, &b);
    return t;
}

// eigen_polarizations (birefringence.rs:542-553) of channel b at wave normal w:
// (d_slow, d_fast).
__device__ __noinline__ void bx_eigen_polarizations(const Bx &b, const BxFrame &f, V3 w, V3 *d_slow,
                                                    V3 *d_fast) {
    const V3 k = normalize_or_zero(w);
    const V3 l = bx_local(f, k);
    float ns, nf;
    bx_wave_indices(b, l, &ns, &nf);
    const float p = 1.0f / (b.na * b.na), q = 1.0f / (b.nb * b.nb), r = 1.0f / (b.ng * b.ng);
    const float x_slow = 1.0f / (ns * ns);
    const float x_fast = 1.0f / (nf * nf);
    *d_slow = bx_eigenvector(f, l, p, q, r, x_slow, k);
    *d_fast = bx_eigenvector(f, l, p, q, r, x_fast, k);
}

// mode_poynting_dir (birefringence.rs:831-836): the walked-off energy direction of one
// mode at wave normal w (want_slow: the n_slow root) -- eigen_polarizations (of which only
// the requested eigenvector is formed), d_to_e_direction (:763-771) and
// poynting_direction (:855-865).
__device__ __noinline__ V3 bx_mode_poynting_dir(const Bx &b, const BxFrame &f, V3 w, bool want_slow) {
    const V3 k = normalize_or_zero(w);
    const V3 l = bx_local(f, k);
    float ns, nf;
    bx_wave_indices(b, l, &ns, &nf);
    const float na2 = b.na * b.na, nb2 = b.nb * b.nb, ng2 = b.ng * b.ng;
    const float n = want_slow ? ns : nf;
    const V3 d = bx_eigenvector(f, l, 1.0f / na2, 1.0f / nb2, 1.0f / ng2, 1.0f / (n * n), k);
    const V3 dl = bx_local(f, d);
    const V3 e = normalize_or_zero(bx_world(f, v3(dl.x / na2, dl.y / nb2, dl.z / ng2)));
    const V3 eh = normalize_or_zero(e);
    const V3 perp = sub(k, mul(eh, dot(eh, k)));
    const float p2 = len2(perp);
    if (p2 > 1e-10f) {
        const float len = sqrtf(p2);
        return v3(perp.x / len, perp.y / len, perp.z / len);
    }
    return k;
}

// resolve_entry_mode (birefringence.rs:795-818): the refracted wave normal of one mode at
// an air -> crystal entry, by two fixed-point steps from the seed index.
__device__ __forceinline__ V3 bx_entry_wave_dir(const Bx &b, const BxFrame &f, V3 incident, V3 normal,
                                                float cos_i, float one_m_c2, float n_seed, bool want_slow) {
    float n_guess = n_seed;
    V3 wave = incident;
    for (int it = 0; it < 2; ++it) {
        const float eta = 1.0f / n_guess;
        const float sin2 = eta * eta * one_m_c2;
        if (sin2 > 1.0f) break;
        const float cos_t = sqrtf(fmaxf(1.0f - sin2, 0.0f));
        wave = normalize(add(mul(incident, eta), mul(normal, fmaf(eta, cos_i, -cos_t))));
        float ns, nf;
        bx_wave_indices(b, bx_local(f, normalize_or_zero(wave)), &ns, &nf);
        n_guess = want_slow ? ns : nf;
    }
    return wave;
}

}  // namespace gr
