// transport.hip -- the gfx950 spectral transport megakernel and its film reduction.
//
// One lane traces one (pixel, absolute sample) tuple: the full estimator of
// gemray's transport_main (crates/gemray/src/renderer/shaders/spectral_transport.wgsl:672)
// whose CPU statement is trace_spectral_ray (optics/raytracer/transport.rs:617-872).
// Facet planes are staged in LDS once per workgroup; the material, camera and
// environment constants arrive as one by-value kernel argument (scalar loads); per-path
// state (8 Stokes vectors, 8 path pdfs, 8 radiances, the wavelength comb and its
// dispersion/absorption caches) lives in VGPRs.  No MFMA: there is no dense contraction.
#include "transport.h"

#include "device_physics.h"

namespace gr {

// ---------------------------------------------------------------------------------
// Environment radiance for an escaping direction (transport.rs:67-90).
// ---------------------------------------------------------------------------------

// blackbody_spectrum (optics/raytracer/environment.rs:12-21) with the lambda-independent
// 560 nm denominator precomputed on the host.
__device__ __forceinline__ float blackbody(float lambda_nm, const DevScene &sc) {
    float exp_val = expf(fminf(14388000.0f / (lambda_nm * sc.temp_k_clamped), 80.0f));
    float denom = fmaxf(exp_val - 1.0f, 1e-6f);
    float ratio = sc.denom_560 / denom;
    return clampf(powi_cr<5>(560.0f / lambda_nm) * ratio, 0.01f, 20.0f);
}

// rgb_to_spectral_radiance (renderer/env_map_spectrum.rs:60-82).
__device__ __forceinline__ float asym_gauss(float x, float mu, float lo, float hi) {
    float sigma = x < mu ? lo : hi;
    float t = (x - mu) / sigma;
    return expf(-0.5f * t * t);
}
__device__ __forceinline__ float rgb_to_spectral(V3 rgb, float l) {
    float r = fmaxf(rgb.x, 0.0f), g = fmaxf(rgb.y, 0.0f), b = fmaxf(rgb.z, 0.0f);
    return fmaf(r, asym_gauss(l, 615.0f, 45.0f, 65.0f),
                fmaf(g, asym_gauss(l, 545.0f, 45.0f, 45.0f), b * asym_gauss(l, 465.0f, 40.0f, 45.0f)));
}

// EnvironmentMap::radiance_rgb (renderer/env_map.rs:231-319): lat-long bilinear lookup.
__device__ __forceinline__ V3 map_rgb(V3 dir, const DevScene &sc) {
    V3 d = normalize_or_zero(dir);
    float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    float phi = atan2f(d.x, d.z);
    float v = theta / kPi;
    float u = rem_euclid(phi / (2.0f * kPi), 1.0f);
    const int w = sc.map_w, h = sc.map_h;
    const float *px = sc.map_rgb;
    if (w == 1 && h == 1) return v3(px[0], px[1], px[2]);
    float fx = fmaf(rem_euclid(u, 1.0f), (float)w, -0.5f);
    float fy = fmaf(clampf(v, 0.0f, 1.0f), (float)h, -0.5f);
    float x0 = floorf(fx), y0 = floorf(fy);
    float tx = fx - x0, ty = fy - y0;
    int xi = (int)x0, yi = (int)y0;
    int x0i = xi % w;
    if (x0i < 0) x0i += w;
    int x1i = x0i + 1 == w ? 0 : x0i + 1;
    int y0i = yi < 0 ? 0 : (yi > h - 1 ? h - 1 : yi);
    int y1i = yi + 1 < 0 ? 0 : (yi + 1 > h - 1 ? h - 1 : yi + 1);
    const float *p00 = px + 3 * (y0i * w + x0i);
    const float *p10 = px + 3 * (y0i * w + x1i);
    const float *p01 = px + 3 * (y1i * w + x0i);
    const float *p11 = px + 3 * (y1i * w + x1i);
    float out[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float top = fmaf(p10[c], tx, p00[c] * (1.0f - tx));
        float bottom = fmaf(p11[c], tx, p01[c] * (1.0f - tx));
        out[c] = fmaf(bottom, ty, top * (1.0f - ty));
    }
    return v3(out[0], out[1], out[2]);
}

// ---------------------------------------------------------------------------------
// The estimator for one tuple.
// ---------------------------------------------------------------------------------

// Column j of the absorption tensor's principal frame (a1, a2, c).
__device__ __forceinline__ V3 afr(const DevScene &sc, int j) {
    return v3(sc.abs_frame[j][0], sc.abs_frame[j][1], sc.abs_frame[j][2]);
}

struct PathOut {
    V3 xyz;
    uint32_t segments, termination, facet_hash, fate_hash;
};

__device__ __forceinline__ void fnv(uint32_t &h, uint32_t v) {
    h ^= v;
    h *= 16777619u;
}

template <bool kAniso, bool kAbsorb, int kEnv>
__device__ __forceinline__ PathOut trace_path(V3 ray_o, V3 ray_d, uint32_t rng_seed, float hero_rand,
                              const DevScene &sc, const float4 *__restrict__ planes,
                              uint32_t *ev) {
    // wrapped_hero_wavelengths (transport.rs:45-53), evaluated on demand instead of held
    // in 8 registers.  offset = k * 50 + (lambda_hero - 380) lies in [0, 750], so
    // rem_euclid(offset, 400) is offset or offset - 400, and that subtraction is exact
    // (Sterbenz): the same bits as the reference's fmod.
    const float hero_off = fmaf(hero_rand, 780.0f - 380.0f, 380.0f) - 380.0f;
    auto lam = [hero_off](int k) -> float {
        float offset = fmaf((float)k, (780.0f - 380.0f) / 8.0f, hero_off);
        return 380.0f + (offset >= 400.0f ? offset - 400.0f : offset);
    };
    // Per-ray caches (refraction.rs:127-174): ordinary index per channel and absorption
    // coefficients per channel and eigen-axis.  Both are pure functions of the comb, so
    // they are filled lazily: the indices at the first facet hit, the absorption at the
    // first interior segment (a path that misses the stone never pays for either).
    float n_o[8], ao[8], ae[8];
    bool have_n = false, have_abs = false;
    const V3 c_axis = v3(sc.c_axis[0], sc.c_axis[1], sc.c_axis[2]);
    const float dn = sc.birefringence_delta;

    Stokes st[8];
    float pdf[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        st[k] = Stokes{1.0f, 0.0f, 0.0f, 0.0f};
        pdf[k] = 1.0f;
    }
    V3 cur_k = ray_d;
    bool inside = false, is_extra = false, have_prev = false, escaped = false;
    V3 prev_pn = v3(0.0f, 0.0f, 0.0f);
    uint32_t term_bounce = sc.max_bounces, term_reason = 3u;
    uint32_t fh = 2166136261u, fate_h = 2166136261u;
    const int n_planes = (int)sc.n_planes;

    for (uint32_t bounce = 0; bounce < sc.max_bounces; ++bounce) {
        ev[GR_EV_SEGMENT]++;
        // ---- convex facet-plane clip (simd/slab.rs:94-152, intersect.rs:93-132) ----
        // The reference divides t = -side / denom for every plane and keeps the first
        // strict maximum (entering) / minimum (exiting).  Here a = -side * rcp(denom)
        // (v_rcp_f32, 1 ulp) screens each plane: |a - t| <= |a| * 2^-19, so a plane whose
        // exact t could beat the running extreme always passes the screen, and only those
        // planes take the correctly rounded division and the exact comparison (a 2^-120
        // absolute slack covers denormal products).  The selected planes and t values are
        // therefore exactly the reference's.
        float t_near = -1e30f, t_far = 1e30f;
        int near_idx = -1, far_idx = -1;
        bool outside = false;
        for (int i = 0; i < n_planes; ++i) {
            float4 p = planes[i];
            float denom = (p.x * ray_d.x + p.y * ray_d.y) + p.z * ray_d.z;
            float side = p.w + ((p.x * ray_o.x + p.y * ray_o.y) + p.z * ray_o.z);
            float a = (-side) * __builtin_amdgcn_rcpf(denom);
            if (fabsf(denom) > 1e-7f) {
                if (denom < 0.0f) {
                    if (fmaf(fabsf(a), 0x1p-19f, a) + 0x1p-120f > t_near) {
                        float t = (-side) / denom;
                        if (t > t_near) {
                            t_near = t;
                            near_idx = i;
                        }
                    }
                } else if (fmaf(-fabsf(a), 0x1p-19f, a) - 0x1p-120f < t_far) {
                    float t = (-side) / denom;
                    if (t < t_far) {
                        t_far = t;
                        far_idx = i;
                    }
                }
            } else if (side > 0.0f) {
                outside = true;
                break;
            }
        }
        bool hit = false;
        float hit_t = 0.0f;
        int facet = 0;
        if (!outside && !(t_near > t_far)) {
            if (t_near > 1e-4f) {
                hit = true;
                hit_t = t_near;
                facet = near_idx;
            } else if (t_far > 1e-4f) {
                hit = true;
                hit_t = t_far;
                facet = far_idx;
            }
        }
        if (!hit) {
            // Escape: the environment is sampled after the loop (the path's only radiance
            // contribution; transport.rs:307-318).
            ev[GR_EV_MISS]++;
            escaped = true;
            fnv(fh, 0xFFFFu);
            term_bounce = bounce;
            term_reason = 0u;
            break;
        }
        ev[GR_EV_HIT]++;
        if (!have_n) {
            // DispersionModel::evaluate (optics/dispersion.rs:12-37) for the comb.
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                float l_um = lam(k) * 1e-3f;
                float l2 = l_um * l_um;
                float n;
                if (sc.disp_kind == 0) {
                    float n2 = 1.0f + (sc.disp[0] * l2) / (l2 - sc.disp[1]);
                    n = sqrtf(fmaxf(n2, 1.0f));
                } else if (sc.disp_kind == 1) {
                    float n2 = 1.0f;
                    n2 += (sc.disp[0] * l2) / (l2 - sc.disp[3]);
                    n2 += (sc.disp[1] * l2) / (l2 - sc.disp[4]);
                    n2 += (sc.disp[2] * l2) / (l2 - sc.disp[5]);
                    n = sqrtf(fmaxf(n2, 1.0f));
                } else {
                    float l4 = l2 * l2;
                    n = sc.disp[0] + (sc.disp[1] / l2) + (sc.disp[2] / l4);
                }
                n_o[k] = n;
            }
            have_n = true;
        }
        // An unbounded side (no plane on it) reports facet 0 with a zero normal, as the
        // reference's HitRecord does for a negative index (intersect.rs:105-126).
        V3 normal = v3(0.0f, 0.0f, 0.0f);
        if (facet >= 0) {
            float4 hp = planes[facet];
            normal = v3(hp.x, hp.y, hp.z);
        } else {
            facet = 0;
        }
        V3 hit_point = add(ray_o, mul(ray_d, hit_t));

        // ---- Stokes frame rotation into the new plane of incidence
        //      (raytracer/absorption.rs:44-89) ----
        V3 cpn = normalize_or_zero(cross(cur_k, normal));
        if (have_prev && len2(cpn) > 1e-6f && len2(prev_pn) > 1e-6f) {
            ev[GR_EV_ROTATE]++;
            float cos_psi = clampf(dot(prev_pn, cpn), -1.0f, 1.0f);
            float sin_psi = dot(cross(prev_pn, cpn), normalize_or_zero(cur_k));
            float psi = atan2f(sin_psi, cos_psi);
            float s2, c2;
            sincosf(2.0f * psi, &s2, &c2);
#pragma unroll
            for (int k = 0; k < 8; ++k) st[k] = apply_rotation(st[k], c2, s2);
        }
        prev_pn = cpn;
        have_prev = true;

        // ---- interior segment: pleochroic Beer-Lambert (raytracer/absorption.rs:115-216,
        //      birefringence.rs:252-286, polarization.rs:172-192) ----
        if (inside) {
            normal = neg(normal);
            ev[GR_EV_ABSORB]++;
            if (kAbsorb) {
                if (!have_abs) {
                    // spectral_absorption per eigen-axis (raytracer/absorption.rs:28-30,
                    // optics/absorption.rs:68-71), summed from -0.0 like Iterator::sum.
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        float so = -0.0f, se = -0.0f;
                        const float lk = lam(k);
                        for (int b = 0; b < sc.n_o_bands; ++b) {
                            float t = (lk - sc.o_bands[b][0]) / sc.o_bands[b][1];
                            so = so + sc.o_bands[b][2] * expf(-0.5f * t * t);
                        }
                        for (int b = 0; b < sc.n_e_bands; ++b) {
                            float t = (lk - sc.e_bands[b][0]) / sc.e_bands[b][1];
                            se = se + sc.e_bands[b][2] * expf(-0.5f * t * t);
                        }
                        ao[k] = so;
                        ae[k] = se;
                    }
                    have_abs = true;
                }
                V3 ea = ordinary_eigen_polarization(cur_k, c_axis);
                V3 eb = extraordinary_eigen_polarization(cur_k, c_axis);
                // Channel-invariant parts of electric_field_direction and of the
                // quadratic forms: the absorption frame (a1, a2, c), the s/p axes, and
                // the eigenmode projections.
                V3 kk = normalize_or_zero(cur_k);
                V3 s_raw = sub(cpn, mul(kk, dot(kk, cpn)));
                V3 s_hat = len2(s_raw) > 1e-8f ? normalize(s_raw) : arbitrary_perpendicular(kk);
                V3 p_hat = cross(kk, s_hat);
                V3 ea_l = v3(dot(afr(sc, 0), ea), dot(afr(sc, 1), ea), dot(afr(sc, 2), ea));
                V3 eb_l = v3(dot(afr(sc, 0), eb), dot(afr(sc, 1), eb), dot(afr(sc, 2), eb));
                float spl = hit_t * sc.absorption_path_scale;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    float psi = 0.5f * atan2f(st[k].u, st[k].q);
                    float sp, cp;
                    sincosf(psi, &sp, &cp);
                    V3 e = add(mul(s_hat, cp), mul(p_hat, sp));
                    V3 e_hat = len2(e) > 1e-8f ? normalize(e) : s_hat;
                    V3 el = v3(dot(afr(sc, 0), e_hat), dot(afr(sc, 1), e_hat), dot(afr(sc, 2), e_hat));
                    float qf_e = (ao[k] * (el.x * el.x) + ao[k] * (el.y * el.y)) + ae[k] * (el.z * el.z);
                    float qf_a = (ao[k] * (ea_l.x * ea_l.x) + ao[k] * (ea_l.y * ea_l.y)) + ae[k] * (ea_l.z * ea_l.z);
                    float qf_b = (ao[k] * (eb_l.x * eb_l.x) + ao[k] * (eb_l.y * eb_l.y)) + ae[k] * (eb_l.z * eb_l.z);
                    float p = clampf(degree_of_polarization(st[k]), 0.0f, 1.0f);
                    float unpol = midpoint(qf_a, qf_b);
                    float alpha = fmaf(p, qf_e - unpol, unpol);
                    st[k] = scale(st[k], exp_poly(-alpha * spl));
                }
            }
        }

        // ---- per-bounce refraction geometry (refraction.rs:235-547) ----
        float cos_i = clampf(dot(neg(cur_k), normal), 0.0f, 1.0f);
        float sin_i = sqrtf(fmaxf(fmaf(cos_i, -cos_i, 1.0f), 0.0f));
        float one_m_c2 = fmaf(cos_i, -cos_i, 1.0f);
        float n_med[8];
        if (kAniso) {
            float theta_c;
            if (!inside) {
                // theta_c_for_bounce's two-step fixed point (refraction.rs:248-269).
                float n_seed = n_o[0];
                float n_e_seed = n_seed + dn;
                float n_guess = n_seed;
                theta_c = 0.0f;
                for (int it = 0; it < 2; ++it) {
                    float eta_g = 1.0f / n_guess;
                    float sin2_g = eta_g * eta_g * one_m_c2;
                    if (sin2_g > 1.0f) break;
                    float cos_t_g = sqrtf(fmaxf(1.0f - sin2_g, 0.0f));
                    V3 wave = normalize(add(mul(cur_k, eta_g), mul(normal, fmaf(eta_g, cos_i, -cos_t_g))));
                    theta_c = acosf(fabsf(clampf(dot(wave, c_axis), -1.0f, 1.0f)));
                    n_guess = effective_extraordinary_index(n_seed, n_e_seed, theta_c);
                }
            } else {
                theta_c = acosf(fabsf(clampf(dot(cur_k, c_axis), -1.0f, 1.0f)));
            }
            bool use_o = inside && !is_extra;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                n_med[k] = use_o ? n_o[k] : effective_extraordinary_index(n_o[k], n_o[k] + dn, theta_c);
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) n_med[k] = n_o[k];
        }
        float n1_ch[8], n2_ch[8], s2t_ch[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float n1k = inside ? n_med[k] : 1.0f;
            float n2k = inside ? 1.0f : n_med[k];
            float etak = n1k / n2k;
            n1_ch[k] = n1k;
            n2_ch[k] = n2k;
            s2t_ch[k] = etak * etak * one_m_c2;
        }
        const float n1 = n1_ch[0], n2 = n2_ch[0];
        const float sin2_t = s2t_ch[0];
        const float n_o_hero = n_o[0], n_e_hero = n_o[0] + dn;

        // ---- TIR / partial-reflect / refract dispatch (transport.rs:128-193) ----
        V3 new_k, new_s;
        bool new_inside = inside, has_update = false, update = false;
        uint32_t event;
        // Hero event: TIR when sin2_t > 1, else a Fresnel-weighted reflect/refract draw
        // (transport.rs:158-176, refraction.rs:1036-1086).
        const bool tir = sin2_t > 1.0f;
        float r_unpol = 1.0f;
        bool reflect = tir;
        if (!tir) {
            float cos_t = sqrtf(1.0f - sin2_t);
            float r_s, r_p;
            fresnel_rs_rp(n1, n2, cos_i, cos_t, &r_s, &r_p);
            r_unpol = clampf(0.5f * fmaf(r_p, r_p, r_s * r_s), 1e-4f, 1.0f - 1e-4f);
            reflect = stream_draw(rng_seed, bounce, FRESNEL_BRANCH_STREAM) < r_unpol;
        }
        if (reflect) {
            // apply_tir_bounce (refraction.rs:599-637) and apply_partial_reflect_bounce
            // (:650-690) share their per-channel work: retardation for channels past their
            // own critical angle, Fresnel reflection (and its pdf factor) otherwise.  The
            // partial reflection then divides by the hero's r_unpol; TIR scales by 1.0,
            // which is exact.
            if (tir) {
                ev[GR_EV_TIR]++;
                event = 1u;
            } else {
                ev[GR_EV_REFLECT]++;
                event = 2u;
            }
            const float inv_r = tir ? 1.0f : 1.0f / r_unpol;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (s2t_ch[k] > 1.0f) {
                    float sd, cd;
                    sincosf(tir_phase_delta(n1_ch[k], cos_i, sin_i), &sd, &cd);
                    st[k] = apply_retardation(st[k], cd, sd);
                } else {
                    float cos_t_k = sqrtf(fmaxf(1.0f - s2t_ch[k], 0.0f));
                    float rs, rp;
                    fresnel_rs_rp(n1_ch[k], n2_ch[k], cos_i, cos_t_k, &rs, &rp);
                    pdf[k] *= clampf(0.5f * fmaf(rp, rp, rs * rs), 1e-4f, 1.0f - 1e-4f);
                    st[k] = apply_fresnel_reflection(st[k], rs, rp);
                }
                st[k] = scale(st[k], inv_r);
            }
            new_k = sub(cur_k, mul(normal, 2.0f * dot(cur_k, normal)));
            new_s = new_k;
            if (kAniso && inside && is_extra) {
                ev[GR_EV_WALKOFF]++;
                new_s = extraordinary_poynting_dir(new_k, c_axis, n_o_hero, n_e_hero);
            }
        } else {
            {
                // apply_refract_bounce + apply_refract_channel (refraction.rs:748-1011).
                ev[GR_EV_REFRACT]++;
                event = 3u;
                const float DIRECTION_MATCH_COS_TOL = 1.0f - 1e-6f;
                bool entering = kAniso && !inside;
                bool use_extra = is_extra;
                if (entering) {
                    ev[GR_EV_ENTRY_ANISO]++;
                    use_extra = stream_draw(rng_seed, bounce, BIREFRINGENT_SPLIT_STREAM) < 0.5f;
                }
                float n2_dir = (entering && !use_extra) ? n_o_hero : n2;
                float eta_dir = n1 / n2_dir;
                float sin2_dir = fminf(eta_dir * eta_dir * one_m_c2, 1.0f);
                float cos_t_dir = sqrtf(fmaxf(1.0f - sin2_dir, 0.0f));
                V3 refr = normalize(add(mul(cur_k, eta_dir), mul(normal, fmaf(eta_dir, cos_i, -cos_t_dir))));
                V3 final_dir = refr;
                if (entering && use_extra) {
                    ev[GR_EV_WALKOFF]++;
                    final_dir = extraordinary_poynting_dir(refr, c_axis, n_o_hero, n_e_hero);
                }
                float inv_t = 1.0f / (1.0f - r_unpol);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    float n1k = n1_ch[k];
                    float n2k = (entering && !use_extra) ? n_o[k] : n2_ch[k];
                    float q = n1k / n2k;
                    float s2k = q * q * one_m_c2;
                    bool alive = !(s2k > 1.0f);
                    float cos_t_k = sqrtf(fmaxf(1.0f - s2k, 0.0f));
                    if (alive) {
                        V3 dk = normalize(add(mul(cur_k, q), mul(normal, fmaf(q, cos_i, -cos_t_k))));
                        if (entering && use_extra)
                            dk = extraordinary_poynting_dir(dk, c_axis, n_o[k], n_o[k] + dn);
                        alive = dot(dk, final_dir) >= DIRECTION_MATCH_COS_TOL;
                    }
                    if (alive) {
                        float num = 2.0f * n1k * cos_i;
                        float t_s = num / fmaf(n2k, cos_t_k, n1k * cos_i);
                        float t_p = num / fmaf(n1k, cos_t_k, n2k * cos_i);
                        st[k] = scale(apply_fresnel_transmission(st[k], n1k, n2k, cos_i, cos_t_k, t_s, t_p), inv_t);
                        float rs, rp;
                        fresnel_rs_rp(n1k, n2k, cos_i, cos_t_k, &rs, &rp);
                        pdf[k] *= 1.0f - clampf(0.5f * fmaf(rp, rp, rs * rs), 1e-4f, 1.0f - 1e-4f);
                    } else {
                        st[k] = scale(st[k], 0.0f);
                        pdf[k] = 0.0f;
                    }
                }
                new_k = refr;
                new_s = final_dir;
                new_inside = !inside;
                has_update = entering;
                update = use_extra;
            }
        }
        bool internal_reflection = inside && !has_update && new_inside == inside;
        ray_d = new_s;
        ray_o = add(hit_point, mul(new_s, 1e-4f));
        cur_k = new_k;
        inside = new_inside;
        if (has_update) is_extra = update;
        if (kAniso && internal_reflection)
            is_extra = stream_draw(rng_seed, bounce, MODE_COUPLING_STREAM) < 0.5f;
        fnv(fh, (uint32_t)facet);
        fnv(fh, event | (is_extra ? 16u : 0u));
        {
            // Per-channel fates: past the channel's critical angle / pdf zero.
            uint32_t fate = 0u;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                fate |= (s2t_ch[k] > 1.0f ? 1u : 0u) << k;
                fate |= (pdf[k] == 0.0f ? 1u : 0u) << (8 + k);
            }
            fnv(fate_h, fate);
        }

        // ---- Russian roulette (transport.rs:204-220) ----
        if (bounce > 4) {
            ev[GR_EV_RR]++;
            float max_i = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; ++k) max_i = fmaxf(max_i, intensity(st[k]));
            float q = clampf(max_i, 0.05f, 1.0f);
            if (stream_draw(rng_seed, bounce, RUSSIAN_ROULETTE_STREAM) > q) {
                term_bounce = bounce;
                term_reason = 2u;
                break;
            }
            float inv_q = 1.0f / q;
#pragma unroll
            for (int k = 0; k < 8; ++k) st[k] = scale(st[k], inv_q);
        }
    }

    PathOut out;
    // A path that never escaped carries zero radiance: the reference's integration and
    // white balance then produce exactly +0 in every component (every product is +-0 and
    // every sum has a +0 operand), so that work is skipped.
    out.xyz = v3(0.0f, 0.0f, 0.0f);
    if (escaped) {
        // ---- environment radiance of the escape direction (transport.rs:67-90) ----
        float rad[8];
        if (kEnv == 0) {
            // sample_studio_environment_with_rig (environment.rs:280-321): the
            // direction-only terms are formed once, the fma chain runs per channel.
            V3 d = normalize(ray_d);
            float bg = fmaxf(fmaf(0.012f, fmaf(d.y, 0.5f, 0.5f), 0.015f), 0.005f) * sc.exposure;
            float terms[18];
            bool act[18];
            float kd = fmaxf(dot(d, v3(sc.key[0], sc.key[1], sc.key[2])), 0.0f);
            act[0] = kd > 0.0f;
            terms[0] = powi_cr<28>(kd) * 12.0f * sc.spot_mult * sc.exposure;
            float fd = fmaxf(dot(d, v3(sc.fill[0], sc.fill[1], sc.fill[2])), 0.0f);
            act[1] = fd > 0.0f;
            terms[1] = powi_cr<18>(fd) * 4.5f * sc.exposure;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                float rd = fmaxf(dot(d, v3(sc.ring[r][0], sc.ring[r][1], sc.ring[r][2])), 0.0f);
                act[2 + r] = rd > 0.96f;
                float spark = (rd - 0.96f) / 0.04f;
                terms[2 + r] = powi_cr<6>(spark) * 22.0f * sc.spot_mult * sc.exposure;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                float spec = blackbody(lam(k), sc);
                float r = bg * spec;
#pragma unroll
                for (int j = 0; j < 18; ++j)
                    if (act[j]) r = fmaf(terms[j], spec, r);
                rad[k] = fmaf(intensity(st[k]), r, 0.0f);
            }
        } else {
            V3 rgb = map_rgb(ray_d, sc);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                rad[k] = fmaf(intensity(st[k]), rgb_to_spectral(rgb, lam(k)), 0.0f);
        }
        // ---- spectral MIS + CIE XYZ (color.rs:121-164) + von Kries (color.rs:215-251) --
        float sum_pdf = -0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) sum_pdf = sum_pdf + pdf[k];
        float mis = sum_pdf <= 1e-12f ? 1.0f : (8.0f * pdf[0]) / sum_pdf;
        const float norm_factor = (400.0f / 8.0f) / 106.856f;
        V3 xyz = v3(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            V3 c = cmf_x8_channel(lam(k));
            float w = rad[k] * mis * norm_factor;
            xyz = add(xyz, mul(c, w));
        }
        V3 lms = v3(fmaf(0.8951f, xyz.x, fmaf(0.2664f, xyz.y, -0.1614f * xyz.z)),
                    fmaf(-0.7502f, xyz.x, fmaf(1.7135f, xyz.y, 0.0367f * xyz.z)),
                    fmaf(0.0389f, xyz.x, fmaf(-0.0685f, xyz.y, 1.0296f * xyz.z)));
        lms = v3(lms.x * sc.wb[0], lms.y * sc.wb[1], lms.z * sc.wb[2]);
        out.xyz = v3(fmaf(0.986993f, lms.x, fmaf(-0.147054f, lms.y, 0.159963f * lms.z)),
                     fmaf(0.432305f, lms.x, fmaf(0.518360f, lms.y, 0.049291f * lms.z)),
                     fmaf(-0.008529f, lms.x, fmaf(0.040043f, lms.y, 0.968487f * lms.z)));
    }
    out.segments = term_reason == 3u ? sc.max_bounces : term_bounce + 1u;
    out.termination = term_reason;
    out.facet_hash = fh;
    out.fate_hash = fate_h;
    return out;
}

// ---------------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------------

constexpr int kBlock = 256;

// Occupancy target of the trace kernel (waves per SIMD); 0 leaves the register budget to
// the compiler.  A tuning parameter chosen by measurement: on the single-job workload
// 3 waves/SIMD (168 VGPRs) ran 1.36x faster than the compiler's 2 and 1.03x faster than 4.
#ifndef GR_TRACE_WAVES_PER_EU
#define GR_TRACE_WAVES_PER_EU 3
#endif
#if GR_TRACE_WAVES_PER_EU > 0
#define GR_TRACE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(GR_TRACE_WAVES_PER_EU, GR_TRACE_WAVES_PER_EU)))
#else
#define GR_TRACE_OCCUPANCY
#endif

// One lane per (pixel, sample) tuple t of the launch: pixel = pixel_begin + t / spp,
// sample = sample_offset + t % spp.  Writes the tuple's XYZ (and, with kInfo, its
// segment count, termination and facet-sequence hash); accumulates per-class event
// counters into stats (one atomic per counter per workgroup).
template <bool kAniso, bool kAbsorb, int kEnv, bool kInfo>
__global__ __launch_bounds__(kBlock) GR_TRACE_OCCUPANCY void trace_kernel(const DevScene sc, const float4 *__restrict__ planes_g,
                                                       uint32_t pixel_begin, uint32_t sample_offset,
                                                       uint32_t spp, uint32_t n_tuples,
                                                       float *__restrict__ xyz_out,
                                                       uint32_t *__restrict__ info_out,
                                                       unsigned long long *__restrict__ stats) {
    __shared__ float4 planes[GR_MAX_PLANES];
    __shared__ uint32_t block_ev[GR_EV_COUNT];
    for (uint32_t i = threadIdx.x; i < sc.n_planes; i += kBlock) planes[i] = planes_g[i];
    if (threadIdx.x < GR_EV_COUNT) block_ev[threadIdx.x] = 0u;
    __syncthreads();

    uint32_t ev[GR_EV_COUNT];
#pragma unroll
    for (int i = 0; i < GR_EV_COUNT; ++i) ev[i] = 0u;

    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n_tuples) {
        const uint32_t local_pixel = t / spp;
        const uint32_t s = t - local_pixel * spp;
        const uint32_t pixel = pixel_begin + local_pixel;
        const uint32_t sample_num = sample_offset + s;
        const uint32_t x = pixel % sc.width, y = pixel / sc.width;
        // pixel_rotations + sample_draws (sampling.rs:232-271).
        float rot_x = low_discrepancy_base2(hash_u32(pixel ^ PIXEL_JITTER_X_ROTATION_STREAM));
        float rot_y = low_discrepancy_base2(hash_u32(pixel ^ PIXEL_JITTER_Y_ROTATION_STREAM));
        float rot_h = low_discrepancy_base2(hash_u32(pixel ^ HERO_WAVELENGTH_ROTATION_STREAM));
        uint32_t seed = hash_u32((pixel * 0x9e3779b9u) ^ (sample_num * 0x85ebca6bu));
        float jx = cranley_patterson(low_discrepancy_base2(sample_num), rot_x) - 0.5f;
        float jy = cranley_patterson(radical_inverse(sample_num, 3), rot_y) - 0.5f;
        float hero = cranley_patterson(radical_inverse(sample_num, 5), rot_h);
        // Camera::generate_ray (camera.rs:96-114).
        float w = (float)sc.width, h = (float)sc.height;
        float aspect = w / h;
        float u = (((float)x + jx) / w - 0.5f) * 2.0f * aspect * sc.fov_tan;
        float v = (0.5f - ((float)y + jy) / h) * 2.0f * sc.fov_tan;
        V3 dir = normalize(add(add(v3(sc.cam_f[0], sc.cam_f[1], sc.cam_f[2]),
                                   mul(v3(sc.cam_r[0], sc.cam_r[1], sc.cam_r[2]), u)),
                               mul(v3(sc.cam_u[0], sc.cam_u[1], sc.cam_u[2]), v)));
        V3 org = v3(sc.cam_o[0], sc.cam_o[1], sc.cam_o[2]);
        PathOut r = trace_path<kAniso, kAbsorb, kEnv>(org, dir, seed, hero, sc, planes, ev);
        xyz_out[3 * (size_t)t + 0] = r.xyz.x;
        xyz_out[3 * (size_t)t + 1] = r.xyz.y;
        xyz_out[3 * (size_t)t + 2] = r.xyz.z;
        if (kInfo) {
            info_out[4 * (size_t)t + 0] = r.segments;
            info_out[4 * (size_t)t + 1] = r.termination;
            info_out[4 * (size_t)t + 2] = r.facet_hash;
            info_out[4 * (size_t)t + 3] = r.fate_hash;
        }
    }
    // Per-class counters: wave-level sum, one LDS atomic per wave, one global atomic per
    // counter per workgroup.
#pragma unroll
    for (int i = 0; i < GR_EV_COUNT; ++i) {
        uint32_t v = ev[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&block_ev[i], v);
    }
    __syncthreads();
    if (threadIdx.x < GR_EV_COUNT && block_ev[threadIdx.x])
        atomicAdd(&stats[threadIdx.x], (unsigned long long)block_ev[threadIdx.x]);
}

// Each pixel's spp tuples summed in sample order and ADDED into the film
// (frame.rs:799-814: sum = 0; sum += xyz[s] for s in order; accum[pixel] += sum).
__global__ __launch_bounds__(kBlock) void film_reduce_kernel(const float *__restrict__ xyz,
                                                             uint32_t n_pixels, uint32_t spp,
                                                             uint32_t pixel_begin,
                                                             float *__restrict__ film) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n_pixels) return;
    const float *src = xyz + 3 * (size_t)p * spp;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (uint32_t s = 0; s < spp; ++s) {
        sx = sx + src[3 * s + 0];
        sy = sy + src[3 * s + 1];
        sz = sz + src[3 * s + 2];
    }
    float *dst = film + 3 * (size_t)(pixel_begin + p);
    dst[0] = dst[0] + sx;
    dst[1] = dst[1] + sy;
    dst[2] = dst[2] + sz;
}

}  // namespace gr

// ---------------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------------

namespace {
template <bool A, bool B, int E>
hipError_t launch_trace_t(const gr::DevScene &sc, const float4 *planes, uint32_t pixel_begin,
                          uint32_t sample_offset, uint32_t spp, uint32_t n_tuples, float *xyz,
                          uint32_t *info, unsigned long long *stats, hipStream_t stream) {
    dim3 grid((n_tuples + gr::kBlock - 1) / gr::kBlock);
    if (info)
        hipLaunchKernelGGL((gr::trace_kernel<A, B, E, true>), grid, dim3(gr::kBlock), 0, stream, sc, planes,
                           pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats);
    else
        hipLaunchKernelGGL((gr::trace_kernel<A, B, E, false>), grid, dim3(gr::kBlock), 0, stream, sc, planes,
                           pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats);
    return hipGetLastError();
}
}  // namespace

hipError_t gr_launch_trace(const gr::DevScene &sc, const float4 *planes, uint32_t pixel_begin,
                           uint32_t sample_offset, uint32_t spp, uint32_t n_tuples, float *xyz,
                           uint32_t *info, unsigned long long *stats, hipStream_t stream) {
    const bool a = sc.anisotropic != 0, b = sc.has_absorption != 0;
    const int e = sc.env_kind;
    if (a && b && e == 0) return launch_trace_t<true, true, 0>(sc, planes, pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats, stream);
    if (a && b && e == 1) return launch_trace_t<true, true, 1>(sc, planes, pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats, stream);
    if (a && !b && e == 0) return launch_trace_t<true, false, 0>(sc, planes, pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats, stream);
    if (a && !b && e == 1) return launch_trace_t<true, false, 1>(sc, planes, pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats, stream);
    if (!a && b && e == 0) return launch_trace_t<false, true, 0>(sc, planes, pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats, stream);
    if (!a && b && e == 1) return launch_trace_t<false, true, 1>(sc, planes, pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats, stream);
    if (!a && !b && e == 0) return launch_trace_t<false, false, 0>(sc, planes, pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats, stream);
    return launch_trace_t<false, false, 1>(sc, planes, pixel_begin, sample_offset, spp, n_tuples, xyz, info, stats, stream);
}

hipError_t gr_launch_film_reduce(const float *xyz, uint32_t n_pixels, uint32_t spp,
                                 uint32_t pixel_begin, float *film, hipStream_t stream) {
    dim3 grid((n_pixels + gr::kBlock - 1) / gr::kBlock);
    hipLaunchKernelGGL(gr::film_reduce_kernel, grid, dim3(gr::kBlock), 0, stream, xyz, n_pixels, spp,
                       pixel_begin, film);
    return hipGetLastError();
}
