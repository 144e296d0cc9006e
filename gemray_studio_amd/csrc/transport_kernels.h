// transport_kernels.h -- the gfx950 spectral transport kernels (device functions and the
// kernel templates), included by the per-class translation units that instantiate them
// and by transport.hip for the shared non-template kernels.
//
// The estimator is gemray's transport_main
// (crates/gemray/src/renderer/shaders/spectral_transport.wgsl:672), whose CPU statement is
// trace_spectral_ray (optics/raytracer/transport.rs:617-872).  It is split into
//   init_path    -- draws, camera ray and per-path state for one (pixel, sample) tuple;
//   bounce_step  -- one bounce-loop iteration: facet-plane clip plus the 8-channel event;
//   finish_path  -- environment radiance at the escape, spectral MIS, CIE XYZ, white balance.
// Two trace kernels drive them:
//   trace_kernel        -- one lane per tuple, bounce loop then finish (the plain path);
//   trace_wave_kernel   -- persistent wavefronts: a lane whose path ends takes the next
//                          tuple of its wave's range (ballot + popcount compaction), and
//                          the finished path is written as an escape record that
//                          finish_kernel turns into XYZ in a separate, fully coherent pass.
// Facet planes are staged in LDS once per workgroup; material, camera and environment
// constants arrive as one by-value kernel argument; per-path state lives in VGPRs.
// No MFMA: there is no dense contraction.
#pragma once

#include "transport.h"

#include <cstddef>

#include "device_physics.h"
#include "biaxial.h"
#include "clip.h"

namespace gr {

// ---------------------------------------------------------------------------------
// Environment radiance (transport.rs:67-90).
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
    const bool below = x < mu;
    float t = div_by_const(x - mu, below ? lo : hi, below ? 1.0f / lo : 1.0f / hi);
    return exp_lobe(-0.5f * t * t);
}
// The red, green and blue lobes at l, and their weighted sum.
__device__ __forceinline__ V3 rgb_lobes(float l) {
    return v3(asym_gauss(l, 615.0f, 45.0f, 65.0f), asym_gauss(l, 545.0f, 45.0f, 45.0f),
              asym_gauss(l, 465.0f, 40.0f, 45.0f));
}
__device__ __forceinline__ float rgb_to_spectral_lobes(V3 rgb, V3 lobes) {
    float r = fmaxf(rgb.x, 0.0f), g = fmaxf(rgb.y, 0.0f), b = fmaxf(rgb.z, 0.0f);
    return fmaf(r, lobes.x, fmaf(g, lobes.y, b * lobes.z));
}
__device__ __forceinline__ float rgb_to_spectral(V3 rgb, float l) { return rgb_to_spectral_lobes(rgb, rgb_lobes(l)); }

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

// Column j of the absorption tensor's principal frame (a1, a2, c).
__device__ __forceinline__ V3 afr(const DevScene &sc, int j) {
    return v3(sc.abs_frame[j][0], sc.abs_frame[j][1], sc.abs_frame[j][2]);
}

__device__ __forceinline__ void fnv(uint32_t &h, uint32_t v) {
    h ^= v;
    h *= 16777619u;
}

// wrapped_hero_wavelengths (transport.rs:45-53) for channel k, from
// hero_off = lambda_hero - 380.  offset = k * 50 + hero_off lies in [0, 750], so
// rem_euclid(offset, 400) is offset or offset - 400, and that subtraction is exact
// (Sterbenz): the same bits as the reference's fmod.
__device__ __forceinline__ float comb_lambda(float hero_off, int k) {
    float offset = fmaf((float)k, (780.0f - 380.0f) / 8.0f, hero_off);
    return 380.0f + (offset >= 400.0f ? offset - 400.0f : offset);
}

// The hero offset as the lazy fills read it: an identity DPP move (quad_perm 0,1,2,3, each
// lane its own value) is convergent, so the compiler cannot hoist the fill's evaluation (a
// pure function of the offset) out of the bounce loop into init_path, where its results
// would stay live across the loop -- in the kernels without a material table that hoist
// spilled the eight indices.  One VALU per fill.
__device__ __forceinline__ float fill_offset(float h) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(h), 0xE4, 0xF, 0xF, false));
}

// The hero offset of entry e of a hero table (transport.h), and the entry of hero offset h:
// kHeroEntries unless h is a key (every offset init_path forms is one; the check costs a
// few instructions once per path and keeps any other value on the direct evaluation).
__device__ __forceinline__ float hero_key(uint32_t e) {
    return e < kHeroLow ? (float)e * 0x1p-15f : 132.0f + (float)(e - kHeroLow) * 0x1p-14f;
}
__device__ __forceinline__ uint32_t hero_entry(float h) {
    if (!(h >= 0.0f && h <= 400.0f)) return kHeroEntries;
    const uint32_t e = h < 132.0f ? (uint32_t)(h * 32768.0f) : kHeroLow + (uint32_t)((h - 132.0f) * 16384.0f);
    return hero_key(e) == h ? e : kHeroEntries;
}

// DispersionModel::evaluate (optics/dispersion.rs:12-37) at comb wavelength lk (nm).
__device__ __forceinline__ float dispersion_index(float lk, const DevScene &sc) {
    float l_um = lk * 1e-3f;
    float l2 = l_um * l_um;
    if (sc.disp_kind == 0) {
        float n2 = 1.0f + (sc.disp[0] * l2) / (l2 - sc.disp[1]);
        return sqrtf(fmaxf(n2, 1.0f));
    } else if (sc.disp_kind == 1) {
        float n2 = 1.0f;
        n2 += (sc.disp[0] * l2) / (l2 - sc.disp[3]);
        n2 += (sc.disp[1] * l2) / (l2 - sc.disp[4]);
        n2 += (sc.disp[2] * l2) / (l2 - sc.disp[5]);
        return sqrtf(fmaxf(n2, 1.0f));
    }
    float l4 = l2 * l2;
    return sc.disp[0] + (sc.disp[1] / l2) + (sc.disp[2] / l4);
}

// spectral_absorption (raytracer/absorption.rs:28-30, optics/absorption.rs:68-71) of band
// set kSet (0 ordinary, 1 extraordinary, 2 beta axis) at lk, summed from -0.0 like
// Iterator::sum.
template <int kSet>
__device__ __forceinline__ float band_sum(float lk, const DevScene &sc) {
    const int n = kSet == 0 ? sc.n_o_bands : kSet == 1 ? sc.n_e_bands : sc.n_beta_bands;
    float s = -0.0f;
    for (int b = 0; b < n; ++b) {
        const float c = kSet == 0 ? sc.o_bands[b][0] : kSet == 1 ? sc.e_bands[b][0] : sc.beta_bands[b][0];
        const float w = kSet == 0 ? sc.o_bands[b][1] : kSet == 1 ? sc.e_bands[b][1] : sc.beta_bands[b][1];
        const float p = kSet == 0 ? sc.o_bands[b][2] : kSet == 1 ? sc.e_bands[b][2] : sc.beta_bands[b][2];
        float t = (lk - c) / w;
        s = s + p * expf(-0.5f * t * t);
    }
    return s;
}

// The per-path dispersion and absorption caches (refraction.rs:127-174): 24 floats per
// lane in LDS (32 in the kernels built for biaxial stones, whose trichroic tensor adds the
// beta-axis absorption), channel-major so that a wave's accesses hit distinct banks.  They
// are filled lazily and read a few times per bounce; keeping them out of VGPRs leaves
// registers to the Stokes state.
constexpr int kBlock = 256;
#define LC_NO(k) lcache[(k) * kBlock + threadIdx.x]
#define LC_AO(k) lcache[(8 + (k)) * kBlock + threadIdx.x]
#define LC_AE(k) lcache[(16 + (k)) * kBlock + threadIdx.x]
#define LC_AB(k) lcache[(24 + (k)) * kBlock + threadIdx.x]

// Per-lane event counters, two 16-bit counts per word (a lane counts at most kWaveTuples
// paths x GR_MAX_BOUNCES segments, 32768, before flush_events), in the lane's own LDS
// words (word w at ev[w * kBlock], ev = the block's counter array + threadIdx.x): an event
// is one LDS add without return.  Kept in registers, the eight words were live across the
// bounce loop's divergent exits, and in the biaxial rare-feature kernel the compiler's copies
// of them at a join read a register it leaves undefined on some incoming paths whenever the
// kernel's register allocation changed (counts off, films unaffected).
constexpr int kEvWords = (GR_EV_COUNT + 1) / 2;
__device__ __forceinline__ void ev_inc(uint32_t *ev, int i) {
    __hip_atomic_fetch_add(ev + (i >> 1) * kBlock, 1u << (16 * (i & 1)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <int kAniso>
struct CacheFloats {
    static constexpr int value = (kAniso == 2 ? 32 : 24) * kBlock;
};

// ---------------------------------------------------------------------------------
// Per-path state.
// ---------------------------------------------------------------------------------

struct PathState {
    V3 ray_o, ray_d, cur_k, prev_pn;
    Stokes st[8];
    float pdf[8];
    float hero_off;
    uint32_t seed, bounce, last_facet, cert;
    uint32_t term_bounce, term_reason, fh, fate_h;
    // Channels zeroed by a refraction whose direction they did not share (bit k): their
    // Stokes vector is +-0 and their pdf 0, and every later operation of the estimator keeps
    // them so, so their per-channel updates are skipped -- when no lane of the wave needs a
    // channel, the wave skips its work.  The hero (channel 0) is always traced.
    uint32_t dead;
    bool inside, is_extra, have_prev, have_n, have_abs, escaped;
};

#define GR_LIVE(ps, k) ((k) == 0 || !(((ps).dead >> (k)) & 1u))

// pixel_rotations + sample_draws (sampling.rs:232-271) with the sample-only radical
// inverses (bases 3 and 5) taken from a host-built table, then Camera::generate_ray
// (camera.rs:96-114) and the initial state of trace_spectral_ray_inner
// (transport.rs:651-722).
__device__ __forceinline__ void sample_draws(uint32_t pixel, uint32_t sample_num, float2 ri35, uint32_t *seed,
                                             float *jx, float *jy, float *hero) {
    const float rot_x = low_discrepancy_base2(hash_u32(pixel ^ PIXEL_JITTER_X_ROTATION_STREAM));
    const float rot_y = low_discrepancy_base2(hash_u32(pixel ^ PIXEL_JITTER_Y_ROTATION_STREAM));
    const float rot_h = low_discrepancy_base2(hash_u32(pixel ^ HERO_WAVELENGTH_ROTATION_STREAM));
    *seed = hash_u32((pixel * 0x9e3779b9u) ^ (sample_num * 0x85ebca6bu));
    *jx = cranley_patterson(low_discrepancy_base2(sample_num), rot_x) - 0.5f;
    *jy = cranley_patterson(ri35.x, rot_y) - 0.5f;
    *hero = cranley_patterson(ri35.y, rot_h);
}
__device__ __forceinline__ void init_path(PathState &ps, uint32_t pixel, uint32_t sample_num,
                                          float2 ri35, uint32_t cert, const DevScene &sc) {
    const uint32_t x = pixel % sc.width, y = pixel / sc.width;
    float jx, jy, hero;
    sample_draws(pixel, sample_num, ri35, &ps.seed, &jx, &jy, &hero);
    float w = (float)sc.width, h = (float)sc.height;
    float aspect = w / h;
    float u = (((float)x + jx) / w - 0.5f) * 2.0f * aspect * sc.fov_tan;
    float v = (0.5f - ((float)y + jy) / h) * 2.0f * sc.fov_tan;
    ps.ray_d = normalize(add(add(v3(sc.cam_f[0], sc.cam_f[1], sc.cam_f[2]),
                                 mul(v3(sc.cam_r[0], sc.cam_r[1], sc.cam_r[2]), u)),
                             mul(v3(sc.cam_u[0], sc.cam_u[1], sc.cam_u[2]), v)));
    ps.ray_o = v3(sc.cam_o[0], sc.cam_o[1], sc.cam_o[2]);
    ps.cur_k = ps.ray_d;
    ps.prev_pn = v3(0.0f, 0.0f, 0.0f);
    ps.hero_off = fmaf(hero, 780.0f - 380.0f, 380.0f) - 380.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        ps.st[k] = Stokes{1.0f, 0.0f, 0.0f, 0.0f};
        ps.pdf[k] = 1.0f;
    }
    ps.bounce = 0u;
    ps.last_facet = 0u;
    ps.cert = cert;
    ps.term_bounce = sc.max_bounces;
    ps.term_reason = 3u;
    ps.fh = 2166136261u;
    ps.fate_h = 2166136261u;
    ps.dead = 0u;
    ps.inside = ps.is_extra = ps.have_prev = ps.have_n = ps.have_abs = ps.escaped = false;
}


// Miss certificate of a camera ray: planes i (entering) and j (exiting) whose exact t
// values, formed with the scan's own operations, satisfy t_i > t_j.  The scan's entering
// extreme is the maximum over entering planes, so at least t_i, and its exiting extreme at
// most t_j: the scan reports a miss, which these two planes prove without the others.
// cert packs i | j << 16 (from the pixel's central ray, see cert_kernel); ~0 = none.
__device__ __forceinline__ bool camera_miss_certified(const float4 *__restrict__ planes,
                                                      const float *__restrict__ cam_side, uint32_t cert, V3 rd) {
    if (cert == 0xFFFFFFFFu) return false;
    const uint32_t i = cert & 0xFFFFu, j = cert >> 16;
    const float4 pi = planes[i], pj = planes[j];
    const float di = (pi.x * rd.x + pi.y * rd.y) + pi.z * rd.z;
    const float dj = (pj.x * rd.x + pj.y * rd.y) + pj.z * rd.z;
    if (!(fabsf(di) > 1e-7f && di < 0.0f && fabsf(dj) > 1e-7f && dj > 0.0f)) return false;
    const float ti = (-cam_side[i]) / di;
    const float tj = (-cam_side[j]) / dj;
    return ti > tj;
}

// spectral_absorption per eigen-axis (raytracer/absorption.rs:28-30, optics/absorption.rs:68-71)
// into the lane's LDS caches, summed from -0.0 like Iterator::sum; the kernels built for
// biaxial stones add the beta band set of a trichroic stone (refraction.rs:151-166; otherwise
// their tensor is uniaxial, (ao, ao, ae)).  Called once per path (PathState::have_abs) by
// the scatter step; the interior Beer-Lambert block spells the same loop out.
// With the scene's material table, the sums are read from the path's entry.
template <bool kBiax>
__device__ __forceinline__ void fill_absorption_cache(float hero_off, const DevScene &sc, float *__restrict__ lcache) {
    // From the material table the sums were read with the dispersion indices at the path's
    // first hit (every interior segment comes after it); they are evaluated otherwise.
    if (sc.mat_tab && hero_entry(hero_off) < kHeroEntries) return;
    const float h = fill_offset(hero_off);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float lk = comb_lambda(h, k);
        const float so = band_sum<0>(lk, sc);
        LC_AO(k) = so;
        LC_AE(k) = band_sum<1>(lk, sc);
        if (kBiax) LC_AB(k) = sc.has_beta_bands ? band_sum<2>(lk, sc) : so;
    }
}

// channel_absorption_alphas (raytracer/absorption.rs:174-214): each channel's pleochroic
// coefficient p alpha_pol + (1 - p) alpha_unpol for wave normal k and plane-of-incidence
// normal s_axis, the eigenmodes being the hero's biaxial indicatrix's or the uniaxial
// ordinary / extraordinary pair.  The channel-invariant parts of electric_field_direction
// and of the quadratic forms (the s/p axes and their projections on the absorption frame
// (a1, a2, c)) are formed once; per channel the form follows from (Q, U) (qf_channel).
// Each coefficient is handed to use(c, alpha) as soon as it is formed, so that Beer-Lambert
// consumes it at once instead of holding all eight (register pressure of the hot kernels).
template <bool kBiax, class Use>
__device__ __forceinline__ void channel_alphas(const PathState &ps, const DevScene &sc,
                                               const float *__restrict__ lcache, V3 k, V3 s_axis, Use &&use) {
    V3 ea, eb;
    if (kBiax) {
        const BxFrame fr{afr(sc, 0), afr(sc, 1), afr(sc, 2)};
        bx_eigen_polarizations(bx_of(LC_NO(0), sc.delta_beta_alpha, sc.birefringence_delta), fr, k, &ea, &eb);
    } else {
        const V3 c_axis = v3(sc.c_axis[0], sc.c_axis[1], sc.c_axis[2]);
        uniaxial_eigen_polarizations_hw(k, c_axis, &ea, &eb);
    }
    // These directions only orient the absorption frame: hardware normalizations.
    V3 kk = normalize_or_zero_hw(k);
    V3 s_raw = sub(s_axis, mul(kk, dot(kk, s_axis)));
    V3 s_hat = len2(s_raw) > 1e-8f ? normalize_hw(s_raw) : arbitrary_perpendicular(kk);
    V3 p_hat = cross(kk, s_hat);
    V3 ea_l = v3(dot(afr(sc, 0), ea), dot(afr(sc, 1), ea), dot(afr(sc, 2), ea));
    V3 eb_l = v3(dot(afr(sc, 0), eb), dot(afr(sc, 1), eb), dot(afr(sc, 2), eb));
    V3 s_l = v3(dot(afr(sc, 0), s_hat), dot(afr(sc, 1), s_hat), dot(afr(sc, 2), s_hat));
    V3 p_l = v3(dot(afr(sc, 0), p_hat), dot(afr(sc, 1), p_hat), dot(afr(sc, 2), p_hat));
    if (kBiax) {
        // The trichroic tensor diag(ao, ab, ae) (birefringence.rs:238-286): the same sums as
        // below with the a1 and a2 axes weighted apart.
        const QfInv3 qi = qf_invariants3(s_l, p_l);
        const float u_x = (ea_l.x * ea_l.x + eb_l.x * eb_l.x) * 0.5f;
        const float u_y = (ea_l.y * ea_l.y + eb_l.y * eb_l.y) * 0.5f;
        const float u_z = (ea_l.z * ea_l.z + eb_l.z * eb_l.z) * 0.5f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (!GR_LIVE(ps, c)) continue;
            const float ao = LC_AO(c), ab = LC_AB(c), ae = LC_AE(c);
            const float qf_e = qf_channel3(qi, ps.st[c].q, ps.st[c].u, ao, ab, ae);
            const float unpol = fmaf(ao, u_x, fmaf(ab, u_y, ae * u_z));
            const float p = clampf(degree_of_polarization(ps.st[c]), 0.0f, 1.0f);
            use(c, fmaf(p, qf_e - unpol, unpol));
        }
    } else {
        const QfInv qi = qf_invariants(s_l, p_l);
        // The unpolarized term midpoint(ea.A.ea, eb.A.eb) with its channel-invariant parts
        // gathered: ao (|ea_xy|^2 + |eb_xy|^2) / 2 + ae (ea_z^2 + eb_z^2) / 2.
        const float u_xy = ((ea_l.x * ea_l.x + ea_l.y * ea_l.y) + (eb_l.x * eb_l.x + eb_l.y * eb_l.y)) * 0.5f;
        const float u_z = (ea_l.z * ea_l.z + eb_l.z * eb_l.z) * 0.5f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (!GR_LIVE(ps, c)) continue;
            const float ao = LC_AO(c), ae = LC_AE(c);
            const float qf_e = qf_channel(qi, ps.st[c].q, ps.st[c].u, ao, ae);
            const float unpol = fmaf(ao, u_xy, ae * u_z);
            const float p = clampf(degree_of_polarization(ps.st[c]), 0.0f, 1.0f);
            use(c, fmaf(p, qf_e - unpol, unpol));
        }
    }
}

// The kernels whose interior clip takes the screened scan: those without the rare features
// (frosted facets, volume scattering), where it measured slower.
template <bool kAbsorb, bool kFrost>
constexpr bool kScreenedClip = !kFrost;

// One iteration of the bounce loop (transport.rs:781-929).  Returns true when the path
// has ended (escape, Russian roulette, or the bounce cap).  kAniso: 0 isotropic, 1
// uniaxial, 2 biaxial (an anisotropic stone with a biaxial indicatrix).  kInfo: keep the
// two path hashes of the per-tuple record (facet/event sequence and per-channel fates);
// the kernels that write no record skip them, and with them the dead channels' refraction
// indices, which only the fate hash reads.
template <int kAniso, bool kAbsorb, bool kFrost, bool kInfo>
__device__ __forceinline__ bool bounce_step(PathState &ps, const DevScene &sc, const float4 *__restrict__ planes,
                                            const float *__restrict__ cam_side, float *__restrict__ lcache,
                                            uint32_t *ev) {
    constexpr bool kBiax = kAniso == 2;
    const V3 c_axis = v3(sc.c_axis[0], sc.c_axis[1], sc.c_axis[2]);
    const float dn = sc.birefringence_delta;
    const uint32_t bounce = ps.bounce;
    ev_inc(ev, GR_EV_SEGMENT);

    // ---- convex facet-plane clip (simd/slab.rs:94-152, intersect.rs:93-132) ----
    // The entering extreme starts at the hit threshold 1e-4 instead of -1e30.  The scan's
    // outcome reads t_near only through t_near > 1e-4 (then it is the hit, with its facet)
    // and t_near > t_far; planes whose exact t is at most 1e-4 can change neither (if no
    // entering plane exceeds 1e-4 the hit is t_far > 1e-4 or a miss, whatever t_near is), so
    // the screen skips them: a ray inside the stone, whose entering planes all lie behind it,
    // takes no division for them.  Same hit, t and facet as the reference's scan.
    float t_near = 1e-4f, t_far = 1e30f;
    int near_idx = -1, far_idx = -1;
    bool outside = false;
    const V3 ro = ps.ray_o, rd = ps.ray_d;
    if (bounce == 0u) {
        // Camera rays share their origin: the origin's signed distances are staged per
        // workgroup (same operations, same values).  Most background rays are proved to
        // miss by their pixel's certificate.
        if (camera_miss_certified(planes, cam_side, ps.cert, rd)) {
            ev_inc(ev, GR_EV_CERT_MISS);
            outside = true;
        } else
            clip_scan<true>(planes, cam_side, sc.n_planes, ro, rd, t_near, near_idx, t_far, far_idx, outside);
    } else if (!ps.inside && exterior_miss(planes[ps.last_facet], ro, rd)) {
        // The ray left the stone's surface outwards (a refraction out, or a reflection off
        // the outside): it starts outside the facet it just met and moves away from it, so
        // that facet alone makes the scan's far extreme negative and the scan a miss.
        ev_inc(ev, GR_EV_EXTERIOR_MISS);
        outside = true;
    } else if (kScreenedClip<kAbsorb, kFrost>) {
        // The screened scan (the planes padded with copies of plane 0, which it rules out);
        // in the rare-feature kernels it measured slower, the register allocation of the
        // longer scan costing more than the screen saves.
        clip_scan_screened(planes, sc.n_planes, sc.clip_w_max, sc.clip_b_max, ro, rd, t_near, near_idx, t_far, far_idx,
                           outside);
    } else {
        clip_scan<false>(planes, cam_side, sc.n_planes, ro, rd, t_near, near_idx, t_far, far_idx, outside);
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
        // Escape: the environment is sampled by finish_path (the path's only radiance
        // contribution; transport.rs:307-318).
        ev_inc(ev, GR_EV_MISS);
        if (bounce == 0u) ev_inc(ev, GR_EV_FIRST_MISS);
        ps.escaped = true;
        if (kInfo) fnv(ps.fh, 0xFFFFu);
        ps.term_bounce = bounce;
        ps.term_reason = 0u;
        return true;
    }
    ev_inc(ev, GR_EV_HIT);
    if (!ps.have_n) {
        // The comb's dispersion indices: the material table's entry, or evaluated.  From the
        // table, an absorbing stone's band sums come in the same round trip (the rest of the
        // entry), instead of at the path's first interior segment.
        const uint32_t e = sc.mat_tab ? hero_entry(ps.hero_off) : kHeroEntries;
        if (e < kHeroEntries) {
            const float *m = sc.mat_tab + (size_t)e * mat_floats(kBiax);
#pragma unroll
            for (int k = 0; k < 8; ++k) LC_NO(k) = m[k];
            if (kAbsorb) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    LC_AO(k) = m[8 + k];
                    LC_AE(k) = m[16 + k];
                    if (kBiax) LC_AB(k) = m[24 + k];
                }
            }
        } else {
            const float h = fill_offset(ps.hero_off);
#pragma unroll
            for (int k = 0; k < 8; ++k) LC_NO(k) = dispersion_index(comb_lambda(h, k), sc);
        }
        ps.have_n = true;
    }
    // ---- volume scattering (try_scatter_step + maybe_scatter_or_extinguish,
    //      scattering.rs:132-336; call site transport.rs:805-848) ----
    // Inside a scattering stone, before the facet hit is processed: an exponential free path
    // on the hero's extinction alpha_hero + sigma_s, the alphas taken in the previous plane
    // of incidence (none after a scatter).  A scatter before the facet depolarizes and
    // reweights every channel, turns the ray by the Henyey-Greenstein phase function, rolls
    // the roulette past bounce 4 and ends the iteration; otherwise every channel takes its
    // extinction over the segment relative to the hero's, in place of Beer-Lambert.  Traced
    // by the kernels built for the rare features (kFrost).
    const bool scattering = kFrost && !(sc.sigma_s <= 0.0f);
    if (scattering && ps.inside) {
        ev_inc(ev, GR_EV_SCATTER);
        const float sigma_s = sc.sigma_s;
        float alphas[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) alphas[k] = 0.0f;  // a dead channel's stays unused
        if (kAbsorb) {
            if (!ps.have_abs) {
                ev_inc(ev, GR_EV_ABSORPTION_FILL);
                fill_absorption_cache<kBiax>(ps.hero_off, sc, lcache);
                ps.have_abs = true;
            }
            channel_alphas<kBiax>(ps, sc, lcache, ps.cur_k, ps.have_prev ? ps.prev_pn : v3(0.0f, 0.0f, 0.0f),
                                  [&](int c, float a) { alphas[c] = a; });
        }  // no band: every coefficient is zero
        const float sigma_t_hero = alphas[0] + sigma_s;
        const float one_minus_u = fmaxf(1.0f - stream_draw(ps.seed, bounce, DISTANCE_SAMPLE_STREAM), 1e-7f);
        const float t_free = -(logf(one_minus_u)) / sigma_t_hero;
        const float hit_t_scaled = hit_t * sc.absorption_path_scale;
        if (t_free < hit_t_scaled) {
            const float pdf_hero = sigma_t_hero * one_minus_u;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (!GR_LIVE(ps, k)) continue;
                const float sigma_t_k = alphas[k] + sigma_s;
                const float tr_k = exp_lobe(-(alphas[k] + sigma_s) * t_free);
                const float weight = tr_k * sigma_s / pdf_hero;
                ps.st[k] = Stokes{intensity(ps.st[k]) * weight, 0.0f, 0.0f, 0.0f};
                ps.pdf[k] *= sigma_t_k * tr_k;
            }
            const float u1 = stream_draw(ps.seed, bounce, PHASE_DIR_U_STREAM);
            const float u2 = stream_draw(ps.seed, bounce, PHASE_DIR_V_STREAM);
            const V3 new_dir = sample_henyey_greenstein_direction(u1, u2, sc.scattering_g, rd);
            ps.ray_o = add(ro, mul(rd, t_free / sc.absorption_path_scale));
            ps.ray_d = new_dir;
            ps.cur_k = new_dir;
            ps.have_prev = false;
            if (kInfo) fnv(ps.fh, 0xFFFEu);
            if (bounce > 4) {
                // apply_russian_roulette (transport.rs:204-220); an ended path is
                // PathTermination::ScatterAbsorbed.
                ev_inc(ev, GR_EV_RR);
                float max_i = 0.0f;
#pragma unroll
                for (int k = 0; k < 8; ++k) max_i = fmaxf(max_i, intensity(ps.st[k]));
                const float q = clampf(max_i, 0.05f, 1.0f);
                if (stream_draw(ps.seed, bounce, RUSSIAN_ROULETTE_STREAM) > q) {
                    ps.term_bounce = bounce;
                    ps.term_reason = 1u;
                    return true;
                }
                const float inv_q = 1.0f / q;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (GR_LIVE(ps, k)) ps.st[k] = scale(ps.st[k], inv_q);
            }
            ps.bounce = bounce + 1u;
            return ps.bounce >= sc.max_bounces;
        }
        float survive[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) survive[k] = GR_LIVE(ps, k) ? exp_lobe(-(alphas[k] + sigma_s) * hit_t_scaled) : 0.0f;
        const float survive_hero = fmaxf(survive[0], 1e-30f);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!GR_LIVE(ps, k)) continue;
            ps.st[k] = scale(ps.st[k], survive[k] / survive_hero);
            ps.pdf[k] *= survive[k];
        }
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
    V3 hit_point = add(ro, mul(rd, hit_t));
    // Facet edge rounding (transport.rs:850-857); off (the flat normal) for radius <= 0.
    if (sc.edge_rounding_radius > 0.0f)
        normal = shading_normal_near_edge(planes, sc.n_planes, hit_point, facet, normal, sc.edge_rounding_radius);
    V3 cur_k = ps.cur_k;
    const bool inside = ps.inside;

    // ---- Stokes frame rotation into the new plane of incidence
    //      (raytracer/absorption.rs:44-89) ----
    // The plane-of-incidence normal only orients Stokes frames and absorption axes.
    V3 cpn = normalize_or_zero_hw(cross(cur_k, normal));
    if (ps.have_prev && len2(cpn) > 1e-6f && len2(ps.prev_pn) > 1e-6f) {
        ev_inc(ev, GR_EV_ROTATE);
        float c2, s2;
        frame_rotation_cs(ps.prev_pn, cpn, cur_k, &c2, &s2);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (GR_LIVE(ps, k)) ps.st[k] = apply_rotation(ps.st[k], c2, s2);
    }
    ps.prev_pn = cpn;
    ps.have_prev = true;

    // ---- interior segment: pleochroic Beer-Lambert (raytracer/absorption.rs:115-216,
    //      birefringence.rs:252-286, polarization.rs:172-192); a scattering stone's
    //      extinction was taken by the scatter step above (transport.rs:586-606) ----
    if (inside) {
        normal = neg(normal);
        if (!scattering) {
            ev_inc(ev, GR_EV_ABSORB);
            if (kAbsorb) {
                // The cache fill is written out here rather than calling
                // fill_absorption_cache: with the call, the compiler's allocation of this
                // (hot) block spilled 178 VGPRs instead of 30 (same arithmetic).
                if (!ps.have_abs) {
                    ev_inc(ev, GR_EV_ABSORPTION_FILL);
                    fill_absorption_cache<kBiax>(ps.hero_off, sc, lcache);
                    ps.have_abs = true;
                }
                const float spl = hit_t * sc.absorption_path_scale;
                channel_alphas<kBiax>(ps, sc, lcache, cur_k, cpn,
                                      [&](int c, float a) { ps.st[c] = scale(ps.st[c], exp_lobe(-a * spl)); });
            }
        }
    }

    // ---- per-bounce refraction geometry (refraction.rs:235-547) ----
    float cos_i = clampf(dot(neg(cur_k), normal), 0.0f, 1.0f);
    float sin_i = sqrtf(fmaxf(fmaf(cos_i, -cos_i, 1.0f), 0.0f));
    float one_m_c2 = fmaf(cos_i, -cos_i, 1.0f);
    float n_med[8];
    float n_biax_a[8];  // biaxial entry: each channel's mode-A (fast) index
    if (kBiax) {
        // hero_biaxial_wave_dirs + per_channel_biaxial_indices (refraction.rs:345-400): one
        // wave normal per mode from the hero's indicatrix -- at an entry the two-step
        // fixed point of resolve_entry_mode seeded with n_beta, inside the stone the
        // tracked wave normal -- and each channel's own mode indices there.  Inside, the
        // medium is mode A (fast) unless the path took mode B; at an entry, mode B.
        const BxFrame fr{afr(sc, 0), afr(sc, 1), afr(sc, 2)};
        const float dba = sc.delta_beta_alpha;
        if (inside) {
            const bool use_a = !ps.is_extra;
            const V3 l = bx_local(fr, normalize_or_zero(cur_k));
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (!kInfo && !GR_LIVE(ps, k)) continue;
                float ns, nf;
                bx_wave_indices(bx_of(LC_NO(k), dba, dn), l, &ns, &nf);
                n_med[k] = use_a ? nf : ns;
                n_biax_a[k] = nf;
            }
        } else {
            const Bx hero = bx_of(LC_NO(0), dba, dn);
            const V3 dir_a = bx_entry_wave_dir(hero, fr, cur_k, normal, cos_i, one_m_c2, LC_NO(0), false);
            const V3 dir_b = bx_entry_wave_dir(hero, fr, cur_k, normal, cos_i, one_m_c2, LC_NO(0), true);
            const V3 la = bx_local(fr, normalize_or_zero(dir_a));
            const V3 lb = bx_local(fr, normalize_or_zero(dir_b));
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (!kInfo && !GR_LIVE(ps, k)) continue;
                const Bx bk = bx_of(LC_NO(k), dba, dn);
                float ns, nf;
                bx_wave_indices(bk, la, &ns, &nf);
                n_biax_a[k] = nf;
                bx_wave_indices(bk, lb, &ns, &nf);
                n_med[k] = ns;
            }
        }
    } else if (kAniso) {
        const bool use_o = inside && !ps.is_extra;
        if (use_o) {
#pragma unroll
            for (int k = 0; k < 8; ++k) n_med[k] = LC_NO(k);
        } else {
            // theta_c, the angle between the wave normal and the optic axis, enters only
            // through its sine and cosine.  The reference forms theta_c = acos(|cos|) and
            // takes sin/cos of it (refraction.rs:248-269, birefringence.rs:27-40); here they
            // come from the cosine directly, sin = sqrt(1 - cos^2), shared by the 8 channels.
            float ct = 1.0f, st = 0.0f;  // theta_c = 0
            if (!inside) {
                // theta_c_for_bounce's two-step fixed point.
                const float n_seed = LC_NO(0);
                const float n_e_seed = n_seed + dn;
                float n_guess = n_seed;
                for (int it = 0; it < 2; ++it) {
                    float eta_g = 1.0f / n_guess;
                    float sin2_g = eta_g * eta_g * one_m_c2;
                    if (sin2_g > 1.0f) break;
                    float cos_t_g = sqrtf(fmaxf(1.0f - sin2_g, 0.0f));
                    V3 wave = normalize(add(mul(cur_k, eta_g), mul(normal, fmaf(eta_g, cos_i, -cos_t_g))));
                    ct = fabsf(clampf(dot(wave, c_axis), -1.0f, 1.0f));
                    st = sqrtf(fmaxf(fmaf(ct, -ct, 1.0f), 0.0f));
                    n_guess = effective_extraordinary_index_sc(n_seed, n_e_seed, st, ct);
                }
            } else {
                ct = fabsf(clampf(dot(cur_k, c_axis), -1.0f, 1.0f));
                st = sqrtf(fmaxf(fmaf(ct, -ct, 1.0f), 0.0f));
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (!kInfo && !GR_LIVE(ps, k)) continue;
                n_med[k] = effective_extraordinary_index_sc(LC_NO(k), LC_NO(k) + dn, st, ct);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) n_med[k] = LC_NO(k);
    }
    // A dead channel's index and sin^2(theta_t) are read only by the fate hash (the
    // reflection and refraction loops skip the channel).
    float s2t_ch[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (!kInfo && !GR_LIVE(ps, k)) continue;
        float n1k = inside ? n_med[k] : 1.0f;
        float n2k = inside ? 1.0f : n_med[k];
        float etak = n1k / n2k;
        s2t_ch[k] = etak * etak * one_m_c2;
    }
    const float n1 = inside ? n_med[0] : 1.0f, n2 = inside ? 1.0f : n_med[0];
    const float sin2_t = s2t_ch[0];
    const float n_o_hero = LC_NO(0), n_e_hero = LC_NO(0) + dn;

    // ---- TIR / partial-reflect / refract dispatch (transport.rs:128-193) ----
    V3 new_k, new_s;
    bool new_inside = inside, has_update = false, update = false;
    uint32_t event;
    // Hero event: TIR when sin2_t > 1, else a Fresnel-weighted reflect/refract draw
    // (transport.rs:158-176, refraction.rs:1036-1086).
    const bool tir = sin2_t > 1.0f;
    float r_unpol = 1.0f;
    bool reflect = tir;
    // The hero's Fresnel amplitudes (set when !tir).  The channel loops below form channel
    // 0's from the same operands (n1, n2, cos_i and cos_t = sqrt(1 - sin2_t), whose fmaxf
    // with 0 is a no-op here), so they take these instead.
    float r_s = 0.0f, r_p = 0.0f;
    if (!tir) {
        float cos_t = sqrtf(1.0f - sin2_t);
        fresnel_rs_rp(n1, n2, cos_i, cos_t, &r_s, &r_p);
        r_unpol = clampf(0.5f * fmaf(r_p, r_p, r_s * r_s), 1e-4f, 1.0f - 1e-4f);
        reflect = stream_draw(ps.seed, bounce, FRESNEL_BRANCH_STREAM) < r_unpol;
    }
    if (kFrost && sc.n_frosted != 0u && sc.finishes[facet] == GR_FINISH_FROSTED) {
        // apply_frosted_bounce (scattering.rs:522-590): the hero's TIR or Fresnel draw
        // (above) picks a diffuse reflection or transmission, cosine-distributed about the
        // facet normal; every channel is depolarized, its intensity divided by the hero's
        // branch probability, and k' = S'.
        const float u1 = stream_draw(ps.seed, bounce, FROSTED_DIR_U_STREAM);
        const float u2 = stream_draw(ps.seed, bounce, FROSTED_DIR_V_STREAM);
        if (reflect) {
            if (tir) {
                ev_inc(ev, GR_EV_TIR);
                event = 4u;
            } else {
                ev_inc(ev, GR_EV_REFLECT);
                event = 5u;
            }
            new_k = cosine_weighted_hemisphere(u1, u2, normal);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (!GR_LIVE(ps, k)) continue;
                const float i = intensity(ps.st[k]);
                ps.st[k] = Stokes{tir ? i : i / r_unpol, 0.0f, 0.0f, 0.0f};
                if (!tir) ps.pdf[k] *= r_unpol;
            }
        } else {
            ev_inc(ev, GR_EV_REFRACT);
            event = 6u;
            new_k = cosine_weighted_hemisphere(u1, u2, neg(normal));
            const bool entering = kAniso && !inside;
            bool use_extra = ps.is_extra;
            if (entering) {
                ev_inc(ev, GR_EV_ENTRY_ANISO);
                use_extra = stream_draw(ps.seed, bounce, BIREFRINGENT_SPLIT_STREAM) < 0.5f;
            }
            const float t_unpol = 1.0f - r_unpol;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (!GR_LIVE(ps, k)) continue;
                ps.st[k] = Stokes{intensity(ps.st[k]) / t_unpol, 0.0f, 0.0f, 0.0f};
                ps.pdf[k] *= t_unpol;
            }
            new_inside = !inside;
            has_update = entering;
            update = use_extra;
        }
        new_s = new_k;
    } else if (reflect) {
        // apply_tir_bounce (refraction.rs:599-637) and apply_partial_reflect_bounce
        // (:650-690) share their per-channel work: retardation for channels past their
        // own critical angle, Fresnel reflection (and its pdf factor) otherwise.  The
        // partial reflection then divides by the hero's r_unpol; TIR scales by 1.0, which
        // is exact.
        if (tir) {
            ev_inc(ev, GR_EV_TIR);
            event = 1u;
        } else {
            ev_inc(ev, GR_EV_REFLECT);
            event = 2u;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!GR_LIVE(ps, k)) continue;
            const float n1k = inside ? n_med[k] : 1.0f, n2k = inside ? 1.0f : n_med[k];
            if (s2t_ch[k] > 1.0f) {
                float sd, cd;
                tir_retardation_cs(n1k, cos_i, sin_i, &cd, &sd);
                ps.st[k] = apply_retardation(ps.st[k], cd, sd);
            } else {
                float rs = r_s, rp = r_p;  // channel 0 (not past its critical angle: !tir)
                if (k != 0) {
                    // A companion's amplitudes only scale its own Stokes vector and weight.
                    float cos_t_k = sqrt_hw(fmaxf(1.0f - s2t_ch[k], 0.0f));
                    fresnel_rs_rp_hw(n1k, n2k, cos_i, cos_t_k, &rs, &rp);
                }
                ps.pdf[k] *= clampf(0.5f * fmaf(rp, rp, rs * rs), 1e-4f, 1.0f - 1e-4f);
                ps.st[k] = apply_fresnel_reflection(ps.st[k], rs, rp);
            }
        }
        // The partial reflection divides every channel by the hero's r_unpol; TIR's factor is
        // exactly 1.0, and skipping it leaves every bit as it is.
        if (!tir) {
            const float inv_r = 1.0f / r_unpol;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (GR_LIVE(ps, k)) ps.st[k] = scale(ps.st[k], inv_r);
        }
        new_k = sub(cur_k, mul(normal, 2.0f * dot(cur_k, normal)));
        new_s = new_k;
        if (kBiax && inside) {
            // Both biaxial modes walk off (refraction.rs:574-576).
            ev_inc(ev, GR_EV_WALKOFF);
            const BxFrame fr{afr(sc, 0), afr(sc, 1), afr(sc, 2)};
            new_s = bx_mode_poynting_dir(bx_of(LC_NO(0), sc.delta_beta_alpha, dn), fr, new_k, ps.is_extra);
        } else if (kAniso && inside && ps.is_extra) {
            ev_inc(ev, GR_EV_WALKOFF);
            new_s = extraordinary_poynting_dir(new_k, c_axis, n_o_hero, n_e_hero);
        }
    } else {
        // apply_refract_bounce + apply_refract_channel (refraction.rs:748-1011).
        ev_inc(ev, GR_EV_REFRACT);
        event = 3u;
        const float DIRECTION_MATCH_COS_TOL = 1.0f - 1e-6f;
        bool entering = kAniso && !inside;
        bool use_extra = ps.is_extra;
        if (entering) {
            ev_inc(ev, GR_EV_ENTRY_ANISO);
            use_extra = stream_draw(ps.seed, bounce, BIREFRINGENT_SPLIT_STREAM) < 0.5f;
        }
        // At a biaxial entry the hero refracts with its own mode index (mode B: n_med, mode
        // A: n_biax_a) and both modes walk off (refraction.rs:803-818).
        const BxFrame fr{afr(sc, 0), afr(sc, 1), afr(sc, 2)};
        const float dba = sc.delta_beta_alpha;
        float n2_dir;
        if (kBiax && entering)
            n2_dir = use_extra ? n_med[0] : n_biax_a[0];
        else
            n2_dir = (entering && !use_extra) ? n_o_hero : n2;
        float eta_dir = n1 / n2_dir;
        float sin2_dir = fminf(eta_dir * eta_dir * one_m_c2, 1.0f);
        float cos_t_dir = sqrtf(fmaxf(1.0f - sin2_dir, 0.0f));
        V3 refr = normalize(add(mul(cur_k, eta_dir), mul(normal, fmaf(eta_dir, cos_i, -cos_t_dir))));
        V3 final_dir = refr;
        if (kBiax && entering) {
            ev_inc(ev, GR_EV_WALKOFF);
            final_dir = bx_mode_poynting_dir(bx_of(LC_NO(0), dba, dn), fr, refr, use_extra);
        } else if (entering && use_extra) {
            ev_inc(ev, GR_EV_WALKOFF);
            final_dir = extraordinary_poynting_dir(refr, c_axis, n_o_hero, n_e_hero);
        }
        float inv_t = 1.0f / (1.0f - r_unpol);
        // Channel 0's medium indices are the hero's (n1, n2_dir), so its q, cos_t_k and
        // refracted (and walked-off) direction are eta_dir, cos_t_dir and final_dir, formed
        // above by the same operations (its s2k never exceeds 1 here, so the fminf is a
        // no-op); unless it enters in the ordinary mode (n2k = n_o instead of n2), its
        // Fresnel amplitudes are the hero event's r_s, r_p.
        const bool hero_rs = !(entering && !use_extra);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!GR_LIVE(ps, k)) continue;
            float n1k = inside ? n_med[k] : 1.0f;
            float n2k = (entering && !use_extra) ? (kBiax ? n_biax_a[k] : LC_NO(k)) : (inside ? 1.0f : n_med[k]);
            float cos_t_k;
            bool alive;
            if (k == 0) {
                cos_t_k = cos_t_dir;
                alive = dot(final_dir, final_dir) >= DIRECTION_MATCH_COS_TOL;
            } else {
                float q = n1k / n2k;
                float s2k = q * q * one_m_c2;
                alive = !(s2k > 1.0f);
                cos_t_k = sqrtf(fmaxf(1.0f - s2k, 0.0f));
                if (alive) {
                    V3 dk = normalize(add(mul(cur_k, q), mul(normal, fmaf(q, cos_i, -cos_t_k))));
                    if (kBiax && entering)
                        dk = bx_mode_poynting_dir(bx_of(LC_NO(k), dba, dn), fr, dk, use_extra);
                    else if (entering && use_extra)
                        dk = extraordinary_poynting_dir(dk, c_axis, LC_NO(k), LC_NO(k) + dn);
                    alive = dot(dk, final_dir) >= DIRECTION_MATCH_COS_TOL;
                }
            }
            if (alive) {
                // The transmission and the weight's amplitudes only scale this channel's
                // Stokes vector and weight (the draw used the hero's r_unpol above).
                float num = 2.0f * n1k * cos_i;
                float t_s = div_hw(num, fmaf(n2k, cos_t_k, n1k * cos_i));
                float t_p = div_hw(num, fmaf(n1k, cos_t_k, n2k * cos_i));
                ps.st[k] = scale(apply_fresnel_transmission(ps.st[k], n1k, n2k, cos_i, cos_t_k, t_s, t_p), inv_t);
                float rs = r_s, rp = r_p;
                if (k != 0 || !hero_rs) fresnel_rs_rp_hw(n1k, n2k, cos_i, cos_t_k, &rs, &rp);
                ps.pdf[k] *= 1.0f - clampf(0.5f * fmaf(rp, rp, rs * rs), 1e-4f, 1.0f - 1e-4f);
            } else {
                ps.st[k] = scale(ps.st[k], 0.0f);
                ps.pdf[k] = 0.0f;
                ps.dead |= 1u << k;
            }
        }
        new_k = refr;
        new_s = final_dir;
        new_inside = !inside;
        has_update = entering;
        update = use_extra;
    }
    bool internal_reflection = inside && !has_update && new_inside == inside;
    ps.ray_d = new_s;
    ps.ray_o = add(hit_point, mul(new_s, 1e-4f));
    ps.cur_k = new_k;
    ps.inside = new_inside;
    if (has_update) ps.is_extra = update;
    if (kAniso && internal_reflection)
        ps.is_extra = stream_draw(ps.seed, bounce, MODE_COUPLING_STREAM) < 0.5f;
    ps.last_facet = (uint32_t)facet;
    if (kInfo) {
        fnv(ps.fh, (uint32_t)facet);
        fnv(ps.fh, event | (ps.is_extra ? 16u : 0u));
        // Per-channel fates: past the channel's critical angle / pdf zero.
        uint32_t fate = 0u;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            fate |= (s2t_ch[k] > 1.0f ? 1u : 0u) << k;
            fate |= (ps.pdf[k] == 0.0f ? 1u : 0u) << (8 + k);
        }
        fnv(ps.fate_h, fate);
    }

    // ---- Russian roulette (transport.rs:204-220) ----
    if (bounce > 4) {
        ev_inc(ev, GR_EV_RR);
        float max_i = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) max_i = fmaxf(max_i, intensity(ps.st[k]));
        float q = clampf(max_i, 0.05f, 1.0f);
        if (stream_draw(ps.seed, bounce, RUSSIAN_ROULETTE_STREAM) > q) {
            ps.term_bounce = bounce;
            ps.term_reason = 2u;
            return true;
        }
        float inv_q = 1.0f / q;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (GR_LIVE(ps, k)) ps.st[k] = scale(ps.st[k], inv_q);
    }
    ps.bounce = bounce + 1u;
    return ps.bounce >= sc.max_bounces;
}

// Environment radiance of the escape direction, spectral MIS, CIE XYZ and white balance
// (transport.rs:67-90, color.rs:121-164, 215-251).  A path that never escaped carries
// zero radiance: the reference's integration and white balance then produce exactly +0
// in every component (every product is +-0 and every sum has a +0 operand), so that work
// is skipped by the callers.
template <int kEnv>
__device__ __forceinline__ V3 finish_escaped(V3 dir, const float *inten, const float *pdf,
                                             float hero_off, const DevScene &sc) {
    float rad[8];
    // The spectral table's entry for this hero offset (kHeroEntries: evaluate instead), read
    // whole before the environment lookup: one round trip to memory, overlapping that work,
    // instead of one for the lobes after it and one per live channel's CIE values.
    const uint32_t he = sc.spec_tab ? hero_entry(hero_off) : kHeroEntries;
    float spec[kSpecFloats];
    if (he < kHeroEntries) {
        const float4 *e4 = reinterpret_cast<const float4 *>(sc.spec_tab + (size_t)he * kSpecFloats);
#pragma unroll
        for (int i = 0; i < kSpecFloats / 4; ++i) {
            const float4 v = e4[i];
            spec[4 * i] = v.x;
            spec[4 * i + 1] = v.y;
            spec[4 * i + 2] = v.z;
            spec[4 * i + 3] = v.w;
        }
    }
    if (kEnv == 0) {
        // sample_studio_environment_with_rig (environment.rs:280-321): the direction-only
        // terms are formed once, the fma chain runs per channel.
        V3 d = normalize(dir);
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
            // Used only when rd > 0.96, where div_by_const is the correctly rounded quotient.
            float spark = div_by_const(rd - 0.96f, 0.04f, 1.0f / 0.04f);
            terms[2 + r] = powi_cr<6>(spark) * 22.0f * sc.spot_mult * sc.exposure;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float spec = blackbody(comb_lambda(hero_off, k), sc);
            float r = bg * spec;
#pragma unroll
            for (int j = 0; j < 18; ++j)
                if (act[j]) r = fmaf(terms[j], spec, r);
            rad[k] = fmaf(inten[k], r, 0.0f);
        }
    } else {
        V3 rgb = map_rgb(dir, sc);
        if (he < kHeroEntries) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const V3 lobes = v3(spec[24 + 3 * k], spec[25 + 3 * k], spec[26 + 3 * k]);
                rad[k] = fmaf(inten[k], rgb_to_spectral_lobes(rgb, lobes), 0.0f);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) rad[k] = fmaf(inten[k], rgb_to_spectral(rgb, comb_lambda(hero_off, k)), 0.0f);
        }
    }
    float sum_pdf = -0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) sum_pdf = sum_pdf + pdf[k];
    float mis = sum_pdf <= 1e-12f ? 1.0f : (8.0f * pdf[0]) / sum_pdf;
    const float norm_factor = (400.0f / 8.0f) / 106.856f;
    // A channel whose weight is +-0 (a dead channel: zero intensity) adds +-0 to each
    // component: the CIE lobes are finite, and a sum that starts at +0 never holds -0, so
    // x + (+-0) == x bit for bit.  Its colour-matching evaluation is skipped.
    V3 xyz = v3(0.0f, 0.0f, 0.0f);
    if (he < kHeroEntries) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float w = rad[k] * mis * norm_factor;
            if (w == 0.0f) continue;
            xyz = add(xyz, mul(v3(spec[3 * k], spec[3 * k + 1], spec[3 * k + 2]), w));
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float w = rad[k] * mis * norm_factor;
            if (w == 0.0f) continue;
            xyz = add(xyz, mul(cmf_x8_channel(comb_lambda(hero_off, k)), w));
        }
    }
    V3 lms = v3(fmaf(0.8951f, xyz.x, fmaf(0.2664f, xyz.y, -0.1614f * xyz.z)),
                fmaf(-0.7502f, xyz.x, fmaf(1.7135f, xyz.y, 0.0367f * xyz.z)),
                fmaf(0.0389f, xyz.x, fmaf(-0.0685f, xyz.y, 1.0296f * xyz.z)));
    lms = v3(lms.x * sc.wb[0], lms.y * sc.wb[1], lms.z * sc.wb[2]);
    return v3(fmaf(0.986993f, lms.x, fmaf(-0.147054f, lms.y, 0.159963f * lms.z)),
              fmaf(0.432305f, lms.x, fmaf(0.518360f, lms.y, 0.049291f * lms.z)),
              fmaf(-0.008529f, lms.x, fmaf(0.040043f, lms.y, 0.968487f * lms.z)));
}

__device__ __forceinline__ uint32_t path_segments(const PathState &ps, const DevScene &sc) {
    return ps.term_reason == 3u ? sc.max_bounces : ps.term_bounce + 1u;
}

// Continuation records (kContFields, transport.h): the complete state of a path still
// running at the split bounce, field-major (field f of record r at cont[f * cap + r]).
// Fields: 0 tuple index, 1-12 ray_o, ray_d, cur_k, prev_pn, 13-44 Stokes, 45-52 pdfs,
// 53-76 the LDS caches (n_o, ao, ae), 77 hero offset, 78-84 seed, bounce, last facet,
// termination bounce and reason and the two path hashes, 85 the mode flags (the dead-channel
// mask in bits 8-15), and in the
// kernels built for biaxial stones (kBx) 86-93 the beta-axis absorption cache.
//
// A dead channel's fields are neither written nor read: its Stokes vector is +-0 and its
// pdf +0, and every later use of them gives the same result for +0 (the roulette's maximum
// intensity is clamped up to 0.05, the radiance fma(+-0, r, +0) is +0 and a +-0 weight adds
// nothing), so a continued path reloads them as +0; its absorption caches are read by no
// later step, and its dispersion index only by the fate hash (kInfo).  A path parked deep
// in the stone, whose companions have mostly died at the entry refraction, moves about half
// of its record.

template <bool kBx, bool kInfo>
__device__ __forceinline__ void save_path(const PathState &ps, uint32_t t, const float *__restrict__ lcache,
                                          float *__restrict__ cont, uint32_t cap, uint32_t r) {
    float *c = cont + r;
    auto put = [&](int f, float v) { c[(size_t)f * cap] = v; };
    auto putu = [&](int f, uint32_t v) { c[(size_t)f * cap] = __uint_as_float(v); };
    putu(0, t);
    const V3 v[4] = {ps.ray_o, ps.ray_d, ps.cur_k, ps.prev_pn};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        put(1 + 3 * j, v[j].x);
        put(2 + 3 * j, v[j].y);
        put(3 + 3 * j, v[j].z);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const bool live = GR_LIVE(ps, k);
        if (kInfo || live) put(53 + k, LC_NO(k));
        if (!live) continue;
        put(13 + 4 * k, ps.st[k].i);
        put(14 + 4 * k, ps.st[k].q);
        put(15 + 4 * k, ps.st[k].u);
        put(16 + 4 * k, ps.st[k].v);
        put(45 + k, ps.pdf[k]);
        put(61 + k, LC_AO(k));
        put(69 + k, LC_AE(k));
        if (kBx) put(86 + k, LC_AB(k));
    }
    put(77, ps.hero_off);
    putu(78, ps.seed);
    putu(79, ps.bounce);
    putu(80, ps.last_facet);
    if (kInfo) {  // read only by the per-path record
        putu(81, ps.term_bounce);
        putu(82, ps.term_reason);
        putu(83, ps.fh);
        putu(84, ps.fate_h);
    }
    putu(85, (ps.inside ? 1u : 0u) | (ps.is_extra ? 2u : 0u) | (ps.have_prev ? 4u : 0u) | (ps.have_n ? 8u : 0u) |
                 (ps.have_abs ? 16u : 0u) | (ps.escaped ? 32u : 0u) | (ps.dead << 8));
}

template <bool kBx, bool kInfo>
__device__ __forceinline__ uint32_t load_path(PathState &ps, float *__restrict__ lcache,
                                              const float *__restrict__ cont, uint32_t cap, uint32_t r) {
    const float *c = cont + r;
    auto get = [&](int f) { return c[(size_t)f * cap]; };
    auto getu = [&](int f) { return __float_as_uint(c[(size_t)f * cap]); };
    ps.ray_o = v3(get(1), get(2), get(3));
    ps.ray_d = v3(get(4), get(5), get(6));
    ps.cur_k = v3(get(7), get(8), get(9));
    ps.prev_pn = v3(get(10), get(11), get(12));
    const uint32_t fl = getu(85);
    ps.dead = (fl >> 8) & 0xFFu;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const bool live = GR_LIVE(ps, k);
        if (kInfo || live) LC_NO(k) = get(53 + k);
        if (!live) {
            ps.st[k] = Stokes{0.0f, 0.0f, 0.0f, 0.0f};
            ps.pdf[k] = 0.0f;
            continue;
        }
        ps.st[k] = Stokes{get(13 + 4 * k), get(14 + 4 * k), get(15 + 4 * k), get(16 + 4 * k)};
        ps.pdf[k] = get(45 + k);
        LC_AO(k) = get(61 + k);
        LC_AE(k) = get(69 + k);
        if (kBx) LC_AB(k) = get(86 + k);
    }
    ps.hero_off = get(77);
    ps.seed = getu(78);
    ps.bounce = getu(79);
    ps.last_facet = getu(80);
    ps.cert = 0xFFFFFFFFu;  // used at bounce 0 only
    if (kInfo) {
        ps.term_bounce = getu(81);
        ps.term_reason = getu(82);
        ps.fh = getu(83);
        ps.fate_h = getu(84);
    } else {
        ps.term_bounce = 0u;
        ps.term_reason = 3u;
        ps.fh = ps.fate_h = 0u;
    }
    ps.inside = fl & 1u;
    ps.is_extra = fl & 2u;
    ps.have_prev = fl & 4u;
    ps.have_n = fl & 8u;
    ps.have_abs = fl & 16u;
    ps.escaped = fl & 32u;
    return getu(0);
}

// Reserves a continuation record for every calling lane (one atomic per wave) and saves
// the path into it; false when the records have run out (the path then carries on).
// Called by all live lanes of a wave together: in the plain and continuation kernels
// every live lane reaches a given bounce count in the same loop iteration.
template <bool kBx, bool kInfo>
__device__ __forceinline__ bool park(const PathState &ps, uint32_t t, const float *__restrict__ lcache,
                                     float *__restrict__ cont, uint32_t cap, uint32_t *__restrict__ count) {
    const uint64_t live = __ballot(1);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)live) - 1u;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(live));
    base = __shfl(base, (int)leader, 64);
    const uint32_t r = base + (uint32_t)__popcll(live & ((1ull << lane) - 1ull));
    if (r >= cap) return false;
    save_path<kBx, kInfo>(ps, t, lcache, cont, cap, r);
    return true;
}

// Whether a wave parks its live paths after `it` iterations of this launch: at the level's
// cap of `cap` iterations (0: never, the last level), or earlier, from park_min iterations
// on, once fewer than park_lanes of its lanes are still tracing (0: no early parking).  A
// wave whose paths have mostly ended would otherwise run its few long paths with most lanes
// idle; parked, they continue in fully populated waves of the next level.  Every live lane
// of a wave has run the same iterations of the launch and the ballot counts them all, so the
// decision is the wave's.
__device__ __forceinline__ bool park_due(uint32_t it, uint32_t cap, uint32_t park_min, uint32_t park_lanes) {
    if (cap == 0u) return false;
    if (it >= cap) return true;
    return park_lanes != 0u && it >= park_min && (uint32_t)__popcll(__ballot(1)) < park_lanes;
}

// XYZ (and, with kInfo, the per-path record) of a finished path into tuple t's slot.
template <int kEnv, bool kInfo>
__device__ __forceinline__ void write_result(const PathState &ps, const DevScene &sc, uint32_t t,
                                             float *__restrict__ xyz_out, uint32_t *__restrict__ info_out) {
    V3 xyz = v3(0.0f, 0.0f, 0.0f);
    if (ps.escaped) {
        float inten[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) inten[k] = intensity(ps.st[k]);
        xyz = finish_escaped<kEnv>(ps.ray_d, inten, ps.pdf, ps.hero_off, sc);
    }
    xyz_out[3 * (size_t)t + 0] = xyz.x;
    xyz_out[3 * (size_t)t + 1] = xyz.y;
    xyz_out[3 * (size_t)t + 2] = xyz.z;
    if (kInfo) {
        info_out[4 * (size_t)t + 0] = path_segments(ps, sc);
        info_out[4 * (size_t)t + 1] = ps.term_reason;
        info_out[4 * (size_t)t + 2] = ps.fh;
        info_out[4 * (size_t)t + 3] = ps.fate_h;
    }
}

// ---------------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------------

// Occupancy target of the trace kernels (waves per SIMD), which sets their register budget
// (2: 256 VGPRs, 3: 168, 4: 128).  The plain trace and continuation kernels are built at 2,
// 3 and 4 waves/SIMD and the autotuner picks one per scene class and plane count; the
// default, chosen by measurement, is 4 for the isotropic and uniaxial kernels (with the
// event counters in LDS and the planes sized to the stone, they fit 128 VGPRs and four
// workgroups' LDS per CU: 17 % faster than 3 on the single-job workload) and 3 for the
// biaxial ones, whose 32-float lane caches leave LDS for three workgroups, and for the
// rare-feature kernels, which spill at 128 VGPRs (7-15 % slower at 4).  The batched and
// persistent kernels take the default.
#ifndef GR_TRACE_WAVES_PER_EU
#define GR_TRACE_WAVES_PER_EU 4
#endif
#ifndef GR_HEAVY_WAVES_PER_EU
#define GR_HEAVY_WAVES_PER_EU 3
#endif
#define GR_WPE(kAniso, kFrost) (((kAniso) == 2 || (kFrost)) ? GR_HEAVY_WAVES_PER_EU : GR_TRACE_WAVES_PER_EU)
#define GR_OCCUPANCY(w) __attribute__((amdgpu_waves_per_eu((w), (w))))

// Tuples per wavefront of trace_wave_kernel: 64 lanes x this many rounds.
#ifndef GR_WAVE_ROUNDS
#define GR_WAVE_ROUNDS 8
#endif
constexpr uint32_t kWaveTuples = 64u * GR_WAVE_ROUNDS;

// The trace kernels keep the facet planes and the camera origin's sides in the launch's
// dynamic LDS, sized to its plane count (plane_lds_bytes; for a batched launch, its largest
// stone's): 20 bytes per plane rather than a fixed GR_MAX_PLANES array's 20 KB, which beside
// the biaxial lane caches and the event counters left room for two workgroups per CU.
__host__ __device__ constexpr uint32_t plane_slots(uint32_t n) { return (n + 3u) & ~3u; }
__host__ __device__ constexpr size_t plane_lds_bytes(uint32_t n) {
    return (size_t)plane_slots(n) * (sizeof(float4) + sizeof(float));
}
extern __shared__ float4 gr_plane_lds[];
#define GR_PLANE_LDS(n)                  \
    float4 *const planes = gr_plane_lds; \
    float *const cam_side = reinterpret_cast<float *>(gr_plane_lds + plane_slots(n))

// Facet planes into LDS, with the camera origin's signed distance to each
// (p.w + n . origin, the clip's operations), padded to a multiple of four for the scan's
// four-plane reads: with inert planes (normal 0, d = -1: never steep, never outside), or with
// kCopyPad copies of plane 0.  A copy is inert too: it has plane 0's t, side and steepness,
// and comes after it, so it never wins a strict extreme (equal t keeps the first index) and
// sets the outside flag only where plane 0 has set it; unlike a zero-normal plane it is steep
// wherever plane 0 is, so the interior scan's first pass can rule it out
// (clip_scan_screened).
template <bool kCopyPad = false>
__device__ __forceinline__ void stage_planes(float4 *lds, float *cam_side, const float4 *__restrict__ g,
                                             const DevScene &sc) {
    const V3 o = v3(sc.cam_o[0], sc.cam_o[1], sc.cam_o[2]);
    const uint32_t n4 = (sc.n_planes + 3u) & ~3u;
    for (uint32_t i = threadIdx.x; i < n4; i += kBlock) {
        const float4 p = i < sc.n_planes ? g[i] : (kCopyPad ? g[0] : make_float4(0.0f, 0.0f, 0.0f, -1.0f));
        lds[i] = p;
        cam_side[i] = p.w + ((p.x * o.x + p.y * o.y) + p.z * o.z);
    }
}

// Per-class counters, once the workgroup's lanes have all finished: thread (w, g) =
// (threadIdx.x / 32, threadIdx.x % 32) sums counter word w of lanes 8g..8g+7 packed, splits
// the sum into its two 16-bit counts, sums those over the 32 groups with five shuffles and
// adds them to the workgroup's counters with one LDS atomic per class; then one global
// atomic per class per workgroup.  The packed sum cannot carry: the eight lanes belong to
// one wave, and a class is counted at most once per segment, so a wave's count is at most
// kWaveTuples x GR_MAX_BOUNCES (the persistent kernel's wave; one path per lane elsewhere).
static_assert(kEvWords * 32 == kBlock, "one counter word per 32 threads");
static_assert(kWaveTuples * GR_MAX_BOUNCES < 65536u, "a wave's counts fit 16 bits");
__device__ __forceinline__ void flush_events(const uint32_t *ev, uint32_t *block_ev,
                                             unsigned long long *stats) {
    __syncthreads();
    const uint32_t w = threadIdx.x >> 5, g = threadIdx.x & 31u;
    const uint32_t *src = ev - threadIdx.x + w * kBlock + 8u * g;
    uint32_t packed = 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) packed += src[j];
    uint32_t lo = packed & 0xFFFFu, hi = packed >> 16;
    for (int off = 16; off > 0; off >>= 1) {
        lo += __shfl_xor(lo, off, 64);
        hi += __shfl_xor(hi, off, 64);
    }
    if (g == 0u) {
        if (lo) atomicAdd(&block_ev[2 * w], lo);
        if (hi && 2 * w + 1 < GR_EV_COUNT) atomicAdd(&block_ev[2 * w + 1], hi);
    }
    __syncthreads();
    if (threadIdx.x < GR_EV_COUNT && block_ev[threadIdx.x])
        atomicAdd(&stats[threadIdx.x], (unsigned long long)block_ev[threadIdx.x]);
}

// The plain path: one lane per (pixel, sample) tuple t of the launch, pixel =
// pixel_begin + t / spp, sample = sample_offset + t % spp.  Writes the tuple's XYZ (and,
// with kInfo, segments, termination and the two path hashes).
//
// Path-length split: a wave runs until its longest path ends, and stone paths are short
// with a long tail (on the oval-brilliant workload 69 % end after 3 segments, but most
// 64-sample waves hold one path that runs much longer).  With split_bounce > 0, a wave's
// paths still running after split_bounce segments -- or, with park_lanes > 0, after
// park_min segments once fewer than park_lanes of its lanes are still tracing (park_due)
// -- are parked as continuation records (one atomic per wave reserves the records) and
// cont_kernel continues the parked paths in fully populated waves, parking again by the
// same rule for kContLevels levels.  Parking moves state, not arithmetic: every path's
// result is bit-identical to an unsplit run.  When the records run out, the lane carries on.
template <int kAniso, bool kAbsorb, int kEnv, bool kFrost, bool kInfo, int kW>
__global__ __launch_bounds__(kBlock) GR_OCCUPANCY(kW) void trace_kernel(
    const DevScene sc, const float4 *__restrict__ planes_g, const float2 *__restrict__ ri35,
    const uint32_t *__restrict__ cert, uint32_t pixel_begin, uint32_t sample_offset, uint32_t spp,
    uint32_t n_tuples, float *__restrict__ xyz_out, uint32_t *__restrict__ info_out,
    unsigned long long *__restrict__ stats, uint32_t split_bounce, uint32_t park_min, uint32_t park_lanes,
    float *__restrict__ cont, uint32_t cont_cap, uint32_t *__restrict__ cont_count) {
    GR_PLANE_LDS(sc.n_planes);
    __shared__ float lcache[CacheFloats<kAniso>::value];
    __shared__ uint32_t block_ev[GR_EV_COUNT];
    stage_planes<kScreenedClip<kAbsorb, kFrost>>(planes, cam_side, planes_g, sc);
    if (threadIdx.x < GR_EV_COUNT) block_ev[threadIdx.x] = 0u;
    __syncthreads();

    __shared__ uint32_t lane_ev[kEvWords * kBlock];
    uint32_t *const ev = lane_ev + threadIdx.x;
#pragma unroll
    for (int i = 0; i < kEvWords; ++i) ev[i * kBlock] = 0u;

    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n_tuples) {
        const uint32_t local_pixel = t / spp;
        const uint32_t s = t - local_pixel * spp;
        PathState ps;
        init_path(ps, pixel_begin + local_pixel, sample_offset + s, ri35[s], cert[pixel_begin + local_pixel], sc);
        bool parked = false;
        uint32_t it = 0u;
        while (!bounce_step<kAniso, kAbsorb, kFrost, kInfo>(ps, sc, planes, cam_side, lcache, ev)) {
            if (park_due(++it, split_bounce, park_min, park_lanes) &&
                park<kAniso == 2, kInfo>(ps, t, lcache, cont, cont_cap, cont_count)) {
                parked = true;
                break;
            }
        }
        if (!parked) write_result<kEnv, kInfo>(ps, sc, t, xyz_out, info_out);
    }
    flush_events(ev, block_ev, stats);
}

// Continues the paths the previous launch parked: record r per lane, for r < *cont_count
// (capped at cont_cap); workgroups past the count return at once.  With split_iters > 0 a
// wave parks its paths again, into next / next_count, for the next level, after split_iters
// iterations of this launch or earlier by the park_lanes rule (park_due).
template <int kAniso, bool kAbsorb, int kEnv, bool kFrost, bool kInfo, int kW>
__global__ __launch_bounds__(kBlock) GR_OCCUPANCY(kW) void cont_kernel(
    const DevScene sc, const float4 *__restrict__ planes_g, const float *__restrict__ cont, uint32_t cont_cap,
    const uint32_t *__restrict__ cont_count, float *__restrict__ xyz_out, uint32_t *__restrict__ info_out,
    unsigned long long *__restrict__ stats, uint32_t split_iters, uint32_t park_min, uint32_t park_lanes,
    float *__restrict__ next, uint32_t *__restrict__ next_count) {
    const uint32_t count = min(*cont_count, cont_cap);
    if (blockIdx.x * kBlock >= count) return;
    GR_PLANE_LDS(sc.n_planes);
    __shared__ float lcache[CacheFloats<kAniso>::value];
    __shared__ uint32_t block_ev[GR_EV_COUNT];
    stage_planes<kScreenedClip<kAbsorb, kFrost>>(planes, cam_side, planes_g, sc);
    if (threadIdx.x < GR_EV_COUNT) block_ev[threadIdx.x] = 0u;
    __syncthreads();

    __shared__ uint32_t lane_ev[kEvWords * kBlock];
    uint32_t *const ev = lane_ev + threadIdx.x;
#pragma unroll
    for (int i = 0; i < kEvWords; ++i) ev[i * kBlock] = 0u;

    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r < count) {
        PathState ps;
        const uint32_t t = load_path<kAniso == 2, kInfo>(ps, lcache, cont, cont_cap, r);
        bool parked = false;
        uint32_t it = 0u;
        while (!bounce_step<kAniso, kAbsorb, kFrost, kInfo>(ps, sc, planes, cam_side, lcache, ev)) {
            if (park_due(++it, split_iters, park_min, park_lanes) &&
                park<kAniso == 2, kInfo>(ps, t, lcache, next, cont_cap, next_count)) {
                parked = true;
                break;
            }
        }
        if (!parked) write_result<kEnv, kInfo>(ps, sc, t, xyz_out, info_out);
    }
    flush_events(ev, block_ev, stats);
}

// The batched kernels' view of BatchSlice: the same layout with every pointer typed by the
// memory it addresses.  The scene block is constant for the launch (address space 4), so its
// uniform fields are read with scalar loads exactly as the plain kernel reads its by-value
// scene from the kernel arguments; the buffers are global memory (address space 1), so the
// kernels issue global rather than flat loads and stores.  Through generic pointers loaded
// from the slice array the compiler could prove neither and fell back to per-lane flat
// accesses.
#define GR_AS_CONST __attribute__((address_space(4)))
#define GR_AS_GLOBAL __attribute__((address_space(1)))
struct BatchSliceView {
    const GR_AS_CONST DevScene *scene;
    const GR_AS_CONST float4 *planes;
    const GR_AS_CONST float *ri35;  // float2 pairs
    const GR_AS_CONST uint32_t *cert;
    GR_AS_GLOBAL float *xyz;
    GR_AS_GLOBAL float *film;
    GR_AS_GLOBAL unsigned long long *stats;
    uint32_t pixel_begin, n_pixels, sample_offset, spp;
    uint32_t trace_block_begin;
    uint32_t reduce_block_begin;
    uint32_t split_bounce;
    uint32_t park_lanes;
    uint32_t cont_cap;
    uint32_t cont_block_begin;
    GR_AS_GLOBAL float *cont[2];
    GR_AS_GLOBAL uint32_t *cont_count;
};
static_assert(sizeof(BatchSliceView) == sizeof(BatchSlice), "BatchSliceView mirrors BatchSlice");
static_assert(offsetof(BatchSliceView, xyz) == offsetof(BatchSlice, xyz), "BatchSliceView mirrors BatchSlice");
static_assert(offsetof(BatchSliceView, pixel_begin) == offsetof(BatchSlice, pixel_begin),
              "BatchSliceView mirrors BatchSlice");
static_assert(offsetof(BatchSliceView, split_bounce) == offsetof(BatchSlice, split_bounce),
              "BatchSliceView mirrors BatchSlice");
static_assert(offsetof(BatchSliceView, cont) == offsetof(BatchSlice, cont), "BatchSliceView mirrors BatchSlice");
static_assert(offsetof(BatchSliceView, cont_count) == offsetof(BatchSlice, cont_count),
              "BatchSliceView mirrors BatchSlice");

// Index of the slice that owns workgroup b: the last slice whose first block is <= b
// (slices are in block order and the first starts at block 0).
enum SliceBlocks { kTraceBlocks = 0, kReduceBlocks = 1, kContBlocks = 2 };
__device__ __forceinline__ uint32_t find_slice(const BatchSliceView *__restrict__ slices, uint32_t n, uint32_t b,
                                               int which) {
    uint32_t lo = 0u, hi = n;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t first = which == kTraceBlocks    ? slices[mid].trace_block_begin
                               : which == kReduceBlocks ? slices[mid].reduce_block_begin
                                                        : slices[mid].cont_block_begin;
        if (first <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The batched path: one launch traces the pieces of many jobs of one material class (the
// reference's per-class pipeline specialisation, frame.rs:514-524).  Every workgroup
// belongs to one slice, stages that job's planes and reads that job's scene block; the
// per-lane work is the plain kernel's, so each tuple's result is bit-identical to a plain
// launch of the same job.
template <int kAniso, bool kAbsorb, int kEnv, bool kFrost>
__global__ __launch_bounds__(kBlock) GR_OCCUPANCY(GR_WPE(kAniso, kFrost)) void trace_batch_kernel(
    const BatchSliceView *__restrict__ slices, uint32_t n_slices) {
    __shared__ float lcache[CacheFloats<kAniso>::value];
    __shared__ uint32_t block_ev[GR_EV_COUNT];
    const BatchSliceView &J = slices[find_slice(slices, n_slices, blockIdx.x, kTraceBlocks)];
    const DevScene &sc = *(const DevScene *)J.scene;
    GR_PLANE_LDS(sc.n_planes);
    stage_planes<kScreenedClip<kAbsorb, kFrost>>(planes, cam_side, (const float4 *)J.planes, sc);
    if (threadIdx.x < GR_EV_COUNT) block_ev[threadIdx.x] = 0u;
    __syncthreads();

    __shared__ uint32_t lane_ev[kEvWords * kBlock];
    uint32_t *const ev = lane_ev + threadIdx.x;
#pragma unroll
    for (int i = 0; i < kEvWords; ++i) ev[i * kBlock] = 0u;

    const uint32_t spp = J.spp;
    const uint32_t t = (blockIdx.x - J.trace_block_begin) * kBlock + threadIdx.x;
    if (t < J.n_pixels * spp) {
        const uint32_t local_pixel = t / spp;
        const uint32_t s = t - local_pixel * spp;
        PathState ps;
        init_path(ps, J.pixel_begin + local_pixel, J.sample_offset + s, make_float2(J.ri35[2 * s], J.ri35[2 * s + 1]),
                  J.cert[J.pixel_begin + local_pixel], sc);
        bool parked = false;
        uint32_t it = 0u;
        while (!bounce_step<kAniso, kAbsorb, kFrost, false>(ps, sc, planes, cam_side, lcache, ev)) {
            if (park_due(++it, J.split_bounce, kParkMin, J.park_lanes) &&
                park<kAniso == 2, false>(ps, t, lcache, (float *)J.cont[0], J.cont_cap, (uint32_t *)J.cont_count)) {
                parked = true;
                break;
            }
        }
        if (!parked) write_result<kEnv, false>(ps, sc, t, (float *)J.xyz, nullptr);
    }
    flush_events(ev, block_ev, (unsigned long long *)J.stats);
}

// Continuation level `level` (1..kContLevels) of a batched class launch: each workgroup
// belongs to one slice (binary search of cont_block_begin) and continues that slice's
// records of the previous level, parking again for the next level except at the last.
template <int kAniso, bool kAbsorb, int kEnv, bool kFrost>
__global__ __launch_bounds__(kBlock) GR_OCCUPANCY(GR_WPE(kAniso, kFrost)) void cont_batch_kernel(
    const BatchSliceView *__restrict__ slices, uint32_t n_slices, int level) {
    const BatchSliceView &J = slices[find_slice(slices, n_slices, blockIdx.x, kContBlocks)];
    const uint32_t count = min(J.cont_count[level - 1], J.cont_cap);
    const uint32_t r = (blockIdx.x - J.cont_block_begin) * kBlock + threadIdx.x;
    if ((blockIdx.x - J.cont_block_begin) * kBlock >= count) return;
    __shared__ float lcache[CacheFloats<kAniso>::value];
    __shared__ uint32_t block_ev[GR_EV_COUNT];
    const DevScene &sc = *(const DevScene *)J.scene;
    GR_PLANE_LDS(sc.n_planes);
    stage_planes<kScreenedClip<kAbsorb, kFrost>>(planes, cam_side, (const float4 *)J.planes, sc);
    if (threadIdx.x < GR_EV_COUNT) block_ev[threadIdx.x] = 0u;
    __syncthreads();

    __shared__ uint32_t lane_ev[kEvWords * kBlock];
    uint32_t *const ev = lane_ev + threadIdx.x;
#pragma unroll
    for (int i = 0; i < kEvWords; ++i) ev[i * kBlock] = 0u;

    if (r < count) {
        PathState ps;
        const uint32_t t = load_path<kAniso == 2, false>(ps, lcache, (const float *)J.cont[(level - 1) & 1], J.cont_cap, r);
        const uint32_t split_iters = level < kContLevels ? J.split_bounce : 0u;
        bool parked = false;
        uint32_t it = 0u;
        while (!bounce_step<kAniso, kAbsorb, kFrost, false>(ps, sc, planes, cam_side, lcache, ev)) {
            if (park_due(++it, split_iters, kParkMin, J.park_lanes) &&
                park<kAniso == 2, false>(ps, t, lcache, (float *)J.cont[level & 1], J.cont_cap, (uint32_t *)(J.cont_count + level))) {
                parked = true;
                break;
            }
        }
        if (!parked) write_result<kEnv, false>(ps, sc, t, (float *)J.xyz, nullptr);
    }
    flush_events(ev, block_ev, (unsigned long long *)J.stats);
}

// Escape records: field-major (record field f of tuple t at rec[f * cap + t]).
// Fields: 0-2 escape direction, 3-10 channel intensities, 11-18 channel pdfs,
// 19 hero offset (negative: the path never escaped).
constexpr int kRecFields = 20;

// Persistent wavefronts.  Wave w owns tuples [w * kWaveTuples, (w+1) * kWaveTuples) of
// the launch.  Every lane starts on one tuple; whenever lanes finish their paths, the
// wave hands them the next unstarted tuples of its range in lane order (ballot +
// popcount), so lanes stay busy until the range is exhausted instead of idling behind the
// wave's longest path.  Each finished path writes its escape record; every loop
// iteration advances every busy lane by one bounce, and every path ends within
// max_bounces iterations, so the loop ends after at most kWaveTuples * max_bounces
// iterations.
template <int kAniso, bool kAbsorb, bool kInfo>
__global__ __launch_bounds__(kBlock) GR_OCCUPANCY(GR_WPE(kAniso, true)) void trace_wave_kernel(
    const DevScene sc, const float4 *__restrict__ planes_g, const float2 *__restrict__ ri35,
    const uint32_t *__restrict__ cert, uint32_t pixel_begin, uint32_t sample_offset, uint32_t spp, uint32_t n_tuples, uint32_t rec_cap,
    float *__restrict__ rec, uint32_t *__restrict__ info_out, unsigned long long *__restrict__ stats) {
    GR_PLANE_LDS(sc.n_planes);
    __shared__ float lcache[CacheFloats<kAniso>::value];
    __shared__ uint32_t block_ev[GR_EV_COUNT];
    stage_planes<kScreenedClip<kAbsorb, true>>(planes, cam_side, planes_g, sc);
    if (threadIdx.x < GR_EV_COUNT) block_ev[threadIdx.x] = 0u;
    __syncthreads();

    __shared__ uint32_t lane_ev[kEvWords * kBlock];
    uint32_t *const ev = lane_ev + threadIdx.x;
#pragma unroll
    for (int i = 0; i < kEvWords; ++i) ev[i * kBlock] = 0u;

    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t begin = wave * kWaveTuples;
    const uint32_t end = min(begin + kWaveTuples, n_tuples);
    const uint64_t lanes_below = (1ull << lane) - 1ull;

    PathState ps;
    uint32_t t = begin + lane;
    bool busy = t < end;
    uint32_t next = begin + 64u;  // wave-uniform
    if (busy) {
        const uint32_t lp = t / spp, s = t - lp * spp;
        init_path(ps, pixel_begin + lp, sample_offset + s, ri35[s], cert[pixel_begin + lp], sc);
    }
    while (__ballot(busy) != 0ull) {
        if (busy && bounce_step<kAniso, kAbsorb, true, kInfo>(ps, sc, planes, cam_side, lcache, ev)) {
            // Path ended: escape record (and per-tuple info).
            float *r = rec + t;
            r[0 * (size_t)rec_cap] = ps.ray_d.x;
            r[1 * (size_t)rec_cap] = ps.ray_d.y;
            r[2 * (size_t)rec_cap] = ps.ray_d.z;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                r[(3 + k) * (size_t)rec_cap] = intensity(ps.st[k]);
                r[(11 + k) * (size_t)rec_cap] = ps.pdf[k];
            }
            r[19 * (size_t)rec_cap] = ps.escaped ? ps.hero_off : -1.0f;
            if (kInfo) {
                info_out[4 * (size_t)t + 0] = path_segments(ps, sc);
                info_out[4 * (size_t)t + 1] = ps.term_reason;
                info_out[4 * (size_t)t + 2] = ps.fh;
                info_out[4 * (size_t)t + 3] = ps.fate_h;
            }
            busy = false;
        }
        const uint64_t idle = __ballot(!busy);
        const uint32_t slot = next + (uint32_t)__popcll(idle & lanes_below);
        next += (uint32_t)__popcll(idle);
        if (!busy && slot < end) {
            t = slot;
            const uint32_t lp = t / spp, s = t - lp * spp;
            init_path(ps, pixel_begin + lp, sample_offset + s, ri35[s], cert[pixel_begin + lp], sc);
            busy = true;
        }
    }
    flush_events(ev, block_ev, stats);
}

// Escape records -> per-tuple XYZ (the finish of trace_spectral_ray_inner).
template <int kEnv>
__global__ __launch_bounds__(kBlock) void finish_kernel(const DevScene sc, uint32_t n_tuples, uint32_t rec_cap,
                                                        const float *__restrict__ rec,
                                                        float *__restrict__ xyz_out) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_tuples) return;
    const float *r = rec + t;
    const float hero_off = r[19 * (size_t)rec_cap];
    V3 xyz = v3(0.0f, 0.0f, 0.0f);
    if (hero_off >= 0.0f) {
        float inten[8], pdf[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            inten[k] = r[(3 + k) * (size_t)rec_cap];
            pdf[k] = r[(11 + k) * (size_t)rec_cap];
        }
        V3 dir = v3(r[0], r[(size_t)rec_cap], r[2 * (size_t)rec_cap]);
        xyz = finish_escaped<kEnv>(dir, inten, pdf, hero_off, sc);
    }
    xyz_out[3 * (size_t)t + 0] = xyz.x;
    xyz_out[3 * (size_t)t + 1] = xyz.y;
    xyz_out[3 * (size_t)t + 2] = xyz.z;
}


}  // namespace gr
