// clip.h -- the convex facet-plane clip shared by the transport kernels and the
// gemological metrics kernel: bit-identical to the reference's intersect_polyhedron /
// intersect_polyhedron_soa over planes staged in LDS.  The scans are host-callable as well,
// so that a host build can check the screened scan against the plain one.
#pragma once

#include <hip/hip_runtime.h>

#include "device_physics.h"

namespace gr {

// The convex facet-plane scan (simd/slab.rs:94-152, intersect.rs:93-132).  The reference
// divides t = -side / denom for every plane and keeps the first strict maximum over
// entering planes (denom < -1e-7) and the first strict minimum over exiting planes
// (denom > 1e-7).  A plane can only change an extreme t_ext if its exact quotient is
// beyond it: -side / denom > t_near with denom < 0, or < t_far with denom > 0, which are
// both side + t_ext * denom > 0.  One fma evaluates that sum with a single rounding, which
// keeps its sign (the operands' magnitudes keep the sum far from the denormal range), so
// the screen is exact: a plane that fails it has a rounded quotient no better than t_ext,
// and a plane that passes takes the correctly rounded division and the reference's strict
// comparison.  A parallel plane (|denom| <= 1e-7) with the origin outside makes the
// reference stop scanning (the ray misses whatever the other planes give); here the flag
// is kept and the scan runs on, to the same outcome.
// kCam: the origin is the camera's, whose per-plane sides are staged in cam_side.
template <bool kCam>
__host__ __device__ __forceinline__ void clip_plane(float4 p, float side_cam, int i, V3 ro, V3 rd, float &t_near, int &near_idx,
                                           float &t_far, int &far_idx, bool &outside) {
    const float denom = (p.x * rd.x + p.y * rd.y) + p.z * rd.z;
    const float side = kCam ? side_cam : p.w + ((p.x * ro.x + p.y * ro.y) + p.z * ro.z);
    const bool steep = fabsf(denom) > 1e-7f;
    const bool entering = denom < 0.0f;
    const float t_ext = entering ? t_near : t_far;
    // Selects rather than branches on steepness and side: the wave branches only where a
    // plane can change an extreme, and the update inside is branch-free too.
    const float screen = steep ? fmaf(t_ext, denom, side) : side;
    outside = outside || (screen > 0.0f && !steep);
    if (screen > 0.0f && steep) {
        const float t = (-side) / denom;
        const bool upd_near = entering && t > t_near;
        const bool upd_far = !entering && t < t_far;
        t_near = upd_near ? t : t_near;
        near_idx = upd_near ? i : near_idx;
        t_far = upd_far ? t : t_far;
        far_idx = upd_far ? i : far_idx;
    }
}
// Planes are read four at a time (the staged array is padded to a multiple of four with
// inert planes: zero normals with d = -1, never steep, never outside, never winning, or
// copies of plane 0 -- see stage_planes), so one LDS wait covers four planes.
template <bool kCam>
__host__ __device__ __forceinline__ void clip_scan(const float4 *__restrict__ planes, const float *__restrict__ cam_side,
                                          uint32_t n, V3 ro, V3 rd, float &t_near, int &near_idx, float &t_far,
                                          int &far_idx, bool &outside) {
    const int n4 = (int)((n + 3u) & ~3u);
    for (int i = 0; i < n4; i += 4) {
        const float4 p0 = planes[i], p1 = planes[i + 1], p2 = planes[i + 2], p3 = planes[i + 3];
        float4 cs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (kCam) cs = *reinterpret_cast<const float4 *>(cam_side + i);
        clip_plane<kCam>(p0, cs.x, i, ro, rd, t_near, near_idx, t_far, far_idx, outside);
        clip_plane<kCam>(p1, cs.y, i + 1, ro, rd, t_near, near_idx, t_far, far_idx, outside);
        clip_plane<kCam>(p2, cs.z, i + 2, ro, rd, t_near, near_idx, t_far, far_idx, outside);
        clip_plane<kCam>(p3, cs.w, i + 3, ro, rd, t_near, near_idx, t_far, far_idx, outside);
    }
}

// The interior scan with a first-pass screen.  The reference's step for a plane changes
// something only if the plane is steep (|denom| > 1e-7) and its exact screen
// S_ext = side + t_ext * denom is positive (t_ext = t_near for denom < 0, t_far for
// denom > 0), or if it is not steep and side > 0 (the outside flag).  The first pass forms
// the two dot products in FMA form (dq, sq: three operations each instead of five and six)
// and rules the plane out when it proves both
//   S_n = side + t_near denom < -1e-7 |t_near|   and   S_f = side + t_far denom < -1e-7 |t_far|.
// That settles every case without knowing denom's sign or steepness: a steep plane's S_ext
// is one of the two, so it is negative; a plane with |denom| <= 1e-7 has
// side = S_n - t_near denom < -1e-7 |t_near| + |t_near| 1e-7 = 0.  Only a plane the first
// pass cannot rule out goes through clip_plane, which recomputes everything in the
// reference's operation order -- so the scan's result is bit-identical to clip_scan.  The
// proof of S < -1e-7 |t| from fl(t dq + sq) < -M(t):
//   * the reference's denom and side are within gamma_3 sum|n_i d_i| and
//     gamma_4 (|w| + sum|n_i o_i|) of the exact dot products, and dq and sq within gamma_3 of
//     them, so |dq - denom| <= 6.01 u B |d|_inf and |sq - side| <= 7.01 u (W + B |o|_inf),
//     where u = 2^-24, B bounds |n_x| + |n_y| + |n_z| and W bounds |w| over the staged planes;
//   * S <= fl(t dq + sq) + u |t dq + sq| + |t| |dq - denom| + |sq - side|, so
//     M(t) = |t| (2^-20 B |d|_inf + 1e-7) + 2^-20 (W + B |o|_inf) -- the bounds taken at
//     2^-20 = 16 u, twice over, which also covers the fma's own rounding and the margins'
//     -- leaves S < -1e-7 |t|.
// The margins of the two extremes are formed once and again only when clip_plane moves an
// extreme.  The comparisons are strict, so that a NaN, or an overflow to -inf against an
// infinite margin, fails the proof and takes the exact path.
struct ClipMargins {
    float base;  // 2^-20 (W + B |o|_inf)
    float md;    // 2^-20 B |d|_inf + 1e-7
};
__host__ __device__ __forceinline__ ClipMargins clip_margins(float w_max, float b_max, V3 ro, V3 rd) {
    const float om = fmaxf(fmaxf(fabsf(ro.x), fabsf(ro.y)), fabsf(ro.z));
    const float dm = fmaxf(fmaxf(fabsf(rd.x), fabsf(rd.y)), fabsf(rd.z));
    ClipMargins m;
    m.base = 0x1p-20f * fmaf(b_max, om, w_max);
    m.md = fmaf(0x1p-20f, b_max * dm, 1e-7f);
    return m;
}
// M(t) of the running extremes (after the scan starts and whenever clip_plane moves one).
__host__ __device__ __forceinline__ void clip_extreme_margins(const ClipMargins &m, float t_near, float t_far,
                                                               float &m_near, float &m_far) {
    m_near = fmaf(fabsf(t_near), m.md, m.base);
    m_far = fmaf(fabsf(t_far), m.md, m.base);
}
__host__ __device__ __forceinline__ void clip_plane_screened(float4 p, int i, V3 ro, V3 rd, const ClipMargins &m,
                                                             float &m_near, float &m_far, float &t_near,
                                                             int &near_idx, float &t_far, int &far_idx,
                                                             bool &outside) {
    const float dq = fmaf(p.z, rd.z, fmaf(p.y, rd.y, p.x * rd.x));
    const float sq = fmaf(p.z, ro.z, fmaf(p.y, ro.y, fmaf(p.x, ro.x, p.w)));
    const bool ruled_out = fmaf(t_near, dq, sq) < -m_near && fmaf(t_far, dq, sq) < -m_far;
    if (!ruled_out) {
        clip_plane<false>(p, 0.0f, i, ro, rd, t_near, near_idx, t_far, far_idx, outside);
        clip_extreme_margins(m, t_near, t_far, m_near, m_far);
    }
}
__host__ __device__ __forceinline__ void clip_scan_screened(const float4 *__restrict__ planes, uint32_t n, float w_max,
                                                   float b_max, V3 ro, V3 rd, float &t_near, int &near_idx,
                                                   float &t_far, int &far_idx, bool &outside) {
    const ClipMargins m = clip_margins(w_max, b_max, ro, rd);
    float m_near, m_far;
    clip_extreme_margins(m, t_near, t_far, m_near, m_far);
    const int n4 = (int)((n + 3u) & ~3u);
    for (int i = 0; i < n4; i += 4) {
        const float4 p0 = planes[i], p1 = planes[i + 1], p2 = planes[i + 2], p3 = planes[i + 3];
        clip_plane_screened(p0, i, ro, rd, m, m_near, m_far, t_near, near_idx, t_far, far_idx, outside);
        clip_plane_screened(p1, i + 1, ro, rd, m, m_near, m_far, t_near, near_idx, t_far, far_idx, outside);
        clip_plane_screened(p2, i + 2, ro, rd, m, m_near, m_far, t_near, near_idx, t_far, far_idx, outside);
        clip_plane_screened(p3, i + 3, ro, rd, m, m_near, m_far, t_near, near_idx, t_far, far_idx, outside);
    }
}

// True when the scan is certain to miss because of plane p alone: the origin is outside p
// and the ray moves away from it, so p's exact t = -side / denom is negative, the far
// extreme is at most that, and neither extreme can exceed the 1e-4 hit threshold.
__device__ __forceinline__ bool exterior_miss(float4 p, V3 ro, V3 rd) {
    const float denom = (p.x * rd.x + p.y * rd.y) + p.z * rd.z;
    const float side = p.w + ((p.x * ro.x + p.y * ro.y) + p.z * ro.z);
    return fabsf(denom) > 1e-7f && denom > 0.0f && side > 0.0f;
}

// shading_normal_near_edge (optics/raytracer/intersect.rs:196-229): within
// rounding_radius of the edge shared with the nearest other facet, the flat normal is
// blended (smoothstep) toward the bisector of the two facet normals.  The caller skips it
// for radius <= 0, where the reference returns the flat normal unchanged.
__device__ __forceinline__ V3 shading_normal_near_edge(const float4 *__restrict__ planes, uint32_t n, V3 hit_point,
                                                    int hit_facet, V3 hit_normal, float rounding_radius) {
    float nearest_dist = __builtin_huge_valf();
    V3 nearest_normal = hit_normal;
    for (uint32_t i = 0; i < n; ++i) {
        if ((int)i == hit_facet) continue;
        const float4 p = planes[i];
        const float dist = -(p.w + ((p.x * hit_point.x + p.y * hit_point.y) + p.z * hit_point.z));
        if (dist < nearest_dist) {
            nearest_dist = dist;
            nearest_normal = v3(p.x, p.y, p.z);
        }
    }
    if (nearest_dist >= rounding_radius) return hit_normal;
    float t = 1.0f - nearest_dist / rounding_radius;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    const float smooth_t = t * t * fmaf(2.0f, -t, 3.0f);
    const V3 bisector = normalize_or_zero(add(hit_normal, nearest_normal));
    return normalize_or_zero(add(mul(hit_normal, 1.0f - smooth_t), mul(bisector, smooth_t)));
}

}  // namespace gr
