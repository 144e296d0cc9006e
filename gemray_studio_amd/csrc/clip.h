// clip.h -- the convex facet-plane clip shared by the transport kernels and the
// gemological metrics kernel: bit-identical to the reference's intersect_polyhedron /
// intersect_polyhedron_soa over planes staged in LDS.
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
__device__ __forceinline__ void clip_plane(float4 p, float side_cam, int i, V3 ro, V3 rd, float &t_near, int &near_idx,
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
// planes 0.x + 0.y + 0.z - 1 <= 0, which are never steep, never outside and never win), so
// one LDS wait covers four planes.
template <bool kCam>
__device__ __forceinline__ void clip_scan(const float4 *__restrict__ planes, const float *__restrict__ cam_side,
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
