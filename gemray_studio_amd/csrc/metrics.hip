// metrics.hip -- the gemological metrics of a cut stone (brilliance, fire, scintillation,
// windowing, extinction) on the device: evaluate_gem_optical_metrics
// (crates/gemray/src/color/metrics.rs:1082-1248) for many poses in one call.
//
// One evaluation fires an 18 x 18 grid of cells over the stone's face (cells inside the
// perimeter u^2 + v^2 <= 0.7), five sub-aperture rays per cell traced at the d-line index
// (and, when returned to the observer, replayed at the F and C lines for fire), and per
// cell five single rays from camera poses turned by -3..3 degrees of yaw (temporal
// scintillation).  Every ray is independent, so metrics_items_kernel gives each of a
// pose's 1620 aperture rays and 1620 temporal rays its own lane (13 workgroups per pose)
// and writes one record per ray; metrics_reduce_kernel then walks a pose's records in the
// reference's (ix, iz, sample) order, so the fma chain of the fire sum and the running
// cell-fraction sums round exactly as the reference's loop does.  Host-side constants
// (d/F/C indices, the six camera bases, the studio rig) are computed per pose on the host
// with the reference's operations, as the reference computes them once per call.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/gemray_hip.h"
#include "clip.h"
#include "device_physics.h"
#include "runtime_internal.h"

namespace gr {

constexpr int kMxGrid = 18;
constexpr int kMxSamples = 5;
constexpr int kMxCells = kMxGrid * kMxGrid;
constexpr int kMxItems = kMxCells * kMxSamples;  // aperture rays (and temporal rays) per pose
constexpr int kMxBlock = 256;
constexpr int kMxBlocksPerPose = (2 * kMxItems + kMxBlock - 1) / kMxBlock;
constexpr uint32_t kMxMaxPoses = 1u << 16;

// Ray classes (RayClassification, metrics.rs:505-517; a miss is not counted at all).
enum : uint8_t { MX_NONE = 0, MX_BLOCKED = 1, MX_WINDOWED = 2, MX_EXTINCT = 3, MX_RETURNED = 4, MX_RETURNED_FIRE = 5 };

// Per-pose constants: refractive indices at 589.3 / 486.1 / 656.3 nm, the studio rig's key
// and fill directions and sin(light pitch), and the (forward, right, up) bases of the pose
// and of its five temporal yaw offsets.
struct MetricsPose {
    uint32_t plane_off, n_planes;  // the pose's stone: planes [plane_off, plane_off + n_planes)
    float nd, nf, nc, sin_lp;
    float key[3], fill[3];
    float basis[1 + kMxSamples][9];
};

// f32::clamp (NaN passes through, unlike fmin/fmax).
__device__ __forceinline__ float clamp_rs(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

__device__ __forceinline__ V3 basis_v(const MetricsPose &p, int b, int k) {
    return v3(p.basis[b][3 * k + 0], p.basis[b][3 * k + 1], p.basis[b][3 * k + 2]);
}

// intersect_polyhedron_soa (optics/raytracer/intersect.rs:93-132) over the staged planes.
__device__ __forceinline__ bool mx_intersect(const float4 *__restrict__ planes, uint32_t n, V3 ro, V3 rd, float &t,
                                             V3 &normal, int &facet) {
    float t_near = -1e30f, t_far = 1e30f;
    int near_idx = -1, far_idx = -1;
    bool outside = false;
    clip_scan<false>(planes, nullptr, n, ro, rd, t_near, near_idx, t_far, far_idx, outside);
    if (outside || t_near > t_far) return false;
    int idx;
    if (t_near > 1e-4f) {
        t = t_near;
        idx = near_idx;
    } else if (t_far > 1e-4f) {
        t = t_far;
        idx = far_idx;
    } else {
        return false;
    }
    if (idx >= 0) {
        const float4 p = planes[idx];
        normal = v3(p.x, p.y, p.z);
        facet = idx;
    } else {
        normal = v3(0.0f, 0.0f, 0.0f);
        facet = 0;
    }
    return true;
}

// fresnel_transmittance (metrics.rs:117-131); f32::midpoint is the f64 mean rounded once.
__device__ __forceinline__ float mx_fresnel_t(float n1, float n2, float cos_i, float cos_t) {
    const float denom_s = fmaf(n2, cos_t, n1 * cos_i);
    const float denom_p = fmaf(n2, cos_i, n1 * cos_t);
    if (fabsf(denom_s) < 1e-6f || fabsf(denom_p) < 1e-6f) return 0.0f;
    float rs = fmaf(n2, -cos_t, n1 * cos_i) / denom_s;
    rs = rs * rs;
    float rp = fmaf(n2, -cos_i, n1 * cos_t) / denom_p;
    rp = rp * rp;
    const float r = (float)(((double)rs + (double)rp) * 0.5);
    return clamp_rs(1.0f - r, 0.0f, 1.0f);
}

enum { FATE_BLOCKED = 0, FATE_LEAKED = 1, FATE_UP = 2, FATE_ABSORBED = 3 };
struct MxFate {
    int kind;
    V3 dir;
    float transmittance, exit_cos;
    int facet;
    uint32_t bounces;
};

// trace_wavelength (metrics.rs:138-222): refract in at `index`, then up to 10 internal
// bounces, each either an exit (pavilion leak, upward exit, or sideways) or a total
// internal reflection.
__device__ MxFate mx_trace(V3 entry, V3 incoming, V3 n_entry, float cos_i, const float4 *__restrict__ planes,
                           uint32_t n, float index) {
    MxFate f{FATE_ABSORBED, v3(0.0f, 0.0f, 0.0f), 0.0f, 0.0f, -1, 0u};
    const float sin2_t = (1.0f / (index * index)) * fmaf(cos_i, -cos_i, 1.0f);
    if (sin2_t > 1.0f) {
        f.kind = FATE_BLOCKED;
        return f;
    }
    const float cos_t = sqrtf(1.0f - sin2_t);
    const float inv = 1.0f / index;
    V3 dir = normalize(add(mul(incoming, inv), mul(n_entry, fmaf(inv, cos_i, -cos_t))));
    V3 hit_point = add(entry, mul(dir, 1e-4f));
    const float sin_crit = 1.0f / index;
    const float entry_t = mx_fresnel_t(1.0f, index, cos_i, cos_t);
    for (uint32_t bounce = 0; bounce < 10u; ++bounce) {
        float t;
        V3 n_out;
        int facet;
        if (!mx_intersect(planes, n, hit_point, dir, t, n_out, facet)) break;
        const V3 next_point = add(hit_point, mul(dir, t));
        const float cos_theta = clamp_rs(dot(dir, n_out), 0.0f, 1.0f);
        const float sin_theta = sqrtf(fmaxf(fmaf(cos_theta, -cos_theta, 1.0f), 0.0f));
        if (sin_theta < sin_crit) {
            const float sin2_out = (index * index) * fmaf(cos_theta, -cos_theta, 1.0f);
            if (sin2_out <= 1.0f) {
                const float cos_out = sqrtf(1.0f - sin2_out);
                const V3 out_dir = normalize(add(mul(dir, index), mul(n_out, fmaf(index, -cos_theta, cos_out))));
                if (n_out.y < -0.05f) {
                    f.kind = FATE_LEAKED;
                } else if (out_dir.y > 0.05f) {
                    f.kind = FATE_UP;
                    f.dir = out_dir;
                    f.transmittance = entry_t * mx_fresnel_t(index, 1.0f, cos_theta, cos_out);
                    f.exit_cos = cos_out;
                    f.facet = facet;
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
__device__ __forceinline__ bool mx_returned(V3 exit_dir, V3 fwd, const MetricsPose &p) {
    const bool head_shadow = dot(neg(exit_dir), fwd) > 0.96f;
    const float key_dot = fmaxf(dot(exit_dir, v3(p.key[0], p.key[1], p.key[2])), 0.0f);
    const float fill_dot = fmaxf(dot(exit_dir, v3(p.fill[0], p.fill[1], p.fill[2])), 0.0f);
    const bool ring = fabsf(fmaf(p.sin_lp, -0.8f, exit_dir.y)) < 0.35f;
    const bool lit = key_dot > 0.70f || fill_dot > 0.75f || (ring && exit_dir.y > 0.2f);
    return !head_shadow && lit;
}

__device__ __forceinline__ float mx_grid(int i) { return fmaf(((float)i + 0.5f) / (float)kMxGrid, 2.0f, -1.0f); }

// One lane per ray.  Lane item i < kMxItems is aperture ray (cell i / 5, sample i % 5)
// (classify_aperture_sample, metrics.rs:519-771); kMxItems <= i < 2 kMxItems is the
// temporal ray (cell, yaw offset) (cell_returned_at_yaw_offset, :309-334).
__global__ __launch_bounds__(kMxBlock) void metrics_items_kernel(const float4 *__restrict__ planes_all,
                                                                 const MetricsPose *__restrict__ poses, uint8_t *cls,
                                                                 float2 *fire, uint8_t *temporal) {
    __shared__ float4 planes[GR_MAX_PLANES];
    const uint32_t pose_i = blockIdx.x / kMxBlocksPerPose;
    const uint32_t item = (blockIdx.x % kMxBlocksPerPose) * kMxBlock + threadIdx.x;
    const MetricsPose &p = poses[pose_i];
    const uint32_t n_planes = p.n_planes;
    const float4 *__restrict__ planes_g = planes_all + p.plane_off;
    const uint32_t n4 = (n_planes + 3u) & ~3u;
    for (uint32_t i = threadIdx.x; i < n4; i += kMxBlock)
        planes[i] = i < n_planes ? planes_g[i] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    __syncthreads();
    if (item >= 2u * kMxItems) return;
    const bool aperture = item < (uint32_t)kMxItems;
    const uint32_t j = aperture ? item : item - kMxItems;
    const int cell = (int)(j / kMxSamples), k = (int)(j % kMxSamples);
    const float u = mx_grid(cell / kMxGrid), v = mx_grid(cell % kMxGrid);
    const size_t out = (size_t)pose_i * kMxItems + j;
    const bool inside = !(fmaf(v, v, u * u) > 0.70f);
    if (aperture) {
        uint8_t c = MX_NONE;
        float2 fr = make_float2(0.0f, 0.0f);
        if (inside) {
            const float dx[kMxSamples] = {0.0f, 0.08f, -0.08f, 0.0f, 0.0f};
            const float dz[kMxSamples] = {0.0f, 0.0f, 0.0f, 0.08f, -0.08f};
            const V3 fwd = basis_v(p, 0, 0), right = basis_v(p, 0, 1), up = basis_v(p, 0, 2);
            const V3 dir = normalize(add(add(fwd, mul(right, dx[k])), mul(up, dz[k])));
            const V3 org = add(add(mul(neg(fwd), 2.5f), mul(right, u * 0.95f)), mul(up, v * 0.95f));
            float t;
            V3 nrm;
            int facet;
            if (mx_intersect(planes, n_planes, org, dir, t, nrm, facet)) {
                const V3 entry = add(org, mul(dir, t));
                const float cos_i = clamp_rs(dot(neg(dir), nrm), 0.0f, 1.0f);
                const MxFate fd = mx_trace(entry, dir, nrm, cos_i, planes, n_planes, p.nd);
                c = MX_EXTINCT;
                if (fd.kind == FATE_BLOCKED) {
                    c = MX_BLOCKED;
                } else if (fd.kind == FATE_LEAKED) {
                    c = MX_WINDOWED;
                } else if (fd.kind == FATE_UP && mx_returned(fd.dir, fwd, p)) {
                    c = MX_RETURNED;
                    const MxFate ff = mx_trace(entry, dir, nrm, cos_i, planes, n_planes, p.nf);
                    const MxFate fc = mx_trace(entry, dir, nrm, cos_i, planes, n_planes, p.nc);
                    if (ff.kind == FATE_UP && fc.kind == FATE_UP && ff.facet == fc.facet && ff.bounces == fc.bounces) {
                        const float cos_sep = clamp_rs(dot(ff.dir, fc.dir), -1.0f, 1.0f);
                        // f32::acos, correctly rounded, then f32::to_degrees.
                        fr.x = acos_cr(cos_sep) * 57.2957795130823208767981548141051703f;
                        const float energy = fd.transmittance * ff.transmittance * fc.transmittance;
                        const float radiance = fmaxf(ff.exit_cos, 0.0f) * fmaxf(fc.exit_cos, 0.0f);
                        fr.y = energy * radiance;
                        c = MX_RETURNED_FIRE;
                    }
                }
            }
        }
        cls[out] = c;
        fire[out] = fr;
    } else {
        uint8_t r = 0;
        if (inside) {
            const V3 fwd = basis_v(p, 1 + k, 0), right = basis_v(p, 1 + k, 1), up = basis_v(p, 1 + k, 2);
            const V3 org = add(add(mul(neg(fwd), 2.5f), mul(right, u * 0.95f)), mul(up, v * 0.95f));
            float t;
            V3 nrm;
            int facet;
            if (mx_intersect(planes, n_planes, org, fwd, t, nrm, facet)) {
                const V3 entry = add(org, mul(fwd, t));
                const float cos_i = clamp_rs(dot(neg(fwd), nrm), 0.0f, 1.0f);
                const MxFate fd = mx_trace(entry, fwd, nrm, cos_i, planes, n_planes, p.nd);
                r = (fd.kind == FATE_UP && mx_returned(fd.dir, fwd, p)) ? 1 : 0;
            }
        }
        temporal[out] = r;
    }
}

// The accumulation and final metrics (metrics.rs:1143-1248 with :366-450), one lane per
// pose, over that pose's records in the reference's order.
__global__ __launch_bounds__(64) void metrics_reduce_kernel(const uint8_t *__restrict__ cls,
                                                            const float2 *__restrict__ fire,
                                                            const uint8_t *__restrict__ temporal, uint32_t n_poses,
                                                            float *__restrict__ out) {
    const uint32_t pose_i = blockIdx.x * 64u + threadIdx.x;
    if (pose_i >= n_poses) return;
    const uint8_t *c8 = cls + (size_t)pose_i * kMxItems;
    const float2 *fr = fire + (size_t)pose_i * kMxItems;
    const uint8_t *t8 = temporal + (size_t)pose_i * kMxItems;
    uint32_t total = 0, windowed = 0, extinct = 0, returned = 0, cell_count = 0;
    float fire_sum = 0.0f, cfs = 0.0f, cfs2 = 0.0f, tvs = 0.0f;
    for (int cell = 0; cell < kMxCells; ++cell) {
        uint32_t cell_total = 0, cell_returned = 0;
        for (int k = 0; k < kMxSamples; ++k) {
            const uint8_t r = c8[cell * kMxSamples + k];
            if (r == MX_NONE) continue;
            ++total;
            ++cell_total;
            if (r == MX_WINDOWED) {
                ++windowed;
            } else if (r == MX_EXTINCT) {
                ++extinct;
            } else if (r == MX_RETURNED || r == MX_RETURNED_FIRE) {
                ++returned;
                ++cell_returned;
                if (r == MX_RETURNED_FIRE) {
                    const float2 f = fr[cell * kMxSamples + k];
                    fire_sum = fmaf(f.x, f.y, fire_sum);
                }
            }
        }
        if (cell_total > 0) {
            const float frac = (float)cell_returned / (float)cell_total;
            cfs += frac;
            cfs2 = fmaf(frac, frac, cfs2);
            ++cell_count;
            uint32_t tr = 0;
            for (int k = 0; k < kMxSamples; ++k) tr += t8[cell * kMxSamples + k] ? 1u : 0u;
            const float pr = (float)tr / 5.0f;
            tvs += fmaf(pr, -pr, pr);
        }
    }
    const float n_total = (float)(total > 1u ? total : 1u);
    const float windowing = clamp_rs((float)windowed / n_total * 100.0f, 0.0f, 100.0f);
    const float extinction = clamp_rs((float)extinct / n_total * 100.0f, 0.0f, 100.0f);
    const float brilliance = clamp_rs((float)returned / n_total * 100.0f, 0.0f, 100.0f);
    const float fire_index = fmaxf(fire_sum / n_total * 275.0f, 0.1f);
    float spatial = 0.0f, temporal_pct = 0.0f;
    if (cell_count > 0) {
        const float mean = cfs / (float)cell_count;
        if (mean > 1e-4f) {
            const float var = fmaxf(fmaf(mean, -mean, cfs2 / (float)cell_count), 0.0f);
            const float cv = sqrtf(var) / mean;
            spatial = clamp_rs(cv / (1.0f + cv) * 100.0f, 0.0f, 100.0f);
        }
        temporal_pct = clamp_rs((tvs / (float)cell_count) / 0.24f * 100.0f, 0.0f, 100.0f);
    }
    const float scint = clamp_rs(fmaf(0.4f, temporal_pct, 0.6f * spatial), 0.0f, 100.0f);
    float *o = out + 5 * (size_t)pose_i;
    o[0] = brilliance;
    o[1] = fire_index;
    o[2] = scint;
    o[3] = windowing;
    o[4] = extinction;
}

}  // namespace gr

namespace {

constexpr float kPiF = 3.14159265358979323846f;

struct M3 {
    float x, y, z;
};
inline float mdot(M3 a, M3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
inline M3 mscale(M3 a, float s) { return M3{a.x * s, a.y * s, a.z * s}; }
inline M3 mcross(M3 a, M3 b) { return M3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline M3 mnormalize(M3 a) { return mscale(a, 1.0f / std::sqrt(mdot(a, a))); }
inline float to_radians(float deg) { return deg * (kPiF / 180.0f); }

// DispersionModel::evaluate (optics/dispersion.rs:12-37), host evaluation.
float h_dispersion(const gr_material *m, float lambda_nm) {
    const float l_um = lambda_nm * 1e-3f;
    const float l2 = l_um * l_um;
    if (m->disp_kind == GR_DISP_SELLMEIER1) return std::sqrt(std::fmax(1.0f + (m->disp[0] * l2) / (l2 - m->disp[1]), 1.0f));
    if (m->disp_kind == GR_DISP_SELLMEIER3) {
        float n2 = 1.0f;
        n2 += (m->disp[0] * l2) / (l2 - m->disp[3]);
        n2 += (m->disp[1] * l2) / (l2 - m->disp[4]);
        n2 += (m->disp[2] * l2) / (l2 - m->disp[5]);
        return std::sqrt(std::fmax(n2, 1.0f));
    }
    const float l4 = l2 * l2;
    return m->disp[0] + (m->disp[1] / l2) + (m->disp[2] / l4);
}

// camera_view_basis (metrics.rs:15-30) into basis[0..9) = forward, right, up.
void h_view_basis(float yaw, float pitch, float *basis) {
    const float cos_cp = std::cos(pitch), sin_cp = std::sin(pitch);
    const float cos_cy = std::cos(yaw), sin_cy = std::sin(yaw);
    const M3 f = mnormalize(M3{-cos_cp * sin_cy, -sin_cp, -cos_cp * cos_cy});
    const M3 world_up = std::fabs(cos_cp) < 1e-4f ? M3{0.0f, 0.0f, -1.0f} : M3{0.0f, 1.0f, 0.0f};
    const M3 r = mnormalize(mcross(f, world_up));
    const M3 u = mnormalize(mcross(r, f));
    const M3 v[3] = {f, r, u};
    for (int i = 0; i < 3; ++i) {
        basis[3 * i + 0] = v[i].x;
        basis[3 * i + 1] = v[i].y;
        basis[3 * i + 2] = v[i].z;
    }
}

// build_grid_eval_setup (metrics.rs:846-910) for one pose; StudioRig::new
// (optics/studio_rig.rs:50-80) for the key and fill directions.
void h_metrics_pose(const gr_material *m, const gr_metrics_pose &q, gr::MetricsPose *p) {
    std::memset(p, 0, sizeof(*p));
    p->nd = std::fmax(h_dispersion(m, 589.3f), 1.1f);
    p->nf = std::fmax(h_dispersion(m, 486.1f), 1.001f);
    p->nc = std::fmax(h_dispersion(m, 656.3f), 1.001f);
    const float cos_lp = std::cos(q.light_pitch), sin_lp = std::sin(q.light_pitch);
    const float cos_ly = std::cos(q.light_yaw), sin_ly = std::sin(q.light_yaw);
    const M3 key = mnormalize(M3{cos_lp * sin_ly, sin_lp, cos_lp * cos_ly});
    const float fill_yaw = std::fmaf(kPiF, 0.78f, q.light_yaw);
    float fill_pitch = q.light_pitch * 0.65f;
    fill_pitch = fill_pitch < 0.15f ? 0.15f : (fill_pitch > 1.2f ? 1.2f : fill_pitch);
    const M3 fill = mnormalize(M3{std::cos(fill_pitch) * std::sin(fill_yaw), std::sin(fill_pitch),
                                  std::cos(fill_pitch) * std::cos(fill_yaw)});
    p->sin_lp = sin_lp;
    p->key[0] = key.x;
    p->key[1] = key.y;
    p->key[2] = key.z;
    p->fill[0] = fill.x;
    p->fill[1] = fill.y;
    p->fill[2] = fill.z;
    h_view_basis(q.cam_yaw, q.cam_pitch, p->basis[0]);
    const float offsets_deg[gr::kMxSamples] = {-3.0f, -1.5f, 0.0f, 1.5f, 3.0f};
    for (int k = 0; k < gr::kMxSamples; ++k)
        h_view_basis(q.cam_yaw + to_radians(offsets_deg[k]), q.cam_pitch, p->basis[1 + k]);
}

bool finite4(const float *v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]) && std::isfinite(v[3]);
}

}  // namespace

hipError_t gr_launch_metrics(const float4 *planes, const gr::MetricsPose *poses, uint32_t n_poses, uint8_t *cls,
                             float2 *fire, uint8_t *temporal, float *out, hipStream_t stream) {
    hipLaunchKernelGGL(gr::metrics_items_kernel, dim3(n_poses * gr::kMxBlocksPerPose), dim3(gr::kMxBlock), 0, stream,
                       planes, poses, cls, fire, temporal);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gr::metrics_reduce_kernel, dim3((n_poses + 63u) / 64u), dim3(64), 0, stream, cls, fire,
                       temporal, n_poses, out);
    return hipGetLastError();
}

// Every (stone, pose) of the call in one launch: the stones' planes concatenated, each pose
// pointing at its stone's.  Stones without facets get the reference's placeholders.
static gr_status optical_metrics(gr_ctx *ctx, const gr_metrics_stone *stones, uint32_t n_stones, gr_gem_metrics *out) {
    if (!ctx || !stones || !out || n_stones == 0) return GR_ERR_INVALID;
    uint64_t total_poses = 0, total_planes = 0;
    for (uint32_t s = 0; s < n_stones; ++s) {
        const gr_metrics_stone &st = stones[s];
        if (!st.material || !st.poses || st.n_poses == 0) return GR_ERR_INVALID;
        if (st.n_planes > GR_MAX_PLANES || (st.n_planes > 0 && !st.planes)) return GR_ERR_INVALID;
        for (uint32_t i = 0; i < st.n_planes; ++i) {
            const float *p = st.planes + 4 * i;
            if (!finite4(p)) return GR_ERR_INVALID;
            const double len = std::sqrt((double)p[0] * p[0] + (double)p[1] * p[1] + (double)p[2] * p[2]);
            if (len <= 1e-6) return GR_ERR_INVALID;
        }
        if (st.material->disp_kind < 0 || st.material->disp_kind > 2) return GR_ERR_INVALID;
        for (uint32_t i = 0; i < st.n_poses; ++i)
            if (!finite4(&st.poses[i].cam_yaw)) return GR_ERR_INVALID;
        for (float l : {486.1f, 589.3f, 656.3f}) {
            const float n = h_dispersion(st.material, l);
            if (!std::isfinite(n) || n < 1.0f || n > 6.0f) return GR_ERR_INVALID;
        }
        total_poses += st.n_poses;
        if (st.n_planes > 0) total_planes += st.n_planes;
    }
    if (total_poses > gr::kMxMaxPoses) return GR_ERR_INVALID;
    std::vector<gr::MetricsPose> host;
    std::vector<float> planes;
    std::vector<uint32_t> slot(total_poses, 0xFFFFFFFFu);  // launched pose of each output, or none
    host.reserve(total_poses);
    planes.reserve(4 * total_planes);
    uint64_t o = 0;
    for (uint32_t s = 0; s < n_stones; ++s) {
        const gr_metrics_stone &st = stones[s];
        const uint32_t off = (uint32_t)(planes.size() / 4);
        if (st.n_planes > 0) planes.insert(planes.end(), st.planes, st.planes + 4 * (size_t)st.n_planes);
        for (uint32_t i = 0; i < st.n_poses; ++i, ++o) {
            // No facets: the reference's placeholder values (metrics.rs:1090-1098).
            if (st.n_planes == 0) {
                out[o] = gr_gem_metrics{85.0f, 25.0f, 75.0f, 5.0f, 5.0f};
                continue;
            }
            gr::MetricsPose p;
            h_metrics_pose(st.material, st.poses[i], &p);
            p.plane_off = off;
            p.n_planes = st.n_planes;
            slot[o] = (uint32_t)host.size();
            host.push_back(p);
        }
    }
    const uint32_t n_poses = (uint32_t)host.size();
    if (n_poses == 0) return GR_OK;

    std::lock_guard<std::mutex> g(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return GR_ERR_DEVICE;
    // One buffer, kept by the context between calls: planes, poses, per-ray records, results.
    const size_t items = (size_t)n_poses * gr::kMxItems;
    const size_t off_poses = ((planes.size() / 4) * sizeof(float4) + 255) & ~(size_t)255;
    const size_t off_fire = (off_poses + n_poses * sizeof(gr::MetricsPose) + 255) & ~(size_t)255;
    const size_t off_cls = off_fire + items * sizeof(float2);
    const size_t off_temporal = off_cls + items;
    const size_t off_out = (off_temporal + items + 255) & ~(size_t)255;
    const size_t bytes = off_out + (size_t)n_poses * 5 * sizeof(float);
    if (bytes > ctx->mx_cap) {
        if (ctx->mx_buf) (void)hipFree(ctx->mx_buf);
        ctx->mx_buf = nullptr;
        ctx->mx_cap = 0;
        if (hipMalloc(&ctx->mx_buf, bytes) != hipSuccess) return GR_ERR_DEVICE;
        ctx->mx_cap = bytes;
    }
    char *buf = ctx->mx_buf;
    std::vector<float> res((size_t)n_poses * 5);
    gr_status st = GR_OK;
    if (hipMemcpyAsync(buf, planes.data(), planes.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipMemcpyAsync(buf + off_poses, host.data(), n_poses * sizeof(gr::MetricsPose), hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess ||
        gr_launch_metrics(reinterpret_cast<const float4 *>(buf),
                          reinterpret_cast<const gr::MetricsPose *>(buf + off_poses), n_poses,
                          reinterpret_cast<uint8_t *>(buf + off_cls), reinterpret_cast<float2 *>(buf + off_fire),
                          reinterpret_cast<uint8_t *>(buf + off_temporal), reinterpret_cast<float *>(buf + off_out),
                          ctx->stream) != hipSuccess ||
        hipMemcpyAsync(res.data(), buf + off_out, res.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = GR_ERR_DEVICE;
    if (st != GR_OK) return st;
    for (uint64_t i = 0; i < total_poses; ++i) {
        const uint32_t k = slot[i];
        if (k != 0xFFFFFFFFu)
            out[i] = gr_gem_metrics{res[5 * k + 0], res[5 * k + 1], res[5 * k + 2], res[5 * k + 3], res[5 * k + 4]};
    }
    return GR_OK;
}

extern "C" {

gr_status gr_optical_metrics(gr_ctx *ctx, const float *planes, uint32_t n_planes, const gr_material *material,
                             const gr_metrics_pose *poses, uint32_t n_poses, gr_gem_metrics *out) {
    const gr_metrics_stone stone{planes, n_planes, material, poses, n_poses};
    return optical_metrics(ctx, &stone, 1, out);
}

gr_status gr_optical_metrics_batch(gr_ctx *ctx, const gr_metrics_stone *stones, uint32_t n_stones,
                                   gr_gem_metrics *out) {
    return optical_metrics(ctx, stones, n_stones, out);
}

}  // extern "C"
