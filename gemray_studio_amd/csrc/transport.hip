// transport.hip -- the shared non-template kernels (miss certificates, guide pass, probes,
// film reduction) and the dispatch of trace launches to the per-class translation units.
#include "transport_kernels.h"

namespace gr {

// Per-pixel camera-ray miss certificates: the scan of the ray through the pixel's centre
// (jitter 0); when it misses because its entering extreme exceeds its exiting one, the two
// extreme planes are the pixel's certificate, which every sample of the pixel then checks
// exactly for its own ray (camera_miss_certified).
__device__ __forceinline__ uint32_t pixel_cert(const DevScene &sc, const float4 *__restrict__ planes,
                                               const float *__restrict__ cam_side, uint32_t pixel) {
    const uint32_t x = pixel % sc.width, y = pixel / sc.width;
    const float w = (float)sc.width, h = (float)sc.height;
    const float aspect = w / h;
    const float u = ((float)x / w - 0.5f) * 2.0f * aspect * sc.fov_tan;
    const float v = (0.5f - (float)y / h) * 2.0f * sc.fov_tan;
    const V3 rd = normalize(add(add(v3(sc.cam_f[0], sc.cam_f[1], sc.cam_f[2]),
                                    mul(v3(sc.cam_r[0], sc.cam_r[1], sc.cam_r[2]), u)),
                                mul(v3(sc.cam_u[0], sc.cam_u[1], sc.cam_u[2]), v)));
    const V3 ro = v3(sc.cam_o[0], sc.cam_o[1], sc.cam_o[2]);
    float t_near = -1e30f, t_far = 1e30f;
    int near_idx = -1, far_idx = -1;
    bool outside = false;
    clip_scan<true>(planes, cam_side, sc.n_planes, ro, rd, t_near, near_idx, t_far, far_idx, outside);
    const bool proof = !outside && t_near > t_far && near_idx >= 0 && far_idx >= 0;
    return proof ? ((uint32_t)near_idx | ((uint32_t)far_idx << 16)) : 0xFFFFFFFFu;
}
__global__ __launch_bounds__(kBlock) void cert_kernel(const DevScene sc, const float4 *__restrict__ planes_g,
                                                      uint32_t n_pixels, uint32_t *__restrict__ cert) {
    __shared__ float4 planes[GR_MAX_PLANES];
    __shared__ __align__(16) float cam_side[GR_MAX_PLANES];
    stage_planes(planes, cam_side, planes_g, sc);
    __syncthreads();
    const uint32_t pixel = blockIdx.x * kBlock + threadIdx.x;
    if (pixel >= n_pixels) return;
    cert[pixel] = pixel_cert(sc, planes, cam_side, pixel);
}
// The certificates of many jobs' frames in one launch (the job loop's admission round):
// workgroup b belongs to the last slice whose first block is <= b.
__global__ __launch_bounds__(kBlock) void cert_batch_kernel(const CertSlice *__restrict__ slices, uint32_t n_slices) {
    uint32_t lo = 0u, hi = n_slices;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (slices[mid].block_begin <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const CertSlice &J = slices[lo];
    const DevScene &sc = *J.scene;
    __shared__ float4 planes[GR_MAX_PLANES];
    __shared__ __align__(16) float cam_side[GR_MAX_PLANES];
    stage_planes(planes, cam_side, J.planes, sc);
    __syncthreads();
    const uint32_t pixel = (blockIdx.x - J.block_begin) * kBlock + threadIdx.x;
    if (pixel >= J.n_pixels) return;
    J.cert[pixel] = pixel_cert(sc, planes, cam_side, pixel);
}

// The denoiser's primary-hit guide buffers (apps/diagram-gui/src/bridge/guide_pass.rs:60-110):
// per pixel, the camera ray through the pixel's corner (jitter 0, the ray cert_kernel traces)
// against the stone (intersect_polyhedron, intersect.rs:9-78): the hit distance, the facet's
// normal and index, or 1e6, a zero normal and -1 on a miss.  An unbounded side (no plane on
// it) reports facet 0 with a zero normal, as HitRecord does.
__global__ __launch_bounds__(kBlock) void guide_kernel(const DevScene sc, const float4 *__restrict__ planes_g,
                                                       uint32_t n_pixels, float *__restrict__ depth,
                                                       float *__restrict__ normal, int32_t *__restrict__ facet) {
    __shared__ float4 planes[GR_MAX_PLANES];
    __shared__ __align__(16) float cam_side[GR_MAX_PLANES];
    stage_planes(planes, cam_side, planes_g, sc);
    __syncthreads();
    const uint32_t pixel = blockIdx.x * kBlock + threadIdx.x;
    if (pixel >= n_pixels) return;
    const uint32_t x = pixel % sc.width, y = pixel / sc.width;
    const float w = (float)sc.width, h = (float)sc.height;
    const float aspect = w / h;
    const float u = ((float)x / w - 0.5f) * 2.0f * aspect * sc.fov_tan;
    const float v = (0.5f - (float)y / h) * 2.0f * sc.fov_tan;
    const V3 rd = normalize(add(add(v3(sc.cam_f[0], sc.cam_f[1], sc.cam_f[2]),
                                    mul(v3(sc.cam_r[0], sc.cam_r[1], sc.cam_r[2]), u)),
                                mul(v3(sc.cam_u[0], sc.cam_u[1], sc.cam_u[2]), v)));
    const V3 ro = v3(sc.cam_o[0], sc.cam_o[1], sc.cam_o[2]);
    float t_near = -1e30f, t_far = 1e30f;
    int near_idx = -1, far_idx = -1;
    bool outside = false;
    clip_scan<true>(planes, cam_side, sc.n_planes, ro, rd, t_near, near_idx, t_far, far_idx, outside);
    float t = 1.0e6f;
    int idx = -1;
    bool hit = false;
    if (!outside && !(t_near > t_far)) {
        if (t_near > 1e-4f) {
            hit = true;
            t = t_near;
            idx = near_idx;
        } else if (t_far > 1e-4f) {
            hit = true;
            t = t_far;
            idx = far_idx;
        }
    }
    V3 n = v3(0.0f, 0.0f, 0.0f);
    if (hit && idx >= 0) {
        const float4 p = planes[idx];
        n = v3(p.x, p.y, p.z);
    }
    depth[pixel] = t;
    normal[3 * (size_t)pixel + 0] = n.x;
    normal[3 * (size_t)pixel + 1] = n.y;
    normal[3 * (size_t)pixel + 2] = n.z;
    facet[pixel] = hit ? (idx >= 0 ? idx : 0) : -1;
}

// Per-function probes: one device function of the transport path per lane, on
// caller-given inputs (9 floats in, 3 out per item), for the per-function ULP tests.
__global__ __launch_bounds__(kBlock) void probe_kernel(int fn, const float *__restrict__ in,
                                                       float *__restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float *a = in + 9 * (size_t)i;
    float *o = out + 3 * (size_t)i;
    o[0] = o[1] = o[2] = 0.0f;
    switch (fn) {
        case GR_PROBE_NEFF:
            o[0] = neff_from_cos(a[0], a[1], a[2]);
            break;
        case GR_PROBE_POYNTING: {
            const V3 d = extraordinary_poynting_dir(v3(a[0], a[1], a[2]), v3(a[3], a[4], a[5]), a[6], a[7]);
            o[0] = d.x;
            o[1] = d.y;
            o[2] = d.z;
            break;
        }
        case GR_PROBE_TIR_RETARDATION:
            tir_retardation_cs(a[0], a[1], a[2], &o[0], &o[1]);
            break;
        case GR_PROBE_FRAME_ROTATION:
            frame_rotation_cs(v3(a[0], a[1], a[2]), v3(a[3], a[4], a[5]), v3(a[6], a[7], a[8]), &o[0], &o[1]);
            break;
        case GR_PROBE_EXP:
            o[0] = exp_lobe(a[0]);
            break;
        case GR_PROBE_ACOS:
            o[0] = acos_cr(a[0]);
            break;
        case GR_PROBE_CMF: {
            const V3 c = cmf_x8_channel(a[0]);
            o[0] = c.x;
            o[1] = c.y;
            o[2] = c.z;
            break;
        }
        case GR_PROBE_FIELD_QFORM: {
            // a = (q, u, S.x, S.y, S.z, P.x, P.y, P.z, ao); ae = 1 - ao.
            const QfInv qi = qf_invariants(v3(a[2], a[3], a[4]), v3(a[5], a[6], a[7]));
            o[0] = qf_channel(qi, a[0], a[1], a[8], 1.0f - a[8]);
            break;
        }
        case GR_PROBE_BX_WAVE_INDICES:
        case GR_PROBE_BX_EIGEN_SLOW:
        case GR_PROBE_BX_EIGEN_FAST:
        case GR_PROBE_BX_POYNTING_SLOW:
        case GR_PROBE_BX_POYNTING_FAST: {
            // a = (n_alpha, n_beta, n_gamma, w.xyz, gamma axis.xyz); the frame is the
            // gamma axis normalized and its stable orthonormal completion.
            BxFrame fr;
            fr.az = normalize_or_zero(v3(a[6], a[7], a[8]));
            stable_orthonormal_basis(fr.az, &fr.ax, &fr.ay);
            const Bx b{a[0], a[1], a[2]};
            const V3 w = v3(a[3], a[4], a[5]);
            V3 r = v3(0.0f, 0.0f, 0.0f);
            if (fn == GR_PROBE_BX_WAVE_INDICES) {
                bx_wave_indices(b, bx_local(fr, normalize_or_zero(w)), &r.x, &r.y);
            } else if (fn == GR_PROBE_BX_EIGEN_SLOW || fn == GR_PROBE_BX_EIGEN_FAST) {
                V3 ds, df;
                bx_eigen_polarizations(b, fr, w, &ds, &df);
                r = fn == GR_PROBE_BX_EIGEN_SLOW ? ds : df;
            } else {
                r = bx_mode_poynting_dir(b, fr, w, fn == GR_PROBE_BX_POYNTING_SLOW);
            }
            o[0] = r.x;
            o[1] = r.y;
            o[2] = r.z;
            break;
        }
        case GR_PROBE_SAMPLE_DRAWS: {
            // a = (pixel, sample_num as u32 bits, ri3, ri5 from the host's sample table).
            uint32_t seed;
            sample_draws(__float_as_uint(a[0]), __float_as_uint(a[1]), make_float2(a[2], a[3]), &seed, &o[0], &o[1],
                         &o[2]);
            break;
        }
        case GR_PROBE_SEED: {
            // a = (pixel, sample_num, x as u32 bits).
            uint32_t seed;
            float jx, jy, hero;
            sample_draws(__float_as_uint(a[0]), __float_as_uint(a[1]), make_float2(0.0f, 0.0f), &seed, &jx, &jy, &hero);
            o[0] = __uint_as_float(seed);
            o[1] = __uint_as_float(hash_u32(__float_as_uint(a[2])));
            o[2] = __uint_as_float(__builtin_bitreverse32(__float_as_uint(a[2])));
            break;
        }
        case GR_PROBE_STREAM_DRAW: {
            // a = (seed, bounce, salt0, salt1, salt2 as u32 bits).
            const uint32_t seed = __float_as_uint(a[0]), bounce = __float_as_uint(a[1]);
            for (int k = 0; k < 3; ++k) o[k] = stream_draw(seed, bounce, __float_as_uint(a[2 + k]));
            break;
        }
        case GR_PROBE_RGB_SPECTRAL:
            o[0] = rgb_to_spectral(v3(a[0], a[1], a[2]), a[3]);
            break;
        case GR_PROBE_DOP:
            o[0] = degree_of_polarization(Stokes{a[0], a[1], a[2], a[3]});
            break;
        case GR_PROBE_FRESNEL_COMPANION: {
            // a = (n1, n2, cos_i, cos_t): the refraction loop's companion arithmetic.
            const float n1 = a[0], n2 = a[1], cos_i = a[2], cos_t = a[3];
            fresnel_rs_rp_hw(n1, n2, cos_i, cos_t, &o[0], &o[1]);
            const float num = 2.0f * n1 * cos_i;
            const float t_s = div_hw(num, fmaf(n2, cos_t, n1 * cos_i));
            const float t_p = div_hw(num, fmaf(n1, cos_t, n2 * cos_i));
            o[2] = apply_fresnel_transmission(Stokes{1.0f, 0.0f, 0.0f, 0.0f}, n1, n2, cos_i, cos_t, t_s, t_p).i;
            break;
        }
        default:
            break;
    }
}

// The spectral hero table (transport.h): entry e holds, for hero offset hero_key(e), each comb
// channel's CIE colour-matching values (cmf_x8_channel) and map-spectrum lobes (rgb_lobes),
// the calls finish_escaped makes per path.
__global__ __launch_bounds__(kBlock) void spec_table_kernel(float *__restrict__ tab) {
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= kHeroEntries) return;
    const float h = hero_key(e);
    float *out = tab + (size_t)e * kSpecFloats;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float l = comb_lambda(h, k);
        const V3 c = cmf_x8_channel(l);
        const V3 lobes = rgb_lobes(l);
        out[3 * k] = c.x;
        out[3 * k + 1] = c.y;
        out[3 * k + 2] = c.z;
        out[24 + 3 * k] = lobes.x;
        out[25 + 3 * k] = lobes.y;
        out[26 + 3 * k] = lobes.z;
    }
}

// A material's hero table (transport.h): entry e holds, for hero offset hero_key(e), each comb
// channel's dispersion index and absorption sums, as the trace kernels' lazy fills evaluate
// them (dispersion_index, band_sum; a biaxial stone without a beta band set repeats the
// ordinary sum).
template <bool kBiax>
__global__ __launch_bounds__(kBlock) void mat_table_kernel(const DevScene sc, float *__restrict__ tab) {
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= kHeroEntries) return;
    const float h = hero_key(e);
    float *out = tab + (size_t)e * mat_floats(kBiax);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float l = comb_lambda(h, k);
        const float so = band_sum<0>(l, sc);
        out[k] = dispersion_index(l, sc);
        out[8 + k] = so;
        out[16 + k] = band_sum<1>(l, sc);
        if (kBiax) out[24 + k] = sc.has_beta_bands ? band_sum<2>(l, sc) : so;
    }
}

// The sample table of a call on the device: ri35[i] = the radical inverses (bases 3 and 5) of
// sample offset + i (radical_inverse_base, sampling.rs:170-181), the operations the host's
// table takes.
__global__ __launch_bounds__(kBlock) void ri35_kernel(uint32_t offset, uint32_t n, float2 *__restrict__ ri35) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    ri35[i] = make_float2(radical_inverse(offset + i, 3u), radical_inverse(offset + i, 5u));
}

// One pixel's spp tuples (XYZ, 3 floats each) summed in sample order.  The lanes of a wave
// read 3*spp floats apart, so every load touches a cache line of its own; with spp a
// multiple of 4 and 16-byte aligned tuples (by4), four samples come in as three float4 --
// the same additions in the same order from a quarter of the load instructions.
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <typename P4, typename P>
__device__ __forceinline__ void sum_pixel_tuples(P src, uint32_t spp, bool by4, float &sx, float &sy, float &sz) {
    sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (by4) {
        P4 q = reinterpret_cast<P4>(src);
        for (uint32_t s = 0; s < spp; s += 4, q += 3) {
            const f32x4 a = q[0], b = q[1], c = q[2];  // x0 y0 z0 x1 | y1 z1 x2 y2 | z2 x3 y3 z3
            sx = sx + a.x; sy = sy + a.y; sz = sz + a.z;
            sx = sx + a.w; sy = sy + b.x; sz = sz + b.y;
            sx = sx + b.z; sy = sy + b.w; sz = sz + c.x;
            sx = sx + c.y; sy = sy + c.z; sz = sz + c.w;
        }
        return;
    }
    for (uint32_t s = 0; s < spp; ++s) {
        sx = sx + src[3 * s + 0];
        sy = sy + src[3 * s + 1];
        sz = sz + src[3 * s + 2];
    }
}

// Each pixel's spp tuples summed in sample order and ADDED into the film
// (frame.rs:799-814: sum = 0; sum += xyz[s] for s in order; accum[pixel] += sum).
__global__ __launch_bounds__(kBlock) void film_reduce_kernel(const float *__restrict__ xyz,
                                                             uint32_t n_pixels, uint32_t spp,
                                                             uint32_t pixel_begin,
                                                             float *__restrict__ film) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n_pixels) return;
    const bool by4 = (spp & 3u) == 0 && (reinterpret_cast<uintptr_t>(xyz) & 15u) == 0;
    float sx, sy, sz;
    sum_pixel_tuples<const f32x4 *>(xyz + 3 * (size_t)p * spp, spp, by4, sx, sy, sz);
    float *dst = film + 3 * (size_t)(pixel_begin + p);
    dst[0] = dst[0] + sx;
    dst[1] = dst[1] + sy;
    dst[2] = dst[2] + sz;
}

// film_reduce_kernel over the slices of a batched tick: each slice's pixels sum their
// spp staged tuples in sample order and add the sum into the slice's job film.
__global__ __launch_bounds__(kBlock) void film_reduce_batch_kernel(const BatchSliceView *__restrict__ slices,
                                                                   uint32_t n_slices) {
    const BatchSliceView &J = slices[find_slice(slices, n_slices, blockIdx.x, kReduceBlocks)];
    const uint32_t p = (blockIdx.x - J.reduce_block_begin) * kBlock + threadIdx.x;
    if (p >= J.n_pixels) return;
    const uint32_t spp = J.spp;
    const bool by4 = (spp & 3u) == 0 && ((uint64_t)J.xyz & 15u) == 0;
    float sx, sy, sz;
    sum_pixel_tuples<const GR_AS_GLOBAL f32x4 *>(J.xyz + 3 * (size_t)p * spp, spp, by4, sx, sy, sz);
    GR_AS_GLOBAL float *dst = J.film + 3 * (size_t)(J.pixel_begin + p);
    dst[0] = dst[0] + sx;
    dst[1] = dst[1] + sy;
    dst[2] = dst[2] + sz;
}

}  // namespace gr

// ---------------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------------

#define GR_CLASS_CASES(X)                                                                                    \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(17) X(19) X(21) \
        X(23) X(25) X(27) X(29) X(31)

hipError_t gr_launch_trace(const gr::DevScene &sc, const gr::TraceArgs &a, hipStream_t stream) {
    switch (gr::scene_class(sc)) {
#define GR_TRACE_CASE(C) \
    case C: return gr::launch_trace_class<C>(sc, a, stream);
        GR_CLASS_CASES(GR_TRACE_CASE)
#undef GR_TRACE_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t gr_launch_trace_batch(const gr::BatchSlice *slices, uint32_t n_slices, uint32_t n_blocks,
                                 uint32_t cont_blocks, uint32_t max_planes, int cls, hipStream_t stream) {
    if (n_slices == 0 || n_blocks == 0) return hipSuccess;
    switch (cls) {
#define GR_BATCH_CASE(C) \
    case C: return gr::launch_batch_class<C>(slices, n_slices, n_blocks, cont_blocks, max_planes, stream);
        GR_CLASS_CASES(GR_BATCH_CASE)
#undef GR_BATCH_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t gr_launch_film_reduce_batch(const gr::BatchSlice *slices, uint32_t n_slices, uint32_t n_blocks,
                                       hipStream_t stream) {
    if (n_slices == 0 || n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(gr::film_reduce_batch_kernel, dim3(n_blocks), dim3(gr::kBlock), 0, stream,
                       reinterpret_cast<const gr::BatchSliceView *>(slices), n_slices);
    return hipGetLastError();
}

hipError_t gr_launch_probe(int fn, const float *in, float *out, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(gr::probe_kernel, dim3((n + gr::kBlock - 1) / gr::kBlock), dim3(gr::kBlock), 0, stream, fn,
                       in, out, n);
    return hipGetLastError();
}

hipError_t gr_launch_cert(const gr::DevScene &sc, const float4 *planes, uint32_t *cert, hipStream_t stream) {
    const uint32_t n = sc.width * sc.height;
    hipLaunchKernelGGL(gr::cert_kernel, dim3((n + gr::kBlock - 1) / gr::kBlock), dim3(gr::kBlock), 0, stream, sc,
                       planes, n, cert);
    return hipGetLastError();
}

hipError_t gr_launch_cert_batch(const gr::CertSlice *slices, uint32_t n_slices, uint32_t n_blocks, hipStream_t stream) {
    if (n_slices == 0 || n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(gr::cert_batch_kernel, dim3(n_blocks), dim3(gr::kBlock), 0, stream, slices, n_slices);
    return hipGetLastError();
}

hipError_t gr_launch_guide(const gr::DevScene &sc, const float4 *planes, float *depth, float *normal, int32_t *facet,
                           hipStream_t stream) {
    const uint32_t n = sc.width * sc.height;
    hipLaunchKernelGGL(gr::guide_kernel, dim3((n + gr::kBlock - 1) / gr::kBlock), dim3(gr::kBlock), 0, stream, sc,
                       planes, n, depth, normal, facet);
    return hipGetLastError();
}

hipError_t gr_launch_film_reduce(const float *xyz, uint32_t n_pixels, uint32_t spp,
                                 uint32_t pixel_begin, float *film, hipStream_t stream) {
    dim3 grid((n_pixels + gr::kBlock - 1) / gr::kBlock);
    hipLaunchKernelGGL(gr::film_reduce_kernel, grid, dim3(gr::kBlock), 0, stream, xyz, n_pixels, spp,
                       pixel_begin, film);
    return hipGetLastError();
}

hipError_t gr_launch_spec_table(float *tab, hipStream_t stream) {
    hipLaunchKernelGGL(gr::spec_table_kernel, dim3((gr::kHeroEntries + gr::kBlock - 1) / gr::kBlock), dim3(gr::kBlock),
                       0, stream, tab);
    return hipGetLastError();
}

hipError_t gr_launch_mat_table(const gr::DevScene &sc, float *tab, hipStream_t stream) {
    const dim3 grid((gr::kHeroEntries + gr::kBlock - 1) / gr::kBlock);
    if (sc.biaxial)
        hipLaunchKernelGGL(gr::mat_table_kernel<true>, grid, dim3(gr::kBlock), 0, stream, sc, tab);
    else
        hipLaunchKernelGGL(gr::mat_table_kernel<false>, grid, dim3(gr::kBlock), 0, stream, sc, tab);
    return hipGetLastError();
}

hipError_t gr_launch_ri35(uint32_t offset, uint32_t n, float2 *ri35, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gr::ri35_kernel, dim3((n + gr::kBlock - 1) / gr::kBlock), dim3(gr::kBlock), 0, stream, offset,
                       n, ri35);
    return hipGetLastError();
}
