// trace_launch.h -- the launchers of the trace kernels for one scene class, instantiated by
// the per-class translation units (trace_<anisotropy>_<environment>.hip) so that the
// kernel variants compile in parallel.  transport.hip dispatches to them by scene class.
#pragma once

#include "transport_kernels.h"

namespace gr {
namespace detail {

template <int A, bool B, int E, bool F, bool I, int W>
hipError_t launch_plain_i(const DevScene &sc, const TraceArgs &a, hipStream_t stream) {
    dim3 grid((a.n_tuples + kBlock - 1) / kBlock);
    const bool split = a.split_bounce != 0 && a.cont != nullptr && a.cont_cap != 0;
    if (split) {
        hipError_t e = hipMemsetAsync(a.cont_count, 0, (kContLevels + 1) * sizeof(uint32_t), stream);
        if (e != hipSuccess) return e;
    }
    const size_t lds = plane_lds_bytes(sc.n_planes);
    hipLaunchKernelGGL((trace_kernel<A, B, E, F, I, W>), grid, dim3(kBlock), lds, stream, sc, a.planes, a.ri35,
                       a.cert, a.pixel_begin, a.sample_offset, a.spp, a.n_tuples, a.xyz, a.info, a.stats,
                       split ? a.split_bounce : 0u, a.park_min, a.park_lanes, a.cont, a.cont_cap, a.cont_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.trace_done) {
        e = hipEventRecord(a.trace_done, stream);
        if (e != hipSuccess) return e;
    }
    if (!split) return hipSuccess;
    // Level k (k = 1..kContLevels) continues the records of level k-1 and, except the
    // last, parks again after at most split_bounce iterations (earlier by the park_lanes
    // rule); records alternate between the two halves of cont.
    dim3 cgrid((a.cont_cap + kBlock - 1) / kBlock);
    for (int k = 1; k <= kContLevels; ++k) {
        float *in = a.cont + (size_t)((k - 1) & 1) * kContFields * a.cont_cap;
        float *out = a.cont + (size_t)(k & 1) * kContFields * a.cont_cap;
        const uint32_t split_iters = k < kContLevels ? a.split_bounce : 0u;
        hipLaunchKernelGGL((cont_kernel<A, B, E, F, I, W>), cgrid, dim3(kBlock), lds, stream, sc, a.planes, in,
                           a.cont_cap, a.cont_count + (k - 1), a.xyz, a.info, a.stats, split_iters, a.park_min,
                           a.park_lanes, out, a.cont_count + k);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int A, bool B, int E>
hipError_t launch_wave(const DevScene &sc, const TraceArgs &a, hipStream_t stream) {
    const uint32_t waves = (a.n_tuples + kWaveTuples - 1) / kWaveTuples;
    dim3 grid((waves + kBlock / 64 - 1) / (kBlock / 64));
    const size_t lds = plane_lds_bytes(sc.n_planes);
    if (a.info)
        hipLaunchKernelGGL((trace_wave_kernel<A, B, true>), grid, dim3(kBlock), lds, stream, sc, a.planes,
                           a.ri35, a.cert, a.pixel_begin, a.sample_offset, a.spp, a.n_tuples, a.rec_cap, a.rec,
                           a.info, a.stats);
    else
        hipLaunchKernelGGL((trace_wave_kernel<A, B, false>), grid, dim3(kBlock), lds, stream, sc, a.planes,
                           a.ri35, a.cert, a.pixel_begin, a.sample_offset, a.spp, a.n_tuples, a.rec_cap, a.rec,
                           a.info, a.stats);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 fgrid((a.n_tuples + kBlock - 1) / kBlock);
    hipEvent_t mark = a.trace_done;
    if (mark) {
        e = hipEventRecord(mark, stream);
        if (e != hipSuccess) return e;
    }
    if (E == 0)
        hipLaunchKernelGGL((finish_kernel<0>), fgrid, dim3(kBlock), 0, stream, sc, a.n_tuples, a.rec_cap,
                           a.rec, a.xyz);
    else
        hipLaunchKernelGGL((finish_kernel<1>), fgrid, dim3(kBlock), 0, stream, sc, a.n_tuples, a.rec_cap,
                           a.rec, a.xyz);
    return hipGetLastError();
}


template <int C>
hipError_t launch_batch_class_impl(const BatchSlice *slices, uint32_t n_slices, uint32_t n_blocks,
                              uint32_t cont_blocks, uint32_t max_planes, hipStream_t stream) {
    constexpr int kA = (C & 16) ? 2 : (C & 1);
    constexpr bool kB = (C & 2) != 0, kF = (C & 8) != 0;
    constexpr int kE = (C & 4) ? 1 : 0;
    const auto *view = reinterpret_cast<const BatchSliceView *>(slices);
    const size_t lds = plane_lds_bytes(max_planes);
    hipLaunchKernelGGL((trace_batch_kernel<kA, kB, kE, kF>), dim3(n_blocks), dim3(kBlock), lds, stream, view,
                       n_slices);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || cont_blocks == 0) return e;
    for (int level = 1; level <= kContLevels; ++level) {
        hipLaunchKernelGGL((cont_batch_kernel<kA, kB, kE, kF>), dim3(cont_blocks), dim3(kBlock), lds, stream,
                           view, n_slices, level);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace detail

// Scene class C (scene_class bits): the plain or wavefront launch of one trace call, and the
// batched launch of one class's slices.
template <int C>
hipError_t launch_trace_class(const DevScene &sc, const TraceArgs &a, hipStream_t stream) {
    constexpr int kA = (C & 16) ? 2 : (C & 1);
    constexpr bool kB = (C & 2) != 0, kF = (C & 8) != 0;
    constexpr int kE = (C & 4) ? 1 : 0;
    constexpr int kW = GR_WPE(kA, kF);
    if (a.wavefront) return detail::launch_wave<kA, kB, kE>(sc, a, stream);
    // The per-path record kernels (parity tests) are built at the default occupancy only.
    if (a.info) return detail::launch_plain_i<kA, kB, kE, kF, true, kW>(sc, a, stream);
    switch (a.waves_per_eu) {
        case 2: return detail::launch_plain_i<kA, kB, kE, kF, false, 2>(sc, a, stream);
        case 3: return detail::launch_plain_i<kA, kB, kE, kF, false, 3>(sc, a, stream);
        case 4: return detail::launch_plain_i<kA, kB, kE, kF, false, 4>(sc, a, stream);
        default: return detail::launch_plain_i<kA, kB, kE, kF, false, kW>(sc, a, stream);
    }
}

template <int C>
hipError_t launch_batch_class(const BatchSlice *slices, uint32_t n_slices, uint32_t n_blocks, uint32_t cont_blocks,
                              uint32_t max_planes, hipStream_t stream) {
    return detail::launch_batch_class_impl<C>(slices, n_slices, n_blocks, cont_blocks, max_planes, stream);
}

}  // namespace gr

// Instantiates both launchers for the four classes with anisotropy bits A and environment
// bit E (absorbing or not, rare features or not).
#define GR_INSTANTIATE_TRACE_CLASS(C)                                                                      \
    template hipError_t gr::launch_trace_class<C>(const gr::DevScene &, const gr::TraceArgs &, hipStream_t); \
    template hipError_t gr::launch_batch_class<C>(const gr::BatchSlice *, uint32_t, uint32_t, uint32_t, uint32_t, \
                                                  hipStream_t);
#define GR_INSTANTIATE_TRACE_CLASSES(A, E) \
    GR_INSTANTIATE_TRACE_CLASS((A) | (E))     \
    GR_INSTANTIATE_TRACE_CLASS((A) | (E) | 2) \
    GR_INSTANTIATE_TRACE_CLASS((A) | (E) | 8) \
    GR_INSTANTIATE_TRACE_CLASS((A) | (E) | 10)
