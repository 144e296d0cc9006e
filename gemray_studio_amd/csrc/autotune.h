// autotune.h -- the configuration search of the plain trace path, host C++ only (no HIP), so
// that the selection rule can be tested on its own.
//
// The reference fixes its dispatch shape (@workgroup_size(64), spectral_transport.wgsl:672)
// and its chunk budget (CHUNK_BUDGET_BYTES, frame.rs:192).  Here the knobs that change speed
// but never a result bit -- the trace kernels' occupancy variant, the path-length split
// bounce and early-parking threshold, the continuation-record capacity, the launch size and
// the trace mode -- are picked per scene class and size by timing probe renders.
#pragma once

#include <stdint.h>

#include <map>
#include <tuple>
#include <vector>

namespace gr {

struct TuneConfig {
    uint32_t waves_per_eu = 4;            // occupancy variant of the plain trace kernels: 2, 3, 4
    uint32_t split_bounce = 6;            // park paths running after this many segments (0: never)
    uint32_t cont_div = 4;                // continuation records for 1 / cont_div of a launch's tuples
    uint32_t trace_mode = 0;              // 0 plain, 1 persistent wavefronts
    uint64_t chunk_tuples = 1ull << 25;   // tuples per launch
    uint32_t park_lanes = 16;             // early parking below this many live lanes (0: none)
    std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint64_t, uint32_t> key() const {
        return std::make_tuple(waves_per_eu, split_bounce, cont_div, trace_mode, chunk_tuples, park_lanes);
    }
    bool operator==(const TuneConfig &o) const { return key() == o.key(); }
};

// The values each knob is searched over, the current value first.
struct TuneSpace {
    std::vector<uint32_t> waves_per_eu{4, 3, 2};
    std::vector<uint32_t> split_bounce{6, 3, 4, 5, 7, 8, 10, 0};
    std::vector<uint32_t> cont_div{4, 2, 1, 8};
    std::vector<uint64_t> chunk_tuples{1ull << 25, 1ull << 24, 1ull << 23};
    std::vector<uint32_t> trace_mode{0, 1};
    std::vector<uint32_t> park_lanes{16, 0, 8, 24, 32};
};

// Coordinate descent from `start`: each knob in turn (occupancy, split bounce, early-parking
// threshold, continuation capacity, launch size, trace mode) takes the value of its list whose probe is fastest with
// the other knobs at the current best, for up to `passes` passes or until a pass changes
// nothing.  A value replaces the current one only when its probe is faster by more than the
// fraction `margin`, so equal timings keep the earlier choice.  time(config) returns the
// milliseconds of one probe; each configuration is timed once.  The result is a pure
// function of the timings: the same timings always give the same choice.
template <class TimeFn>
TuneConfig coordinate_descent(const TuneConfig &start, const TuneSpace &space, double margin, int passes,
                              TimeFn &&time, double *best_ms_out = nullptr, int *n_timed = nullptr) {
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint64_t, uint32_t>, double> memo;
    auto timed = [&](const TuneConfig &c) {
        auto it = memo.find(c.key());
        if (it != memo.end()) return it->second;
        const double t = time(c);
        memo.emplace(c.key(), t);
        return t;
    };
    TuneConfig best = start;
    double best_ms = timed(best);
    for (int pass = 0; pass < passes; ++pass) {
        const TuneConfig before = best;
        auto try_knob = [&](auto member, const auto &values) {
            for (const auto v : values) {
                TuneConfig c = best;
                c.*member = v;
                if (c == best) continue;
                const double t = timed(c);
                if (t < best_ms * (1.0 - margin)) {
                    best = c;
                    best_ms = t;
                }
            }
        };
        try_knob(&TuneConfig::waves_per_eu, space.waves_per_eu);
        try_knob(&TuneConfig::split_bounce, space.split_bounce);
        try_knob(&TuneConfig::park_lanes, space.park_lanes);
        try_knob(&TuneConfig::cont_div, space.cont_div);
        try_knob(&TuneConfig::chunk_tuples, space.chunk_tuples);
        try_knob(&TuneConfig::trace_mode, space.trace_mode);
        if (best == before) break;
    }
    if (best_ms_out) *best_ms_out = best_ms;
    if (n_timed) *n_timed = (int)memo.size();
    return best;
}

// The cache key of a tuned configuration: scene class, plane-count and launch-size buckets
// (powers of two).
inline uint64_t tune_key(int scene_class, uint32_t n_planes, uint64_t tuples) {
    auto log2_ceil = [](uint64_t x) {
        uint32_t b = 0;
        while ((1ull << b) < x && b < 63) ++b;
        return (uint64_t)b;
    };
    return (uint64_t)scene_class | (log2_ceil(n_planes) << 8) | (log2_ceil(tuples) << 16);
}

}  // namespace gr
