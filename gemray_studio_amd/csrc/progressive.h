// progressive.h -- the worker's progressive-delivery rules, host C++ only (no HIP), shared
// by the job loop (serve.hip) and the host-side tests that restate the reference's own.
//
//   next_batch_size  -- apps/gemray-worker/src/stream_emit/sizing.rs:21-30: the next
//                       sub-batch's sample count, scaled toward the target duration from how
//                       long the previous one took, growing at most x4 and never 0;
//   PendingDelta     -- apps/gemray-worker/src/stream_emit/emitter.rs:46-95: the sum of the
//                       sub-batches finished since the last take, as one contiguous sample
//                       range (the first sub-batch after a take is copied, later ones added).
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace gr {

inline uint32_t next_batch_size(uint32_t prev, double elapsed_ms, uint32_t target_ms) {
    const uint64_t grow64 = std::max<uint64_t>(std::min<uint64_t>((uint64_t)prev * 4u, 0xFFFFFFFFull), 1u);
    if (!(elapsed_ms > 0.0)) return (uint32_t)grow64;
    const double scaled = (double)prev * ((double)target_ms / elapsed_ms);
    const double c = std::min(std::max(scaled, 1.0), (double)grow64);
    return (uint32_t)std::round(c);
}

class PendingDelta {
public:
    explicit PendingDelta(size_t n_floats = 0) : sum_(n_floats, 0.0f) {}
    void resize(size_t n_floats) { sum_.assign(n_floats, 0.0f); samples_ = 0; }
    // Folds a finished sub-batch of samples [first, first + samples) into the delta.
    void add(uint32_t first, uint32_t samples, const float *batch) {
        if (samples_ == 0) {
            std::memcpy(sum_.data(), batch, sum_.size() * sizeof(float));
            first_ = first;
        } else {
            for (size_t k = 0; k < sum_.size(); ++k) sum_[k] += batch[k];
        }
        samples_ += samples;
    }
    bool empty() const { return samples_ == 0; }
    // Take: false (out untouched) when nothing is pending; otherwise writes the delta and its
    // sample range and leaves the delta empty.
    bool take(uint32_t *first, uint32_t *samples, float *out) {
        if (samples_ == 0) return false;
        std::memcpy(out, sum_.data(), sum_.size() * sizeof(float));
        *first = first_;
        *samples = samples_;
        samples_ = 0;
        return true;
    }

private:
    std::vector<float> sum_;
    uint32_t first_ = 0, samples_ = 0;
};

}  // namespace gr
