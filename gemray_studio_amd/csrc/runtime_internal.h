// runtime_internal.h -- the context object and the host helpers shared by the C ABI's
// translation units (runtime.hip: plain calls; serve.hip: the batched job loop).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <map>
#include <mutex>
#include <vector>

#include "autotune.h"

#include "../../include/gemray_hip.h"
#include "transport.h"

// A device-resident environment map with its reference count (gr_env_map_upload).
struct gr_env_map {
    float *d = nullptr;
    int32_t w = 0, h = 0;
    std::atomic<int> refs{1};
};

namespace gr {

struct ServeState;

// References to device-resident environment maps: a scene or job that shows one holds a
// reference; the last release frees the texels, stream-ordered on `stream`.
void env_map_retain(const gr_env_map *m);
void env_map_release(const gr_env_map *m, hipStream_t stream);

// validate_scene (apps/gemray-worker/src/validate/mod.rs:74-192) plus the decline rules of
// frame.rs:664-700 for the branches this build does not implement.
gr_status validate_scene(const gr_scene *s);
// Host-derived scene constants (studio rig, white balance, absorption frame).
void fill_dev_scene(const gr_scene *s, DevScene *d);
// radical_inverse_base (optics/raytracer/sampling.rs:170-181), host evaluation.
float h_radical_inverse(uint32_t n, uint32_t base);
// A material hero table costs about as much to fill as the lazy fills of kHeroEntries stone
// paths (each evaluates the eight channels once); a trace (or job) of fewer tuples than this
// builds none in the default mode.
constexpr uint64_t kMatTableMinTuples = 1ull << 26;
// Whether a trace or job of `tuples` tuples of scene sc gets a material table in hero-table
// mode `mode` (gr_set_hero_tables).
inline bool wants_mat_table(const DevScene &sc, uint64_t tuples, int mode) {
    return sc.has_absorption && (mode == 2 || (mode == 1 && tuples >= kMatTableMinTuples));
}
// Stops the job loop of ctx (if started), cancels what is left and frees its memory.
void serve_shutdown(gr_ctx *ctx);

}  // namespace gr

struct gr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // The second lane of a multi-chunk trace (trace_chunks), joined back into stream at its
    // end through ev_join.  Default priority, like stream.
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr, ev_join = nullptr;
    // Timing events of the chunks in flight (trace start, trace end, finish end), and kt, the
    // end of a plain chunk's trace kernel before its continuation levels, which the next
    // chunk's trace waits for on the other lane: a call queues up to kChunkRing chunks ahead
    // of the one whose times it reads.
    static constexpr uint32_t kChunkRing = 4;
    struct ChunkEvents {
        hipEvent_t k0 = nullptr, k1 = nullptr, k2 = nullptr, kt = nullptr;
    } chunk_ev[kChunkRing];
    std::mutex mu;
    uint64_t chunk_tuples = 0;
    int trace_mode = GR_TRACE_PLAIN;
    float *xyz = nullptr;
    uint64_t xyz_cap = 0;  // tuples
    uint32_t *info = nullptr;
    uint64_t info_cap = 0;
    float *rec = nullptr;  // escape records, 20 floats x rec_cap
    uint64_t rec_cap = 0;
    float2 *ri35 = nullptr;  // radical inverses (3, 5) of the call's sample range
    uint32_t ri35_cap = 0, ri35_offset = 0, ri35_n = 0;
    std::vector<float2> ri35_host;
    unsigned long long *stats = nullptr;
    // Hero tables (transport.h): the spectral table, built at gr_open (null when its memory
    // could not be had: the kernels then evaluate per path), and the gr_set_hero_tables mode
    // (0 off, 1 on, 2 on with a material table for every absorbing scene's trace; results
    // are the same bits in every mode).
    float *spec_tab = nullptr;
    std::atomic<int> hero_tables{1};
    // Plain mode: park paths running after this many segments.  Atomic: the job loop's
    // scheduler thread reads it without the context mutex, which a render holds throughout.
    std::atomic<uint32_t> split_bounce{6};
    // Early parking: from kParkMin iterations on, a wave with fewer live lanes than this
    // parks them (0: only at split_bounce).  Read like split_bounce.
    std::atomic<uint32_t> park_lanes{16};
    float *cont = nullptr;      // continuation records, kContFields x cont_cap
    uint32_t cont_cap = 0;
    uint32_t *cont_count = nullptr;
    // This device's copy of the library's fixed scratch array (gr_render_noalloc), from
    // hipGetSymbolAddress at gr_open with the device current.
    unsigned char *scratch = nullptr;
    char *mx_buf = nullptr;     // gr_optical_metrics: planes, poses, per-ray records, results
    size_t mx_cap = 0;          // bytes
    gr_stats last{};
    // Autotuning: when on, gr_render and gr_accumulate trace with the configuration tuned
    // for the scene's class and size, tuning it on first use.
    bool autotune = false;
    std::map<uint64_t, gr::TuneConfig> tuned;
    char label[256] = {0};
    uint64_t hbm_budget = 0;  // admission budget of the job loop; 0 = 3/4 of free HBM
    std::mutex serve_mu;      // guards creation of serve
    gr::ServeState *serve = nullptr;
};
