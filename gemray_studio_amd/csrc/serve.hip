// serve.hip -- the device-resident job loop behind gr_job_* (batched and served rendering).
//
// The reference renders one request at a time: gemray-worker's render path hands each
// RenderRequest{request_id, scene, first_sample, samples} (messages/render.rs:42-70) to
// GpuFrameRenderer::accumulate, which re-uploads the scene per chunk and reads the film
// back per chunk (frame.rs:375-426, 799-814).  Here a job is admitted once: its scene,
// planes, environment map, sample table and film live in HBM until it completes, and one
// scheduler thread per context runs "ticks".  A tick gives every admitted job a piece of
// its frame (a contiguous pixel range, all of the job's samples), traces the pieces of
// each material class in one batched launch, adds each pixel's ordered sample sum into
// its job's film, and retires the jobs whose last piece it traced.  Jobs that arrive while
// a tick runs are admitted at the next one as HBM frees up (continuous batching).
//
// Determinism: a piece is whole pixels with all of their samples, so every pixel's film
// value is (0 + sum over samples in order), the same value gr_accumulate produces for the
// same (first_sample, samples); how jobs are mixed, sliced into pieces or scheduled never
// changes a bit of any job's film.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gemray_hip.h"
#include "progressive.h"
#include "runtime_internal.h"
#include "transport.h"

namespace gr {

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t kDefaultTickTuples = 1ull << 24;
constexpr uint64_t kMinTickTuples = GR_MAX_SAMPLES_PER_CALL;  // one pixel always fits
// The deadline policy's default slack, measured on the served 40-job mix: 0.25 beats the
// plain first-come-first-served path at both p50 and p99 (0.5 had the lower p50 but a p99
// 8 % above it once the kernels' service rate rose).
constexpr double kDefaultSlack = 0.25;
constexpr double kDefaultRate = 2.0e6;  // tuples per ms before any tick has completed
constexpr uint64_t kMaxTickTuples = 1ull << 28;
constexpr uint32_t kBlockTuples = 256;  // workgroup size of the batched kernels

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// One device block for the jobs admitted together: their scene blocks, planes, facet
// finishes, sample tables and certificate table, then their counters, certificates and
// films.  An admission round makes one allocation, one upload (from pinned staging), one
// clear and one certificate launch for all of its jobs instead of a dozen calls per job; the
// block is freed, stream-ordered, when the last of its jobs retires.
struct Arena {
    void *d = nullptr;
    hipStream_t stream = nullptr;
    ~Arena() {
        if (d) (void)hipFreeAsync(d, stream);
    }
};

// A job whose outputs exceed this gets a block of its own, so that a long job does not
// hold the memory of the short jobs admitted with it.
constexpr uint64_t kOwnArenaBytes = 32ull << 20;
// Films up to this size are read back through pinned staging.
constexpr size_t kStagedFilmBytes = 2ull << 20;

struct Job {
    uint64_t id = 0;
    int state = GR_JOB_QUEUED;
    gr_status error = GR_OK;
    // Host copies of the request (inputs are borrowed only during gr_job_submit).
    std::vector<float> planes;
    std::vector<uint32_t> finishes;  // only when some facet is frosted
    std::vector<float> map;              // a map the request passed as host texels, or
    const gr_env_map *env_map = nullptr;  // the device-resident map it references
    DevScene host_sc{};
    int cls = 0;
    uint32_t width = 0, height = 0;
    // The requested range, and the sub-batch being traced: pieces cover samples
    // [sample_offset, sample_offset + spp) of their pixels.  A one-batch job has
    // sample_offset = first_sample and spp = total_spp.
    uint32_t first_sample = 0, total_spp = 0, sample_offset = 0, spp = 0;
    // Progressive delivery: sub-batch sizing, samples folded so far, and the host-side
    // pending delta and running total (the worker's PendingDelta and running_total).
    bool progressive = false, fixed_batches = false;
    uint32_t subbatch_ms = 0, produced = 0, subbatches = 0;
    Clock::time_point t_batch;
    PendingDelta pending;
    bool delta_taken = false;
    uint64_t n_pixels = 0;
    uint64_t footprint = 0;
    // Device copies (from admission to completion): in the admission round's block, except
    // a map the job brought as host texels.
    std::shared_ptr<Arena> arena;
    DevScene *d_scene = nullptr;
    float4 *d_planes = nullptr;
    uint32_t *d_finishes = nullptr;
    float *d_map = nullptr;
    // The material hero table of a large absorbing job (its own allocation, like a map), or
    // null; decided at submission (mat_table) and counted in the footprint.
    bool mat_table = false;
    float *d_mat = nullptr;
    float2 *d_ri35 = nullptr;
    uint32_t *d_cert = nullptr;
    float *d_film = nullptr;
    unsigned long long *d_stats = nullptr;
    // Progress (owned by the scheduler thread).
    uint64_t next_pixel = 0;
    uint32_t inflight = 0;  // pieces in ticks that have not completed
    // Results.
    std::vector<float> film;
    unsigned long long events[GR_EV_COUNT] = {0};
    uint64_t paths_done = 0;
    Clock::time_point t_submit, t_admit, t_done;
    // Deadline policy: submit time (ms from the loop's start) + slack x estimated service.
    double deadline_ms = 0.0;
    // Set by gr_job_cancel (under the loop's mutex), read by the scheduler thread also
    // outside it.
    std::atomic<bool> cancel_req{false};
};

struct Piece {
    Job *job;
    uint64_t pixel_begin, n_pixels;
    uint32_t sample_offset, spp;  // the job's sub-batch when the piece was planned
};

constexpr int kTickSlots = 2;

struct TickSlot {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float *staging = nullptr;
    uint64_t staging_cap = 0;
    BatchSlice *d_slices = nullptr;
    uint32_t slices_cap = 0;
    // Continuation records of the path-length split: per slice two halves of
    // kContFields x ceil(tuples / 4) floats, and kContLevels + 1 counters.
    float *cont = nullptr;
    uint64_t cont_cap = 0;  // floats
    uint32_t *cont_counts = nullptr;
    uint32_t counts_cap = 0;  // slices
    std::vector<BatchSlice> h_slices;  // kept until the tick completes
    std::vector<Piece> pieces;
    uint32_t launches = 0;
    uint64_t tuples = 0;
    Clock::time_point t_launch;
    bool busy = false;
};

}  // namespace

struct ServeState {
    gr_ctx *ctx = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<std::shared_ptr<Job>> queue;    // submitted, not admitted
    std::vector<std::shared_ptr<Job>> active;  // admitted, not finished
    std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
    uint64_t budget = 0, in_use = 0;
    uint64_t tick_tuples = kDefaultTickTuples;
    int policy = GR_SERVE_DEADLINE;
    double slack = kDefaultSlack;
    // Service-rate estimate for deadlines: tuples of completed ticks over the wall time the
    // loop spent with ticks in flight (kDefaultRate until the first tick completes).
    Clock::time_point t_start = Clock::now(), t_last_complete;
    uint64_t rate_tuples = 0;
    double rate_ms = 0.0;
    // Device resources of the loop: a control stream for uploads and read-backs, and two
    // tick slots with their own streams and buffers so that one tick is queued while the
    // other runs.
    hipStream_t stream = nullptr;
    TickSlot slot[kTickSlots];
    // Pinned host staging of the control stream's uploads and read-backs (reused once the
    // stream has been synchronized, which the loop does after every admission and read-back
    // round).
    unsigned char *h_stage = nullptr, *h_back = nullptr;
    size_t stage_cap = 0, back_cap = 0;
    std::vector<unsigned char *> retired_pinned;
    uint64_t ticks_launched = 0, ticks_done = 0;
    size_t rr = 0;  // round-robin start of the next tick
    // Metrics since the last reset.
    gr_serve_metrics m{};
    std::vector<double> latencies;
    bool have_first = false;
    Clock::time_point t_first, t_last;
};

namespace {

#define SV_CHECK(call)                                                                           \
    do {                                                                                         \
        hipError_t err_ = (call);                                                                \
        if (err_ != hipSuccess) {                                                                \
            std::fprintf(stderr, "gemray_hip serve: %s failed: %s\n", #call, hipGetErrorString(err_)); \
            return GR_ERR_DEVICE;                                                                \
        }                                                                                        \
    } while (0)

uint64_t job_footprint(const Job &j) {
    uint64_t b = j.n_pixels * 3 * sizeof(float);             // film
    b += j.planes.size() * sizeof(float);                    // planes
    b += j.finishes.size() * sizeof(uint32_t);               // facet finishes
    b += j.map.size() * sizeof(float);                       // environment map (a shared one is the map's own)
    b += (uint64_t)j.total_spp * sizeof(float2);             // sample table
    b += j.n_pixels * sizeof(uint32_t);                      // miss certificates
    b += sizeof(DevScene) + GR_EV_COUNT * sizeof(unsigned long long);
    if (j.mat_table) b += (uint64_t)kHeroEntries * mat_floats(j.host_sc.biaxial != 0) * sizeof(float);
    return b;
}

// Device memory of the loop is stream-ordered (hipMallocAsync / hipFreeAsync on the control
// stream): a job's buffers are freed only after its last tick completed and its film was read
// back, and a stream-ordered free does not wait for the device, so retiring a job never
// stalls the tick still running on the other slot (hipFree would synchronize the device).
void drop_env_map(ServeState &S, Job &j) {
    if (j.env_map) env_map_release(j.env_map, S.stream);
    j.env_map = nullptr;
}

void free_job_device(ServeState &S, Job &j) {
    if (j.d_map) (void)hipFreeAsync(j.d_map, S.stream);
    if (j.d_mat) (void)hipFreeAsync(j.d_mat, S.stream);
    j.d_mat = nullptr;
    j.arena.reset();
    drop_env_map(S, j);
    j.d_scene = nullptr;
    j.d_planes = nullptr;
    j.d_finishes = nullptr;
    j.d_map = nullptr;
    j.d_ri35 = nullptr;
    j.d_cert = nullptr;
    j.d_film = nullptr;
    j.d_stats = nullptr;
}

// Grows a pinned host buffer to at least `want` bytes, geometrically from `min_bytes`.
// Freeing pinned memory waits for the device, so an outgrown buffer is kept in `retired`
// until the loop's resources are released (with doubling, at most the final size more).
// Called only while no copy from or into the buffer is outstanding.
gr_status reserve_pinned(unsigned char *&buf, size_t &cap, size_t want, size_t min_bytes,
                         std::vector<unsigned char *> &retired) {
    if (want <= cap) return GR_OK;
    const size_t c = std::max({want, 2 * cap, min_bytes});
    if (buf) retired.push_back(buf);
    buf = nullptr;
    cap = 0;
    SV_CHECK(hipHostMalloc((void **)&buf, c, hipHostMallocDefault));
    cap = c;
    return GR_OK;
}

constexpr size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

// Output bytes of a job in its block: counters, certificates, film.
size_t job_out_bytes(const Job &j) {
    return align_up(GR_EV_COUNT * sizeof(unsigned long long)) + align_up(j.n_pixels * sizeof(uint32_t)) +
           align_up(j.n_pixels * 3 * sizeof(float));
}
// Input bytes of a job in its block: scene block, planes, finishes, sample table.
size_t job_in_bytes(const Job &j) {
    return align_up(sizeof(DevScene)) + align_up(j.planes.size() * sizeof(float)) +
           align_up(j.finishes.size() * sizeof(uint32_t)) + align_up((size_t)j.total_spp * sizeof(float2));
}

// Queues the admission of `jobs` on the control stream: the jobs' blocks (the small jobs
// share one, each large one has its own), the upload of their inputs from pinned staging,
// the clear of their counters and films and one certificate launch per block; a map brought
// as host texels is allocated and copied per job.  st[i] receives job i's status.  The
// caller synchronizes the control stream once for the whole round.
void admit_round(ServeState &S, const std::vector<std::shared_ptr<Job>> &jobs, std::vector<gr_status> &st) {
    std::vector<std::vector<size_t>> groups(1);
    for (size_t i = 0; i < jobs.size(); ++i) {
        if (job_out_bytes(*jobs[i]) > kOwnArenaBytes) groups.push_back({i});
        else groups[0].push_back(i);
    }
    // Pinned staging for every block's inputs and certificate table.
    std::vector<size_t> in_bytes(groups.size(), 0), stage_off(groups.size(), 0);
    size_t stage_total = 0;
    for (size_t g = 0; g < groups.size(); ++g) {
        for (size_t i : groups[g]) in_bytes[g] += job_in_bytes(*jobs[i]);
        in_bytes[g] += align_up(groups[g].size() * sizeof(CertSlice));
        stage_off[g] = stage_total;
        stage_total += in_bytes[g];
    }
    const gr_status rs = reserve_pinned(S.h_stage, S.stage_cap, stage_total, (size_t)4 << 20, S.retired_pinned);
    for (size_t g = 0; g < groups.size(); ++g) {
        if (groups[g].empty()) continue;
        auto fail = [&](gr_status e) {
            for (size_t i : groups[g]) st[i] = e;
        };
        if (rs != GR_OK) {
            fail(rs);
            continue;
        }
        size_t out_bytes = 0;
        for (size_t i : groups[g]) out_bytes += job_out_bytes(*jobs[i]);
        auto arena = std::make_shared<Arena>();
        arena->stream = S.stream;
        if (hipMallocAsync(&arena->d, in_bytes[g] + out_bytes, S.stream) != hipSuccess) {
            arena->d = nullptr;
            fail(GR_ERR_DEVICE);
            continue;
        }
        unsigned char *dev = (unsigned char *)arena->d;
        unsigned char *host = S.h_stage + stage_off[g];
        size_t in_off = 0, out_off = in_bytes[g];
        std::vector<CertSlice> certs;
        uint32_t blocks = 0;
        gr_status gst = GR_OK;
        for (size_t i : groups[g]) {
            Job &j = *jobs[i];
            j.arena = arena;
            auto take_in = [&](size_t bytes) {
                const size_t o = in_off;
                in_off += align_up(bytes);
                return o;
            };
            auto take_out = [&](size_t bytes) {
                const size_t o = out_off;
                out_off += align_up(bytes);
                return o;
            };
            const size_t o_scene = take_in(sizeof(DevScene));
            const size_t o_planes = take_in(j.planes.size() * sizeof(float));
            const size_t o_fin = take_in(j.finishes.size() * sizeof(uint32_t));
            const size_t o_ri = take_in((size_t)j.total_spp * sizeof(float2));
            j.d_scene = (DevScene *)(dev + o_scene);
            j.d_planes = (float4 *)(dev + o_planes);
            j.d_finishes = j.finishes.empty() ? nullptr : (uint32_t *)(dev + o_fin);
            j.d_ri35 = (float2 *)(dev + o_ri);
            j.d_stats = (unsigned long long *)(dev + take_out(GR_EV_COUNT * sizeof(unsigned long long)));
            j.d_cert = (uint32_t *)(dev + take_out(j.n_pixels * sizeof(uint32_t)));
            j.d_film = (float *)(dev + take_out(j.n_pixels * 3 * sizeof(float)));
            if (!j.map.empty()) {
                if (hipMallocAsync((void **)&j.d_map, j.map.size() * sizeof(float), S.stream) != hipSuccess ||
                    hipMemcpyAsync(j.d_map, j.map.data(), j.map.size() * sizeof(float), hipMemcpyHostToDevice,
                                   S.stream) != hipSuccess)
                    gst = GR_ERR_DEVICE;
            }
            DevScene sc = j.host_sc;
            sc.map_rgb = j.env_map ? j.env_map->d : j.d_map;
            sc.finishes = j.d_finishes;
            sc.spec_tab = S.ctx->hero_tables.load() != 0 ? S.ctx->spec_tab : nullptr;
            if (j.mat_table) {
                // Filled on the control stream ahead of the job's first tick; without the
                // memory the job evaluates per path (same bits).
                const size_t bytes = (size_t)kHeroEntries * mat_floats(sc.biaxial != 0) * sizeof(float);
                if (hipMallocAsync((void **)&j.d_mat, bytes, S.stream) != hipSuccess) {
                    (void)hipGetLastError();
                    j.d_mat = nullptr;
                } else if (gr_launch_mat_table(sc, j.d_mat, S.stream) != hipSuccess) {
                    gst = GR_ERR_DEVICE;
                }
                sc.mat_tab = j.d_mat;
            }
            std::memcpy(host + o_scene, &sc, sizeof(DevScene));
            std::memcpy(host + o_planes, j.planes.data(), j.planes.size() * sizeof(float));
            if (!j.finishes.empty()) std::memcpy(host + o_fin, j.finishes.data(), j.finishes.size() * sizeof(uint32_t));
            float2 *ri = (float2 *)(host + o_ri);
            for (uint32_t k = 0; k < j.total_spp; ++k)
                ri[k] = make_float2(h_radical_inverse(j.first_sample + k, 3), h_radical_inverse(j.first_sample + k, 5));
            certs.push_back(CertSlice{j.d_scene, j.d_planes, j.d_cert, (uint32_t)j.n_pixels, blocks});
            blocks += (uint32_t)((j.n_pixels + kBlockTuples - 1) / kBlockTuples);
        }
        const size_t o_table = in_off;
        std::memcpy(host + o_table, certs.data(), certs.size() * sizeof(CertSlice));
        if (gst == GR_OK &&
            (hipMemcpyAsync(dev, host, in_bytes[g], hipMemcpyHostToDevice, S.stream) != hipSuccess ||
             hipMemsetAsync(dev + in_bytes[g], 0, out_bytes, S.stream) != hipSuccess ||
             gr_launch_cert_batch((const CertSlice *)(dev + o_table), (uint32_t)certs.size(), blocks, S.stream) !=
                 hipSuccess))
            gst = GR_ERR_DEVICE;
        if (gst != GR_OK) fail(gst);
    }
}

// After the control stream has run a job's admission: its host copy of the environment
// map is no longer needed.
void uploaded(Job &j) { std::vector<float>().swap(j.map); }

gr_status ensure_staging(TickSlot &T, uint64_t tuples, uint32_t slices) {
    if (tuples > T.staging_cap) {
        if (T.staging) SV_CHECK(hipFreeAsync(T.staging, T.stream));
        T.staging = nullptr;
        T.staging_cap = 0;
        SV_CHECK(hipMallocAsync((void **)&T.staging, tuples * 3 * sizeof(float), T.stream));
        T.staging_cap = tuples;
    }
    if (slices > T.slices_cap) {
        if (T.d_slices) SV_CHECK(hipFreeAsync(T.d_slices, T.stream));
        T.d_slices = nullptr;
        T.slices_cap = 0;
        uint32_t cap = std::max<uint32_t>(slices, 64);
        SV_CHECK(hipMallocAsync((void **)&T.d_slices, cap * sizeof(BatchSlice), T.stream));
        T.slices_cap = cap;
    }
    return GR_OK;
}

gr_status ensure_cont(TickSlot &T, uint64_t floats, uint32_t slices) {
    if (floats > T.cont_cap) {
        if (T.cont) SV_CHECK(hipFreeAsync(T.cont, T.stream));
        T.cont = nullptr;
        T.cont_cap = 0;
        SV_CHECK(hipMallocAsync((void **)&T.cont, floats * sizeof(float), T.stream));
        T.cont_cap = floats;
    }
    if (slices > T.counts_cap) {
        if (T.cont_counts) SV_CHECK(hipFreeAsync(T.cont_counts, T.stream));
        T.cont_counts = nullptr;
        T.counts_cap = 0;
        uint32_t cap = std::max<uint32_t>(slices, 64);
        SV_CHECK(hipMallocAsync((void **)&T.cont_counts, (size_t)cap * (kContLevels + 1) * sizeof(uint32_t), T.stream));
        T.counts_cap = cap;
    }
    return GR_OK;
}

// Pieces of one tick.  A job has at most one piece per tick, so no two reduce threads of a
// tick touch the same film pixel.
//  * GR_SERVE_EQUAL_SHARE: every active job gets an equal share of the tick's tuples (at
//    least one pixel), then the unused budget extends the pieces in round-robin order.
//  * GR_SERVE_DEADLINE: the tick's tuples go to jobs in order of deadline (submit time +
//    slack x estimated service time): slack 0 is first-come-first-served, which bounds the
//    worst latency best; a larger slack lets short jobs overtake long ones.  The first job
//    always gets at least one pixel.
std::vector<Piece> plan_tick(ServeState &S, std::vector<Job *> jobs, uint64_t budget, int policy) {
    std::vector<Piece> pieces;
    const size_t n = jobs.size();
    if (n == 0) return pieces;
    if (policy == GR_SERVE_DEADLINE) {
        std::stable_sort(jobs.begin(), jobs.end(),
                         [](const Job *a, const Job *b) { return a->deadline_ms < b->deadline_ms; });
        uint64_t used = 0;
        for (Job *j : jobs) {
            uint64_t px = std::min<uint64_t>(j->n_pixels - j->next_pixel, (budget - used) / j->spp);
            if (pieces.empty()) px = std::max<uint64_t>(px, 1);
            if (px == 0) break;
            pieces.push_back(Piece{j, j->next_pixel, px, j->sample_offset, j->spp});
            j->next_pixel += px;
            used += px * j->spp;
            if (used >= budget) break;
        }
        return pieces;
    }
    const uint64_t share = std::max<uint64_t>(budget / n, 1);
    uint64_t used = 0;
    std::vector<uint64_t> take(n, 0);
    for (size_t k = 0; k < n; ++k) {
        const size_t i = (S.rr + k) % n;
        Job &j = *jobs[i];
        uint64_t left = j.n_pixels - j.next_pixel;
        uint64_t px = std::min<uint64_t>(left, std::max<uint64_t>(share / j.spp, 1));
        px = std::min<uint64_t>(px, (budget - used) / j.spp);
        take[i] = px;
        used += px * j.spp;
    }
    for (size_t k = 0; k < n && used < budget; ++k) {
        const size_t i = (S.rr + k) % n;
        Job &j = *jobs[i];
        uint64_t left = j.n_pixels - j.next_pixel - take[i];
        uint64_t more = std::min<uint64_t>(left, (budget - used) / j.spp);
        take[i] += more;
        used += more * j.spp;
    }
    S.rr = (S.rr + 1) % n;
    for (size_t i = 0; i < n; ++i) {
        if (take[i] == 0) continue;
        pieces.push_back(Piece{jobs[i], jobs[i]->next_pixel, take[i], jobs[i]->sample_offset, jobs[i]->spp});
        jobs[i]->next_pixel += take[i];
    }
    return pieces;
}

// Queues one tick's pieces on slot T: the batched trace launches (one per material class,
// each followed by its continuation launches when split_bounce > 0) and the film reduce,
// bracketed by the slot's events.  Returns without waiting.
gr_status launch_tick(TickSlot &T, std::vector<Piece> &&pieces, uint32_t split_bounce, uint32_t park_lanes) {
    std::stable_sort(pieces.begin(), pieces.end(),
                     [](const Piece &a, const Piece &b) { return a.job->cls < b.job->cls; });
    uint64_t tuples = 0, cont_floats = 0;
    for (const Piece &p : pieces) {
        const uint64_t n = p.n_pixels * p.spp;
        tuples += n;
        cont_floats += 2 * ((n + 3) / 4) * kContFields;
    }
    gr_status st = ensure_staging(T, std::max<uint64_t>(tuples, 1), (uint32_t)pieces.size());
    if (st != GR_OK) return st;
    if (split_bounce) {
        st = ensure_cont(T, cont_floats, (uint32_t)pieces.size());
        if (st != GR_OK) return st;
    }
    std::vector<BatchSlice> &sl = T.h_slices;
    sl.assign(pieces.size(), BatchSlice{});
    uint64_t off = 0, cont_off = 0;
    uint32_t reduce_blocks = 0;
    for (size_t i = 0; i < pieces.size(); ++i) {
        const Piece &p = pieces[i];
        Job &j = *p.job;
        BatchSlice &b = sl[i];
        b.scene = j.d_scene;
        b.planes = j.d_planes;
        b.ri35 = j.d_ri35 + (p.sample_offset - j.first_sample);
        b.cert = j.d_cert;
        b.xyz = T.staging + 3 * off;
        b.film = j.d_film;
        b.stats = j.d_stats;
        b.pixel_begin = (uint32_t)p.pixel_begin;
        b.n_pixels = (uint32_t)p.n_pixels;
        b.sample_offset = p.sample_offset;
        b.spp = p.spp;
        b.reduce_block_begin = reduce_blocks;
        reduce_blocks += (uint32_t)((p.n_pixels + kBlockTuples - 1) / kBlockTuples);
        if (split_bounce) {
            const uint64_t cap = (p.n_pixels * p.spp + 3) / 4;
            b.split_bounce = split_bounce;
            b.park_lanes = park_lanes;
            b.cont_cap = (uint32_t)cap;
            b.cont[0] = T.cont + cont_off;
            b.cont[1] = T.cont + cont_off + cap * kContFields;
            b.cont_count = T.cont_counts + i * (kContLevels + 1);
            cont_off += 2 * cap * kContFields;
        }
        off += p.n_pixels * p.spp;
    }
    // Class launches: slices of one class are contiguous; block numbering restarts per class.
    struct ClassRange {
        int cls;
        uint32_t first, count, blocks, cont_blocks, max_planes;
    };
    std::vector<ClassRange> classes;
    for (size_t i = 0; i < sl.size(); ++i) {
        const int c = pieces[i].job->cls;
        if (classes.empty() || classes.back().cls != c) classes.push_back(ClassRange{c, (uint32_t)i, 0, 0, 0, 0});
        ClassRange &r = classes.back();
        r.max_planes = std::max(r.max_planes, pieces[i].job->host_sc.n_planes);
        sl[i].trace_block_begin = r.blocks;
        r.blocks += (uint32_t)(((uint64_t)sl[i].n_pixels * sl[i].spp + kBlockTuples - 1) / kBlockTuples);
        sl[i].cont_block_begin = r.cont_blocks;
        r.cont_blocks += (sl[i].cont_cap + kBlockTuples - 1) / kBlockTuples;
        r.count++;
    }
    T.launches = 0;
    SV_CHECK(hipMemcpyAsync(T.d_slices, sl.data(), sl.size() * sizeof(BatchSlice), hipMemcpyHostToDevice, T.stream));
    SV_CHECK(hipEventRecord(T.ev0, T.stream));
    if (split_bounce)
        SV_CHECK(hipMemsetAsync(T.cont_counts, 0, sl.size() * (kContLevels + 1) * sizeof(uint32_t), T.stream));
    for (const ClassRange &r : classes) {
        SV_CHECK(gr_launch_trace_batch(T.d_slices + r.first, r.count, r.blocks, split_bounce ? r.cont_blocks : 0u,
                                       r.max_planes, r.cls, T.stream));
        T.launches += split_bounce ? 1 + kContLevels : 1;
    }
    SV_CHECK(gr_launch_film_reduce_batch(T.d_slices, (uint32_t)sl.size(), reduce_blocks, T.stream));
    ++T.launches;
    SV_CHECK(hipEventRecord(T.ev1, T.stream));
    for (const Piece &p : pieces) p.job->inflight++;
    T.pieces = std::move(pieces);
    T.tuples = tuples;
    T.t_launch = Clock::now();
    T.busy = true;
    return GR_OK;
}

void percentiles(std::vector<double> v, gr_serve_metrics *m) {
    if (v.empty()) return;
    std::sort(v.begin(), v.end());
    auto at = [&](double q) {
        // Nearest-rank percentile.
        size_t r = (size_t)std::ceil(q * (double)v.size());
        if (r < 1) r = 1;
        return v[std::min(r, v.size()) - 1];
    };
    m->latency_p50_ms = at(0.50);
    m->latency_p99_ms = at(0.99);
    m->latency_max_ms = v.back();
    double s = 0.0;
    for (double x : v) s += x;
    m->latency_mean_ms = s / (double)v.size();
}

void retire(ServeState &S, Job &j, int state, gr_status err) {
    free_job_device(S, j);
    S.in_use -= j.footprint;
    j.state = state;
    j.error = err;
    j.t_done = Clock::now();
    if (state == GR_JOB_DONE) {
        double lat = ms_between(j.t_submit, j.t_done);
        S.latencies.push_back(lat);
        S.m.completed++;
        S.m.paths += j.paths_done;
        S.m.segments += j.events[GR_EV_SEGMENT];
        S.t_last = j.t_done;
    } else if (state == GR_JOB_CANCELLED) {
        S.m.cancelled++;
    } else {
        S.m.failed++;
    }
}

// Waits for the tick on slot T and books its pieces; the jobs it completed (all pieces
// traced, none cancelled) are appended to `finished` for finish_jobs.
gr_status complete_tick(ServeState &S, TickSlot &T, std::vector<Job *> &finished) {
    gr_status st = GR_OK;
    float gpu_ms = 0.0f;
    if (hipEventSynchronize(T.ev1) != hipSuccess || hipEventElapsedTime(&gpu_ms, T.ev0, T.ev1) != hipSuccess)
        st = GR_ERR_DEVICE;
    for (const Piece &p : T.pieces) {
        p.job->inflight--;
        if (st == GR_OK && p.job->inflight == 0 && p.job->next_pixel == p.job->n_pixels && !p.job->cancel_req)
            finished.push_back(p.job);
    }
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> g(S.mu);
    for (const Piece &p : T.pieces) p.job->paths_done += p.n_pixels * p.spp;
    if (st == GR_OK) {
        S.rate_tuples += T.tuples;
        S.rate_ms += ms_between(std::max(T.t_launch, S.t_last_complete), now);
    }
    S.t_last_complete = now;
    S.m.ticks++;
    S.m.launches += T.launches;
    S.m.busy_ms += gpu_ms;
    S.ticks_done++;
    T.busy = false;
    T.pieces.clear();
    return st;
}

// Reads back the films and counters of completed jobs and retires them.  The loop calls
// this after queuing the next tick, so the read-backs overlap the GPU's work instead of
// leaving it idle; until retire() marks a job done no API call reads its film.
void finish_jobs(ServeState &S, std::vector<Job *> &finished) {
    if (finished.empty()) return;
    // Read-back round, with one synchronization: every finished (sub-)batch film -- a small
    // one through pinned staging (a pageable copy costs more per call than its bytes), a
    // large one straight into the job's host film -- and after a job's last batch its
    // counters; before a further sub-batch the device film is cleared.
    std::vector<gr_status> fin(finished.size(), GR_OK);
    std::vector<std::vector<float>> batch(finished.size());
    std::vector<size_t> off(finished.size());
    std::vector<bool> last(finished.size()), staged(finished.size());
    size_t total = 0;
    for (size_t i = 0; i < finished.size(); ++i) {
        const size_t film_bytes = finished[i]->n_pixels * 3 * sizeof(float);
        staged[i] = film_bytes <= kStagedFilmBytes;
        off[i] = total;
        total += (staged[i] ? align_up(film_bytes) : 0) + align_up(sizeof(finished[i]->events));
    }
    const gr_status rs = reserve_pinned(S.h_back, S.back_cap, total, (size_t)64 << 20, S.retired_pinned);
    for (size_t i = 0; i < finished.size(); ++i) {
        Job &j = *finished[i];
        last[i] = j.produced + j.spp >= j.total_spp;
        if (rs != GR_OK) {
            fin[i] = rs;
            continue;
        }
        const size_t film_bytes = j.n_pixels * 3 * sizeof(float);
        std::vector<float> &dst = j.progressive ? batch[i] : j.film;
        if (!staged[i]) dst.resize(j.n_pixels * 3);
        unsigned char *h = S.h_back + off[i];
        void *film_dst = staged[i] ? (void *)h : (void *)dst.data();
        unsigned char *ev_dst = h + (staged[i] ? align_up(film_bytes) : 0);
        if (hipMemcpyAsync(film_dst, j.d_film, film_bytes, hipMemcpyDeviceToHost, S.stream) != hipSuccess ||
            (last[i] && hipMemcpyAsync(ev_dst, j.d_stats, sizeof(j.events), hipMemcpyDeviceToHost, S.stream) !=
                            hipSuccess) ||
            (!last[i] && hipMemsetAsync(j.d_film, 0, film_bytes, S.stream) != hipSuccess))
            fin[i] = GR_ERR_DEVICE;
    }
    if (hipStreamSynchronize(S.stream) != hipSuccess)
        for (gr_status &f : fin) f = GR_ERR_DEVICE;
    for (size_t i = 0; i < finished.size(); ++i) {
        if (fin[i] != GR_OK) continue;
        Job &j = *finished[i];
        const size_t film_bytes = j.n_pixels * 3 * sizeof(float);
        if (staged[i]) {
            std::vector<float> &dst = j.progressive ? batch[i] : j.film;
            dst.resize(j.n_pixels * 3);
            std::memcpy(dst.data(), S.h_back + off[i], film_bytes);
        }
        if (last[i])
            std::memcpy(j.events, S.h_back + off[i] + (staged[i] ? align_up(film_bytes) : 0), sizeof(j.events));
    }
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> g(S.mu);
    for (size_t i = 0; i < finished.size(); ++i) {
        Job &j = *finished[i];
        if (fin[i] != GR_OK) {
            retire(S, j, GR_JOB_FAILED, fin[i]);
            continue;
        }
        j.subbatches++;
        if (j.progressive) {
            // PendingDelta::add and the running total (tracer.rs:93-103): the first
            // sub-batch since a take is copied, later ones are added in sample order.
            const std::vector<float> &b = batch[i];
            j.pending.add(j.sample_offset, j.spp, b.data());
            for (size_t k = 0; k < b.size(); ++k) j.film[k] += b[k];
        }
        j.produced += j.spp;
        if (j.produced >= j.total_spp) {
            retire(S, j, GR_JOB_DONE, GR_OK);
            continue;
        }
        // The next sub-batch: sized from how long this one took, over the whole frame.
        const uint32_t want = j.fixed_batches ? j.spp : next_batch_size(j.spp, ms_between(j.t_batch, now), j.subbatch_ms);
        j.sample_offset += j.spp;
        j.spp = std::min(want, j.total_spp - j.produced);
        j.next_pixel = 0;
        j.t_batch = now;
    }
    S.active.erase(std::remove_if(S.active.begin(), S.active.end(),
                                  [](const std::shared_ptr<Job> &j) { return j->state != GR_JOB_RUNNING; }),
                   S.active.end());
    S.cv_done.notify_all();
    finished.clear();
}

void loop(ServeState *Sp) {
    ServeState &S = *Sp;
    (void)hipSetDevice(S.ctx->device);
    for (;;) {
        const bool in_flight = S.slot[0].busy || S.slot[1].busy;
        std::vector<std::shared_ptr<Job>> admit;
        {
            std::unique_lock<std::mutex> lk(S.mu);
            if (!in_flight)
                S.cv_work.wait(lk, [&] { return S.stop || !S.queue.empty() || !S.active.empty(); });
            if (S.stop) return;
            // Cancelled jobs leave: queued ones at once, running ones once none of their
            // pieces is in flight.
            for (auto it = S.queue.begin(); it != S.queue.end();) {
                if ((*it)->cancel_req) {
                    drop_env_map(S, **it);
                    (*it)->state = GR_JOB_CANCELLED;
                    (*it)->t_done = Clock::now();
                    S.m.cancelled++;
                    it = S.queue.erase(it);
                } else {
                    ++it;
                }
            }
            for (auto it = S.active.begin(); it != S.active.end();) {
                if ((*it)->cancel_req && (*it)->inflight == 0) {
                    retire(S, **it, GR_JOB_CANCELLED, GR_OK);
                    it = S.active.erase(it);
                } else {
                    ++it;
                }
            }
            // Admission by HBM footprint, in submission order.
            while (!S.queue.empty() && S.in_use + S.queue.front()->footprint <= S.budget) {
                S.in_use += S.queue.front()->footprint;
                admit.push_back(S.queue.front());
                S.queue.pop_front();
            }
            S.cv_done.notify_all();
        }
        std::vector<gr_status> ust(admit.size(), GR_OK);
        if (!admit.empty()) admit_round(S, admit, ust);
        if (!admit.empty() && hipStreamSynchronize(S.stream) != hipSuccess)
            for (gr_status &u : ust) u = GR_ERR_DEVICE;
        for (size_t i = 0; i < admit.size(); ++i) {
            auto &j = admit[i];
            uploaded(*j);
            const gr_status st = ust[i];
            std::lock_guard<std::mutex> g(S.mu);
            if (st != GR_OK) {
                retire(S, *j, GR_JOB_FAILED, st);
                S.cv_done.notify_all();
                continue;
            }
            j->state = GR_JOB_RUNNING;
            j->t_admit = Clock::now();
            j->t_batch = j->t_admit;
            if (!S.have_first) {
                S.have_first = true;
                S.t_first = j->t_admit;
            }
            S.m.admitted++;
            S.active.push_back(j);
        }
        // Plan the next tick over the jobs that still have pixels to hand out.
        std::vector<Job *> schedulable;
        uint64_t tick_tuples = 0;
        uint32_t split_bounce = 0, park_lanes = 0;
        int policy = GR_SERVE_DEADLINE;
        {
            std::lock_guard<std::mutex> g(S.mu);
            for (auto &j : S.active)
                if (!j->cancel_req && j->next_pixel < j->n_pixels) schedulable.push_back(j.get());
            tick_tuples = S.tick_tuples;
            split_bounce = S.ctx->split_bounce.load(std::memory_order_relaxed);
            park_lanes = S.ctx->park_lanes.load(std::memory_order_relaxed);
            policy = S.policy;
        }
        TickSlot &next = S.slot[S.ticks_launched % kTickSlots];
        gr_status st = GR_OK;
        std::vector<Job *> finished;
        if (!schedulable.empty()) {
            // The slot's previous tick must be complete before its buffers are reused.  Jobs
            // that tick completed hold no pixels still to hand out, so the plan skips them.
            if (next.busy) st = complete_tick(S, next, finished);
            if (st == GR_OK) {
                st = launch_tick(next, plan_tick(S, schedulable, tick_tuples, policy), split_bounce, park_lanes);
                if (st == GR_OK) S.ticks_launched++;
            }
        } else {
            // Nothing new to launch: complete the oldest tick in flight.
            TickSlot &oldest = S.slot[S.ticks_done % kTickSlots];
            if (oldest.busy) st = complete_tick(S, oldest, finished);
        }
        if (st == GR_OK) finish_jobs(S, finished);
        if (st != GR_OK) {
            // A device error fails every running job.
            for (TickSlot &T : S.slot) {
                if (T.busy) (void)hipStreamSynchronize(T.stream);
                T.busy = false;
                T.pieces.clear();
            }
            std::lock_guard<std::mutex> g(S.mu);
            for (auto &j : S.active) {
                j->inflight = 0;
                retire(S, *j, GR_JOB_FAILED, st);
            }
            S.active.clear();
            S.cv_done.notify_all();
        }
    }
}

// Streams, events and buffers of the loop (all optional: partially created state too).
void release_device(ServeState &S) {
    for (TickSlot &T : S.slot) {
        if (T.stream) (void)hipStreamSynchronize(T.stream);
        for (void *p : {(void *)T.staging, (void *)T.d_slices, (void *)T.cont, (void *)T.cont_counts})
            if (p) (void)hipFreeAsync(p, T.stream);
        if (T.stream) (void)hipStreamSynchronize(T.stream);
        if (T.ev0) (void)hipEventDestroy(T.ev0);
        if (T.ev1) (void)hipEventDestroy(T.ev1);
        if (T.stream) (void)hipStreamDestroy(T.stream);
        T = TickSlot{};
    }
    if (S.stream) (void)hipStreamDestroy(S.stream);
    S.stream = nullptr;
    S.retired_pinned.push_back(S.h_stage);
    S.retired_pinned.push_back(S.h_back);
    for (unsigned char *p : S.retired_pinned)
        if (p) (void)hipHostFree(p);
    S.retired_pinned.clear();
    S.h_stage = S.h_back = nullptr;
    S.stage_cap = S.back_cap = 0;
}

// Creates the loop on first use.
gr_status get_serve(gr_ctx *ctx, ServeState **out) {
    std::lock_guard<std::mutex> g(ctx->serve_mu);
    if (!ctx->serve) {
        if (hipSetDevice(ctx->device) != hipSuccess) return GR_ERR_DEVICE;
        auto *S = new (std::nothrow) ServeState();
        if (!S) return GR_ERR_DEVICE;
        S->ctx = ctx;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            delete S;
            return GR_ERR_DEVICE;
        }
        S->budget = ctx->hbm_budget ? ctx->hbm_budget : free_b / 4 * 3;
        bool ok = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) == hipSuccess;
        for (TickSlot &T : S->slot)
            ok = ok && hipStreamCreateWithFlags(&T.stream, hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreate(&T.ev0) == hipSuccess && hipEventCreate(&T.ev1) == hipSuccess;
        if (!ok) {
            release_device(*S);
            delete S;
            return GR_ERR_DEVICE;
        }
        S->th = std::thread(loop, S);
        ctx->serve = S;
    }
    *out = ctx->serve;
    return GR_OK;
}

std::shared_ptr<Job> find_job(ServeState &S, uint64_t id) {
    auto it = S.jobs.find(id);
    return it == S.jobs.end() ? nullptr : it->second;
}

gr_status job_status(const Job &j) {
    switch (j.state) {
        case GR_JOB_DONE: return GR_OK;
        case GR_JOB_CANCELLED: return GR_CANCELLED;
        case GR_JOB_FAILED: return j.error == GR_OK ? GR_ERR_DEVICE : j.error;
        default: return GR_PENDING;
    }
}

}  // namespace

void serve_shutdown(gr_ctx *ctx) {
    ServeState *S = nullptr;
    {
        std::lock_guard<std::mutex> g(ctx->serve_mu);
        S = ctx->serve;
        ctx->serve = nullptr;
    }
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->th.joinable()) S->th.join();
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(S->stream);
    for (TickSlot &T : S->slot)
        if (T.stream) (void)hipStreamSynchronize(T.stream);
    for (auto &j : S->active) free_job_device(*S, *j);
    for (auto &j : S->queue) drop_env_map(*S, *j);
    (void)hipStreamSynchronize(S->stream);
    release_device(*S);
    delete S;
}

}  // namespace gr

using gr::ServeState;

extern "C" {

gr_status gr_job_submit(gr_ctx *ctx, const gr_job_desc *desc, uint64_t *job_id) {
    if (!ctx || !desc || !job_id) return GR_ERR_INVALID;
    *job_id = 0;
    const gr_scene *s = desc->scene;
    gr_status st = gr::validate_scene(s);
    if (st != GR_OK) return st;
    if (desc->spp == 0 || desc->spp > GR_MAX_SAMPLES_PER_CALL) return GR_ERR_INVALID;
    if ((uint64_t)desc->sample_offset + desc->spp > 0xFFFFFFFFull) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    auto j = std::make_shared<gr::Job>();
    gr::fill_dev_scene(s, &j->host_sc);
    j->cls = gr::scene_class(j->host_sc);
    j->planes.assign(s->planes, s->planes + 4 * (size_t)s->n_planes);
    if (j->host_sc.n_frosted) j->finishes.assign(s->finishes, s->finishes + s->n_planes);
    if (s->env->kind == GR_ENV_MAP && s->env->map) {
        gr::env_map_retain(s->env->map);
        j->env_map = s->env->map;
    } else if (s->env->kind == GR_ENV_MAP) {
        j->map.assign(s->env->map_rgb, s->env->map_rgb + (size_t)s->env->map_w * s->env->map_h * 3);
    }
    j->width = s->width;
    j->height = s->height;
    j->n_pixels = (uint64_t)s->width * s->height;
    j->first_sample = j->sample_offset = desc->sample_offset;
    j->total_spp = j->spp = desc->spp;
    if (desc->subbatch_ms > 0) {
        j->progressive = true;
        j->subbatch_ms = desc->subbatch_ms;
        j->fixed_batches = desc->subbatch_fixed != 0;
        j->spp = std::min(std::max(desc->subbatch_samples, 1u), j->total_spp);
        j->film.assign(j->n_pixels * 3, 0.0f);
        j->pending.resize(j->n_pixels * 3);
    }
    j->mat_table = gr::wants_mat_table(j->host_sc, j->n_pixels * j->total_spp, ctx->hero_tables.load());
    j->footprint = gr::job_footprint(*j);
    j->t_submit = gr::Clock::now();
    std::lock_guard<std::mutex> g(S->mu);
    if (j->footprint > S->budget) {
        S->m.declined++;
        gr::drop_env_map(*S, *j);
        return GR_DECLINE_BUDGET;
    }
    const double rate = S->rate_ms > 0.0 ? (double)S->rate_tuples / S->rate_ms : gr::kDefaultRate;
    j->deadline_ms =
        gr::ms_between(S->t_start, j->t_submit) + S->slack * (double)(j->n_pixels * j->total_spp) / rate;
    j->id = S->next_id++;
    S->jobs[j->id] = j;
    S->queue.push_back(j);
    S->m.submitted++;
    *job_id = j->id;
    S->cv_work.notify_one();
    return GR_OK;
}

gr_status gr_job_poll(gr_ctx *ctx, uint64_t job_id) {
    if (!ctx) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    auto j = gr::find_job(*S, job_id);
    return j ? gr::job_status(*j) : GR_ERR_NOT_FOUND;
}

gr_status gr_job_wait(gr_ctx *ctx, uint64_t job_id, int64_t timeout_ms) {
    if (!ctx) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::unique_lock<std::mutex> lk(S->mu);
    auto j = gr::find_job(*S, job_id);
    if (!j) return GR_ERR_NOT_FOUND;
    auto settled = [&] { return gr::job_status(*j) != GR_PENDING; };
    if (timeout_ms < 0) S->cv_done.wait(lk, settled);
    else S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), settled);
    return gr::job_status(*j);
}

gr_status gr_job_cancel(gr_ctx *ctx, uint64_t job_id) {
    if (!ctx) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    auto j = gr::find_job(*S, job_id);
    if (!j) return GR_ERR_NOT_FOUND;
    if (j->state == GR_JOB_QUEUED || j->state == GR_JOB_RUNNING) {
        j->cancel_req = true;
        S->cv_work.notify_one();
    }
    return GR_OK;
}

gr_status gr_job_film(gr_ctx *ctx, uint64_t job_id, float *accum_xyz, size_t accum_pixels) {
    if (!ctx || !accum_xyz) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    auto j = gr::find_job(*S, job_id);
    if (!j) return GR_ERR_NOT_FOUND;
    st = gr::job_status(*j);
    if (st != GR_OK) return st;
    if (accum_pixels != j->n_pixels) return GR_ERR_INVALID;
    for (size_t i = 0; i < j->film.size(); ++i) accum_xyz[i] += j->film[i];
    return GR_OK;
}

gr_status gr_job_info(gr_ctx *ctx, uint64_t job_id, gr_job_stats *out) {
    if (!ctx || !out) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    auto j = gr::find_job(*S, job_id);
    if (!j) return GR_ERR_NOT_FOUND;
    std::memset(out, 0, sizeof(*out));
    out->state = (uint32_t)j->state;
    out->width = j->width;
    out->height = j->height;
    out->sample_offset = j->first_sample;
    out->spp = j->total_spp;
    out->samples_done = j->produced;
    out->subbatches = j->subbatches;
    out->footprint_bytes = j->footprint;
    out->paths = j->paths_done;
    if (j->state == GR_JOB_DONE)
        for (int i = 0; i < GR_EV_COUNT; ++i) out->events[i] = j->events[i];
    if (j->state == GR_JOB_DONE || j->state == GR_JOB_FAILED || j->state == GR_JOB_CANCELLED) {
        out->latency_ms = gr::ms_between(j->t_submit, j->t_done);
        if (j->state == GR_JOB_DONE) {
            out->queue_ms = gr::ms_between(j->t_submit, j->t_admit);
            out->run_ms = gr::ms_between(j->t_admit, j->t_done);
        }
    }
    return GR_OK;
}

gr_status gr_job_take_delta(gr_ctx *ctx, uint64_t job_id, uint32_t *first_sample, uint32_t *samples,
                            float *delta_xyz, size_t n_pixels) {
    if (!ctx || !first_sample || !samples || !delta_xyz) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    auto j = gr::find_job(*S, job_id);
    if (!j) return GR_ERR_NOT_FOUND;
    if (n_pixels != j->n_pixels) return GR_ERR_INVALID;
    *first_sample = j->first_sample;
    *samples = 0;
    if (j->progressive) {
        (void)j->pending.take(first_sample, samples, delta_xyz);
        return GR_OK;
    }
    // One batch: the whole film once the job is done.
    if (j->state != GR_JOB_DONE || j->delta_taken) return GR_OK;
    std::memcpy(delta_xyz, j->film.data(), j->film.size() * sizeof(float));
    *samples = j->total_spp;
    j->delta_taken = true;
    return GR_OK;
}

gr_status gr_job_release(gr_ctx *ctx, uint64_t job_id) {
    if (!ctx) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    auto j = gr::find_job(*S, job_id);
    if (!j) return GR_ERR_NOT_FOUND;
    if (gr::job_status(*j) == GR_PENDING) return GR_PENDING;
    S->jobs.erase(job_id);
    return GR_OK;
}

gr_status gr_metrics(gr_ctx *ctx, gr_serve_metrics *out) {
    if (!ctx || !out) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    *out = S->m;
    gr::percentiles(S->latencies, out);
    out->wall_ms = S->have_first && S->m.completed ? gr::ms_between(S->t_first, S->t_last) : 0.0;
    out->hbm_in_use = S->in_use;
    out->hbm_budget = S->budget;
    out->queued = S->queue.size();
    out->running = S->active.size();
    out->tick_tuples = S->tick_tuples;
    out->policy = (uint32_t)S->policy;
    out->slack = S->slack;
    out->rate_tuples_per_ms = S->rate_ms > 0.0 ? (double)S->rate_tuples / S->rate_ms : 0.0;
    return GR_OK;
}

gr_status gr_metrics_reset(gr_ctx *ctx) {
    if (!ctx) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    S->m = gr_serve_metrics{};
    S->latencies.clear();
    S->have_first = false;
    return GR_OK;
}

gr_status gr_serve_set_tick_tuples(gr_ctx *ctx, uint64_t tuples) {
    if (!ctx || tuples < gr::kMinTickTuples || tuples > gr::kMaxTickTuples) return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    S->tick_tuples = tuples;
    return GR_OK;
}

gr_status gr_serve_set_policy(gr_ctx *ctx, int policy, double slack) {
    if (!ctx || (policy != GR_SERVE_EQUAL_SHARE && policy != GR_SERVE_DEADLINE) || !(slack >= 0.0) ||
        slack > 1.0e6)
        return GR_ERR_INVALID;
    ServeState *S = nullptr;
    gr_status st = gr::get_serve(ctx, &S);
    if (st != GR_OK) return st;
    std::lock_guard<std::mutex> g(S->mu);
    S->policy = policy;
    S->slack = slack;
    return GR_OK;
}

}  // extern "C"
