// job.cpp -- one trace's full commitment, uneven batches, and the serving loop.
//
// The reference commits traces one at a time on the CPU (src/m31/stark.zig:
// 335-414 for the trace trees, src/m31/fri.zig:157-233 for FRI) and has no
// batching or serving layer.  Here a trace's commitment is one call
// (commit_job), a batch spreads uneven traces over worker contexts that each
// own a HIP stream and take jobs from a shared queue, and the server admits
// requests against an HBM budget and runs them on its workers while others
// are in flight (continuous batching).  Results do not depend on the mix: every
// job runs the same kernels on its own buffers and scratch.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.h"
#include "serve_core.h"

using namespace zs;

namespace {

constexpr char kCommitSep[] = "zig-stark:hip:commit";

// The channel every commit job starts from: Channel.init("zig-stark:hip:commit")
const B3Hasher &commit_channel() {
    static const B3Hasher st = [] {
        B3Hasher h;
        h.init();
        h.update(reinterpret_cast<const uint8_t *>("zig-stark:channel"), 17);
        h.update(reinterpret_cast<const uint8_t *>(kCommitSep), sizeof(kCommitSep) - 1);
        return h;
    }();
    return st;
}

// the shapes commit_job takes
bool job_shape_ok(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup, uint32_t remainder_log) {
    const uint32_t LT = log_n + log_blowup;
    return !(n_cols == 0 || n_cols > 4096 || log_n < 1 || log_sum_over(log_n, log_blowup, 28) ||
             remainder_log < log_blowup || remainder_log >= LT || remainder_log - log_blowup > kMaxRemainderLog);
}

// staged results of one job: trace root, at most 28 FRI roots, at most 2^kMaxRemainderLog
// remainder coefficients
constexpr size_t kStageBytes = 32 + 28 * 32 + (size_t(16) << kMaxRemainderLog);

// Columns from which the combination is formed before the blowup (see commit_job):
// the 4-column LDE it adds costs about as much as combining 96 columns' LDE rows
// (measured at 2^20 rows: 64 columns ran 0.09 ms slower that way)
constexpr uint32_t kCombineBeforeLdeCols = 96;

// One trace's commitment, queued on the context's stream with no host round
// trip until the end: LDE and row tree; the channel (on the device) absorbs the
// trace root and draws gamma; g = sum_c gamma^c lde_c; FRI on g with challenges
// drawn from the same channel.  Scratch: 7 channel state, 8 the codeword g,
// 9 gamma and its powers; FRI uses 4-6 and the LDE 1-2.
//
// The LDE is linear over M31 and the combination takes each of its four M31
// components separately, so g = LDE(h) with h = sum_c gamma^c col_c over the
// input rows, exactly: from kCombineBeforeLdeCols columns on (when the LDE's
// scratch can hold the planes) the products run over the n input rows
// instead of the 2^log_blowup n LDE rows, and a 4-column LDE of h's planes,
// in the LDE's scratch slots 1-2 once the trace's LDE is done, gives g.
//
// The results go to host memory through a pinned stage (the trace root, then the
// FRI roots, then the remainder words), so no copy into pageable memory blocks the
// host while the chain is being queued.  With `stage` (pinned, kStageBytes) given,
// the call returns without synchronising and trace_root / fri_roots / remainder
// are unused: the caller synchronises the stream before reading the stage.
// Without it the context's own pinned buffer is the stage, and the call returns
// with the results in trace_root / fri_roots / remainder.
//
// With `defer` (not with the plain path) the channel lives in defer->region, the
// results (trace root, FRI roots, remainder) are written to defer->results on the
// device, and the FRI tail is left to the caller (defer->pending, defer->tail): the
// call returns without synchronising or copying.
int32_t commit_job(zs_hip_ctx *h, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_n, uint32_t log_blowup,
                   uint32_t remainder_log, uint32_t *d_lde, uint8_t *d_nodes, uint8_t trace_root[32],
                   uint8_t *fri_roots, uint32_t *remainder, uint32_t flags, uint8_t *stage = nullptr,
                   FriDefer *defer = nullptr) {
    if (!h || !d_cols || !d_lde || !d_nodes || (!stage && !defer && (!trace_root || !fri_roots || !remainder)))
        return ZS_ERR_INPUT;
    if (defer && (stage || (flags & ZS_FLAG_PLAIN))) return ZS_ERR_INPUT;
    const uint32_t LT = log_n + log_blowup;
    if (!job_shape_ok(n_cols, log_n, log_blowup, remainder_log)) return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    const size_t N = size_t(1) << LT;
    uint32_t *nodes = reinterpret_cast<uint32_t *>(d_nodes);
    const uint32_t *d_root = nodes + tree_words(LT) - 8;
    ZS_TRY(ctx->ensure_scratch(7, sizeof(B3Hasher)));
    ZS_TRY(ctx->ensure_scratch(8, N * 16));
    ZS_TRY(ctx->ensure_scratch(9, (size_t(n_cols) + 1) * 16));
    B3Hasher *d_chan = defer ? reinterpret_cast<B3Hasher *>(defer->region) : ctx->scratch[7].as<B3Hasher>();
    const B3Hasher *d_init;
    ZS_TRY(device_commit_channel(ctx, &d_init));
    uint32_t *d_gamma = ctx->scratch[9].as<uint32_t>(), *d_powers = d_gamma + 4;
    if (flags & ZS_FLAG_PLAIN) {
        ZS_TRY(zs_m31_lde_commit_dev(h, d_cols, n_cols, log_n, log_blowup, d_lde, d_nodes, ZS_FLAG_PLAIN));
        ZS_HIP(ctx, hipMemcpyAsync(d_chan, d_init, sizeof(B3Hasher), hipMemcpyDeviceToDevice, ctx->stream));
        ZS_TRY(launch_channel_draw(ctx, d_chan, d_root, d_gamma));
    } else {
        // LDE, row leaves, then the node levels, whose root launch also starts the channel
        // from its initial state, draws gamma and writes gamma's powers
        ZS_TRY(zs_m31_lde_dev(h, d_cols, n_cols, log_n, log_blowup, d_lde));
        ZS_TRY(launch_row_leaves(ctx, d_lde, N, n_cols, N, nodes));
        ZS_TRY(launch_tree_levels(ctx, nodes, LT, 0, d_chan, d_gamma, d_init, d_powers, n_cols));
    }
    if (flags & ZS_FLAG_PLAIN) ZS_TRY(launch_qm31_powers(ctx, d_gamma, d_powers, n_cols));
    // the combination also writes the FRI's layer-0 leaves, where the commit phase's
    // first tree goes (the start of scratch slot 5)
    ZS_TRY(ctx->ensure_scratch(5, fri_tree_words(LT, LT - remainder_log) * 4));
    uint32_t *leaves0 = ctx->scratch[5].as<uint32_t>();
    const size_t n = size_t(1) << log_n;
    const bool before_lde = !(flags & ZS_FLAG_PLAIN) && n_cols >= kCombineBeforeLdeCols &&
                            ctx->scratch[1].bytes >= 8 * n * 4 && ctx->scratch[2].bytes >= 8 * N * 4;
    if (before_lde) {
        uint32_t *h_planes = ctx->scratch[1].as<uint32_t>(), *coef = h_planes + 4 * n;
        uint32_t *big = ctx->scratch[2].as<uint32_t>(), *g_planes = big + 4 * N;
        ZS_TRY(launch_combine_planes(ctx, d_cols, n, n_cols, n, d_powers, h_planes));
        ctx->lde_tiles = lde_tiles_for(ctx->device, 4, log_n, log_blowup);
        ZS_TRY(launch_lde(ctx, h_planes, n, 4, log_n, log_blowup, g_planes, N, coef, big));
        ZS_TRY(launch_interleave4(ctx, g_planes, N, ctx->scratch[8].as<uint32_t>(), leaves0));
    } else {
        ZS_TRY(launch_combine(ctx, d_lde, N, n_cols, N, d_powers, ctx->scratch[8].as<uint32_t>(), leaves0));
    }
    if (defer)
        return fri_commit_impl(h, ctx->scratch[8].as<uint32_t>(), LT, log_blowup, remainder_log, nullptr, d_chan,
                               kFriLeaves0Given, nullptr, nullptr, nullptr, nullptr, d_root, defer);
    const bool own_stage = stage == nullptr;
    if (own_stage) {
        ZS_TRY(ctx->ensure_pinned(kStageBytes));
        stage = static_cast<uint8_t *>(ctx->pinned);
    }
    // the trace root reaches the stage ahead of the FRI roots, in the commit phase's one copy
    ZS_TRY(fri_commit_impl(h, ctx->scratch[8].as<uint32_t>(), LT, log_blowup, remainder_log, nullptr, d_chan,
                           (flags & (ZS_FLAG_NO_GRAPH | ZS_FLAG_GRAPH)) | kFriLeaves0Given, nullptr, nullptr, nullptr, stage, d_root));
    if (!own_stage) return ZS_OK;
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t L = LT - remainder_log;
    std::memcpy(trace_root, stage, 32);
    std::memcpy(fri_roots, stage + 32, size_t(L) * 32);
    std::memcpy(remainder, stage + 32 + size_t(L) * 32, (size_t(16) << (remainder_log - log_blowup)));
    return ZS_OK;
}


// Device memory of one commit job: its input, LDE and row tree, and every scratch
// buffer commit_job uses (slot numbers of Ctx::scratch), each 256-byte aligned.
struct JobLayout {
    enum Part { IN, LDE, TREE, COEF, BIG, FRI_LAYERS, FRI_TREES, ALPHAS, CHAN, COMB, POWERS, NPARTS };
    uint64_t off[NPARTS], len[NPARTS], total;
};
constexpr int kSlotOf[JobLayout::NPARTS] = {-1, -1, -1, 1, 2, 4, 5, 6, 7, 8, 9};

JobLayout job_layout(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup, uint32_t remainder_log) {
    JobLayout j{};
    const uint32_t LT = log_n + log_blowup, L = LT - remainder_log;
    const uint64_t n = uint64_t(1) << log_n, N = uint64_t(1) << LT, c = n_cols;
    size_t coef_words;
    j.len[JobLayout::IN] = 4 * c * n;
    j.len[JobLayout::LDE] = 4 * c * N;
    j.len[JobLayout::TREE] = tree_words(LT) * 4;
    j.len[JobLayout::BIG] = lde_scratch_words(n_cols, log_n, log_blowup, &coef_words) * 4;
    j.len[JobLayout::COEF] = coef_words * 4;
    uint64_t fl = 0, ft = 0;
    for (uint32_t i = 1; i <= L; i++) fl += (N >> i) * 16;  // layers 1..L (layer 0 is read in place)
    for (uint32_t i = 0; i < L; i++) ft += tree_words(LT - i) * 4;
    j.len[JobLayout::FRI_LAYERS] = fl;
    j.len[JobLayout::FRI_TREES] = ft;
    // challenges, the trace root, then the gathered layer roots and the remainder coefficients
    // (fri_commit_impl's slot 6)
    j.len[JobLayout::ALPHAS] = uint64_t(L) * 48 + 32 + (uint64_t(16) << (remainder_log - log_blowup));
    j.len[JobLayout::CHAN] = sizeof(B3Hasher);
    j.len[JobLayout::COMB] = 16 * N;
    j.len[JobLayout::POWERS] = 16 * (c + 1);
    uint64_t o = 0;
    for (int k = 0; k < JobLayout::NPARTS; k++) {
        j.off[k] = o;
        o += (j.len[k] + 255) & ~uint64_t(255);
    }
    j.total = o;
    return j;
}

zs_hip_ctx *new_worker(int device, const std::shared_ptr<DeviceCaches> &caches) {
    zs_hip_ctx *w = nullptr;
    if (zs_hip_init(device, 0, &w) != ZS_OK) return nullptr;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        zs_hip_destroy(w);
        return nullptr;
    }
    w->c.stream = s;
    w->c.own_stream = true;
    if (caches) w->c.caches = caches;
    return w;
}

// The persistent launch's share of a batch: the jobs it takes, in groups whose pool
// fits the device memory free now (largest traces first).  Jobs outside every group
// (other shapes, or a trace whose pool alone does not fit) are left to commit_job.
int32_t persistent_groups(Ctx *ctx, const zs_job *jobs, size_t n_jobs, uint32_t log_blowup, uint32_t remainder_log,
                          std::vector<std::vector<size_t>> *groups) {
    size_t free_b = 0, total_b = 0;
    ZS_HIP(ctx, hipMemGetInfo(&free_b, &total_b));
    const size_t margin = size_t(1) << 30;
    const size_t budget = free_b + ctx->scratch[10].bytes > margin ? free_b + ctx->scratch[10].bytes - margin : 0;
    std::vector<size_t> take;
    for (size_t i = 0; i < n_jobs; i++) {
        const zs_job &j = jobs[i];
        if (job_shape_ok(j.n_cols, j.log_n, log_blowup, remainder_log) &&
            persistent_supported(j.n_cols, j.log_n, log_blowup) && j.d_cols && j.d_lde && j.d_nodes &&
            persistent_pool_bytes(j.n_cols, j.log_n, log_blowup, remainder_log) <= budget)
            take.push_back(i);
    }
    std::stable_sort(take.begin(), take.end(), [&](size_t a, size_t b) {
        return persistent_pool_bytes(jobs[a].n_cols, jobs[a].log_n, log_blowup, remainder_log) >
               persistent_pool_bytes(jobs[b].n_cols, jobs[b].log_n, log_blowup, remainder_log);
    });
    size_t used = 0;
    for (size_t i : take) {
        const size_t b = persistent_pool_bytes(jobs[i].n_cols, jobs[i].log_n, log_blowup, remainder_log);
        if (groups->empty() || used + b > budget) {
            groups->emplace_back();
            used = 0;
        }
        groups->back().push_back(i);
        used += b;
    }
    return ZS_OK;
}

}  // namespace

// the starting channel state on the device, built once per cache (outside its lock, like
// the twiddle tables) so each job starts its channel with a device-to-device copy
int32_t zs::device_commit_channel(Ctx *ctx, const B3Hasher **out) {
    DeviceCaches &dc = *ctx->caches;
    {
        std::lock_guard<std::mutex> lk(dc.mu);
        if (dc.commit_channel) {
            *out = dc.commit_channel->as<B3Hasher>();
            return ZS_OK;
        }
    }
    auto buf = std::make_unique<DevBuf>();
    ZS_HIP(ctx, hipMalloc(&buf->ptr, sizeof(B3Hasher)));
    buf->bytes = sizeof(B3Hasher);
    ZS_HIP(ctx, hipMemcpyAsync(buf->ptr, &commit_channel(), sizeof(B3Hasher), hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> lk(dc.mu);
    if (!dc.commit_channel) dc.commit_channel = std::move(buf);
    *out = dc.commit_channel->as<B3Hasher>();
    return ZS_OK;
}

extern "C" {

int32_t zs_commit_job_dev(zs_hip_ctx *ctx, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_n,
                          uint32_t log_blowup, uint32_t remainder_log, uint32_t *d_lde, uint8_t *d_nodes,
                          uint8_t trace_root[32], uint8_t *fri_roots, uint32_t *remainder, uint32_t flags) {
    return commit_job(ctx, d_cols, n_cols, log_n, log_blowup, remainder_log, d_lde, d_nodes, trace_root, fri_roots,
                      remainder, flags);
}

int32_t zs_commit_batch_dev(zs_hip_ctx *h, zs_job *jobs, size_t n_jobs, uint32_t log_blowup, uint32_t remainder_log,
                            uint32_t n_workers, uint32_t flags) {
    if (!h || (!jobs && n_jobs) || n_workers > 64) return ZS_ERR_INPUT;
    if (n_jobs == 0) return ZS_OK;
    Ctx *ctx = &h->c;
    // ZS_FLAG_PERSISTENT: the whole commitment of the jobs it takes in one persistent
    // launch per memory group; the others below, on the workers
    std::vector<uint8_t> done(n_jobs, 0);
    if ((flags & ZS_FLAG_PERSISTENT) && !(flags & ZS_FLAG_PLAIN)) {
        ZS_HIP(ctx, hipSetDevice(ctx->device));
        ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));  // inputs written on the caller's stream are complete
        std::vector<std::vector<size_t>> groups;
        ZS_TRY(persistent_groups(ctx, jobs, n_jobs, log_blowup, remainder_log, &groups));
        for (const auto &g : groups) {
            const int32_t rc = persistent_commit(ctx, jobs, g, log_blowup, remainder_log);
            for (size_t i : g) {
                if (rc == ZS_OK) done[i] = 1;
                else jobs[i].status = rc;
            }
            if (rc != ZS_OK) return rc;
        }
    }
    auto run_one = [&](zs_hip_ctx *w, size_t i, uint8_t *stage, FriDefer *defer = nullptr) {
        zs_job &j = jobs[i];
        const bool host_out = !stage && !defer;
        return commit_job(w, j.d_cols, j.n_cols, j.log_n, log_blowup, remainder_log, j.d_lde, j.d_nodes,
                          host_out ? j.trace_root : nullptr, host_out ? j.fri_roots : nullptr,
                          host_out ? j.remainder : nullptr, flags, stage, defer);
    };
    std::vector<size_t> rest;
    for (size_t i = 0; i < n_jobs; i++)
        if (!done[i]) rest.push_back(i);
    if (rest.empty()) {
    } else if (n_workers <= 1 || (flags & ZS_FLAG_PLAIN)) {
        for (size_t i : rest) jobs[i].status = run_one(h, i, nullptr);
    } else {
        ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));  // inputs written on the caller's stream are complete
        auto &ws = ctx->workers;
        while (ws.size() < n_workers) {
            zs_hip_ctx *w = new_worker(ctx->device, ctx->caches);
            if (!w) return ZS_ERR_DEVICE;
            ws.push_back(w);
        }
        // largest traces first, so that no large trace starts last and sets the tail
        std::vector<size_t> order = rest;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            const uint64_t ea = uint64_t(jobs[a].n_cols) << std::min(jobs[a].log_n, 40u);
            const uint64_t eb = uint64_t(jobs[b].n_cols) << std::min(jobs[b].log_n, 40u);
            return ea > eb;
        });
        // Every job's results land in pinned staging.  A worker keeps at most kInFlight jobs
        // queued on its stream: it takes the next job once the oldest of them has finished on
        // the device, so the traces are spread over the streams by completion (a worker that
        // drew large traces takes fewer), while the host prepares one job's launches during
        // another's kernels instead of waiting for each result.
        constexpr size_t kInFlight = 2;
        ZS_TRY(ctx->ensure_pinned(n_jobs * kStageBytes));
        uint8_t *stage = static_cast<uint8_t *>(ctx->pinned);
        // Each job's FRI tail (its layers of at most 2^10 values and the remainder, one
        // workgroup's latency-bound chain) is deferred: the jobs' tails run together in one
        // launch after the workers, from per-job regions (slot 12), and every job's results
        // (slot 13, the stage layout) reach the host in one copy.
        std::vector<FriDefer> defers(n_jobs);
        {
            std::vector<size_t> region_off(n_jobs, 0);
            size_t words = 0;
            for (size_t i : rest) {
                region_off[i] = words;
                if (job_shape_ok(jobs[i].n_cols, jobs[i].log_n, log_blowup, remainder_log))
                    words += (fri_defer_words(jobs[i].log_n + log_blowup, remainder_log) + 63) & ~size_t(63);
            }
            const size_t res_bytes = (n_jobs * kStageBytes + 255) & ~size_t(255);
            ZS_TRY(ctx->ensure_scratch(12, std::max<size_t>(words, 64) * 4));
            ZS_TRY(ctx->ensure_scratch(13, res_bytes + n_jobs * sizeof(FriTail)));
            for (size_t i : rest) {
                defers[i].region = ctx->scratch[12].as<uint32_t>() + region_off[i];
                defers[i].results = ctx->scratch[13].as<uint32_t>() + i * (kStageBytes / 4);
            }
        }
        std::atomic<size_t> next{0};
        std::vector<uint32_t> ran_on(n_jobs, 0);
        std::vector<uint8_t> stream_failed(n_workers, 0);
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < n_workers; t++) {
            th.emplace_back([&, t] {
                zs_hip_ctx *w = static_cast<zs_hip_ctx *>(ws[t]);
                (void)hipSetDevice(ctx->device);
                hipEvent_t ev[kInFlight] = {};
                bool ev_ok = true;
                for (auto &e : ev) ev_ok = ev_ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
                size_t queued = 0;  // jobs recorded on ev[], oldest at queued % kInFlight
                for (size_t k;;) {
                    if (ev_ok && queued >= kInFlight && hipEventSynchronize(ev[queued % kInFlight]) != hipSuccess)
                        ev_ok = false;
                    if ((k = next.fetch_add(1)) >= order.size()) break;
                    const size_t i = order[k];
                    ran_on[i] = t;
                    jobs[i].status = run_one(w, i, nullptr, &defers[i]);
                    if (!ev_ok || hipEventRecord(ev[queued % kInFlight], w->c.stream) != hipSuccess) {
                        ev_ok = false;  // no completion events: one job at a time
                        (void)hipStreamSynchronize(w->c.stream);
                    }
                    queued++;
                }
                stream_failed[t] = hipStreamSynchronize(w->c.stream) != hipSuccess;
                for (auto &e : ev)
                    if (e) (void)hipEventDestroy(e);
            });
        }
        for (auto &x : th) x.join();
        // every worker stream has finished: the deferred tails, then one copy of the results
        std::vector<FriTail> tails;
        for (size_t i : rest) {
            if (stream_failed[ran_on[i]]) jobs[i].status = ZS_ERR_DEVICE;  // no result of that stream is complete
            if (jobs[i].status == ZS_OK && defers[i].pending) tails.push_back(defers[i].tail);
        }
        FriTail *d_tails = reinterpret_cast<FriTail *>(ctx->scratch[13].as<uint8_t>() +
                                                       ((n_jobs * kStageBytes + 255) & ~size_t(255)));
        if (!tails.empty()) {
            ZS_HIP(ctx, hipMemcpyAsync(d_tails, tails.data(), tails.size() * sizeof(FriTail), hipMemcpyHostToDevice,
                                       ctx->stream));
            ZS_TRY(launch_fri_tail_multi(ctx, d_tails, (uint32_t)tails.size()));
        }
        ZS_HIP(ctx, hipMemcpyAsync(stage, ctx->scratch[13].ptr, n_jobs * kStageBytes, hipMemcpyDeviceToHost,
                                   ctx->stream));
        ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t i : rest) {
            zs_job &j = jobs[i];
            if (j.status != ZS_OK) continue;
            const uint32_t L = j.log_n + log_blowup - remainder_log;
            const size_t rc = size_t(1) << (remainder_log - log_blowup);
            const uint8_t *st = stage + i * kStageBytes;
            std::memcpy(j.trace_root, st, 32);
            if (j.fri_roots) std::memcpy(j.fri_roots, st + 32, size_t(L) * 32);
            if (j.remainder) std::memcpy(j.remainder, st + 32 + size_t(L) * 32, rc * 16);
        }
    }
    for (size_t i = 0; i < n_jobs; i++)
        if (jobs[i].status != ZS_OK) return jobs[i].status;
    return ZS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// serving loop: the queue, admission, arena, tickets and shutdown are the
// HIP-free ServeCore (serve_core.h); a request's device work is serve_one
// ---------------------------------------------------------------------------
struct ServData {
    zs_request r;
    std::vector<uint32_t> host;
    zs_response resp;
    JobLayout lay;
};
using ServCore = ServeCore<ServData>;

struct zs_server {
    int device = 0;
    DevBuf pool;  // the budget, allocated once; requests are carved from it
    std::vector<zs_hip_ctx *> workers;
    std::unique_ptr<ServCore> core;
};

static void serve_one(zs_server *s, zs_hip_ctx *w, ServCore::Item &it) {
    Ctx *ctx = &w->c;
    ServData &q = it.data;
    const zs_request &r = q.r;
    const JobLayout &j = q.lay;
    char *base = s->pool.as<char>() + it.off;
    auto part = [&](int k) { return reinterpret_cast<void *>(base + j.off[k]); };
    for (int k = 0; k < JobLayout::NPARTS; k++)
        if (kSlotOf[k] >= 0) ctx->lend_scratch(kSlotOf[k], part(k), j.len[k]);
    uint32_t *d_cols = static_cast<uint32_t *>(part(JobLayout::IN));
    const size_t elems = size_t(r.n_cols) << r.log_n;
    int32_t rc;
    if (r.host_cols)
        rc = hipMemcpyAsync(d_cols, q.host.data(), elems * 4, hipMemcpyHostToDevice, ctx->stream) == hipSuccess
                 ? ZS_OK
                 : ZS_ERR_DEVICE;
    else
        rc = zs_fill_m31_dev(w, r.seed, d_cols, elems);
    if (!rc)
        rc = commit_job(w, d_cols, r.n_cols, r.log_n, r.log_blowup, r.remainder_log,
                        static_cast<uint32_t *>(part(JobLayout::LDE)), static_cast<uint8_t *>(part(JobLayout::TREE)),
                        q.resp.trace_root, q.resp.fri_roots, q.resp.remainder, ZS_FLAG_NO_GRAPH);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->release_scratch();  // (each request's buffers differ, so its FRI is not captured as a graph)
    q.resp.status = rc;
    q.resp.fri_layers = r.log_n + r.log_blowup - r.remainder_log;
}

static void take_response(const ServCore::Item &it, zs_response *out) {
    *out = it.data.resp;
    out->t_submit = it.t_submit;
    out->t_admit = it.t_admit;
    out->t_done = it.t_done;
    out->device_bytes = it.bytes;
}

extern "C" {

zs_server *zs_server_new(int32_t device, uint32_t n_workers, uint64_t hbm_budget_bytes) {
    if (n_workers == 0 || n_workers > 64) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    auto *s = new zs_server();
    s->device = device;
    const uint64_t budget = Arena::round(hbm_budget_bytes);
    if (hipMalloc(&s->pool.ptr, budget) != hipSuccess) {
        s->pool.ptr = nullptr;
        delete s;
        return nullptr;
    }
    s->pool.bytes = budget;
    for (uint32_t i = 0; i < n_workers; i++) {
        zs_hip_ctx *w = new_worker(device, s->workers.empty() ? nullptr : s->workers[0]->c.caches);
        if (!w) {
            for (zs_hip_ctx *x : s->workers) zs_hip_destroy(x);
            delete s;
            return nullptr;
        }
        s->workers.push_back(w);
    }
    s->core = std::make_unique<ServCore>(budget, [s](size_t wi, ServCore::Item &it) {
        serve_one(s, s->workers[wi], it);
    });
    s->core->start(n_workers, [device](size_t) { (void)hipSetDevice(device); });
    return s;
}

int32_t zs_server_submit(zs_server *s, const zs_request *req, uint64_t *ticket) {
    if (!s || !req || !ticket) return ZS_ERR_INPUT;
    const uint32_t LT = req->log_n + req->log_blowup;
    if (req->n_cols == 0 || req->n_cols > 4096 || req->log_n < 1 || log_sum_over(req->log_n, req->log_blowup, 28) ||
        req->remainder_log < req->log_blowup ||
        req->remainder_log >= LT || req->remainder_log - req->log_blowup > kMaxRemainderLog)
        return ZS_ERR_INPUT;
    auto it = std::make_unique<ServCore::Item>();
    ServData &q = it->data;
    q.r = *req;
    std::memset(&q.resp, 0, sizeof q.resp);
    if (req->host_cols) q.host.assign(req->host_cols, req->host_cols + (size_t(req->n_cols) << req->log_n));
    q.lay = job_layout(req->n_cols, req->log_n, req->log_blowup, req->remainder_log);
    it->bytes = q.lay.total;
    return s->core->submit(std::move(it), ticket) ? ZS_OK : ZS_ERR_OOM;
}

// A ticket is answered once: the first wait or poll claims it, and any later (or
// concurrent) wait or poll of the same ticket returns ZS_ERR_INPUT.
int32_t zs_server_wait(zs_server *s, uint64_t ticket, zs_response *out) {
    if (!s || !out) return ZS_ERR_INPUT;
    std::unique_ptr<ServCore::Item> it;
    if (!s->core->wait(ticket, &it)) return ZS_ERR_INPUT;
    take_response(*it, out);
    return ZS_OK;
}

int32_t zs_server_poll(zs_server *s, uint64_t ticket, zs_response *out, int32_t *ready) {
    if (!s || !out || !ready) return ZS_ERR_INPUT;
    std::unique_ptr<ServCore::Item> it;
    bool r = false;
    if (!s->core->poll(ticket, &it, &r)) return ZS_ERR_INPUT;
    *ready = r ? 1 : 0;
    if (r) take_response(*it, out);
    return ZS_OK;
}

int32_t zs_server_memory(zs_server *s, uint64_t *in_use, uint64_t *peak) {
    if (!s) return ZS_ERR_INPUT;
    s->core->memory(in_use, peak);
    return ZS_OK;
}

// Every queued request is served before the workers stop; threads blocked in
// zs_server_wait are answered and have left before the server is deleted.  No
// call on the server may start once zs_server_free has been called.
void zs_server_free(zs_server *s) {
    if (!s) return;
    s->core->stop();
    s->core.reset();
    for (zs_hip_ctx *w : s->workers) zs_hip_destroy(w);
    delete s;
}

}  // extern "C"
