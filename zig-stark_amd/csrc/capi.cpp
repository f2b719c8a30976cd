// capi.cpp -- the extern "C" entry points of include/zs_hip.h.
//
// Mirrors the reference's device boundary (src/cuda/circlefft_gpu.zig,
// src/cuda/cuda.zig): a lazily initialised per-device context, host-buffer
// transforms with the CPU transform's signature, and error codes in the
// zig-capi.h convention.  Unlike the reference there is no CPU fallback: a
// missing device or a HIP failure is returned as ZS_ERR_DEVICE.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "blake3.h"
#include "internal.h"

using namespace zs;

namespace zs {
int32_t Ctx::ensure_scratch(int slot, size_t bytes) {
    DevBuf &b = scratch[slot];
    if (b.bytes >= bytes) return ZS_OK;
    if (b.ptr && !b.owned) {
        // a lent buffer (a server request's share of its HBM budget) is never replaced
        // by an allocation outside that budget
        set_error("scratch slot " + std::to_string(slot) + ": lent buffer of " + std::to_string(b.bytes) +
                  " bytes, " + std::to_string(bytes) + " needed");
        return ZS_ERR_GENERIC;
    }
    if (b.ptr) {
        if (stream_ordered) {
            ZS_HIP(this, hipFreeAsync(b.ptr, stream));
        } else {
            ZS_HIP(this, hipStreamSynchronize(stream));
            ZS_HIP(this, hipFree(b.ptr));
        }
        b.ptr = nullptr;
        b.bytes = 0;
    }
    hipError_t e = stream_ordered ? hipMallocAsync(&b.ptr, bytes, stream) : hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        set_error(std::string("scratch allocation: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? ZS_ERR_OOM : ZS_ERR_DEVICE;
    }
    b.bytes = bytes;
    return ZS_OK;
}

int32_t Ctx::ensure_pinned(size_t bytes) {
    if (pinned_bytes >= bytes) return ZS_OK;
    if (pinned) {
        ZS_HIP(this, hipStreamSynchronize(stream));
        (void)hipHostFree(pinned);
        pinned = nullptr;
        pinned_bytes = 0;
    }
    if (hipHostMalloc(&pinned, bytes, hipHostMallocDefault) != hipSuccess) {
        pinned = nullptr;
        set_error("pinned staging allocation failed");
        return ZS_ERR_OOM;
    }
    pinned_bytes = bytes;
    return ZS_OK;
}

// give the scratch arenas back (stream-ordered contexts free them on their stream)
void Ctx::release_scratch() {
    for (DevBuf &b : scratch) {
        if (b.ptr && b.owned) {
            if (stream_ordered) (void)hipFreeAsync(b.ptr, stream);
            else (void)hipFree(b.ptr);
        }
        b.ptr = nullptr;
        b.bytes = 0;
        b.owned = true;
    }
}

void Ctx::lend_scratch(int slot, void *ptr, size_t bytes) {
    DevBuf &b = scratch[slot];
    if (b.ptr && b.owned) {
        if (stream_ordered) (void)hipFreeAsync(b.ptr, stream);
        else {
            (void)hipStreamSynchronize(stream);
            (void)hipFree(b.ptr);
        }
    }
    b.ptr = ptr;
    b.bytes = bytes;
    b.owned = false;
}
}  // namespace zs

static int32_t dev_alloc(Ctx *ctx, DevBuf &b, size_t bytes) {
    hipError_t e = hipMalloc(&b.ptr, bytes ? bytes : 4);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        ctx->set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? ZS_ERR_OOM : ZS_ERR_DEVICE;
    }
    b.bytes = bytes;
    return ZS_OK;
}

static bool is_pow2(size_t n) { return n && !(n & (n - 1)); }

extern "C" {

const char *zs_hip_version(void) { return "zig-stark-hip 0.1.0 (gfx950; M31 circle LDE, BLAKE3 Merkle, QM31 FRI)"; }

int32_t zs_hip_init(int32_t device, uint32_t flags, zs_hip_ctx **out_ctx) {
    (void)flags;
    if (!out_ctx) return ZS_ERR_INPUT;
    *out_ctx = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count) return ZS_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return ZS_ERR_DEVICE;
    noalloc_setup(device);  // once per device: the allocation-free entry point's fixed array
    auto *c = new (std::nothrow) zs_hip_ctx();
    if (!c) return ZS_ERR_OOM;
    c->c.device = device;
    *out_ctx = c;
    return ZS_OK;
}

static void prof_clear(Ctx *ctx);

void zs_hip_destroy(zs_hip_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    for (void *w : ctx->c.workers) zs_hip_destroy(static_cast<zs_hip_ctx *>(w));
    ctx->c.workers.clear();
    (void)hipStreamSynchronize(ctx->c.stream);
    prof_clear(&ctx->c);
    for (auto &g : ctx->c.fri_graphs) (void)hipGraphExecDestroy(g.second);
    ctx->c.fri_graphs.clear();
    if (ctx->c.capture_stream) (void)hipStreamDestroy(ctx->c.capture_stream);
    ctx->c.release_scratch();
    (void)hipStreamSynchronize(ctx->c.stream);
    if (ctx->c.pinned) (void)hipHostFree(ctx->c.pinned);
    if (ctx->c.own_stream) (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

const char *zs_hip_last_error(const zs_hip_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : "no context"; }

int32_t zs_hip_set_stream(zs_hip_ctx *ctx, void *hip_stream) {
    if (!ctx) return ZS_ERR_INPUT;
    ctx->c.stream = static_cast<hipStream_t>(hip_stream);
    return ZS_OK;
}

int32_t zs_hip_synchronize(zs_hip_ctx *ctx) {
    if (!ctx) return ZS_ERR_INPUT;
    ZS_HIP(&ctx->c, hipStreamSynchronize(ctx->c.stream));
    return ZS_OK;
}

static void prof_clear(Ctx *ctx) {
    for (auto &r : ctx->prof) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    ctx->prof.clear();
}

int32_t zs_hip_profile(zs_hip_ctx *h, int32_t enable) {
    if (!h) return ZS_ERR_INPUT;
    ZS_HIP(&h->c, hipStreamSynchronize(h->c.stream));
    prof_clear(&h->c);
    h->c.profiling = enable != 0;
    return ZS_OK;
}

int32_t zs_hip_profile_read(zs_hip_ctx *h, char *names, size_t names_cap, double *ms, uint32_t *counts, uint32_t max,
                            uint32_t *n_out) {
    if (!h || !n_out) return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, uint32_t>> agg;
    for (auto &r : ctx->prof) {
        float t = 0;
        ZS_HIP(ctx, hipEventElapsedTime(&t, r.a, r.b));
        auto it = agg.find(r.name);
        if (it == agg.end()) {
            order.push_back(r.name);
            agg[r.name] = {t, 1};
        } else {
            it->second.first += t;
            it->second.second++;
        }
    }
    uint32_t n = 0;
    size_t off = 0;
    for (auto &name : order) {
        if (n >= max || off + name.size() + 1 > names_cap) break;
        std::memcpy(names + off, name.c_str(), name.size() + 1);
        off += name.size() + 1;
        ms[n] = agg[name].first;
        counts[n] = agg[name].second;
        n++;
    }
    *n_out = n;
    return ZS_OK;
}

void zs_free(zs_host_allocator host, uint8_t *ptr, size_t len) {
    if (ptr && host.free) host.free(host.ctx, ptr, len);
}

// coset.zig:55-65
int32_t zs_circle_coset_canonic_half(uint32_t log_size, zs_circle_coset *out) {
    if (!out || log_size < 1 || log_size > 28) return ZS_ERR_INPUT;
    cpoint g = cp_neg(cpoint{1268011823u, 2u});  // point.zig:49-51, negated
    cpoint off = cp_mul(g, uint64_t(1) << (30 - log_size));
    cpoint st = cp_mul(g, uint64_t(1) << (32 - log_size));
    *out = {off.x, off.y, st.x, st.y, log_size - 1};
    return ZS_OK;
}

// ---- circle FFT, host buffers ------------------------------------------------
static int32_t host_transform(zs_hip_ctx *h, bool inverse, const uint32_t *in, size_t n, const zs_circle_coset *hc,
                              uint32_t *out) {
    if (!h || !in || !out || !hc) return ZS_ERR_INPUT;
    if (n < 2 || !is_pow2(n) || hc->log_size > 28 || (size_t(1) << hc->log_size) != n / 2) return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf a, b;
    ZS_TRY(dev_alloc(ctx, a, n * 4));
    ZS_TRY(dev_alloc(ctx, b, n * 4));
    ZS_HIP(ctx, hipMemcpyAsync(a.ptr, in, n * 4, hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (inverse) ZS_TRY(launch_ifft(ctx, *hc, a.as<uint32_t>(), n, b.as<uint32_t>(), n, 1));
    else ZS_TRY(launch_fft(ctx, *hc, a.as<uint32_t>(), n, b.as<uint32_t>(), n, 1));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(out, b.ptr, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

int32_t zs_m31_circle_fft(zs_hip_ctx *ctx, const uint32_t *coeffs, size_t n, const zs_circle_coset *half_coset,
                          uint32_t *evals) {
    return host_transform(ctx, false, coeffs, n, half_coset, evals);
}

int32_t zs_m31_circle_ifft(zs_hip_ctx *ctx, const uint32_t *evals, size_t n, const zs_circle_coset *half_coset,
                           uint32_t *coeffs) {
    return host_transform(ctx, true, evals, n, half_coset, coeffs);
}

int32_t zs_m31_circle_fft_batch_dev(zs_hip_ctx *h, const uint32_t *d_coeffs, uint32_t *d_evals, uint32_t log_n,
                                    uint32_t n_cols, size_t col_stride) {
    if (!h || log_n < 1 || log_n > 28 || col_stride < (size_t(1) << log_n)) return ZS_ERR_INPUT;
    zs_circle_coset hc;
    zs_circle_coset_canonic_half(log_n, &hc);
    return launch_fft(&h->c, hc, d_coeffs, col_stride, d_evals, col_stride, n_cols);
}

int32_t zs_m31_circle_ifft_batch_dev(zs_hip_ctx *h, const uint32_t *d_evals, uint32_t *d_coeffs, uint32_t log_n,
                                     uint32_t n_cols, size_t col_stride) {
    if (!h || log_n < 1 || log_n > 28 || col_stride < (size_t(1) << log_n)) return ZS_ERR_INPUT;
    zs_circle_coset hc;
    zs_circle_coset_canonic_half(log_n, &hc);
    return launch_ifft(&h->c, hc, d_evals, col_stride, d_coeffs, col_stride, n_cols);
}

// ---- LDE + row commitment ------------------------------------------------------
int32_t zs_m31_lde_commit_dev(zs_hip_ctx *h, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_n,
                              uint32_t log_blowup, uint32_t *d_lde, uint8_t *d_nodes, uint32_t flags) {
    if (!h || !d_cols || !d_lde || !d_nodes || n_cols == 0 || log_n < 1 || log_sum_over(log_n, log_blowup, 28))
        return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    uint32_t LT = log_n + log_blowup;
    size_t n = size_t(1) << log_n, N = size_t(1) << LT;
    if (flags & ZS_FLAG_PLAIN) {
        if (log_n < 3) return ZS_ERR_INPUT;
        ZS_TRY(ctx->ensure_scratch(3, 2 * N * 4));
        ZS_TRY(launch_lde_plain(ctx, d_cols, n, n_cols, log_n, log_blowup, d_lde, N, ctx->scratch[3].as<uint32_t>()));
        return launch_row_commit_plain(ctx, d_lde, N, n_cols, LT, reinterpret_cast<uint32_t *>(d_nodes));
    }
    ZS_TRY(zs_m31_lde_dev(h, d_cols, n_cols, log_n, log_blowup, d_lde));
    return launch_row_commit(ctx, d_lde, N, n_cols, LT, reinterpret_cast<uint32_t *>(d_nodes));
}

int32_t zs_m31_lde_dev(zs_hip_ctx *h, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_n, uint32_t log_blowup,
                       uint32_t *d_lde) {
    if (!h || !d_cols || !d_lde || n_cols == 0 || log_n < 1 || log_sum_over(log_n, log_blowup, 28)) return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ctx->lde_tiles = lde_tiles_for(ctx->device, n_cols, log_n, log_blowup);
    size_t n = size_t(1) << log_n, N = size_t(1) << (log_n + log_blowup);
    size_t coef_words;
    size_t big_words = lde_scratch_words(n_cols, log_n, log_blowup, &coef_words);
    ZS_TRY(ctx->ensure_scratch(1, coef_words * 4));
    if (big_words) ZS_TRY(ctx->ensure_scratch(2, big_words * 4));
    return launch_lde(ctx, d_cols, n, n_cols, log_n, log_blowup, d_lde, N, ctx->scratch[1].as<uint32_t>(),
                      big_words ? ctx->scratch[2].as<uint32_t>() : nullptr);
}

}  // extern "C"

// ---- persisted autotune picks --------------------------------------------------------
// One text line per tuned shape: "<device> <shape key> <6 tile values> <us>", where
// <device> names the architecture, the compute-unit count and this library's
// version.  zs_autotune_lde reads the file before tuning a shape and appends a
// line after tuning one, so a later process skips the tuning.  Path:
// $ZS_AUTOTUNE_CACHE, else $XDG_CACHE_HOME/zig-stark-hip/autotune.txt, else
// $HOME/.cache/zig-stark-hip/autotune.txt; ZS_AUTOTUNE_CACHE=off disables the
// file and ZS_AUTOTUNE_REFRESH=1 re-tunes (the last line of a key wins).
namespace {
std::string autotune_cache_path() {
    const char *e = getenv("ZS_AUTOTUNE_CACHE");
    if (e && *e) return std::string(e) == "off" ? std::string() : std::string(e);
    const char *x = getenv("XDG_CACHE_HOME");
    const char *home = getenv("HOME");
    if (x && *x) return std::string(x) + "/zig-stark-hip/autotune.txt";
    if (home && *home) return std::string(home) + "/.cache/zig-stark-hip/autotune.txt";
    return std::string();
}

std::string device_key(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return std::string();
    std::string arch = prop.gcnArchName;
    for (char &c : arch)
        if (c == ' ' || c == ':') c = '_';
    uint64_t vh = 1469598103934665603ull;  // FNV-1a of the version string
    for (const char *v = zs_hip_version(); *v; v++) vh = (vh ^ (uint8_t)*v) * 1099511628211ull;
    char buf[160];
    snprintf(buf, sizeof buf, "%s-cu%d-v%016llx", arch.c_str(), prop.multiProcessorCount, (unsigned long long)vh);
    return buf;
}

bool autotune_refresh() {
    const char *r = getenv("ZS_AUTOTUNE_REFRESH");
    return r && atoi(r);
}

// every persisted pick of this device into `into` (the last line of a key wins)
void autotune_cache_load(int device, std::map<uint64_t, TunedPick> &into) {
    const std::string path = autotune_cache_path(), dev = device_key(device);
    if (path.empty() || dev.empty() || autotune_refresh()) return;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char d[200];
        unsigned long long k;
        unsigned v[7];
        if (sscanf(line, "%199s %llx %u %u %u %u %u %u %u", d, &k, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6]) != 9)
            continue;
        if (dev != d) continue;
        bool ok = true;
        for (int r = 0; r < 3; r++) ok = ok && v[2 * r] >= 11 && v[2 * r] <= 14 && v[2 * r + 1] >= 1 && v[2 * r + 1] <= 8;
        if (!ok) continue;  // not a configuration the tuner can pick: ignored
        TunedPick tp;
        for (int r = 0; r < 3; r++) tp.t.role[r] = {v[2 * r], v[2 * r + 1]};
        tp.us = v[6];
        into[k] = tp;
    }
    fclose(f);
}

// Tuned launch shapes are per device and shared by every context of the process
// (a context, its batch workers, a server's workers), so a shape tuned by any of
// them serves all.  The first lookup on a device reads that device's persisted picks.
struct TunedTable {
    std::mutex mu;
    bool loaded = false;
    std::map<uint64_t, TunedPick> picks;
};
TunedTable &tuned_table(int device) {
    static std::mutex m;
    static std::map<int, std::unique_ptr<TunedTable>> tables;
    std::lock_guard<std::mutex> lk(m);
    auto &t = tables[device];
    if (!t) t = std::make_unique<TunedTable>();
    return *t;
}
void tuned_load_locked(TunedTable &tb, int device) {
    if (tb.loaded) return;
    tb.loaded = true;
    autotune_cache_load(device, tb.picks);
}
void tuned_set(int device, uint64_t key, const LdeTiles &t, uint32_t us) {
    TunedTable &tb = tuned_table(device);
    std::lock_guard<std::mutex> lk(tb.mu);
    tb.picks[key] = {t, us};
}
void tuned_erase(int device, uint64_t key) {
    TunedTable &tb = tuned_table(device);
    std::lock_guard<std::mutex> lk(tb.mu);
    tb.picks.erase(key);
}
bool tuned_find(int device, uint64_t key, TunedPick *out) {
    TunedTable &tb = tuned_table(device);
    std::lock_guard<std::mutex> lk(tb.mu);
    tuned_load_locked(tb, device);
    auto it = tb.picks.find(key);
    if (it == tb.picks.end()) return false;
    *out = it->second;
    return true;
}

void autotune_cache_store(int device, uint64_t key, const LdeTiles &t, uint32_t us) {
    const std::string path = autotune_cache_path(), dev = device_key(device);
    if (path.empty() || dev.empty()) return;
    for (size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1))
        (void)mkdir(path.substr(0, p).c_str(), 0755);
    char line[512];
    const int n = snprintf(line, sizeof line, "%s %llx %u %u %u %u %u %u %u\n", dev.c_str(), (unsigned long long)key,
                           t.role[0].elems_log, t.role[0].cols_per_wg, t.role[1].elems_log, t.role[1].cols_per_wg,
                           t.role[2].elems_log, t.role[2].cols_per_wg, us);
    const int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_APPEND, 0644);
    if (fd < 0) return;
    if (n > 0 && n < (int)sizeof line) (void)!write(fd, line, (size_t)n);  // one append per line
    close(fd);
}
}  // namespace

// The launch shapes of an LDE: the pick tuned for this shape (columns rounded up
// to a power of two), else the pick of the nearest tuned column count with the
// same log sizes, else the defaults.
LdeTiles zs::lde_tiles_for(int device, uint32_t n_cols, uint32_t log_n, uint32_t log_blowup) {
    const uint64_t key = lde_shape_key(n_cols, log_n, log_blowup);
    TunedTable &tb = tuned_table(device);
    std::lock_guard<std::mutex> lk(tb.mu);
    tuned_load_locked(tb, device);
    auto it = tb.picks.find(key);
    if (it != tb.picks.end()) return it->second.t;
    const TunedPick *near = nullptr;
    int best = 1 << 30;
    for (auto lo = tb.picks.lower_bound(key & ~uint64_t(0xff)); lo != tb.picks.end() && (lo->first >> 8) == (key >> 8); ++lo) {
        const int d = std::abs((int)(lo->first & 0xff) - (int)(key & 0xff));
        if (d <= best) {  // ties: the larger column count
            best = d;
            near = &lo->second;
        }
    }
    return near ? near->t : default_lde_tiles();
}

extern "C" {

// Times each candidate configuration with per-kernel events and keeps, for each
// kind of LDE pass, the candidate under which that pass's kernels ran fastest: the
// passes exchange data in one fixed layout, so their choices are independent.
// A shape tuned by an earlier process on the same device is read from the
// autotune file instead.
int32_t zs_autotune_lde(zs_hip_ctx *h, uint32_t n_cols, uint32_t log_n, uint32_t log_blowup, uint32_t out[7]) {
    if (!h || n_cols == 0 || log_n < 1 || log_sum_over(log_n, log_blowup, 28)) return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipSetDevice(ctx->device));
    {
        // tuned before, in this process or (persisted) by an earlier one
        TunedPick cached;
        if (tuned_find(ctx->device, lde_shape_key(n_cols, log_n, log_blowup), &cached)) {
            if (out) {
                for (int k = 0; k < 3; k++) {
                    out[2 * k] = cached.t.role[k].elems_log;
                    out[2 * k + 1] = cached.t.role[k].cols_per_wg;
                }
                out[6] = cached.us;
            }
            return ZS_OK;
        }
    }
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t LT = log_n + log_blowup;
    const size_t n = size_t(1) << log_n, N = size_t(1) << LT;
    DevBuf cols, lde, nodes;
    ZS_TRY(dev_alloc(ctx, cols, n * n_cols * 4));
    ZS_TRY(dev_alloc(ctx, lde, N * n_cols * 4));
    ZS_TRY(dev_alloc(ctx, nodes, tree_words(LT) * 4));
    ZS_TRY(launch_fill_m31(ctx, 0x5eed, cols.as<uint32_t>(), n * n_cols));
    const uint64_t key = lde_shape_key(n_cols, log_n, log_blowup);
    // the caller's profiling state is set aside while the candidates are timed
    const bool was_profiling = ctx->profiling;
    std::vector<ProfRec> saved_prof;
    saved_prof.swap(ctx->prof);
    LdeTiles best = default_lde_tiles();
    float best_ms[3] = {0, 0, 0};
    uint8_t ref_root[32];
    bool have_ref = false;
    int32_t rc = ZS_OK;
    // the candidate being timed is installed as this shape's pick
    auto set_tuned = [&](const LdeTiles &t) { tuned_set(ctx->device, key, t, 0); };
    for (uint32_t el = 11; el <= 14 && rc == ZS_OK; el++) {
        for (uint32_t cpw = 1; cpw <= 32 && rc == ZS_OK; cpw *= 2) {
            if (cpw > 1 && cpw / 2 >= n_cols) continue;  // the same launch as cpw / 2
            TileConfig cand;
            cand.elems_log = el;
            cand.cols_per_wg = cpw;
            LdeTiles all;
            for (auto &r : all.role) r = cand;
            set_tuned(all);
            // one run for the root (every candidate must give the same one), then the timed runs
            ctx->profiling = false;
            rc = zs_m31_lde_commit_dev(h, cols.as<uint32_t>(), n_cols, log_n, log_blowup, lde.as<uint32_t>(),
                                       nodes.as<uint8_t>(), 0);
            if (rc == ZS_ERR_INPUT) {  // a configuration this shape cannot launch (LDS): not a candidate
                rc = ZS_OK;
                continue;
            }
            if (rc) break;
            uint8_t root[32];
            if (hipMemcpyAsync(root, nodes.as<uint8_t>() + (tree_words(LT) - 8) * 4, 32, hipMemcpyDeviceToHost,
                               ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess) {
                rc = ZS_ERR_DEVICE;
                break;
            }
            if (!have_ref) {
                std::memcpy(ref_root, root, 32);
                have_ref = true;
            } else if (std::memcmp(ref_root, root, 32) != 0) {
                ctx->set_error("autotune: launch configurations disagree on the LDE root");
                rc = ZS_ERR_GENERIC;
                break;
            }
            const int reps = 3;
            ctx->profiling = true;
            for (int r = 0; r < reps && rc == ZS_OK; r++)
                rc = zs_m31_lde_dev(h, cols.as<uint32_t>(), n_cols, log_n, log_blowup, lde.as<uint32_t>());
            ctx->profiling = false;
            if (rc || hipStreamSynchronize(ctx->stream) != hipSuccess) {
                rc = rc ? rc : ZS_ERR_DEVICE;
                break;
            }
            float role_ms[3] = {0, 0, 0};
            for (auto &r : ctx->prof) {
                float t = 0;
                if (r.role >= 0 && r.role < 3 && hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) role_ms[r.role] += t;
            }
            prof_clear(ctx);
            for (int k = 0; k < 3; k++) {
                if (!best_ms[k] || role_ms[k] < best_ms[k]) {
                    best_ms[k] = role_ms[k];
                    best.role[k] = cand;
                }
            }
        }
    }
    // the chosen combination, timed as a whole
    float total_ms = 0;
    if (rc == ZS_OK) {
        set_tuned(best);
        // no early return from here on: the caller's profiling state is restored below
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            ctx->set_error("autotune: hipEventCreate failed");
            rc = ZS_ERR_DEVICE;
        }
        const int reps = 3;
        if (rc == ZS_OK) rc = zs_m31_lde_dev(h, cols.as<uint32_t>(), n_cols, log_n, log_blowup, lde.as<uint32_t>());
        if (rc == ZS_OK) {
            (void)hipEventRecord(a, ctx->stream);
            for (int r = 0; r < reps && rc == ZS_OK; r++)
                rc = zs_m31_lde_dev(h, cols.as<uint32_t>(), n_cols, log_n, log_blowup, lde.as<uint32_t>());
            (void)hipEventRecord(b, ctx->stream);
            if (rc == ZS_OK && hipEventSynchronize(b) == hipSuccess) {
                (void)hipEventElapsedTime(&total_ms, a, b);
                total_ms /= reps;
            } else if (rc == ZS_OK) {
                rc = ZS_ERR_DEVICE;
            }
        }
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    prof_clear(ctx);
    ctx->prof.swap(saved_prof);
    ctx->profiling = was_profiling;
    if (rc == ZS_OK) tuned_set(ctx->device, key, best, (uint32_t)(total_ms * 1000.0f));
    else tuned_erase(ctx->device, key);
    if (rc == ZS_OK) autotune_cache_store(ctx->device, key, best, (uint32_t)(total_ms * 1000.0f));
    if (rc == ZS_OK && out) {
        for (int k = 0; k < 3; k++) {
            out[2 * k] = best.role[k].elems_log;
            out[2 * k + 1] = best.role[k].cols_per_wg;
        }
        out[6] = (uint32_t)(total_ms * 1000.0f);
    }
    return rc;
}

int32_t zs_m31_row_leaves_dev(zs_hip_ctx *h, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_rows,
                              size_t col_stride, uint8_t *d_leaves) {
    if (!h || !d_cols || !d_leaves || n_cols == 0 || log_rows > 28 || col_stride < (size_t(1) << log_rows))
        return ZS_ERR_INPUT;
    return launch_row_leaves(&h->c, d_cols, col_stride, n_cols, uint64_t(1) << log_rows,
                             reinterpret_cast<uint32_t *>(d_leaves));
}

int32_t zs_m31_row_commit_dev(zs_hip_ctx *h, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_rows,
                              size_t col_stride, uint8_t *d_nodes, uint32_t flags) {
    if (!h || !d_cols || !d_nodes || n_cols == 0 || log_rows > 28 || col_stride < (size_t(1) << log_rows))
        return ZS_ERR_INPUT;
    if (flags & ZS_FLAG_PLAIN)
        return launch_row_commit_plain(&h->c, d_cols, col_stride, n_cols, log_rows,
                                       reinterpret_cast<uint32_t *>(d_nodes));
    return launch_row_commit(&h->c, d_cols, col_stride, n_cols, log_rows, reinterpret_cast<uint32_t *>(d_nodes));
}

int32_t zs_m31_lde_commit(zs_hip_ctx *h, const uint32_t *cols, uint32_t n_cols, uint32_t log_n, uint32_t log_blowup,
                          zs_tree **out_tree, uint8_t root[32], uint32_t *lde_out) {
    if (!h || !cols || n_cols == 0 || log_n < 1 || log_sum_over(log_n, log_blowup, 28)) return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t LT = log_n + log_blowup;
    size_t n = size_t(1) << log_n, N = size_t(1) << LT;
    DevBuf d_cols, d_lde;
    auto *t = new (std::nothrow) zs_tree();
    if (!t) return ZS_ERR_OOM;
    t->ctx = ctx;
    t->log_leaves = LT;
    int32_t rc = dev_alloc(ctx, d_cols, n * n_cols * 4);
    if (!rc) rc = dev_alloc(ctx, d_lde, N * n_cols * 4);
    if (!rc) rc = dev_alloc(ctx, t->nodes, tree_words(LT) * 4);
    if (!rc && hipMemcpyAsync(d_cols.ptr, cols, n * n_cols * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = ZS_ERR_DEVICE;
    if (!rc)
        rc = zs_m31_lde_commit_dev(h, d_cols.as<uint32_t>(), n_cols, log_n, log_blowup, d_lde.as<uint32_t>(),
                                   t->nodes.as<uint8_t>(), 0);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ZS_ERR_DEVICE;
    if (!rc && hipMemcpyAsync(t->root, t->nodes.as<uint8_t>() + (tree_words(LT) - 8) * 4, 32, hipMemcpyDeviceToHost,
                              ctx->stream) != hipSuccess)
        rc = ZS_ERR_DEVICE;
    if (!rc && lde_out &&
        hipMemcpyAsync(lde_out, d_lde.ptr, N * n_cols * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = ZS_ERR_DEVICE;
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ZS_ERR_DEVICE;
    if (rc) {
        delete t;
        return rc;
    }
    if (root) std::memcpy(root, t->root, 32);
    if (out_tree) *out_tree = t;
    else delete t;
    return ZS_OK;
}

// ---- trees -----------------------------------------------------------------------
uint32_t zs_tree_log_leaves(const zs_tree *t) { return t ? t->log_leaves : 0; }

int32_t zs_tree_root(const zs_tree *t, uint8_t out[32]) {
    if (!t || !out) return ZS_ERR_INPUT;
    std::memcpy(out, t->root, 32);
    return ZS_OK;
}

int32_t zs_merkle_open(const zs_tree *t, size_t index, uint8_t *path_out) {
    if (!t || !path_out || index >= (size_t(1) << t->log_leaves)) return ZS_ERR_INPUT;
    if (t->log_leaves == 0) return ZS_OK;
    Ctx *ctx = t->ctx;
    DevBuf out;
    ZS_TRY(dev_alloc(ctx, out, size_t(t->log_leaves) * 32));
    std::vector<PathReq> reqs{{0, 0, t->log_leaves, (uint32_t)index}};
    ZS_TRY(launch_gather_paths(ctx, t->nodes.as<uint32_t>(), reqs, out.as<uint32_t>()));
    ZS_HIP(ctx, hipMemcpyAsync(path_out, out.ptr, size_t(t->log_leaves) * 32, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

int32_t zs_tree_read_level(const zs_tree *t, uint32_t level, uint8_t *out) {
    if (!t || !out || level > t->log_leaves) return ZS_ERR_INPUT;
    Ctx *ctx = t->ctx;
    size_t cnt = (size_t(1) << t->log_leaves) >> level;
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(out, t->nodes.as<uint32_t>() + level_offset_words(t->log_leaves, level), cnt * 32,
                          hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

void zs_tree_free(zs_tree *t) { delete t; }

// ---- FRI ---------------------------------------------------------------------------
}  // extern "C"

// univariate.zig:79-139 (O(n^2) Lagrange; the interpolant is unique)
void zs::interpolate_host(const std::vector<qm31> &xs, const std::vector<qm31> &ys, std::vector<qm31> &coeffs) {
    size_t n = xs.size();
    const qm31 zero{{0, 0}, {0, 0}}, one{{1, 0}, {0, 0}};
    std::vector<qm31> p(n + 1, zero), acc(n, zero), tmp(n, zero);
    p[0] = one;
    size_t plen = 1;
    for (size_t i = 0; i < n; i++) {
        for (size_t j = plen; j > 0; j--) p[j] = qm_sub(p[j - 1], qm_mul(p[j], xs[i]));
        p[0] = qm_mul(qm_sub(zero, p[0]), xs[i]);
        plen++;
    }
    for (size_t i = 0; i < n; i++) {
        qm31 w = one;
        for (size_t j = 0; j < n; j++)
            if (i != j) w = qm_mul(w, qm_sub(xs[i], xs[j]));
        qm31 sc = qm_mul(ys[i], qm_inv(w));
        tmp[n - 1] = p[n];
        for (size_t j = n - 1; j > 0; j--) tmp[j - 1] = qm_add(p[j], qm_mul(xs[i], tmp[j]));
        for (size_t k = 0; k < n; k++) acc[k] = qm_add(acc[k], qm_mul(tmp[k], sc));
    }
    coeffs = acc;
}

zs::DeferLayout zs::defer_layout(uint32_t dlog, uint32_t remainder_log) {
    DeferLayout d{};
    const uint32_t L = dlog - remainder_log;
    const uint32_t i_tail = std::min(L, dlog > FRI_TAIL_LOG ? dlog - FRI_TAIL_LOG : 0u);
    auto up = [](size_t w) { return (w + 63) & ~size_t(63); };  // 256-byte aligned parts
    d.alphas = up((sizeof(B3Hasher) + 3) / 4);
    d.layers = d.alphas + up(4 * size_t(L));
    size_t o = 0;
    for (uint32_t i = i_tail; i <= L && i - i_tail < 12; i++) {
        d.lay_off[i - i_tail] = o;
        o += size_t(4) << (dlog - i);
    }
    d.trees = d.layers + up(o);
    o = 0;
    for (uint32_t i = i_tail; i < L && i - i_tail < 11; i++) {
        d.tree_off[i - i_tail] = o;
        o += tree_words(dlog - i);
    }
    d.total = d.trees + up(o);
    return d;
}

int32_t zs::fri_commit_impl(zs_hip_ctx *h, const uint32_t *d_codeword, uint32_t dlog, uint32_t log_blowup,
                            uint32_t remainder_log, const uint32_t *alphas, B3Hasher *d_chan, uint32_t flags,
                            uint8_t *roots_out, uint32_t *remainder_out, zs_fri_result **out_result,
                            uint8_t *stage, const uint32_t *d_prefix, FriDefer *defer) {
    if (defer && (!d_chan || alphas || stage || out_result || !defer->region || !defer->results))
        return ZS_ERR_INPUT;
    if (!h || !d_codeword || (!stage && !defer && (!roots_out || !remainder_out)) || (stage && out_result) ||
        (d_prefix && !stage && !defer))
        return ZS_ERR_INPUT;
    if (dlog > 28 || remainder_log > dlog || remainder_log < log_blowup || dlog - remainder_log < 1)
        return ZS_ERR_INPUT;
    // the remainder is interpolated on the host in O(rc^2): at most kMaxRemainderLog
    // doublings above the blowup (64 coefficients)
    if (remainder_log - log_blowup > kMaxRemainderLog) return ZS_ERR_INPUT;
    if (!alphas && !d_chan) return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipSetDevice(ctx->device));
    auto *r = new (std::nothrow) zs_fri_result();
    if (!r) return ZS_ERR_OOM;
    std::unique_ptr<zs_fri_result> guard(r);
    r->ctx = ctx;
    r->dlog = dlog;
    r->L = dlog - remainder_log;
    r->log_blowup = log_blowup;
    r->remainder_log = remainder_log;
    const uint32_t L = r->L;
    size_t n0 = size_t(1) << dlog;
    size_t lw = 0, tw = 0;
    for (uint32_t i = 0; i <= L; i++) {
        r->layer_off.push_back(lw);
        lw += (n0 >> i) * 4;
    }
    for (uint32_t i = 0; i < L; i++) {
        r->tree_off.push_back(tw);
        tw += tree_words(dlog - i);
    }
    // the caller wrote layer 0's leaves where its tree goes (only into the scratch trees)
    const bool leaves0_given = (flags & kFriLeaves0Given) != 0;
    if (leaves0_given && out_result) return ZS_ERR_INPUT;
    // A kept result owns its layers (layer 0 copied, for the query phase) and trees.
    // Otherwise layer 0 is read in place and layers 1..L and the trees live in the
    // context's scratch.
    uint32_t *layers, *trees, *alpha_dev;
    const uint32_t *lay0;
    if (out_result) {
        ZS_TRY(dev_alloc(ctx, r->layers, lw * 4));
        ZS_TRY(dev_alloc(ctx, r->trees, tw * 4));
        layers = r->layers.as<uint32_t>();
        trees = r->trees.as<uint32_t>();
        ZS_HIP(ctx, hipMemcpyAsync(layers, d_codeword, n0 * 16, hipMemcpyDeviceToDevice, ctx->stream));
        lay0 = layers;
    } else {
        ZS_TRY(ctx->ensure_scratch(4, (lw - 4 * n0) * 4));
        ZS_TRY(ctx->ensure_scratch(5, tw * 4));
        layers = ctx->scratch[4].as<uint32_t>() - 4 * n0;  // layer i >= 1 at layers + layer_off[i]
        trees = ctx->scratch[5].as<uint32_t>();
        lay0 = d_codeword;
    }
    // layers of at most 2^FRI_TAIL_LOG values go to one single-workgroup launch
    const uint32_t i_tail = std::min(L, dlog > FRI_TAIL_LOG ? dlog - FRI_TAIL_LOG : 0u);
    // A deferred tail reads and writes only the caller's region: layer i_tail onwards,
    // the trees from layer i_tail's on, the challenges and the channel.  (With no chain
    // layer before the tail, or no tail, nothing is deferred: the tail, if any, runs here.)
    const bool deferring = defer && i_tail > 0 && i_tail < L;
    DeferLayout dl{};
    if (defer) dl = defer_layout(dlog, remainder_log);
    auto layer = [&](uint32_t i) -> uint32_t * {
        if (deferring && i >= i_tail) return defer->region + dl.layers + dl.lay_off[i - i_tail];
        return i == 0 ? const_cast<uint32_t *>(lay0) : layers + r->layer_off[i];
    };
    // leaves of layer i, where its tree starts
    auto leaves_of = [&](uint32_t i) -> uint32_t * {
        if (deferring && i >= i_tail) return defer->region + dl.trees + dl.tree_off[i - i_tail];
        return trees + r->tree_off[i];
    };
    // slot 6 (or the deferred region and results): the challenges (16 B per layer), 32 B
    // for a prefix digest, then the layer roots gathered for one copy (32 B per layer) and
    // right after them the remainder coefficients (16 B each)
    const uint32_t rc = 1u << (remainder_log - log_blowup);
    uint32_t *roots_dev, *rem_dev;
    if (defer) {
        alpha_dev = defer->region + dl.alphas;
        roots_dev = defer->results + 8;
    } else {
        ZS_TRY(ctx->ensure_scratch(6, size_t(L) * 48 + 32 + size_t(rc) * 16));
        alpha_dev = ctx->scratch[6].as<uint32_t>();
        roots_dev = alpha_dev + 4 * size_t(L) + 8;
    }
    rem_dev = roots_dev + 8 * size_t(L);
    const uint32_t *fri_tab, *rem_mat;
    ZS_TRY(get_fri_tables(ctx, dlog, &fri_tab));  // cached before any capture
    ZS_TRY(get_remainder_matrix(ctx, L, remainder_log, rc, &rem_mat));
    if (alphas)
        ZS_HIP(ctx, hipMemcpyAsync(alpha_dev, alphas, size_t(L) * 16, hipMemcpyHostToDevice, ctx->stream));
    // Leaves of layer 0, then per layer: node levels, (challenge), fold + next
    // layer's leaves.  With the challenges given, no fold waits for a root: the
    // folds run first and the trees of all layers are built together after them.
    auto tail_desc = [&]() -> FriTail {
        FriTail t{};
        t.in = layer(i_tail);
        t.layers = layers;
        t.trees = trees;
        t.tables = fri_tab;
        t.alphas = alpha_dev;
        t.chan = alphas ? nullptr : d_chan;
        t.dlog = dlog;
        t.i0 = i_tail;
        t.L = L;
        for (uint32_t i = i_tail + 1; i <= L; i++) t.lay_off[i] = r->layer_off[i];
        for (uint32_t i = i_tail; i < L; i++) {
            t.tree_off[i] = r->tree_off[i];
            t.inv_off[i] = qm_inv(fri_offset_pow(i));
        }
        if (deferring) {
            // everything the tail touches is in the region; it also writes its roots and
            // the remainder (the finish's work for those layers)
            t.layers = defer->region + dl.layers;
            t.trees = defer->region + dl.trees;
            for (uint32_t i = i_tail + 1; i <= L; i++) t.lay_off[i] = dl.lay_off[i - i_tail];
            for (uint32_t i = i_tail; i < L; i++) t.tree_off[i] = dl.tree_off[i - i_tail];
            t.roots = roots_dev;
            t.rem_mat = rem_mat;
            t.rc = rc;
            t.rem = rem_dev;
        }
        return t;
    };
    auto tail = [&]() -> int32_t {
        if (deferring) {
            defer->tail = tail_desc();
            defer->pending = true;
            return ZS_OK;
        }
        return launch_fri_tail(ctx, tail_desc());
    };
    // with the challenges given and two folds in the first launch, that launch hashes
    // layer 0's leaves too (it reads every layer-0 value)
    const bool leaves0_in_fold2 = alphas && i_tail >= 2;
    auto enqueue = [&]() -> int32_t {
        if (!leaves0_given && !leaves0_in_fold2) ZS_TRY(launch_qm31_leaves(ctx, layer(0), n0, trees));
        if (alphas) {
            std::vector<uint64_t> offs(L);
            std::vector<uint32_t> logs(L);
            // folds two layers per launch while both are below the tail (layer i+1 then
            // has leaves: i + 1 < i_tail <= L)
            for (uint32_t i = 0; i < i_tail;) {
                const uint32_t li = dlog - i;
                if (i + 1 < i_tail) {
                    uint32_t *leaves2 = (i + 2 < L) ? trees + r->tree_off[i + 2] : nullptr;
                    uint32_t *leaves0 = (i == 0 && !leaves0_given) ? trees : nullptr;
                    ZS_TRY(launch_fri_fold2(ctx, layer(i), size_t(1) << li, dlog, i, alpha_dev + 4 * i, leaves0,
                                            layer(i + 1), trees + r->tree_off[i + 1], layer(i + 2), leaves2));
                    i += 2;
                } else {
                    uint32_t *next_leaves = (i + 1 < L) ? trees + r->tree_off[i + 1] : nullptr;
                    ZS_TRY(launch_fri_fold(ctx, layer(i), size_t(1) << li, dlog, i, alpha_dev + 4 * i, layer(i + 1),
                                           next_leaves));
                    i += 1;
                }
            }
            for (uint32_t i = 0; i < L; i++) {
                offs[i] = r->tree_off[i];
                logs[i] = dlog - i;
            }
            if (i_tail < L) ZS_TRY(tail());
            return launch_trees_multi(ctx, trees, offs.data(), logs.data(), L);
        }
        // A layer's fold into the next one goes into the next tree's first launch when
        // that is a workgroup-subtree launch (fewer than 2^kSubtreeMinLog leaves) and the
        // next tree is built here (not in the tail): one launch fewer per layer, and the
        // leaves are hashed where the tree reads them.
        bool folded = false;  // layer i's values and leaves come from its tree's first launch
        for (uint32_t i = 0; i < i_tail; i++) {
            uint32_t li = dlog - i;
            FoldLeaves fl{};
            if (folded) {
                fl.cur = layer(i - 1);
                fl.half = size_t(1) << li;
                fl.layer = i - 1;
                fl.tables = fri_tab;
                fl.alpha = alpha_dev + 4 * (i - 1);
                fl.inv_off = qm_inv(fri_offset_pow(i - 1));
                fl.next = layer(i);
            }
            // the launch that computes this layer's root also draws its challenge on the device
            uint64_t off = r->tree_off[i];
            ZS_TRY(launch_trees_multi(ctx, trees, &off, &li, 1, d_chan, alpha_dev + 4 * i, nullptr, nullptr, 0,
                                      folded ? &fl : nullptr));
            folded = i + 1 < i_tail && li - 1 < kSubtreeMinLog;
            if (folded) continue;
            uint32_t *next_leaves = (i + 1 < L) ? leaves_of(i + 1) : nullptr;
            ZS_TRY(launch_fri_fold(ctx, layer(i), size_t(1) << li, dlog, i, alpha_dev + 4 * i, layer(i + 1),
                                   next_leaves));
        }
        return i_tail < L ? tail() : ZS_OK;
    };
    // every layer's root into roots_dev and the remainder right after them (fri.zig:228-233:
    // the interpolant of the first rc points of layer L), so the host reads both with one copy
    auto enqueue_all = [&]() -> int32_t {
        ZS_TRY(enqueue());
        // a deferred tail gathers its own roots and computes the remainder
        const uint32_t n_roots = deferring ? i_tail : L;
        std::vector<uint64_t> root_off(n_roots);
        for (uint32_t i = 0; i < n_roots; i++) root_off[i] = r->tree_off[i] + tree_words(dlog - i) - 8;
        return launch_fri_finish(ctx, trees, root_off.data(), n_roots, d_prefix, roots_dev, layer(L), rem_mat,
                                 deferring ? 0 : rc, rem_dev);
    };
    if (defer) {  // queued only: the caller runs the tail (if pending) and reads the results
        defer->pending = false;
        return enqueue_all();
    }
    const bool graph = !out_result && !ctx->profiling && (flags & ZS_FLAG_GRAPH) && !(flags & ZS_FLAG_NO_GRAPH);
    if (!graph) {
        ZS_TRY(enqueue_all());
    } else {
        std::array<uint64_t, 7> key = {reinterpret_cast<uint64_t>(lay0), reinterpret_cast<uint64_t>(layers),
                                       reinterpret_cast<uint64_t>(trees), reinterpret_cast<uint64_t>(alpha_dev),
                                       reinterpret_cast<uint64_t>(d_chan),
                                       (uint64_t(leaves0_given) << 48) | (uint64_t(dlog) << 32) | (uint64_t(L) << 16) | rc,
                                       reinterpret_cast<uint64_t>(d_prefix)};
        auto it = ctx->fri_graphs.find(key);
        if (it == ctx->fri_graphs.end()) {
            if (ctx->fri_graphs.size() >= 32) {
                ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
                for (auto &g : ctx->fri_graphs) (void)hipGraphExecDestroy(g.second);
                ctx->fri_graphs.clear();
            }
            if (!ctx->capture_stream)
                ZS_HIP(ctx, hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
            hipStream_t saved = ctx->stream;
            ctx->stream = ctx->capture_stream;
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal);
            int32_t rc = e == hipSuccess ? enqueue_all() : ZS_ERR_DEVICE;
            hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(ctx->capture_stream, &g) : e;
            ctx->stream = saved;
            if (rc != ZS_OK || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (rc == ZS_OK) ctx->set_error(std::string("FRI graph capture: ") + hipGetErrorString(e2));
                return rc != ZS_OK ? rc : ZS_ERR_DEVICE;
            }
            hipGraphExec_t ex = nullptr;
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ctx->set_error(std::string("FRI graph instantiate: ") + hipGetErrorString(e));
                return ZS_ERR_DEVICE;
            }
            it = ctx->fri_graphs.emplace(key, ex).first;
        }
        ZS_HIP(ctx, hipGraphLaunch(it->second, ctx->stream));
    }
    if (stage) {  // the caller synchronises later and reads (the prefix,) the roots, then the remainder
        const size_t pre = d_prefix ? 8 : 0;
        ZS_HIP(ctx, hipMemcpyAsync(stage, roots_dev - pre, pre * 4 + size_t(L) * 32 + size_t(rc) * 16,
                                   hipMemcpyDeviceToHost, ctx->stream));
        return ZS_OK;
    }
    // roots and remainder are contiguous on the device: one copy into the context's
    // pinned buffer (two copies into pageable memory cost ~35 us of host staging)
    const size_t out_bytes = size_t(L) * 32 + size_t(rc) * 16;
    ZS_TRY(ctx->ensure_pinned(out_bytes));
    const uint8_t *pin = static_cast<const uint8_t *>(ctx->pinned);
    ZS_HIP(ctx, hipMemcpyAsync(ctx->pinned, roots_dev, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::memcpy(roots_out, pin, size_t(L) * 32);
    std::memcpy(remainder_out, pin + size_t(L) * 32, size_t(rc) * 16);
    if (out_result) *out_result = guard.release();
    return ZS_OK;
}

extern "C" {

int32_t zs_fri_commit_dev(zs_hip_ctx *h, const uint32_t *d_codeword, uint32_t dlog, uint32_t log_blowup,
                          uint32_t remainder_log, const uint32_t *alphas, zs_channel *chan, uint32_t flags,
                          uint8_t *roots_out, uint32_t *remainder_out, zs_fri_result **out_result) {
    if (!h || (!alphas && !chan)) return ZS_ERR_INPUT;
    flags &= ZS_FLAG_NO_GRAPH | ZS_FLAG_GRAPH;  // the only flags of this entry point (the others are internal)
    if (alphas)
        return fri_commit_impl(h, d_codeword, dlog, log_blowup, remainder_log, alphas, nullptr, flags, roots_out,
                               remainder_out, out_result);
    // the channel travels to the device, draws every challenge there, and comes back
    Ctx *ctx = &h->c;
    ZS_TRY(ctx->ensure_scratch(7, sizeof(B3Hasher)));
    B3Hasher *d_chan = ctx->scratch[7].as<B3Hasher>();
    ZS_HIP(ctx, hipMemcpyAsync(d_chan, &chan->h, sizeof(B3Hasher), hipMemcpyHostToDevice, ctx->stream));
    ZS_TRY(fri_commit_impl(h, d_codeword, dlog, log_blowup, remainder_log, nullptr, d_chan, flags, roots_out,
                           remainder_out, out_result));
    ZS_HIP(ctx, hipMemcpyAsync(&chan->h, d_chan, sizeof(B3Hasher), hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // absorbMany(remainder) (fri.zig:228-233)
    size_t rc = size_t(1) << (remainder_log - log_blowup);
    uint8_t b[16];
    for (size_t k = 0; k < rc; k++) {
        for (int q = 0; q < 4; q++) {
            uint32_t v = remainder_out[4 * k + q];
            b[4 * q] = (uint8_t)v; b[4 * q + 1] = (uint8_t)(v >> 8);
            b[4 * q + 2] = (uint8_t)(v >> 16); b[4 * q + 3] = (uint8_t)(v >> 24);
        }
        zs_channel_absorb(chan, b, 16);
    }
    return ZS_OK;
}

int32_t zs_fri_commit(zs_hip_ctx *h, const uint32_t *codeword, uint32_t dlog, uint32_t log_blowup,
                      uint32_t remainder_log, const uint32_t *alphas, zs_channel *chan, uint8_t *roots_out,
                      uint32_t *remainder_out, zs_fri_result **out_result) {
    if (!h || !codeword || dlog > 28) return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf d;
    size_t bytes = (size_t(1) << dlog) * 16;
    ZS_TRY(dev_alloc(ctx, d, bytes));
    ZS_HIP(ctx, hipMemcpyAsync(d.ptr, codeword, bytes, hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return zs_fri_commit_dev(h, d.as<uint32_t>(), dlog, log_blowup, remainder_log, alphas, chan, 0, roots_out,
                             remainder_out, out_result);
}

int32_t zs_fri_query(zs_fri_result *r, zs_channel *chan, size_t num_queries, size_t *idx_out, uint32_t *vals_out,
                     uint8_t *paths_out) {
    if (!r || !chan || (num_queries && (!idx_out || !vals_out || !paths_out))) return ZS_ERR_INPUT;
    Ctx *ctx = r->ctx;
    const uint32_t L = r->L, dlog = r->dlog;
    size_t n0 = size_t(1) << dlog;
    size_t per_q_digests = 0;
    for (uint32_t i = 0; i < L; i++) per_q_digests += 2 * size_t(dlog - i);
    std::vector<PathReq> preqs;
    std::vector<uint64_t> vreqs;
    for (size_t q = 0; q < num_queries; q++) {
        size_t p = zs_channel_sample_index(chan, n0);  // fri.zig:241
        idx_out[q] = p;
        uint64_t pout = q * per_q_digests * 8;
        for (uint32_t i = 0; i < L; i++) {
            size_t ni = n0 >> i, nn = ni / 2, pn = p % nn;
            uint64_t vout = ((q * L + i) * 2) * 4;
            vreqs.push_back(r->layer_off[i] + pn * 4);
            vreqs.push_back(vout);
            vreqs.push_back(r->layer_off[i] + (pn + nn) * 4);
            vreqs.push_back(vout + 4);
            uint32_t li = dlog - i;
            preqs.push_back({r->tree_off[i], pout, li, (uint32_t)pn});
            pout += uint64_t(li) * 8;
            preqs.push_back({r->tree_off[i], pout, li, (uint32_t)(pn + nn)});
            pout += uint64_t(li) * 8;
            p = pn;
        }
    }
    if (num_queries == 0) return ZS_OK;
    DevBuf dv, dp;
    ZS_TRY(dev_alloc(ctx, dv, num_queries * L * 2 * 16));
    ZS_TRY(dev_alloc(ctx, dp, num_queries * per_q_digests * 32));
    ZS_TRY(launch_gather_words(ctx, r->layers.as<uint32_t>(), vreqs, 4, dv.as<uint32_t>()));
    ZS_TRY(launch_gather_paths(ctx, r->trees.as<uint32_t>(), preqs, dp.as<uint32_t>()));
    ZS_HIP(ctx, hipMemcpyAsync(vals_out, dv.ptr, num_queries * L * 2 * 16, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(paths_out, dp.ptr, num_queries * per_q_digests * 32, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

// Fri.Proof in the wire format of src/core/serialization.zig (docs/wire.md):
// unsigned ints little-endian (usize as 8 bytes), digests verbatim, slices as a
// u64 length then their elements, structs as their fields in order, QM31 as
// its four u32 words.
namespace {
struct WireWriter {
    std::vector<uint8_t> b;
    void bytes(const void *d, size_t n) {
        const uint8_t *p = static_cast<const uint8_t *>(d);
        b.insert(b.end(), p, p + n);
    }
    void uint(uint64_t v, unsigned n) {
        for (unsigned k = 0; k < n; k++) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void qm31(const uint32_t *v) {
        for (int k = 0; k < 4; k++) uint(v[k], 4);
    }
};
}  // namespace

int32_t zs_fri_prove_dev(zs_hip_ctx *h, const uint32_t *d_codeword, uint32_t log_size, uint32_t log_blowup,
                         uint32_t remainder_log, size_t num_queries, zs_channel *chan, zs_host_allocator host,
                         uint8_t **proof_out, size_t *proof_len) {
    if (!h || !d_codeword || !chan || !proof_out || !proof_len || !host.alloc || log_size > 255 ||
        log_blowup > 255 || remainder_log > 255)
        return ZS_ERR_INPUT;
    const uint32_t dlog = log_size + log_blowup;
    if (dlog > 28 || remainder_log < log_blowup || remainder_log >= dlog) return ZS_ERR_INPUT;
    const uint32_t L = dlog - remainder_log;
    const size_t rc = size_t(1) << (remainder_log - log_blowup);
    std::vector<uint8_t> roots(size_t(L) * 32);
    std::vector<uint32_t> rem(4 * rc);
    zs_fri_result *res = nullptr;
    ZS_TRY(zs_fri_commit_dev(h, d_codeword, dlog, log_blowup, remainder_log, nullptr, chan, 0, roots.data(),
                             rem.data(), &res));
    std::unique_ptr<zs_fri_result, void (*)(zs_fri_result *)> guard(res, zs_fri_result_free);
    size_t per_q = 0;
    for (uint32_t i = 0; i < L; i++) per_q += 2 * size_t(dlog - i) * 32;
    std::vector<size_t> idx(num_queries);
    std::vector<uint32_t> vals(num_queries * L * 8);
    std::vector<uint8_t> paths(num_queries * per_q);
    ZS_TRY(zs_fri_query(res, chan, num_queries, idx.data(), vals.data(), paths.data()));
    WireWriter w;
    w.uint(log_size, 1);  // FriParams (fri.zig:38-46)
    w.uint(log_blowup, 1);
    w.uint(num_queries, 8);
    w.uint(remainder_log, 1);
    w.uint(L, 8);  // roots
    w.bytes(roots.data(), roots.size());
    w.uint(rc, 8);  // remainder
    for (size_t k = 0; k < rc; k++) w.qm31(&rem[4 * k]);
    w.uint(num_queries, 8);  // queries
    for (size_t q = 0; q < num_queries; q++) {
        w.uint(idx[q], 8);
        w.uint(L, 8);
        const uint8_t *pp = paths.data() + q * per_q;
        for (uint32_t i = 0; i < L; i++) {
            const size_t plen = dlog - i;
            w.qm31(&vals[8 * (q * L + i)]);
            w.qm31(&vals[8 * (q * L + i) + 4]);
            w.uint(plen, 8);
            w.bytes(pp, plen * 32);
            pp += plen * 32;
            w.uint(plen, 8);
            w.bytes(pp, plen * 32);
            pp += plen * 32;
        }
    }
    uint8_t *out = static_cast<uint8_t *>(host.alloc(host.ctx, w.b.size()));
    if (!out) return ZS_ERR_OOM;
    std::memcpy(out, w.b.data(), w.b.size());
    *proof_out = out;
    *proof_len = w.b.size();
    return ZS_OK;
}

int32_t zs_fri_read_layer(const zs_fri_result *r, uint32_t layer, uint32_t *out) {
    if (!r || !out || layer > r->L) return ZS_ERR_INPUT;
    Ctx *ctx = r->ctx;
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(out, r->layers.as<uint32_t>() + r->layer_off[layer], (size_t(1) << (r->dlog - layer)) * 16,
                          hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

void zs_fri_result_free(zs_fri_result *r) { delete r; }

int32_t zs_fill_m31_dev(zs_hip_ctx *h, uint64_t seed, uint32_t *d_out, size_t n) {
    if (!h || (!d_out && n)) return ZS_ERR_INPUT;
    return launch_fill_m31(&h->c, seed, d_out, n);
}

}  // extern "C"
