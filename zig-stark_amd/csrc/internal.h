// internal.h -- context, device buffers and the launch entry points shared by
// the kernel files and the C ABI (capi.cpp).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <array>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/zs_hip.h"
#include "b3_hasher.h"
#include "field.h"

namespace zs {

// HIP call checking: record the failure on the context and return ZS_ERR_DEVICE.
#define ZS_HIP(ctx, call)                                                              \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            (ctx)->set_error(std::string(#call) + ": " + hipGetErrorString(_e));       \
            return ZS_ERR_DEVICE;                                                      \
        }                                                                              \
    } while (0)

// Launch a hot kernel; when the context is profiling, bracket it with events
// on the launch stream so its duration can be read back (zs_hip_profile_read).
// The events only time: they skip the system-scope fences (no cache flush
// between the kernels they separate).
#define ZS_PROF_EVENT(e) (hipEventCreateWithFlags(&(e), hipEventDisableSystemFence) == hipSuccess)
#define ZS_LAUNCH(ctx, label, ...)                                                 \
    do {                                                                           \
        hipEvent_t _za = nullptr, _zb = nullptr;                                   \
        if ((ctx)->profiling && ZS_PROF_EVENT(_za))                                \
            (void)hipEventRecord(_za, (ctx)->stream);                              \
        hipLaunchKernelGGL(__VA_ARGS__);                                           \
        if (_za && ZS_PROF_EVENT(_zb)) {                                           \
            (void)hipEventRecord(_zb, (ctx)->stream);                              \
            (ctx)->prof.push_back({std::string(label), _za, _zb, (ctx)->lde_role}); \
        }                                                                          \
    } while (0)

#define ZS_TRY(expr)                 \
    do {                             \
        int32_t _rc = (expr);        \
        if (_rc != ZS_OK) return _rc; \
    } while (0)

// Loads and stores of device (global) memory.  Device bodies whose pointers come from
// descriptors in memory rather than from kernel arguments (the tasks of the persistent
// batch kernel) would otherwise get flat accesses, which also count against the LDS
// counter; in a kernel these are plain accesses.
#define ZS_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ uint32_t ldg(const uint32_t *p) { return *(const ZS_GLOBAL uint32_t *)p; }
__device__ __forceinline__ void stg(uint32_t *p, uint32_t v) { *(ZS_GLOBAL uint32_t *)p = v; }
typedef uint32_t zs_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 ldg4(const uint32_t *p) {
    const zs_u32x4 v = *(const ZS_GLOBAL zs_u32x4 *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg4(uint32_t *p, uint4 v) {
    const zs_u32x4 w = {v.x, v.y, v.z, v.w};
    *(ZS_GLOBAL zs_u32x4 *)p = w;
}

// A device allocation that frees itself.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool owned = true;  // false: a slice of memory that someone else allocated and frees
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (ptr && owned) (void)hipFree(ptr);
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

// Twiddle tables for one half-coset: heap layout, stage s of a transform of
// log size L at [2^(L-1-s), 2^(L-s)); forward values and their inverses.
struct TwiddleTable {
    uint32_t log_n = 0;
    DevBuf fwd, inv;
};

struct ProfRec {
    std::string name;
    hipEvent_t a, b;
    int role;  // LdeRole of the LDE pass that launched it, -1 outside an LDE
};

// Launch configuration of the register-group FFT passes: elements of the tile
// set one workgroup transforms (log2) and columns it runs with one set of staged
// twiddles.  Defaults come from the environment (ZS_TILE_ELEMS_LOG,
// ZS_COLS_PER_WG) or measurement; zs_autotune_lde picks one per LDE shape.
struct TileConfig {
    uint32_t elems_log = 13, cols_per_wg = 4;
};
TileConfig default_tile_config();
// An LDE's three kinds of pass (interpolation passes, the fused middle kernel,
// evaluation passes) move different data patterns, so each has its own configuration.
enum LdeRole { LDE_INTERP = 0, LDE_MID = 1, LDE_EVAL = 2 };
struct LdeTiles {
    TileConfig role[3];
};
LdeTiles default_lde_tiles();
// a tuned pick and the LDE time it was measured at (microseconds)
struct TunedPick {
    LdeTiles t;
    uint32_t us = 0;
};
// the launch shapes for an LDE on `device`: tuned, nearest tuned column count, or default
LdeTiles lde_tiles_for(int device, uint32_t n_cols, uint32_t log_n, uint32_t log_blowup);
// a + b > max for log sizes, without the wrap of a 32-bit sum
inline bool log_sum_over(uint32_t a, uint32_t b, uint32_t max) { return a > max || b > max || a + b > max; }
// key of an LDE shape in the tuned table: columns rounded up to a power of two
inline uint64_t lde_shape_key(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup) {
    uint32_t cl = 0;
    while ((1u << cl) < n_cols) cl++;
    return (uint64_t(log_n) << 16) | (uint64_t(log_blowup) << 8) | cl;
}

// Device-resident tables shared by the contexts of one device (a context and its
// batch workers, or a server's workers): built once, read by every stream.
struct DeviceCaches {
    std::mutex mu;
    using CosetKey = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>;
    std::map<CosetKey, std::unique_ptr<TwiddleTable>> twiddles;
    std::deque<CosetKey> other_cosets;  // non-canonic entries of `twiddles`, oldest first
    // FRI inverse-domain tables keyed by dlog: [lo table 2^10 | hi table] QM31 words
    std::map<uint32_t, std::unique_ptr<DevBuf>> fri_tables;
    // QM31 power tables of the trace-commitment DFTs, keyed by (kind << 32) | log length
    std::map<uint64_t, std::unique_ptr<DevBuf>> qtables;
    // FRI remainder interpolation matrices, keyed by (L << 32) | (remainder_log << 16) | coefficients
    std::map<uint64_t, std::unique_ptr<DevBuf>> remainder_mats;
    // a commit job's starting channel state (Channel.init("zig-stark:hip:commit")) on the device
    std::unique_ptr<DevBuf> commit_channel;
};

struct Ctx {
    int device = 0;
    bool own_stream = false;          // the stream was created for this context
    bool stream_ordered = false;      // scratch comes from the stream-ordered allocator
    std::vector<void *> workers;      // worker contexts (zs_hip_ctx *) of batched commits
    bool profiling = false;           // record HIP events around the hot kernels
    std::vector<ProfRec> prof;
    hipStream_t stream = nullptr;  // nullptr = the null stream
    std::string last_error;
    std::shared_ptr<DeviceCaches> caches = std::make_shared<DeviceCaches>();
    TileConfig tile = default_tile_config();  // used by the next transform launches
    LdeTiles lde_tiles = default_lde_tiles();  // used by the next LDE, per kind of pass
    int lde_role = -1;                         // the kind of LDE pass being launched (profiling)
    // grow-only scratch arenas
    DevBuf scratch[14];  // 0 transforms, 1-2 LDE, 3 plain path, 4-6 FRI, 7-9 commit jobs, 10-11 persistent batch,
                         // 12-13 a batch's deferred FRI tails
    // FRI commit phases with challenges given up front, captured once per buffer set and
    // replayed (key: codeword, layers, trees, challenges, channel, dlog / layer count / remainder length)
    hipStream_t capture_stream = nullptr;
    std::map<std::array<uint64_t, 7>, hipGraphExec_t> fri_graphs;

    // grow-only pinned host buffer (staged results of batched jobs)
    void *pinned = nullptr;
    size_t pinned_bytes = 0;

    void set_error(const std::string &s) { last_error = s; }
    int32_t ensure_pinned(size_t bytes);
    int32_t ensure_scratch(int slot, size_t bytes);
    void release_scratch();
    // use [ptr, ptr + bytes) as scratch slot `slot` until release_scratch (not freed here)
    void lend_scratch(int slot, void *ptr, size_t bytes);
};

// ---- the allocation-free LDE's fixed device array (noalloc.hip) ----
// once per device, with the device current: the array's address and its twiddle tables
void noalloc_setup(int device);

// ---- twiddles (fft.hip) ----
int32_t get_twiddles(Ctx *ctx, const zs_circle_coset &hc, const TwiddleTable **out);
int32_t get_canonic_twiddles(Ctx *ctx, uint32_t log_n, const TwiddleTable **out);
cpoint coset_at_host(const zs_circle_coset &c, uint64_t i);

// ---- FFT launches (fft.hip); all device pointers, column-major, async on ctx->stream ----
// Forward circle FFT of n_cols columns: A-then-B coefficients -> natural-order evaluations.
int32_t launch_fft(Ctx *ctx, const zs_circle_coset &hc, const uint32_t *d_in, uint64_t in_stride, uint32_t *d_out,
                   uint64_t out_stride, uint32_t n_cols);
// Inverse circle FFT: natural-order evaluations -> A-then-B coefficients.
int32_t launch_ifft(Ctx *ctx, const zs_circle_coset &hc, const uint32_t *d_in, uint64_t in_stride, uint32_t *d_out,
                    uint64_t out_stride, uint32_t n_cols);
// Twiddle tables of an LDE given by the caller (heap layout, see TwiddleTable): the inverse
// table of canonicHalf(log_n) and the forward table of canonicHalf(log_n + log_blowup).
struct LdeTables {
    const uint32_t *inv_small, *fwd_big;
};
// twiddle_gen_kernel of the half-coset into fwd / inv (2^(log_size+1) words each) on `stream`
hipError_t launch_twiddle_gen(hipStream_t stream, const zs_circle_coset &hc, uint32_t *fwd, uint32_t *inv);
// Blowup LDE on canonic cosets: columns of 2^log_n values -> 2^(log_n+log_blowup) values.
// d_coef: scratch of n_cols * 2^log_n words; d_big: scratch of n_cols * 2^(log_n+log_blowup) words
// (may be nullptr when the transform fits one pass).  tables: null to use the context's
// cached tables (built on first use), or the caller's (no device allocation at all).
int32_t launch_lde(Ctx *ctx, const uint32_t *d_in, uint64_t in_stride, uint32_t n_cols, uint32_t log_n,
                   uint32_t log_blowup, uint32_t *d_out, uint64_t out_stride, uint32_t *d_coef, uint32_t *d_big,
                   const LdeTables *tables = nullptr);
size_t lde_scratch_words(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup, size_t *coef_words);
// Reference-shaped per-column, per-layer path (one launch per butterfly layer).
int32_t launch_lde_plain(Ctx *ctx, const uint32_t *d_in, uint64_t in_stride, uint32_t n_cols, uint32_t log_n,
                         uint32_t log_blowup, uint32_t *d_out, uint64_t out_stride, uint32_t *d_tmp);
int32_t launch_fill_m31(Ctx *ctx, uint64_t seed, uint32_t *d_out, uint64_t n);

// ---- Merkle (merkle.hip) ----
// leaves = hashM31s(row) over n_cols columns at stride col_stride, then every
// node level; d_nodes holds (2 * 2^log_rows - 1) digests, leaves first.
int32_t launch_row_commit(Ctx *ctx, const uint32_t *d_cols, uint64_t col_stride, uint32_t n_cols, uint32_t log_rows,
                          uint32_t *d_nodes);
int32_t launch_row_commit_plain(Ctx *ctx, const uint32_t *d_cols, uint64_t col_stride, uint32_t n_cols,
                                uint32_t log_rows, uint32_t *d_nodes);
// ---- the persistent batch launch (persist.hip) ----
// shapes it takes; device bytes of its pool per trace
bool persistent_supported(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup);
size_t persistent_pool_bytes(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup, uint32_t remainder_log);
// the whole commitment of jobs[idx[..]] (LDE, row tree, gamma, combination, FRI) in one
// persistent launch on ctx->stream; synchronised, results and statuses set on return
int32_t persistent_commit(Ctx *ctx, zs_job *jobs, const std::vector<size_t> &idx, uint32_t log_blowup,
                          uint32_t remainder_log);
// the channel every commit job starts from, Channel.init("zig-stark:hip:commit"), on the device
int32_t device_commit_channel(Ctx *ctx, const B3Hasher **out);

int32_t launch_row_leaves(Ctx *ctx, const uint32_t *d_cols, uint64_t col_stride, uint32_t n_cols, uint64_t n,
                          uint32_t *d_nodes);
// node levels above an already written level 0 of 2^log_leaves digests; with
// draw_chan, the launch that computes the root also absorbs it into that device
// channel and draws a QM31 challenge into draw_alpha (log_leaves - from_level >= 1),
// the channel starting from chan_init when given; with powers, that launch also
// writes the challenge's powers 0 .. n_powers - 1 (QM31, 4 words each)
int32_t launch_tree_levels(Ctx *ctx, uint32_t *d_nodes, uint32_t log_leaves, uint32_t from_level,
                           B3Hasher *draw_chan = nullptr, uint32_t *draw_alpha = nullptr,
                           const B3Hasher *chan_init = nullptr, uint32_t *powers = nullptr, uint32_t n_powers = 0);
int32_t launch_qm31_leaves(Ctx *ctx, const uint32_t *d_vals, uint64_t n, uint32_t *d_nodes);
// Tree levels with at least 2^kSubtreeMinLog children go 3 at a time, one lane per
// 8-child subtree; narrower ones in workgroup subtrees of up to 8 levels.
constexpr uint32_t kSubtreeMinLog = 18;
// An FRI fold whose outputs are the leaves of a tree (fri.zig:211-219): leaf j is
// hashQM31(next[j]), next[j] = fold of cur[j] and cur[j + half] with the challenge at
// alpha; the tree's first workgroup-subtree launch computes them (and writes next and
// the leaves) instead of reading written leaves.
struct FoldLeaves {
    const uint32_t *cur;
    uint64_t half;
    uint32_t layer;
    const uint32_t *tables;  // get_fri_tables(dlog)
    const uint32_t *alpha;
    qm31 inv_off;            // (FRI_OFFSET^(2^layer))^-1
    uint32_t *next;
};
// node levels of n_trees independent trees with written leaves, tree k at base + off_words[k];
// with fold (one tree of fewer than 2^kSubtreeMinLog leaves), the leaves are not written
// yet: the first launch computes them from the fold
int32_t launch_trees_multi(Ctx *ctx, uint32_t *base, const uint64_t *off_words, const uint32_t *log_leaves,
                           uint32_t n_trees, B3Hasher *draw_chan = nullptr, uint32_t *draw_alpha = nullptr,
                           const B3Hasher *chan_init = nullptr, uint32_t *powers = nullptr, uint32_t n_powers = 0,
                           const FoldLeaves *fold = nullptr);
// word offsets of up to 28 tree roots (a gather's argument)
struct RootOffsets {
    uint64_t off[28];
};
// Merkle path gather: for each request, log_leaves sibling digests written at out_off_words.
struct PathReq {
    uint64_t tree_off_words;
    uint64_t out_off_words;
    uint32_t log_leaves;
    uint32_t index;
};
int32_t launch_gather_paths(Ctx *ctx, const uint32_t *d_trees, const std::vector<PathReq> &reqs, uint32_t *d_out);
// copies `words` words from src + pairs[2r] to out + pairs[2r+1] for each r; synchronous
int32_t launch_gather_words(Ctx *ctx, const uint32_t *d_src, const std::vector<uint64_t> &pairs, uint32_t words,
                            uint32_t *d_out);

// ---- FRI (fri.hip) ----
int32_t get_fri_tables(Ctx *ctx, uint32_t dlog, const uint32_t **out);
// next[j] = fold(D_layer[j], cur[j], cur[j + n/2], alpha); also writes the
// hashQM31 leaves of `next` into d_next_leaves when non-null.
int32_t launch_fri_fold(Ctx *ctx, const uint32_t *d_cur, uint64_t n, uint32_t dlog, uint32_t layer,
                        const uint32_t *d_alpha, uint32_t *d_next, uint32_t *d_next_leaves);
// one workgroup-subtree launch of a single tree (bt) whose leaves come from fold (fri.hip)
struct BlockTrees;
int32_t launch_tree_block_fold(Ctx *ctx, uint32_t *base, const BlockTrees &bt, uint32_t wgs, size_t lds_bytes,
                               const FoldLeaves &fold);
// folds of layers `layer` and layer + 1 in one launch (challenges at d_alpha, d_alpha + 4)
// (d_leaves0, when non-null, gets the hashQM31 leaves of layer `layer` itself)
int32_t launch_fri_fold2(Ctx *ctx, const uint32_t *d_cur, uint64_t n, uint32_t dlog, uint32_t layer,
                         const uint32_t *d_alpha, uint32_t *d_leaves0, uint32_t *d_next1, uint32_t *d_leaves1,
                         uint32_t *d_next2, uint32_t *d_leaves2);

// g[i] = sum_c powers[c] * lde_c[i]: QM31 combination of n_cols M31 columns of n values
// device Fiat-Shamir: absorb the digest at d_digest, write a QM31 challenge to d_out
int32_t launch_channel_draw(Ctx *ctx, B3Hasher *d_state, const uint32_t *d_digest, uint32_t *d_out);
// FRI layers i0..L-1 of at most 2^FRI_TAIL_LOG values in one workgroup (fri.hip)
constexpr uint32_t FRI_TAIL_LOG = 10;
struct FriTail {
    const uint32_t *in;      // layer i0 values (4 words per QM31)
    uint32_t *layers;        // layer i > i0 at layers + lay_off[i]
    uint32_t *trees;         // tree of layer i at trees + tree_off[i], leaves first (layer i0's written)
    const uint32_t *tables;  // get_fri_tables(dlog)
    uint32_t *alphas;        // challenge of layer i at alphas + 4 i: read, or written when chan is set
    B3Hasher *chan;          // device channel (the trees are built and the challenges drawn here), or null
    uint32_t dlog, i0, L;
    uint64_t lay_off[29], tree_off[28];
    qm31 inv_off[28];  // (FRI_OFFSET^(2^i))^-1
    // optional, channel mode: the root of layer i (i0 <= i < L) written to roots + 8 i
    uint32_t *roots;
    // optional: the remainder coefficients, rem_mat * (first rc values of layer L), to rem
    const uint32_t *rem_mat;
    uint32_t rc;
    uint32_t *rem;
};
int32_t launch_fri_tail(Ctx *ctx, const FriTail &t);
// the tails of several commit phases in one launch, one workgroup each (d_tails: n
// descriptors in device memory)
int32_t launch_fri_tail_multi(Ctx *ctx, const FriTail *d_tails, uint32_t n);
// A channel-mode commit phase whose tail (the layers of at most 2^FRI_TAIL_LOG values,
// the remainder) is left to a later launch that runs the tails of many commit phases
// at once.  Everything the tail reads or writes lives in `region` (fri_defer_words
// words: the channel, the challenges, the tail's layers and trees) and `results`
// (the stage layout: 8 words of prefix, 8 per root, 4 per remainder coefficient), not
// in the context's scratch, so the context may run other commit phases meanwhile.
struct FriDefer {
    uint32_t *region;   // in: device, fri_defer_words(dlog, remainder_log) words
    uint32_t *results;  // in: device, 8 + 8 L + 4 rc words
    bool pending;       // out: `tail` still has to run (false: results complete)
    FriTail tail;       // out
};
// A deferred region (word offsets): the channel state at 0, the challenges (4 words
// per layer), then layers i_tail..L and the trees of layers i_tail..L-1 (entry k is
// layer i_tail + k; i_tail = min(L, dlog - FRI_TAIL_LOG), where fri_commit_impl splits
// the tail off).
struct DeferLayout {
    size_t alphas, layers, trees, total;
    size_t lay_off[12], tree_off[11];
};
DeferLayout defer_layout(uint32_t dlog, uint32_t remainder_log);
inline size_t fri_defer_words(uint32_t dlog, uint32_t remainder_log) { return defer_layout(dlog, remainder_log).total; }
int32_t launch_qm31_powers(Ctx *ctx, const uint32_t *d_x, uint32_t *d_out, uint32_t n);
// UnivariateQM31.interpolate (univariate.zig:79-139) on the host, O(n^2) Lagrange
void interpolate_host(const std::vector<qm31> &xs, const std::vector<qm31> &ys, std::vector<qm31> &coeffs);
// the FRI remainder as a device map: coefficients = M * (first rc values of layer L)
int32_t get_remainder_matrix(Ctx *ctx, uint32_t L, uint32_t remainder_log, uint32_t rc, const uint32_t **out);
// FRI commit phase on a device codeword; challenges from `alphas` (host) or drawn
// on the device from the channel state at d_chan (left advanced past the last root).
// With `stage` (host memory, pinned for a true asynchronous copy) the L roots and
// then the remainder words are copied there on the stream and the call returns
// without synchronising (roots_out and remainder_out are then unused); otherwise
// they are in roots_out / remainder_out when it returns.
// d_prefix (with stage or defer only): 8 device words copied to the stage ahead of the
// roots (the commit job's trace root), in the same copy.
// With `defer` (channel mode; d_chan must be the region's channel) the call queues
// everything but the tail, writes the prefix and the roots it computed into
// defer->results and returns without synchronising or copying; defer->tail is then
// the tail to launch (launch_fri_tail_multi) before the results are read.
int32_t fri_commit_impl(zs_hip_ctx *h, const uint32_t *d_codeword, uint32_t dlog, uint32_t log_blowup,
                        uint32_t remainder_log, const uint32_t *alphas, B3Hasher *d_chan, uint32_t flags,
                        uint8_t *roots_out, uint32_t *remainder_out, zs_fri_result **out_result,
                        uint8_t *stage = nullptr, const uint32_t *d_prefix = nullptr, FriDefer *defer = nullptr);
// every layer root (tree word offsets root_off) into d_roots, d_prefix's 8 words (if
// given) at d_roots - 8, and the remainder (rc coefficients) into d_rem, in one launch
int32_t launch_fri_finish(Ctx *ctx, const uint32_t *d_trees, const uint64_t *root_off, uint32_t n_roots,
                          const uint32_t *d_prefix, uint32_t *d_roots, const uint32_t *d_last, const uint32_t *d_mat,
                          uint32_t rc, uint32_t *d_rem);
int32_t launch_combine(Ctx *ctx, const uint32_t *d_lde, uint64_t cs, uint32_t n_cols, uint64_t n,
                       const uint32_t *d_powers, uint32_t *d_out, uint32_t *d_leaves = nullptr);
// the same combination with its four M31 components written as planes: component k of row
// i at d_planes[k * n + i]
int32_t launch_combine_planes(Ctx *ctx, const uint32_t *d_cols, uint64_t cs, uint32_t n_cols, uint64_t n,
                              const uint32_t *d_powers, uint32_t *d_planes);
// four planes of n words -> n interleaved QM31 values (d_out[4 i + k] = d_planes[k * n + i]);
// d_leaves, when non-null, gets each value's hashQM31 leaf
int32_t launch_interleave4(Ctx *ctx, const uint32_t *d_planes, uint64_t n, uint32_t *d_out,
                           uint32_t *d_leaves = nullptr);
// fri_commit_impl flag (internal): layer 0's hashQM31 leaves are already at the start of
// scratch slot 5, where its tree goes; the commit phase does not hash them again
constexpr uint32_t kFriLeaves0Given = 1u << 30;

// FRI remainders have at most 2^kMaxRemainderLog coefficients (remainder_log - log_blowup <= 6)
constexpr uint32_t kMaxRemainderLog = 6;
// host-side constants
qm31 fri_offset_pow(uint32_t layer);  // FRI_OFFSET^(2^layer)
qm31 qm31_root_of_unity(uint32_t log_size);

inline size_t tree_words(uint32_t log_leaves) { return ((size_t(2) << log_leaves) - 1) * 8; }
// FRI commit-phase trees of a codeword of 2^dlog values, L layers: their total words
// (layer i's tree at the sum of the earlier layers' tree_words)
inline size_t fri_tree_words(uint32_t dlog, uint32_t L) {
    size_t tw = 0;
    for (uint32_t i = 0; i < L; i++) tw += tree_words(dlog - i);
    return tw;
}
inline size_t level_offset_words(uint32_t log_leaves, uint32_t level) {
    // digests before `level`: sum_{l<level} 2^(log_leaves - l)
    size_t n = size_t(1) << log_leaves;
    size_t off = 0;
    for (uint32_t l = 0; l < level; l++) off += n >> l;
    return off * 8;
}

}  // namespace zs

// ---- handle types of the C ABI (defined here so every translation unit can use them) ----
struct zs_hip_ctx {
    zs::Ctx c;
};

struct zs_tree {
    zs::Ctx *ctx = nullptr;
    zs::DevBuf nodes;
    uint32_t log_leaves = 0;
    uint8_t root[32];
};

struct zs_fri_result {
    zs::Ctx *ctx = nullptr;
    uint32_t dlog = 0, L = 0, log_blowup = 0, remainder_log = 0;
    zs::DevBuf layers;  // layer i at layer_off[i] words, 2^(dlog-i) QM31 values
    zs::DevBuf trees;   // tree i at tree_off[i] words, every level
    std::vector<size_t> layer_off, tree_off;
};

struct zs_channel {
    zs::B3Hasher h;
};
