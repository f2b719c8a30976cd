// persist.hip -- the LDE, row leaves and every Merkle node level of many traces of
// uneven size in one persistent launch on gfx950.
//
// The reference commits a trace's columns one at a time (src/m31/stark.zig:396-414)
// with per-layer launches on the GPU side (src/cuda/circlefft_gpu.zig:87-105).
// Here one launch of (workgroups per CU) x (CUs) workgroups of 512 threads runs the
// whole batch.  Each trace is a sequence of stages -- its LDE passes, its row leaves,
// then its node levels three or eight at a time -- and stage k of every trace forms
// phase k.  A phase is a list of tasks (one workgroup's share of one stage: the tile
// group and column block of a pass, 512 rows of leaves, 512 lanes of 3-level
// subtrees, one 8-level subtree); workgroups claim them with an atomic counter,
// largest stages first, then meet at a grid barrier before the next phase, so a
// stage reads only what earlier phases wrote.  The tasks run the same pass, leaf and
// tree bodies as the single-trace kernels (tile_dev.h, pass_dev.h, merkle_dev.h)
// with the same arithmetic, so every LDE value and node equals theirs.
//
// The launch holds no more workgroups than the device keeps resident at once (the
// occupancy query), so every workgroup reaches every barrier.  A barrier wait is
// bounded by the wall clock all the same: past the bound the workgroup raises an
// abort flag, every workgroup leaves at its next claim or barrier, and the host
// reports a device error instead of waiting on the launch.
#include <algorithm>
#include <cstring>
#include <vector>

#include "merkle_dev.h"
#include "pass_dev.h"
#include "tile_dev.h"

namespace zs {
namespace {

enum PKind : uint32_t {
    PK_INTERP_FIRST = 0,  // pipelined first interpolation pass (natural in, k-space out)
    PK_MID,               // fused interpolation top pass + scale + forward top pass of every coset
    PK_EVAL_LAST,         // pipelined last evaluation pass (k-space in, natural out)
    PK_PASS_INV,          // generic LDS pass, inverse
    PK_PASS_FWD,          // generic LDS pass, forward
    PK_LEAF,              // row leaves (<= 256 columns: one BLAKE3 chunk per row)
    PK_SUBTREE3,          // three node levels, one lane per 8-child subtree
    PK_BLOCK,             // up to eight node levels, one workgroup per subtree
};

constexpr uint32_t kPThreads = 512;
constexpr uint32_t kMaxPhases = 64;

struct LeafArgs {
    const uint32_t *cols;
    uint64_t cs, n_rows;
    uint32_t n_cols;
    uint32_t *nodes;
};
struct SubArgs {
    const uint32_t *child;
    uint32_t *lvl1;
    uint64_t lanes;
};
struct BlockArgs {
    uint32_t *base;
    OneTree bt;
};

struct PStage {
    uint32_t kind, C, gx, gy;  // virtual grid of a pass stage: x fastest, then y, then z
    uint64_t end;              // tasks of this phase's stages up to and including this one
    union {
        FftPass p;
        LdeFused f;
        LeafArgs leaf;
        SubArgs sub;
        BlockArgs blk;
    } u;
};

struct PCtl {
    uint32_t arrive;  // workgroups that reached a barrier, summed over the barriers
    uint32_t abort;   // a barrier wait ran past its bound: the results are incomplete
    unsigned long long claim[kMaxPhases];  // next task of each phase
};

struct PPlan {
    const PStage *stages;
    const uint32_t *phase_first;  // phase k: stages [phase_first[k], phase_first[k + 1])
    PCtl *ctl;
    uint64_t wait_ticks;  // a barrier waits at most this many wall-clock ticks
    uint32_t n_phases, n_wg;
};

__device__ __forceinline__ uint32_t ld_relaxed(uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every workgroup's writes so far are visible to every workgroup after it; k counts the
// barriers (1, 2, ...).  False when the wait gave up or another workgroup's did.
__device__ bool grid_barrier(const PPlan &pl, uint32_t k, uint32_t *s_ok) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&pl.ctl->arrive, 1u);
        const uint32_t want = pl.n_wg * k;
        const uint64_t t0 = wall_clock64();
        uint32_t ok = 1;
        while (ld_relaxed(&pl.ctl->arrive) < want) {
            if (ld_relaxed(&pl.ctl->abort)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > pl.wait_ticks) {
                __hip_atomic_store(&pl.ctl->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
        *s_ok = ok;
    }
    __syncthreads();
    __threadfence();
    return *s_ok != 0;
}

// A task's descriptor with every word made wave-uniform (scalar registers): the task
// functions below are called, not inlined -- one register allocation per body, as in
// the single-trace kernels, where inlining them all into one kernel spilled -- and a
// call passes its arguments in vector registers.
template <class T>
__device__ __forceinline__ T uniform_copy(const T &src) {
    static_assert(sizeof(T) % 4 == 0, "whole words");
    T out;
    const uint32_t *s = reinterpret_cast<const uint32_t *>(&src);
    uint32_t *d = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; i++) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
    return out;
}

__device__ __forceinline__ VB task_vb(const PStage &st, uint64_t u) {
    const uint32_t gx = __builtin_amdgcn_readfirstlane(st.gx), gy = __builtin_amdgcn_readfirstlane(st.gy);
    const uint32_t x = (uint32_t)(u % gx);
    const uint64_t r = u / gx;
    return VB{x, (uint32_t)(r % gy), (uint32_t)(r / gy), gx};
}
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    return (uint64_t(__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32))) << 32) |
           __builtin_amdgcn_readfirstlane((uint32_t)v);
}

#define ZS_TASK __device__ __attribute__((noinline))
template <int C>
ZS_TASK void task_interp_first(const PStage *sp, uint64_t u) {
    lde_interp_first_body<C>(uniform_copy(sp->u.p), task_vb(*sp, uniform_u64(u)));
}
template <int C>
ZS_TASK void task_mid(const PStage *sp, uint64_t u) {
    lde_mid_body<C>(uniform_copy(sp->u.f), task_vb(*sp, uniform_u64(u)));
}
template <int C>
ZS_TASK void task_eval_last(const PStage *sp, uint64_t u) {
    lde_eval_last_body<C>(uniform_copy(sp->u.p), task_vb(*sp, uniform_u64(u)));
}
template <bool INV>
ZS_TASK void task_pass(const PStage *sp, uint64_t u) {
    fft_pass_body<INV>(uniform_copy(sp->u.p), task_vb(*sp, uniform_u64(u)), threadIdx.x, kPThreads);
}
ZS_TASK void task_leaves(const PStage *sp, uint64_t u) {
    const LeafArgs a = uniform_copy(sp->u.leaf);
    const uint64_t row = uniform_u64(u) * kPThreads + threadIdx.x;
    if (row < a.n_rows) row_leaf_body(a.cols, a.cs, a.n_cols, row, a.nodes);
}
ZS_TASK void task_subtree3(const PStage *sp, uint64_t u) {
    const SubArgs a = uniform_copy(sp->u.sub);
    const uint64_t i = uniform_u64(u) * kPThreads + threadIdx.x;
    if (i < a.lanes) subtree_lane<3>(a.child, a.lvl1, a.lanes, i);
}
ZS_TASK void task_block(const PStage *sp, uint64_t u) {
    const BlockArgs a = uniform_copy(sp->u.blk);
    tree_block_body<kPThreads>(a.base, a.bt, (uint32_t)uniform_u64(u), threadIdx.x);
}
#undef ZS_TASK

__device__ __forceinline__ void run_task(const PStage *sp, uint64_t u) {
    const uint32_t kind = sp->kind, C = sp->C;
    switch (kind) {
        case PK_INTERP_FIRST:
            if (C == 8) task_interp_first<8>(sp, u);
            else if (C == 9) task_interp_first<9>(sp, u);
            else task_interp_first<10>(sp, u);
            break;
        case PK_MID:
            if (C == 8) task_mid<8>(sp, u);
            else if (C == 9) task_mid<9>(sp, u);
            else task_mid<10>(sp, u);
            break;
        case PK_EVAL_LAST:
            if (C == 8) task_eval_last<8>(sp, u);
            else if (C == 9) task_eval_last<9>(sp, u);
            else task_eval_last<10>(sp, u);
            break;
        case PK_PASS_INV: task_pass<true>(sp, u); break;
        case PK_PASS_FWD: task_pass<false>(sp, u); break;
        case PK_LEAF: task_leaves(sp, u); break;
        case PK_SUBTREE3: task_subtree3(sp, u); break;
        default: task_block(sp, u); break;  // PK_BLOCK
    }
}

__global__ __launch_bounds__(kPThreads, 4) void persistent_batch_kernel(const PPlan pl) {
    __shared__ uint32_t s_go, s_stage, s_ok;
    __shared__ unsigned long long s_local;
    const uint32_t tid = threadIdx.x;
    for (uint32_t ph = 0; ph < pl.n_phases; ph++) {
        const uint32_t f0 = pl.phase_first[ph], f1 = pl.phase_first[ph + 1];
        const uint64_t total = pl.stages[f1 - 1].end;
        for (;;) {
            if (tid == 0) {
                const unsigned long long t = atomicAdd(&pl.ctl->claim[ph], 1ull);
                uint32_t go = t < total && !ld_relaxed(&pl.ctl->abort);
                if (go) {
                    uint32_t lo = f0, hi = f1 - 1;  // the first stage whose end exceeds t
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) / 2;
                        if (pl.stages[mid].end > t) hi = mid;
                        else lo = mid + 1;
                    }
                    s_stage = lo;
                    s_local = t - (lo > f0 ? pl.stages[lo - 1].end : 0);
                }
                s_go = go;
            }
            __syncthreads();
            const uint32_t go = __builtin_amdgcn_readfirstlane(s_go);
            const uint32_t si = __builtin_amdgcn_readfirstlane(s_stage);
            const uint64_t lo32 = __builtin_amdgcn_readfirstlane((uint32_t)s_local);
            const uint64_t hi32 = __builtin_amdgcn_readfirstlane((uint32_t)(s_local >> 32));
            __syncthreads();  // every thread has read the claim before the next one is written
            if (!go) break;
            run_task(pl.stages + si, (hi32 << 32) | lo32);
            __syncthreads();  // the task's LDS is free for the next one
        }
        if (!grid_barrier(pl, ph + 1, &s_ok)) return;
    }
}

// ---- host: per-trace stage lists ----

struct TraceStages {
    std::vector<PStage> st;
};

uint64_t stage_tasks(const PStage &s) {
    switch (s.kind) {
        case PK_INTERP_FIRST:
        case PK_EVAL_LAST:
        case PK_PASS_INV:
        case PK_PASS_FWD: {
            const FftPass &p = s.u.p;
            const uint64_t gz = (uint64_t(1) << p.beta) >> p.Q_log;
            return uint64_t(s.gx) * s.gy * gz;
        }
        case PK_MID:
            return uint64_t(s.gx) * s.gy;
        case PK_LEAF:
            return (s.u.leaf.n_rows + kPThreads - 1) / kPThreads;
        case PK_SUBTREE3:
            return (s.u.sub.lanes + kPThreads - 1) / kPThreads;
        default:
            return s.u.blk.bt.wg_end[0];
    }
}

uint32_t ncb_for(uint32_t n_cols, uint32_t cols_per_wg) { return std::max<uint32_t>(1, std::min(cols_per_wg, n_cols)); }

// the generic pass (fft.hip run_pass) as a stage: same G_log rule, a 512-thread workgroup
PStage generic_pass(const FftPass &p0, bool inverse, uint32_t n_cols, size_t *lds) {
    PStage s{};
    s.kind = inverse ? PK_PASS_INV : PK_PASS_FWD;
    FftPass p = p0;
    const uint32_t avail = p.s_lo > 0 ? p.s_lo : (p.L - p.C);
    const int32_t want = (int32_t)PASS_ELEMS_LOG - (int32_t)p.C - (int32_t)p.Q_log;
    p.G_log = (uint32_t)std::max<int32_t>(0, std::min<int32_t>((int32_t)avail, want));
    p.n_cols = n_cols;
    s.C = p.C;
    s.gx = (1u << (p.L - p.C)) >> p.G_log;
    s.gy = n_cols;
    s.u.p = p;
    *lds = std::max(*lds, (size_t(1) << (p.G_log + p.Q_log)) * ((size_t(1) << p.C) + 1) * 4);
    return s;
}

// LDE stages of one trace: in (c x n, stride n) -> out (c x N, stride N).  coef holds c x n
// words, big c x N words; they do not overlap each other, and coef may lie in out only when
// the forward transform has more than one pass.
// the pipelined kernels' shapes (two passes per direction of 8..10 stages), with the fused
// middle's tiles and 1 + 2^beta twiddle sets small enough for two workgroups per CU
bool pipelined_lde(uint32_t L, uint32_t beta) {
    if (L < 16 || L > 20 || beta > 8) return false;
    const uint32_t C = plan_passes(L, false, PASS_CMAX)[0].C;
    return (size_t(8) * tile_stride(C, 8) + (((size_t(1) << beta) + 1) << C)) * 4 <= 80 * 1024;
}

void lde_stages(const zs_job &j, uint32_t beta, const uint32_t *tw_inv, const uint32_t *tw_fwd, uint32_t *coef,
                uint32_t *big, std::vector<PStage> &out, size_t *lds) {
    const uint32_t L = j.log_n, LT = L + beta, c = j.n_cols;
    const uint64_t n = uint64_t(1) << L, N = uint64_t(1) << LT;
    const LdeTiles tiles = default_lde_tiles();
    const auto pi = plan_passes(L, true, PASS_CMAX), pf = plan_passes(L, false, PASS_CMAX);
    if (pipelined_lde(L, beta)) {
        // interpolation's first pass: 8 tiles (waves) per workgroup
        {
            PStage s{};
            s.kind = PK_INTERP_FIRST;
            FftPass &p = s.u.p;
            p.in = j.d_cols; p.in_cs = n; p.out = coef; p.out_cs = n; p.tw = tw_inv;
            p.L = L; p.LT = L; p.beta = 0; p.s_lo = pi[0].s_lo; p.C = pi[0].C; p.G_log = 3; p.Q_log = 0;
            p.in_mode = IN_NATURAL; p.out_mode = OUT_KSPACE; p.natural_tiles = 1; p.scale = 0;
            p.n_cols = c; p.ncb = ncb_for(c, tiles.role[LDE_INTERP].cols_per_wg);
            p.TS = tile_stride(p.C, 8);
            s.C = p.C;
            s.gx = (1u << (L - p.C)) >> 3;
            s.gy = (c + p.ncb - 1) / p.ncb;
            const uint32_t set_log = p.C - low_group_stages((int)p.C);
            *lds = std::max(*lds, (size_t(8) * p.TS + (size_t(8) << set_log)) * 4);
            out.push_back(s);
        }
        // the fused middle: interpolation top pass, 1/n, forward top pass of every coset
        {
            PStage s{};
            s.kind = PK_MID;
            LdeFused &f = s.u.f;
            FftPass &p = f.p;
            p.in = coef; p.in_cs = n; p.out = big; p.out_cs = N;
            p.L = L; p.LT = L; p.beta = beta; p.s_lo = pf[0].s_lo; p.C = pf[0].C; p.scale = 0;
            p.G_log = 3; p.Q_log = 0; p.in_mode = IN_KSPACE; p.out_mode = OUT_KSPACE; p.natural_tiles = 0;
            p.n_cols = c; p.ncb = ncb_for(c, tiles.role[LDE_MID].cols_per_wg);
            p.TS = tile_stride(p.C, 8);
            f.tw_inv = tw_inv; f.tw_fwd = tw_fwd; f.inv_n = m31_inv(uint32_t(1) << L);
            s.C = p.C;
            s.gx = (1u << (L - p.C)) >> 3;
            s.gy = (c + p.ncb - 1) / p.ncb;
            *lds = std::max(*lds, (size_t(8) * p.TS + (((size_t(1) << beta) + 1) << p.C)) * 4);
            out.push_back(s);
        }
        // the evaluation's last pass: 2^Q cosets x 2^G tiles = 8 waves per workgroup
        {
            PStage s{};
            s.kind = PK_EVAL_LAST;
            FftPass &p = s.u.p;
            p.in = big; p.in_cs = N; p.out = j.d_lde; p.out_cs = N; p.tw = tw_fwd;
            p.L = L; p.LT = LT; p.beta = beta; p.s_lo = pf[1].s_lo; p.C = pf[1].C;
            p.Q_log = std::min<uint32_t>(beta, 3); p.G_log = 3 - p.Q_log;
            p.in_mode = IN_KSPACE; p.out_mode = OUT_NATURAL; p.natural_tiles = 1; p.scale = 0;
            p.n_cols = c; p.ncb = ncb_for(c, tiles.role[LDE_EVAL].cols_per_wg);
            p.TS = tile_stride(p.C, 8);
            s.C = p.C;
            s.gx = (1u << (L - p.C)) >> p.G_log;
            s.gy = (c + p.ncb - 1) / p.ncb;
            const uint32_t set_log = p.C - low_group_stages((int)p.C);
            *lds = std::max(*lds, (size_t(8) * p.TS + (size_t(8) << set_log)) * 4);
            out.push_back(s);
        }
        return;
    }
    // generic passes (fft.hip launch_lde_impl's untiled branch)
    for (size_t i = 0; i < pi.size(); i++) {
        FftPass p{};
        const bool first = i == 0, last = i + 1 == pi.size();
        p.in = first ? j.d_cols : coef; p.in_cs = n; p.out = coef; p.out_cs = n; p.tw = tw_inv;
        p.L = L; p.LT = L; p.beta = 0; p.s_lo = pi[i].s_lo; p.C = pi[i].C; p.Q_log = 0;
        p.in_mode = first ? IN_NATURAL : IN_KSPACE; p.out_mode = OUT_KSPACE; p.natural_tiles = first ? 1 : 0;
        p.scale = last ? m31_inv((uint32_t)n) : 0;
        out.push_back(generic_pass(p, true, c, lds));
    }
    for (size_t i = 0; i < pf.size(); i++) {
        FftPass p{};
        const bool first = i == 0, last = i + 1 == pf.size();
        p.in = first ? coef : big; p.in_cs = first ? n : N;
        p.out = last ? j.d_lde : big; p.out_cs = N; p.tw = tw_fwd;
        p.L = L; p.LT = LT; p.beta = beta; p.s_lo = pf[i].s_lo; p.C = pf[i].C;
        p.Q_log = last ? std::min<uint32_t>(beta, PASS_ELEMS_LOG > p.C ? PASS_ELEMS_LOG - p.C : 0) : 0;
        p.in_mode = first ? IN_SHARED : IN_KSPACE; p.out_mode = last ? OUT_NATURAL : OUT_KSPACE;
        p.natural_tiles = last ? 1 : 0; p.scale = 0;
        out.push_back(generic_pass(p, false, c, lds));
    }
}

// the row leaves and node levels of one trace (merkle.hip launch_trees_multi per tree:
// 3-level lane subtrees while 2^18 children or more remain, then 8-level workgroup subtrees)
void tree_stages(const zs_job &j, uint32_t beta, std::vector<PStage> &out, size_t *lds) {
    const uint32_t LT = j.log_n + beta;
    uint32_t *nodes = reinterpret_cast<uint32_t *>(j.d_nodes);
    {
        PStage s{};
        s.kind = PK_LEAF;
        s.u.leaf = LeafArgs{j.d_lde, uint64_t(1) << LT, uint64_t(1) << LT, j.n_cols, nodes};
        out.push_back(s);
    }
    uint32_t level = 0;
    while (LT - level >= 18) {
        PStage s{};
        s.kind = PK_SUBTREE3;
        s.u.sub = SubArgs{nodes + level_offset_words(LT, level), nodes + level_offset_words(LT, level + 1),
                          (uint64_t(1) << (LT - level)) >> 3};
        out.push_back(s);
        level += 3;
    }
    while (level < LT) {
        const uint32_t rem = LT - level, K = std::min<uint32_t>(8, rem);
        PStage s{};
        s.kind = PK_BLOCK;
        BlockArgs &b = s.u.blk;
        b.base = nodes;
        b.bt.n = 1;
        b.bt.child[0] = level_offset_words(LT, level);
        b.bt.r[0] = rem;
        b.bt.K[0] = K;
        b.bt.wg_end[0] = 1u << (rem - K);
        b.bt.chan = nullptr;
        b.bt.alpha = nullptr;
        *lds = std::max(*lds, block_lds_words(K, kPThreads) * 4);
        out.push_back(s);
        level += K;
    }
}

}  // namespace

// rows of at most 256 columns (one BLAKE3 chunk: the chunk-tree leaf's stack would give
// every lane of the launch 512 bytes of scratch)
bool persistent_supported(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup) {
    return n_cols >= 1 && n_cols <= 256 && log_n >= 3 && log_blowup <= 8 && !log_sum_over(log_n, log_blowup, 26);
}

size_t persistent_pool_bytes(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup) {
    // the big k-space array of the LDE (or, with one forward pass, the coefficients)
    return ((size_t(n_cols) << (log_n + log_blowup)) * 4 + 255) & ~size_t(255);
}

int32_t persistent_lde_trees(Ctx *ctx, zs_job *jobs, const std::vector<size_t> &idx, uint32_t log_blowup) {
    if (idx.empty()) return ZS_OK;
    // device memory: one pool region per trace, then the plan
    size_t pool = 0;
    for (size_t i : idx) {
        const zs_job &j = jobs[i];
        if (!persistent_supported(j.n_cols, j.log_n, log_blowup) || !j.d_cols || !j.d_lde || !j.d_nodes)
            return ZS_ERR_INPUT;
        pool += persistent_pool_bytes(j.n_cols, j.log_n, log_blowup);
    }
    ZS_TRY(ctx->ensure_scratch(10, pool));
    // twiddle tables of every size in the batch (built and cached before the launch)
    std::vector<const uint32_t *> inv_tw(32, nullptr), fwd_tw(32, nullptr);
    for (size_t i : idx) {
        const uint32_t L = jobs[i].log_n, LT = L + log_blowup;
        const TwiddleTable *t;
        if (!inv_tw[L]) {
            ZS_TRY(get_canonic_twiddles(ctx, L, &t));
            inv_tw[L] = t->inv.as<uint32_t>();
        }
        if (!fwd_tw[LT]) {
            ZS_TRY(get_canonic_twiddles(ctx, LT, &t));
            fwd_tw[LT] = t->fwd.as<uint32_t>();
        }
    }
    // stages per trace
    std::vector<TraceStages> ts(idx.size());
    size_t lds = 0, off = 0;
    uint32_t n_phases = 0;
    uint8_t *pool_base = ctx->scratch[10].as<uint8_t>();
    for (size_t k = 0; k < idx.size(); k++) {
        const zs_job &j = jobs[idx[k]];
        uint32_t *region = reinterpret_cast<uint32_t *>(pool_base + off);
        off += persistent_pool_bytes(j.n_cols, j.log_n, log_blowup);
        const bool one_fwd = !pipelined_lde(j.log_n, log_blowup) && plan_passes(j.log_n, false, PASS_CMAX).size() == 1;
        // coefficients in the output's first c x n words unless one forward pass reads them
        // while writing the output
        uint32_t *coef = one_fwd ? region : j.d_lde;
        lde_stages(j, log_blowup, inv_tw[j.log_n], fwd_tw[j.log_n + log_blowup], coef, region, ts[k].st, &lds);
        tree_stages(j, log_blowup, ts[k].st, &lds);
        n_phases = std::max<uint32_t>(n_phases, (uint32_t)ts[k].st.size());
    }
    if (n_phases > kMaxPhases) return ZS_ERR_INPUT;
    // phase k: stage k of every trace, the stages with the most tasks first
    std::vector<PStage> all;
    std::vector<uint32_t> first(n_phases + 1, 0);
    for (uint32_t ph = 0; ph < n_phases; ph++) {
        first[ph] = (uint32_t)all.size();
        std::vector<PStage> v;
        for (auto &t : ts)
            if (ph < t.st.size()) v.push_back(t.st[ph]);
        std::stable_sort(v.begin(), v.end(),
                         [](const PStage &a, const PStage &b) { return stage_tasks(a) > stage_tasks(b); });
        uint64_t end = 0;
        for (auto &s : v) {
            end += stage_tasks(s);
            s.end = end;
            all.push_back(s);
        }
    }
    first[n_phases] = (uint32_t)all.size();
    // plan and control block on the device
    const size_t st_bytes = all.size() * sizeof(PStage), ph_bytes = first.size() * 4;
    const size_t ctl_off = 0, ph_off = (sizeof(PCtl) + 255) & ~size_t(255),
                 st_off = (ph_off + ph_bytes + 255) & ~size_t(255);
    ZS_TRY(ctx->ensure_scratch(11, st_off + st_bytes));
    uint8_t *plan = ctx->scratch[11].as<uint8_t>();
    ZS_HIP(ctx, hipMemsetAsync(plan + ctl_off, 0, sizeof(PCtl), ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(plan + ph_off, first.data(), ph_bytes, hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(plan + st_off, all.data(), st_bytes, hipMemcpyHostToDevice, ctx->stream));
    // as many workgroups as stay resident together
    if (lds > 160 * 1024) return ZS_ERR_INPUT;
    int per_cu = 0, cus = 0, clock_khz = 0;
    ZS_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_batch_kernel, kPThreads, lds));
    ZS_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    ZS_HIP(ctx, hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
    if (per_cu < 1 || cus < 1 || clock_khz < 1) {
        ctx->set_error("persistent batch: no resident workgroup fits (" + std::to_string(lds) + " bytes of LDS)");
        return ZS_ERR_DEVICE;
    }
    PPlan pl{};
    pl.stages = reinterpret_cast<const PStage *>(plan + st_off);
    pl.phase_first = reinterpret_cast<const uint32_t *>(plan + ph_off);
    pl.ctl = reinterpret_cast<PCtl *>(plan + ctl_off);
    pl.wait_ticks = uint64_t(clock_khz) * 1000 * 20;  // 20 s
    pl.n_phases = n_phases;
    pl.n_wg = (uint32_t)(per_cu * cus);
    ZS_LAUNCH(ctx, "persistent_batch_kernel", persistent_batch_kernel, dim3(pl.n_wg), dim3(kPThreads), lds,
              ctx->stream, pl);
    ZS_HIP(ctx, hipGetLastError());
    uint32_t abort = 0;
    ZS_HIP(ctx, hipMemcpyAsync(&abort, &pl.ctl->abort, 4, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (abort) {
        ctx->set_error("persistent batch kernel: a grid barrier wait ran past its bound; results incomplete");
        return ZS_ERR_DEVICE;
    }
    return ZS_OK;
}

}  // namespace zs
