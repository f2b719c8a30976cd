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
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fri_dev.h"
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
    PK_DRAW,              // start the trace's channel, absorb its root, draw gamma (one workgroup)
    PK_COMBINE,           // gamma^c, then g = sum_c gamma^c lde_c and g's leaves (FRI layer 0)
    PK_FOLD,              // one FRI fold and the next layer's leaves
    PK_REMAINDER,         // the remainder coefficients; the roots gathered for the host
};

constexpr uint32_t kPThreads = 512;
constexpr uint32_t kMaxPhases = 192;
// Workgroups are dealt to the 8 XCDs round-robin by index, and each XCD has its own L2.
// A phase's task list is cut into 8 contiguous ranges, one per XCD: neighbouring tile
// groups, whose 32-byte runs share cache lines, then meet in one L2 (as xcd_block
// arranges for the single-trace kernels).  A workgroup whose XCD's range is used up
// takes tasks from the other ranges in turn.
constexpr uint32_t kXcds = 8;

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
struct DrawArgs {
    B3Hasher *chan;
    const B3Hasher *init;  // or null: continue the channel's transcript
    const uint32_t *digest;
    uint32_t *out;
};
struct CombArgs {
    const uint32_t *lde, *gamma;
    uint64_t cs, n;
    uint32_t *g, *leaves;
    uint32_t n_cols;
};
struct FoldArgs {
    const uint32_t *cur, *tables, *alpha;
    uint32_t *next, *next_leaves;  // next_leaves: null for the last layer
    uint64_t half;
    qm31 inv_off;  // (FRI_OFFSET^(2^layer))^-1
    uint32_t layer;
};
struct RemArgs {
    const uint32_t *last, *mat, *trees, *row_root;
    uint32_t *out;  // trace root, the L layer roots, the rc coefficients
    uint32_t rc, L, dlog;
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
        DrawArgs draw;
        CombArgs comb;
        FoldArgs fold;
        RemArgs rem;
    } u;
};

struct PCtl {
    uint32_t arrive;  // workgroups that reached a barrier, summed over the barriers
    uint32_t abort;   // a barrier wait ran past its bound: the results are incomplete
    unsigned long long claim[kMaxPhases][kXcds];  // next task of each phase's per-XCD range
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
// barriers (1, 2, ...).  False when the wait gave up or another workgroup's did.  Each
// wave waits until its own stores have reached the L2; then one thread per workgroup
// makes them visible device-wide (the release: the XCD's L2 written back), arrives,
// waits, and takes the acquire (its CU's and XCD's caches invalidated) for the whole
// workgroup -- a device-scope fence in every wave made each barrier cost ~0.15 ms.
__device__ bool grid_barrier(const PPlan &pl, uint32_t k, uint32_t *s_ok) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
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
        __threadfence();
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// A task's descriptor with every word made wave-uniform (scalar registers): the task
// functions below are called, not inlined -- one register allocation per body, as in
// the single-trace kernels, where inlining them all into one kernel spilled -- and a
// call passes its arguments in vector registers.  The bodies reach HBM through
// ldg / stg, since pointers read from a descriptor are not known to be global.
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

__device__ __forceinline__ FftPass pass_desc(const PStage *sp) { return uniform_copy(sp->u.p); }

#define ZS_TASK __device__ __attribute__((noinline))
template <int C>
ZS_TASK void task_interp_first(const PStage *sp, uint64_t u) {
    lde_interp_first_body<C>(pass_desc(sp), task_vb(*sp, uniform_u64(u)));
}
template <int C>
ZS_TASK void task_mid(const PStage *sp, uint64_t u) {
    const LdeFused f = uniform_copy(sp->u.f);
    lde_mid_body<C>(f, task_vb(*sp, uniform_u64(u)));
}
template <int C>
ZS_TASK void task_eval_last(const PStage *sp, uint64_t u) {
    lde_eval_last_body<C>(pass_desc(sp), task_vb(*sp, uniform_u64(u)));
}
template <bool INV>
ZS_TASK void task_pass(const PStage *sp, uint64_t u) {
    fft_pass_body<INV>(pass_desc(sp), task_vb(*sp, uniform_u64(u)), threadIdx.x, kPThreads);
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
ZS_TASK void task_draw(const PStage *sp, uint64_t) {
    extern __shared__ uint32_t lds[];
    const DrawArgs a = uniform_copy(sp->u.draw);
    constexpr uint32_t HW = (sizeof(B3Hasher) / 4 + 1) & ~1u;  // hasher states 8-byte aligned
    B3Hasher *h = reinterpret_cast<B3Hasher *>(lds), *snap = reinterpret_cast<B3Hasher *>(lds + HW);
    channel_draw_body(a.chan, a.init, a.digest, a.out, *h, *snap, lds + 2 * HW, lds + 2 * HW + 24, threadIdx.x,
                      kPThreads);
}
// the powers gamma^c (c < n_cols <= 256, square-and-multiply per thread, as
// qm31_powers_kernel) into LDS, then g and its leaves for 512 rows
ZS_TASK void task_combine(const PStage *sp, uint64_t u) {
    extern __shared__ uint32_t lds[];
    const CombArgs a = uniform_copy(sp->u.comb);
    if (threadIdx.x < a.n_cols) {
        qm31 b = ldg_qm(a.gamma), r = {{1, 0}, {0, 0}};
        for (uint32_t e = threadIdx.x; e; e >>= 1) {
            if (e & 1u) r = qm_mul(r, b);
            b = qm_mul(b, b);
        }
        uint32_t *o = lds + 4 * threadIdx.x;
        o[0] = r.a.c0;
        o[1] = r.a.c1;
        o[2] = r.b.c0;
        o[3] = r.b.c1;
    }
    __syncthreads();
    const uint64_t i = uniform_u64(u) * kPThreads + threadIdx.x;
    if (i >= a.n) return;
    const uint4 v = combine_row(a.lde, a.cs, a.n_cols, i, lds);
    stg4(a.g + 4 * i, v);
    store_qm31_leaf(a.leaves + 8 * i, qm31{{v.x, v.y}, {v.z, v.w}});
}
ZS_TASK void task_fold(const PStage *sp, uint64_t u) {
    const FoldArgs a = uniform_copy(sp->u.fold);
    const uint64_t j = uniform_u64(u) * kPThreads + threadIdx.x;
    if (j >= a.half) return;
    const qm31 a_off = qm_half(qm_mul(ldg_qm(a.alpha), a.inv_off));
    const qm31 v = fold_value(a.cur, a.half, j, a.layer, a.tables, a_off);
    stg_qm(a.next + 4 * j, v);
    if (a.next_leaves) store_qm31_leaf(a.next_leaves + 8 * j, v);
}
// threads [0, 8): the trace root; [8, 8 + 8L): layer i's root word w; then the coefficients
// (fri.zig:228-233: the interpolant of layer L's first rc values, a fixed matrix)
ZS_TASK void task_remainder(const PStage *sp, uint64_t) {
    const RemArgs a = uniform_copy(sp->u.rem);
    const uint32_t t = threadIdx.x;
    if (t < 8) {
        stg(a.out + t, ldg(a.row_root + t));
    } else if (t < 8 + 8 * a.L) {
        const uint32_t i = (t - 8) >> 3, w = (t - 8) & 7u;
        uint64_t off = 0;  // words before tree i
        for (uint32_t k = 0; k < i; k++) off += ((uint64_t(2) << (a.dlog - k)) - 1) * 8;
        off += ((uint64_t(2) << (a.dlog - i)) - 1) * 8 - 8;  // its root
        stg(a.out + t, ldg(a.trees + off + w));
    } else if (t < 8 + 8 * a.L + a.rc) {
        const uint32_t k = t - 8 - 8 * a.L;
        qm31 acc = {{0, 0}, {0, 0}};
        for (uint32_t j = 0; j < a.rc; j++) acc = qm_add(acc, qm_mul(ldg_qm(a.mat + 4 * (k * a.rc + j)), ldg_qm(a.last + 4 * j)));
        stg_qm(a.out + 8 + 8 * a.L + 4 * k, acc);
    }
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
        case PK_BLOCK: task_block(sp, u); break;
        case PK_DRAW: task_draw(sp, u); break;
        case PK_COMBINE: task_combine(sp, u); break;
        case PK_FOLD: task_fold(sp, u); break;
        default: task_remainder(sp, u); break;  // PK_REMAINDER
    }
}

__global__ __launch_bounds__(kPThreads, 4) void persistent_batch_kernel(const PPlan pl) {
    __shared__ uint32_t s_go, s_stage, s_ok, s_src, s_cur;
    __shared__ unsigned long long s_local;
    const uint32_t tid = threadIdx.x, xcd = blockIdx.x % kXcds;
    for (uint32_t ph = 0; ph < pl.n_phases; ph++) {
        const uint32_t f0 = pl.phase_first[ph], f1 = pl.phase_first[ph + 1];
        const uint64_t total = pl.stages[f1 - 1].end;
        if (tid == 0) {
            s_src = 0;     // ranges tried so far: own XCD's first
            s_cur = ~0u;   // the stage of this workgroup's last task from range s_src (none yet)
        }
        for (;;) {
            if (tid == 0) {
                uint32_t go = 0;
                unsigned long long t = 0;
                while (!go && s_src < kXcds && !ld_relaxed(&pl.ctl->abort)) {
                    const uint32_t r = (xcd + s_src) % kXcds;
                    const unsigned long long lo = total * r / kXcds, hi = total * (r + 1) / kXcds;
                    if (lo < hi) t = lo + atomicAdd(&pl.ctl->claim[ph][r], 1ull);
                    if (lo < hi && t < hi) {
                        go = 1;
                    } else {
                        s_src++;
                        s_cur = ~0u;
                    }
                }
                if (go) {
                    // the stage of task t: tasks from one range come in increasing order, so
                    // after the first (a binary search) the cursor moves forward
                    uint32_t st = s_cur;
                    if (st == ~0u) {
                        uint32_t lo = f0, hi = f1 - 1;  // the first stage whose end exceeds t
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) / 2;
                            if (pl.stages[mid].end > t) hi = mid;
                            else lo = mid + 1;
                        }
                        st = lo;
                    }
                    while (pl.stages[st].end <= t) st++;
                    s_cur = st;
                    s_stage = st;
                    s_local = t - (st > f0 ? pl.stages[st - 1].end : 0);
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
        case PK_BLOCK:
            return s.u.blk.bt.wg_end[0];
        case PK_COMBINE:
            return (s.u.comb.n + kPThreads - 1) / kPThreads;
        case PK_FOLD:
            return (s.u.fold.half + kPThreads - 1) / kPThreads;
        default:  // PK_DRAW, PK_REMAINDER: one workgroup
            return 1;
    }
}

// columns one pass task runs with one set of staged twiddles (the pass bodies overlap a
// column's loads with the previous column's butterflies, so more columns per task also
// amortise the claim and the call; 4 and 16 measured no better)
constexpr uint32_t kTaskCols = 8;

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

// the pipelined kernels' shapes (two passes per direction of 8..10 stages), with the fused
// middle's tiles and 1 + 2^beta twiddle sets small enough for two workgroups per CU
bool pipelined_lde(uint32_t L, uint32_t beta) {
    if (L < 16 || L > 20 || beta > 8) return false;
    const uint32_t C = plan_passes(L, false, PASS_CMAX)[0].C;
    return (size_t(8) * tile_stride(C, 8) + (((size_t(1) << beta) + 1) << C)) * 4 <= 80 * 1024;
}

// LDE stages of one trace: in (c x n, stride n) -> out (c x N, stride N).  coef holds c x n
// words, big c x N words; they do not overlap each other, and coef may lie in out only when
// the forward transform has more than one pass.
void lde_stages(const zs_job &j, uint32_t beta, const uint32_t *tw_inv, const uint32_t *tw_fwd, uint32_t *coef,
                uint32_t *big, std::vector<PStage> &out, size_t *lds) {
    const uint32_t L = j.log_n, LT = L + beta, c = j.n_cols;
    const uint64_t n = uint64_t(1) << L, N = uint64_t(1) << LT;
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
            p.n_cols = c; p.ncb = ncb_for(c, kTaskCols);
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
            p.n_cols = c; p.ncb = ncb_for(c, kTaskCols);
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
            p.n_cols = c; p.ncb = ncb_for(c, kTaskCols);
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

// the node levels of a tree of 2^LT leaves at `nodes` (merkle.hip launch_trees_multi for
// one tree: 3-level lane subtrees while 2^18 children or more remain, then 8-level
// workgroup subtrees)
// With draw_chan, the last stage (one workgroup: the root) also absorbs the root into
// that channel and draws the next challenge into draw_alpha (tree_block_body's draw).
void level_stages(uint32_t *nodes, uint32_t LT, std::vector<PStage> &out, size_t *lds,
                  B3Hasher *draw_chan = nullptr, uint32_t *draw_alpha = nullptr) {
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
        const bool root = level + K == LT;
        b.bt.chan = root ? draw_chan : nullptr;
        b.bt.alpha = root ? draw_alpha : nullptr;
        *lds = std::max(*lds, (block_lds_words(K, kPThreads) + (root && draw_chan ? kDrawLdsWords : 0)) * 4);
        out.push_back(s);
        level += K;
    }
}

// Device memory of a trace's FRI commitment (job.cpp commit_job's scratch), laid out in
// the trace's pool region once its LDE no longer needs it: the channel, gamma and its
// layers 0..L (layer 0 = g), the L layer trees, the L challenges.
struct FriLayout {
    uint64_t chan, gamma, layer[29], tree[28], alphas, total;
};
FriLayout fri_layout(uint32_t dlog, uint32_t L) {
    FriLayout f{};
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = o;
        o += (bytes + 255) & ~uint64_t(255);
        return at;
    };
    f.chan = take(sizeof(B3Hasher));
    f.gamma = take(16);
    for (uint32_t i = 0; i <= L; i++) f.layer[i] = take((uint64_t(16) << dlog) >> i);
    // the trees back to back (the remainder task finds the roots by their sizes)
    uint64_t tw = 0;
    for (uint32_t i = 0; i < L; i++) tw += tree_words(dlog - i) * 4;
    f.tree[0] = take(tw);
    for (uint32_t i = 1; i < L; i++) f.tree[i] = f.tree[i - 1] + tree_words(dlog - i + 1) * 4;
    f.alphas = take(16 * uint64_t(L));
    f.total = o;
    return f;
}
uint32_t result_words(uint32_t L, uint32_t rc) { return 8 + 8 * L + 4 * rc; }

struct FriShared {
    const B3Hasher *init;  // Channel.init("zig-stark:hip:commit") on the device
    const uint32_t *tables, *mat;
};

// the commitment after the row tree (job.cpp commit_job, fri_commit_impl in channel
// mode): gamma from the trace root, g and its leaves, then per layer its node levels
// (the root's task also draws the layer's challenge) and the fold (with the next
// layer's leaves); last the remainder, with the roots gathered into `res`
void fri_stages(const zs_job &j, uint32_t beta, uint32_t remainder_log, uint8_t *region, const FriShared &sh,
                uint32_t *res, std::vector<PStage> &out, size_t *lds) {
    const uint32_t dlog = j.log_n + beta, L = dlog - remainder_log, rc = 1u << (remainder_log - beta);
    const uint64_t N = uint64_t(1) << dlog;
    const FriLayout f = fri_layout(dlog, L);
    auto at = [&](uint64_t off) { return reinterpret_cast<uint32_t *>(region + off); };
    B3Hasher *chan = reinterpret_cast<B3Hasher *>(region + f.chan);
    const uint32_t *row_root = reinterpret_cast<const uint32_t *>(j.d_nodes) + tree_words(dlog) - 8;
    constexpr size_t draw_lds = (2 * ((sizeof(B3Hasher) / 4 + 1) & ~size_t(1)) + 24 + 16) * 4;
    {  // gamma: the channel starts from Channel.init and absorbs the trace root
        PStage s{};
        s.kind = PK_DRAW;
        s.u.draw = DrawArgs{chan, sh.init, row_root, at(f.gamma)};
        *lds = std::max(*lds, draw_lds);
        out.push_back(s);
    }
    {
        PStage s{};
        s.kind = PK_COMBINE;
        s.u.comb = CombArgs{j.d_lde, at(f.gamma), N, N, at(f.layer[0]), at(f.tree[0]), j.n_cols};
        *lds = std::max(*lds, size_t(16) * j.n_cols);
        out.push_back(s);
    }
    for (uint32_t i = 0; i < L; i++) {
        const uint32_t li = dlog - i;
        level_stages(at(f.tree[i]), li, out, lds, chan, at(f.alphas) + 4 * i);  // the root's task draws
        PStage s{};
        s.kind = PK_FOLD;
        FoldArgs &fa = s.u.fold;
        fa.cur = at(f.layer[i]);
        fa.tables = sh.tables;
        fa.alpha = at(f.alphas) + 4 * i;
        fa.next = at(f.layer[i + 1]);
        fa.next_leaves = i + 1 < L ? at(f.tree[i + 1]) : nullptr;
        fa.half = (N >> i) / 2;
        fa.inv_off = qm_inv(fri_offset_pow(i));
        fa.layer = i;
        out.push_back(s);
    }
    PStage s{};
    s.kind = PK_REMAINDER;
    s.u.rem = RemArgs{at(f.layer[L]), sh.mat, at(f.tree[0]), row_root, res, rc, L, dlog};
    out.push_back(s);
}

}  // namespace

// rows of at most 256 columns (one BLAKE3 chunk: the chunk-tree leaf's stack would give
// every lane of the launch 512 bytes of scratch)
bool persistent_supported(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup) {
    return n_cols >= 1 && n_cols <= 256 && log_n >= 3 && log_blowup <= 8 && !log_sum_over(log_n, log_blowup, 26);
}

// a trace's pool region: the LDE's big k-space array (or, with one forward pass, its
// coefficients), later its FRI commitment's buffers
size_t persistent_pool_bytes(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup, uint32_t remainder_log) {
    const uint32_t dlog = log_n + log_blowup;
    const uint64_t lde = (uint64_t(n_cols) << dlog) * 4;
    const uint64_t fri = fri_layout(dlog, dlog - remainder_log).total;
    return (std::max(lde, fri) + 255) & ~uint64_t(255);
}

int32_t persistent_commit(Ctx *ctx, zs_job *jobs, const std::vector<size_t> &idx, uint32_t log_blowup,
                          uint32_t remainder_log) {
    if (idx.empty()) return ZS_OK;
    const uint32_t rc = 1u << (remainder_log - log_blowup);
    // device memory: one pool region per trace, then the plan and the results
    size_t pool = 0;
    for (size_t i : idx) {
        const zs_job &j = jobs[i];
        if (!persistent_supported(j.n_cols, j.log_n, log_blowup) || !j.d_cols || !j.d_lde || !j.d_nodes ||
            remainder_log < log_blowup || remainder_log >= j.log_n + log_blowup ||
            remainder_log - log_blowup > kMaxRemainderLog)
            return ZS_ERR_INPUT;
        pool += persistent_pool_bytes(j.n_cols, j.log_n, log_blowup, remainder_log);
    }
    ZS_TRY(ctx->ensure_scratch(10, pool));
    // twiddle, FRI and remainder tables of every size in the batch, the channel's start
    // (built and cached before the launch)
    std::vector<const uint32_t *> inv_tw(32, nullptr), fwd_tw(32, nullptr), fri_tab(32, nullptr),
        rem_mat(32, nullptr);
    const B3Hasher *init;
    ZS_TRY(device_commit_channel(ctx, &init));
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
            ZS_TRY(get_fri_tables(ctx, LT, &fri_tab[LT]));
            ZS_TRY(get_remainder_matrix(ctx, LT - remainder_log, remainder_log, rc, &rem_mat[LT]));
        }
    }
    // each trace's words in the results area (its roots and remainder, for one copy out)
    std::vector<uint32_t> res_off(idx.size());
    uint64_t res_words = 0;
    for (size_t k = 0; k < idx.size(); k++) {
        res_off[k] = (uint32_t)res_words;
        res_words += result_words(jobs[idx[k]].log_n + log_blowup - remainder_log, rc);
    }
    // stages per trace
    std::vector<std::vector<PStage>> ts(idx.size());  // LDE, leaves, row tree
    std::vector<uint8_t *> regions(idx.size());
    size_t lds = 0, off = 0;
    uint32_t n_phases = 0;
    uint8_t *pool_base = ctx->scratch[10].as<uint8_t>();
    std::vector<std::vector<PStage>> fri(idx.size());
    for (size_t k = 0; k < idx.size(); k++) {
        const zs_job &j = jobs[idx[k]];
        uint8_t *region = regions[k] = pool_base + off;
        off += persistent_pool_bytes(j.n_cols, j.log_n, log_blowup, remainder_log);
        const uint32_t LT = j.log_n + log_blowup;
        const bool one_fwd = !pipelined_lde(j.log_n, log_blowup) && plan_passes(j.log_n, false, PASS_CMAX).size() == 1;
        // coefficients in the output's first c x n words unless one forward pass reads them
        // while writing the output
        uint32_t *big = reinterpret_cast<uint32_t *>(region);
        uint32_t *coef = one_fwd ? big : j.d_lde;
        lde_stages(j, log_blowup, inv_tw[j.log_n], fwd_tw[LT], coef, big, ts[k], &lds);
        {
            PStage s{};
            s.kind = PK_LEAF;
            s.u.leaf = LeafArgs{j.d_lde, uint64_t(1) << LT, uint64_t(1) << LT, j.n_cols,
                                reinterpret_cast<uint32_t *>(j.d_nodes)};
            ts[k].push_back(s);
        }
        level_stages(reinterpret_cast<uint32_t *>(j.d_nodes), LT, ts[k], &lds);
        n_phases = std::max<uint32_t>(n_phases, (uint32_t)ts[k].size());
    }
    // FRI stages; each trace's last one (the remainder's) gets its results address once
    // the plan buffer exists
    size_t n_st = 0;
    for (size_t k = 0; k < idx.size(); k++) {
        const uint32_t LT = jobs[idx[k]].log_n + log_blowup;
        const FriShared sh{init, fri_tab[LT], rem_mat[LT]};
        fri_stages(jobs[idx[k]], log_blowup, remainder_log, regions[k], sh, nullptr, fri[k], &lds);
        n_st += ts[k].size() + fri[k].size();
        n_phases = std::max<uint32_t>(n_phases, (uint32_t)(ts[k].size() + fri[k].size()));
    }
    if (n_phases > kMaxPhases) return ZS_ERR_INPUT;
    // plan buffer: control block, phase table, stages, results
    const size_t ph_off = (sizeof(PCtl) + 255) & ~size_t(255);
    const size_t st_off = (ph_off + (n_phases + 1) * 4 + 255) & ~size_t(255);
    const size_t res_off_bytes = (st_off + n_st * sizeof(PStage) + 255) & ~size_t(255);
    ZS_TRY(ctx->ensure_scratch(11, res_off_bytes + res_words * 4));
    uint32_t *res_dev = reinterpret_cast<uint32_t *>(ctx->scratch[11].as<uint8_t>() + res_off_bytes);
    for (size_t k = 0; k < idx.size(); k++) fri[k].back().u.rem.out = res_dev + res_off[k];
    // phase k: stage k of every trace, the stages with the most tasks first
    std::vector<uint32_t> first(n_phases + 1, 0);
    std::vector<PStage> all;
    all.reserve(n_st);
    for (uint32_t ph = 0; ph < n_phases; ph++) {
        first[ph] = (uint32_t)all.size();
        std::vector<PStage> v;
        for (size_t k = 0; k < idx.size(); k++) {
            const size_t nl = ts[k].size();
            if (ph < nl) v.push_back(ts[k][ph]);
            else if (ph - nl < fri[k].size()) v.push_back(fri[k][ph - nl]);
        }
        std::stable_sort(v.begin(), v.end(),
                         [](const PStage &x, const PStage &y) { return stage_tasks(x) > stage_tasks(y); });
        uint64_t end = 0;
        for (auto &st : v) {
            end += stage_tasks(st);
            st.end = end;
            all.push_back(st);
        }
    }
    first[n_phases] = (uint32_t)all.size();
    uint8_t *plan = ctx->scratch[11].as<uint8_t>();
    ZS_HIP(ctx, hipMemsetAsync(plan, 0, sizeof(PCtl), ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(plan + ph_off, first.data(), first.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(plan + st_off, all.data(), all.size() * sizeof(PStage), hipMemcpyHostToDevice,
                               ctx->stream));
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
    pl.ctl = reinterpret_cast<PCtl *>(plan);
    // a barrier waits at most 20 s; ZS_PERSIST_WAIT_MS sets another bound (0: give up at
    // once, so that a test can take the abort path with every workgroup still exiting)
    const char *wait_env = getenv("ZS_PERSIST_WAIT_MS");
    const uint64_t wait_ms = wait_env ? strtoull(wait_env, nullptr, 10) : 20000;
    pl.wait_ticks = uint64_t(clock_khz) * wait_ms;
    pl.n_phases = n_phases;
    pl.n_wg = (uint32_t)(per_cu * cus);
    ZS_LAUNCH(ctx, "persistent_batch_kernel", persistent_batch_kernel, dim3(pl.n_wg), dim3(kPThreads), lds,
              ctx->stream, pl);
    ZS_HIP(ctx, hipGetLastError());
    uint32_t abort = 0;
    std::vector<uint32_t> res(res_words);
    ZS_HIP(ctx, hipMemcpyAsync(&abort, &pl.ctl->abort, 4, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipMemcpyAsync(res.data(), plan + res_off_bytes, res_words * 4, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (abort) {
        ctx->set_error("persistent batch kernel: a grid barrier wait ran past its bound; results incomplete");
        return ZS_ERR_DEVICE;
    }
    for (size_t k = 0; k < idx.size(); k++) {
        zs_job &j = jobs[idx[k]];
        const uint32_t L = j.log_n + log_blowup - remainder_log;
        const uint32_t *r = res.data() + res_off[k];
        std::memcpy(j.trace_root, r, 32);
        if (j.fri_roots) std::memcpy(j.fri_roots, r + 8, size_t(L) * 32);
        if (j.remainder) std::memcpy(j.remainder, r + 8 + 8 * L, size_t(rc) * 16);
        j.status = ZS_OK;
    }
    return ZS_OK;
}

}  // namespace zs
