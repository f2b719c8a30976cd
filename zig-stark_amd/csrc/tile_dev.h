// tile_dev.h -- device code of the register-group circle-FFT passes for gfx950
// (the pass bodies and their helpers).  Each pass body takes its workgroup's
// position in the pass's grid as a VB, so the same code runs as its own kernel
// (fft_tile.hip) and as a task of the persistent batch kernel (persist.hip).
//
// A pass covers C consecutive butterfly stages (8 <= C <= 11) of a tile of
// 2^C elements.  The workgroup stages its tiles through LDS with coalesced
// global accesses; then each wave transforms one tile with its 64 lanes
// holding R = 2^(C-6) elements each.  The C stages are split into groups of
// log2 R stages: a group's stages run entirely in registers (the group's bits
// of the element index come from the lane's register index j, the other bits
// from the lane id), and the tile goes through LDS only between groups, so a
// 10-stage pass makes three LDS round trips instead of ten.  Twiddles for a
// stage are loaded once per distinct index (2^r - 1 loads per r-stage group
// per lane, shared by r * 2^(r-1) butterflies).
//
// The fused LDE kernels run the interpolation's top pass, scale by 1/n, and
// then the 2^beta forward transforms' top pass straight from the registers
// (the interpolation's last group and the evaluation's first group cover the
// same index bits), so the interleaved coefficients of those stages never go
// to HBM.  The LDE's first interpolation pass, middle (fused) pass and last
// evaluation pass have pipelined kernels of their own (lde_interp_first_kernel,
// lde_mid_kernel, lde_eval_last_kernel): one wave per tile, the next column
// loaded into registers during the current column's butterflies.
// fft_tile_kernel serves every other pass shape.
#pragma once
#include "fft_common.h"

namespace zs {

// a workgroup's position in a pass's grid: blockIdx x, y, z and gridDim.x
struct VB {
    uint32_t x, y, z, gx;
    __device__ static VB hw() { return {blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x}; }
};


template <int C>
struct TC {
    static constexpr int RB = C - 6;  // log2 elements per lane
    static constexpr int R = 1 << RB;
    static constexpr int T = 1 << C;
    static constexpr int NG = (C + RB - 1) / RB;  // register groups per pass
};

// Twiddle placement in LDS.  Within stage ls's region of the heap, the index
// a lane reads is (lane bits, then k bits that come from its own registers):
// consecutive lanes would read 2^k words apart, up to 8-way bank conflicts.
// The k low bits are rotated to the top of the region's index instead, so the
// lanes of a wave read consecutive words (or the same word).  k is the number of
// the stage's group register bits above ls; see t_of for the lane/register split.
template <int C>
__host__ __device__ constexpr int tw_reg_bits(int ls) {
    constexpr int RB = C - 6;
    const int gf = (C - 1 - ls) / RB;  // the group holding stage ls
    const int hi = C - 1 - gf * RB, lo = C - (gf + 1) * RB;
    const int top = (lo <= 0 && hi + 1 < RB) ? RB - 1 : hi;  // a short last group sits in the low bits
    return top - ls;
}
template <int C>
__host__ __device__ __forceinline__ uint32_t tw_pos(int ls, uint32_t i) {
    const int k = tw_reg_bits<C>(ls), S = C - 1 - ls;
    if (k <= 0 || k >= S) return i;
    return ((i & ((1u << k) - 1)) << (S - k)) | (i >> k);
}

// LDS index of tile element t: one pad word per 16 keeps the group access
// patterns with stride 4 and 16 across lanes free of bank conflicts (32 banks);
// a contiguous run of 32 lanes spans 33 words, one 2-way conflict per half-wave
__device__ __forceinline__ uint32_t lidx(uint32_t t) { return t + (t >> 4); }

// element index of register j of `lane` in a group covering bits [LO, LO + r):
// the group bits are j's low r bits, the other C - r bits are (lane, j >> r)
template <int RB, int LO, int r>
__device__ __forceinline__ uint32_t t_of(uint32_t lane, uint32_t j) {
    const uint32_t g = j & ((1u << r) - 1);
    const uint32_t other = (lane << (RB - r)) | (j >> r);
    return ((other >> LO) << (LO + r)) | (g << LO) | (other & ((1u << LO) - 1));
}

// LDS writes of a wave must land before its lanes read each other's values
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int C, int LO, int r>
__device__ __forceinline__ void g_load(uint32_t (&v)[TC<C>::R], const uint32_t *tl, uint32_t lane) {
#pragma unroll
    for (int j = 0; j < TC<C>::R; j++) v[j] = tl[lidx(t_of<TC<C>::RB, LO, r>(lane, j))];
}
template <int C, int LO, int r>
__device__ __forceinline__ void g_store(const uint32_t (&v)[TC<C>::R], uint32_t *tl, uint32_t lane) {
#pragma unroll
    for (int j = 0; j < TC<C>::R; j++) tl[lidx(t_of<TC<C>::RB, LO, r>(lane, j))] = v[j];
}

// one butterfly stage on local bit LO + b.  twl is the tile set's twiddles in
// LDS, heap layout: local stage ls at [2^(C-1-ls), 2^(C-ls)), entry i being the
// twiddle of the pairs whose bits above ls are i, stored doubled (m31_mul_dbl).
template <bool INV, int C, int LO, int r, int b>
__device__ __forceinline__ void g_stage(uint32_t (&v)[TC<C>::R], uint32_t lane, const uint32_t *twl) {
    constexpr int RB = TC<C>::RB;
    constexpr int ls = LO + b;
    const uint32_t *tws = twl + (1u << (C - 1 - ls));
#pragma unroll
    for (int o = 0; o < (1 << (RB - r)); o++) {
#pragma unroll
        for (int u = 0; u < (1 << (r - 1 - b)); u++) {
            const uint32_t jb = ((uint32_t)o << r) | ((uint32_t)u << (b + 1));
            const uint32_t t0 = t_of<RB, LO, r>(lane, jb);
            const uint32_t w = tws[tw_pos<C>(ls, t0 >> (ls + 1))];
#pragma unroll
            for (int l = 0; l < (1 << b); l++) {
                const int j0 = (int)jb | l, j1 = j0 | (1 << b);
                const uint32_t a = v[j0];
                if (INV) {
                    v[j0] = m31_add(a, v[j1]);
                    v[j1] = m31_mul_dbl(m31_sub(a, v[j1]), w);
                } else {
                    const uint32_t m = m31_mul_dbl(v[j1], w);
                    v[j0] = m31_add(a, m);
                    v[j1] = m31_sub(a, m);
                }
            }
        }
    }
}

template <bool INV, int C, int LO, int r, int i>
__device__ __forceinline__ void g_stages_lds(uint32_t (&v)[TC<C>::R], uint32_t lane, const uint32_t *twl) {
    constexpr int b = INV ? i : (r - 1 - i);
    g_stage<INV, C, LO, r, b>(v, lane, twl);
    if constexpr (i + 1 < r) g_stages_lds<INV, C, LO, r, i + 1>(v, lane, twl);
}

// The same stages with every twiddle of the group read from LDS up front (2^r - 1
// per lane for a full group), before the first butterfly: left to itself the
// compiler issues each read just before its butterfly and waits for it there, one
// LDS latency per butterfly pair.  The scheduling barrier keeps the reads ahead.
template <int C, int LO, int r, int b>
struct TwIdx {
    static constexpr int RB = TC<C>::RB;
    static constexpr int per_o = (1 << (r - 1 - b));                       // distinct twiddles of stage b per o
    static constexpr int base = (1 << (RB - r)) * ((1 << r) - (1 << (r - b)));  // stages below b first
};
template <bool INV, int C, int LO, int r, int b>
__device__ __forceinline__ void g_stage_w(uint32_t (&v)[TC<C>::R], const uint32_t *w) {
    constexpr int RB = TC<C>::RB;
    using TI = TwIdx<C, LO, r, b>;
#pragma unroll
    for (int o = 0; o < (1 << (RB - r)); o++) {
#pragma unroll
        for (int u = 0; u < (1 << (r - 1 - b)); u++) {
            const uint32_t wt = w[TI::base + o * TI::per_o + u];
            const int jb = (o << r) | (u << (b + 1));
#pragma unroll
            for (int l = 0; l < (1 << b); l++) {
                const int j0 = jb | l, j1 = j0 | (1 << b);
                const uint32_t a = v[j0];
                if (INV) {
                    v[j0] = m31_add(a, v[j1]);
                    v[j1] = m31_mul_dbl(m31_sub(a, v[j1]), wt);
                } else {
                    const uint32_t m = m31_mul_dbl(v[j1], wt);
                    v[j0] = m31_add(a, m);
                    v[j1] = m31_sub(a, m);
                }
            }
        }
    }
}
template <int C, int LO, int r, int b>
__device__ __forceinline__ void g_load_tw(uint32_t *w, uint32_t lane, const uint32_t *twl) {
    constexpr int RB = TC<C>::RB, ls = LO + b;
    using TI = TwIdx<C, LO, r, b>;
    const uint32_t *tws = twl + (1u << (C - 1 - ls));
#pragma unroll
    for (int o = 0; o < (1 << (RB - r)); o++) {
#pragma unroll
        for (int u = 0; u < (1 << (r - 1 - b)); u++) {
            const uint32_t jb = ((uint32_t)o << r) | ((uint32_t)u << (b + 1));
            const uint32_t t0 = t_of<RB, LO, r>(lane, jb);
            w[TI::base + o * TI::per_o + u] = tws[tw_pos<C>(ls, t0 >> (ls + 1))];
        }
    }
    if constexpr (b > 0) g_load_tw<C, LO, r, b - 1>(w, lane, twl);
}
template <bool INV, int C, int LO, int r, int i>
__device__ __forceinline__ void g_stages_w(uint32_t (&v)[TC<C>::R], const uint32_t *w) {
    constexpr int b = INV ? i : (r - 1 - i);
    g_stage_w<INV, C, LO, r, b>(v, w);
    if constexpr (i + 1 < r) g_stages_w<INV, C, LO, r, i + 1>(v, w);
}
#ifndef ZS_TW_PRELOAD
#define ZS_TW_PRELOAD 1
#endif
template <bool INV, int C, int LO, int r, int i>
__device__ __forceinline__ void g_stages(uint32_t (&v)[TC<C>::R], uint32_t lane, const uint32_t *twl) {
    static_assert(i == 0, "a group's stages run together");
    // an 11-stage tile holds 32 values per lane: its 31 group twiddles stay in LDS
    if constexpr (ZS_TW_PRELOAD && C <= 10) {
        constexpr int NW = (1 << (TC<C>::RB - r)) * ((1 << r) - 1);
        uint32_t w[NW];
        g_load_tw<C, LO, r, r - 1>(w, lane, twl);
        __builtin_amdgcn_sched_barrier(0);
        g_stages_w<INV, C, LO, r, 0>(v, w);
    } else {
        g_stages_lds<INV, C, LO, r, 0>(v, lane, twl);
    }
}

// group GI in execution order (forward: top bits first; inverse: bottom first)
template <bool INV, int C, int GI>
struct Grp {
    static constexpr int RB = TC<C>::RB, NG = TC<C>::NG;
    static constexpr int GF = INV ? (NG - 1 - GI) : GI;
    static constexpr int HI = C - 1 - GF * RB;
    static constexpr int LO = (C - (GF + 1) * RB) > 0 ? (C - (GF + 1) * RB) : 0;
    static constexpr int r = HI - LO + 1;
};

// groups GI.. of a tile held in LDS
template <bool INV, int C, int GI>
__device__ __forceinline__ void tile_run(uint32_t *tl, uint32_t lane, const uint32_t *twl) {
    using Gp = Grp<INV, C, GI>;
    uint32_t v[TC<C>::R];
    g_load<C, Gp::LO, Gp::r>(v, tl, lane);
    g_stages<INV, C, Gp::LO, Gp::r, 0>(v, lane, twl);
    g_store<C, Gp::LO, Gp::r>(v, tl, lane);
    wave_lds_sync();
    if constexpr (GI + 1 < TC<C>::NG) tile_run<INV, C, GI + 1>(tl, lane, twl);
}

// The lowest register group of a forward tile (LO = 0, RL stages) with its
// twiddles in registers instead of LDS.  Stage b < RL reads entries
// (lane << (RB-1-b)) | (o << (RL-1-b)) | u of its heap level: 2^(RB-1-b)
// consecutive words per lane, the same for every column a workgroup transforms,
// so they are loaded once; the LDS set then holds only heap slots below 2^SLOG.
__host__ __device__ constexpr int low_group_stages(int C) {
    return C - ((C + C - 7) / (C - 6) - 1) * (C - 6);  // C - (NG - 1) RB
}
template <int C>
struct LowTw {
    static constexpr int RB = TC<C>::RB, NG = TC<C>::NG;
    static constexpr int RL = low_group_stages(C);          // stages in the lowest group
    static_assert(RL == C - (NG - 1) * RB, "lowest group");
    static constexpr int N = (1 << RB) - (1 << (RB - RL));  // registers per lane
    static constexpr int SLOG = C - RL;                     // LDS set: heap slots below 2^SLOG
    static constexpr int off(int b) { return (1 << RB) - (1 << (RB - b)); }  // stage b's first register
};

template <bool INV, int C, int b>
__device__ __forceinline__ void g_stage_rt(uint32_t (&v)[TC<C>::R], const uint32_t (&tr)[LowTw<C>::N]) {
    constexpr int RB = TC<C>::RB, r = LowTw<C>::RL;
#pragma unroll
    for (int o = 0; o < (1 << (RB - r)); o++) {
#pragma unroll
        for (int u = 0; u < (1 << (r - 1 - b)); u++) {
            const uint32_t w = tr[LowTw<C>::off(b) + ((o << (r - 1 - b)) | u)];
            const int jb = (o << r) | (u << (b + 1));
#pragma unroll
            for (int l = 0; l < (1 << b); l++) {
                const int j0 = jb | l, j1 = j0 | (1 << b);
                const uint32_t a = v[j0];
                if (INV) {
                    v[j0] = m31_add(a, v[j1]);
                    v[j1] = m31_mul_dbl(m31_sub(a, v[j1]), w);
                } else {
                    const uint32_t m = m31_mul_dbl(v[j1], w);
                    v[j0] = m31_add(a, m);
                    v[j1] = m31_sub(a, m);
                }
            }
        }
    }
}
template <bool INV, int C, int i>
__device__ __forceinline__ void g_stages_rt(uint32_t (&v)[TC<C>::R], const uint32_t (&tr)[LowTw<C>::N]) {
    constexpr int r = LowTw<C>::RL;
    g_stage_rt<INV, C, INV ? i : r - 1 - i>(v, tr);
    if constexpr (i + 1 < r) g_stages_rt<INV, C, i + 1>(v, tr);
}
// groups GI.. of a tile in LDS, the lowest one from register twiddles
template <bool INV, int C, int GI>
__device__ __forceinline__ void tile_run_rt(uint32_t *tl, uint32_t lane, const uint32_t *twl,
                                            const uint32_t (&tr)[LowTw<C>::N]) {
    using Gp = Grp<INV, C, GI>;
    uint32_t v[TC<C>::R];
    g_load<C, Gp::LO, Gp::r>(v, tl, lane);
    if constexpr (Gp::GF + 1 < TC<C>::NG) g_stages<INV, C, Gp::LO, Gp::r, 0>(v, lane, twl);
    else g_stages_rt<INV, C, 0>(v, tr);
    g_store<C, Gp::LO, Gp::r>(v, tl, lane);
    wave_lds_sync();
    if constexpr (GI + 1 < TC<C>::NG) tile_run_rt<INV, C, GI + 1>(tl, lane, twl, tr);
}
// Groups GI.. of a tile whose group-GI values are already in v (a register copy
// of the layout g_load would give), so the tile's first LDS round trip is
// skipped.  KEEP: the last group's values stay in v (nothing stored); otherwise
// they are stored to LDS, as tile_run leaves them.  RT: the lowest group takes
// its twiddles from tr.
template <bool INV, int C, int GI, bool KEEP, bool RT>
__device__ __forceinline__ void tile_regs(uint32_t (&v)[TC<C>::R], uint32_t *tl, uint32_t lane, const uint32_t *twl,
                                          const uint32_t (&tr)[LowTw<C>::N]) {
    using Gp = Grp<INV, C, GI>;
    if constexpr (RT && Gp::GF + 1 == TC<C>::NG) g_stages_rt<INV, C, 0>(v, tr);
    else g_stages<INV, C, Gp::LO, Gp::r, 0>(v, lane, twl);
    if constexpr (GI + 1 < TC<C>::NG) {
        using Gn = Grp<INV, C, GI + 1>;
        g_store<C, Gp::LO, Gp::r>(v, tl, lane);
        wave_lds_sync();
        g_load<C, Gn::LO, Gn::r>(v, tl, lane);
        tile_regs<INV, C, GI + 1, KEEP, RT>(v, tl, lane, twl, tr);
    } else if constexpr (!KEEP) {
        g_store<C, Gp::LO, Gp::r>(v, tl, lane);
        wave_lds_sync();
    }
}
// the same for a tile held in LDS: group 0 is loaded first
template <bool INV, int C, bool KEEP, bool RT>
__device__ __forceinline__ void tile_lds(uint32_t (&v)[TC<C>::R], uint32_t *tl, uint32_t lane, const uint32_t *twl,
                                         const uint32_t (&tr)[LowTw<C>::N]) {
    using G0 = Grp<INV, C, 0>;
    g_load<C, G0::LO, G0::r>(v, tl, lane);
    tile_regs<INV, C, 0, KEEP, RT>(v, tl, lane, twl, tr);
}

// this lane's lowest-group twiddles (doubled) of the tile whose heap entries start at
// heap(ls) + (kb >> (ls + 1)) in the table of log LT
template <int C>
__device__ __forceinline__ void load_low_twiddles(uint32_t (&tr)[LowTw<C>::N], const uint32_t *tw, uint32_t LT,
                                                  uint32_t kb, uint32_t lane) {
    constexpr int RB = TC<C>::RB;
#pragma unroll
    for (int b = 0; b < LowTw<C>::RL; b++) {
        const uint32_t *src = tw + (1u << (LT - 1 - b)) + (kb >> (b + 1)) + (lane << (RB - 1 - b));
#pragma unroll
        for (int k = 0; k < (1 << (RB - 1 - b)); k++) tr[LowTw<C>::off(b) + k] = ldg(src + k) << 1;
    }
}

// count of i >= 0 with first + i * step < end
__device__ __forceinline__ uint32_t strided_count(uint32_t first, uint32_t step, uint32_t end) {
    return first < end ? (end - first + step - 1) / step : 0u;
}

// dst(i) <- src(i) for i < count, in batches: LOAD(i) reads global memory, STORE(i, v) writes LDS
template <uint32_t B, typename LOAD, typename STORE>
__device__ __forceinline__ void batched_copy(uint32_t count, LOAD load, STORE store) {
    for (uint32_t i0 = 0; i0 < count; i0 += B) {
        uint32_t v[B];
#pragma unroll
        for (uint32_t k = 0; k < B; k++)
            if (i0 + k < count) v[k] = load(i0 + k);
#pragma unroll
        for (uint32_t k = 0; k < B; k++)
            if (i0 + k < count) store(i0 + k, v[k]);
    }
}

// Cooperatively copy the 2^C - 1 twiddles a tile set needs into LDS (heap
// layout).  Stage s = s_lo + ls of the global table (log size LT) is read at
// (Kbase >> (s+1)) + i: the tile's fixed bits select one contiguous segment
// per stage, so the copy is coalesced.  Entries are stored doubled, as
// m31_mul_dbl takes them, at the rotated positions tw_pos gives.
template <int C>
__device__ __forceinline__ void stage_twiddles(uint32_t *twl, const uint32_t *tw, uint32_t LT, uint32_t s_lo,
                                               uint32_t Kbase) {
    const uint32_t t0 = threadIdx.x + 1, step = blockDim.x;
    batched_copy<4>(
        strided_count(t0, step, 1u << C),
        [&](uint32_t i) {
            const uint32_t idx = t0 + i * step, lvl = 31 - __builtin_clz(idx), s = s_lo + (C - 1 - lvl);
            return ldg(tw + (1u << (LT - 1 - s)) + (Kbase >> (s + 1)) + (idx - (1u << lvl)));
        },
        [&](uint32_t i, uint32_t v) {
            const uint32_t idx = t0 + i * step, lvl = 31 - __builtin_clz(idx);
            twl[(1u << lvl) + tw_pos<C>(C - 1 - (int)lvl, idx - (1u << lvl))] = v << 1;
        });
}

// Stage nsets twiddle sets at once (set k at twl + (k << C), same layout as
// stage_twiddles).  A thread covers heap slots idx = tid + m * blockDim.x of
// every set, so their stage, entry and LDS position are worked out once per m
// and reused for all sets; the loads of up to 8 sets x 4 slots are in flight
// together.  entry(set, ls, i) reads entry i of local stage ls of the set's table.
template <int C, int SLOG = C, typename ENTRY>
__device__ __forceinline__ void stage_twiddle_sets(uint32_t *twl, uint32_t nsets, ENTRY entry) {
    constexpr uint32_t T = 1u << SLOG, KM = 4, SB = 8;  // the sets keep heap slots below 2^SLOG
    const uint32_t nt = blockDim.x, tid = threadIdx.x;
    for (uint32_t m0 = 0; m0 * nt < T; m0 += KM) {
        uint32_t ls[KM], ent[KM], pos[KM];
        bool ok[KM];
#pragma unroll
        for (uint32_t k = 0; k < KM; k++) {
            const uint32_t idx = tid + (m0 + k) * nt;
            ok[k] = idx != 0 && idx < T;  // heap slot 0 is unused
            const uint32_t x = ok[k] ? idx : 1u, lvl = 31 - __builtin_clz(x);
            ls[k] = C - 1 - lvl;
            ent[k] = x - (1u << lvl);
            pos[k] = (1u << lvl) + tw_pos<C>((int)ls[k], ent[k]);
        }
        for (uint32_t s0 = 0; s0 < nsets; s0 += SB) {
            uint32_t v[SB][KM];
#pragma unroll
            for (uint32_t b = 0; b < SB; b++)
#pragma unroll
                for (uint32_t k = 0; k < KM; k++)
                    v[b][k] = (s0 + b < nsets && ok[k]) ? entry(s0 + b, ls[k], ent[k]) : 0u;
#pragma unroll
            for (uint32_t b = 0; b < SB; b++)
#pragma unroll
                for (uint32_t k = 0; k < KM; k++)
                    if (s0 + b < nsets && ok[k]) twl[((s0 + b) << SLOG) + pos[k]] = v[b][k] << 1;
        }
    }
}

// workgroup -> tile group, remapped so consecutive tile groups share an XCD
// (workgroups are dealt round-robin over the 8 XCDs, each with its own L2)
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblk) {
    if (nblk % 8) return bid;
    return (bid & 7u) * (nblk >> 3) + (bid >> 3);
}

struct TileGeom {
    uint32_t hi, lo0, tau0;
};

__device__ __forceinline__ TileGeom tile_geom(const FftPass &p, uint32_t blk) {
    TileGeom g{0, 0, 0};
    if (p.s_lo > 0) {
        uint32_t lo_groups = (1u << p.s_lo) >> p.G_log;
        g.hi = blk / lo_groups;
        g.lo0 = (blk % lo_groups) << p.G_log;
    } else {
        g.tau0 = blk << p.G_log;
    }
    return g;
}
__device__ __forceinline__ uint32_t tile_fixed(const FftPass &p, const TileGeom &tg, uint32_t g) {
    if (p.s_lo > 0) return (tg.hi << (p.s_lo + p.C)) | (tg.lo0 + g);
    uint32_t h = tg.tau0 + g;
    if (p.natural_tiles) h = bitrev32(h, p.L - p.C);
    return h << p.C;
}

// cooperative global -> LDS load of the workgroup's G * 2^qload tiles
template <int C>
__device__ __forceinline__ void coop_load_generic(const FftPass &p, const TileGeom &tg, uint32_t *lds, uint32_t col,
                                          uint32_t qblk, uint32_t nq_log, bool replicate_q) {
    constexpr uint32_t T = 1u << C;
    const uint32_t TS = p.TS;
    const uint32_t G = 1u << p.G_log, Q = 1u << p.Q_log;
    const bool g_fast = (p.s_lo > 0) || p.natural_tiles;
    const uint32_t half_n = 1u << (p.L - 1);
    const uint32_t *in_col = p.in + (uint64_t)col * p.in_cs;
    const uint32_t nload = (G << nq_log) << C;
    for (uint32_t e = threadIdx.x; e < nload; e += blockDim.x) {
        uint32_t g, t, ql;
        if (g_fast) {
            g = e & (G - 1);
            uint32_t rr = e >> p.G_log;
            ql = rr & ((1u << nq_log) - 1);
            t = rr >> nq_log;
        } else {
            t = e & (T - 1);
            uint32_t rr = e >> C;
            g = rr & (G - 1);
            ql = rr >> p.G_log;
        }
        uint32_t k = tile_fixed(p, tg, g) | (t << p.s_lo);
        uint64_t addr;
        switch (p.in_mode) {
            case IN_KSPACE: addr = ((uint64_t)(qblk + ql) << p.L) | k; break;
            case IN_SHARED: addr = k; break;
            case IN_NATURAL: addr = bitrev32(k, p.L); break;
            default: addr = (k & 1u) ? (half_n + (k >> 1)) : (k >> 1); break;
        }
        uint32_t v = in_col[addr];
        if (replicate_q) {
            for (uint32_t qq = 0; qq < Q; qq++) lds[(qq * G + g) * TS + lidx(t)] = v;
        } else {
            lds[(ql * G + g) * TS + lidx(t)] = v;
        }
    }
}

// cooperative LDS -> global store of nq tiles-per-g (q runs over [qblk, qblk + 2^nq_log))
template <int C>
__device__ __forceinline__ void coop_store_generic(const FftPass &p, const TileGeom &tg, const uint32_t *lds, uint32_t col,
                                           uint32_t qblk, uint32_t nq_log) {
    constexpr uint32_t T = 1u << C;
    const uint32_t TS = p.TS;
    const uint32_t G = 1u << p.G_log, NQ = 1u << nq_log;
    const bool g_fast = (p.s_lo > 0) || p.natural_tiles;
    const uint32_t half_n = 1u << (p.L - 1);
    uint32_t *out_col = p.out + (uint64_t)col * p.out_cs;
    const uint32_t nstore = (G * NQ) << C;
    for (uint32_t e = threadIdx.x; e < nstore; e += blockDim.x) {
        uint32_t g, t, ql;
        if (p.out_mode == OUT_NATURAL) {
            uint32_t rq = e & (NQ - 1);
            g = (e >> nq_log) & (G - 1);
            t = e >> (nq_log + p.G_log);
            ql = bitrev32(rq, nq_log);
        } else if (g_fast) {
            g = e & (G - 1);
            uint32_t rr = e >> p.G_log;
            ql = rr & (NQ - 1);
            t = rr >> nq_log;
        } else {
            t = e & (T - 1);
            uint32_t rr = e >> C;
            g = rr & (G - 1);
            ql = rr >> p.G_log;
        }
        uint32_t v = lds[(ql * G + g) * TS + lidx(t)];
        if (p.scale) v = m31_mul(v, p.scale);
        uint32_t k = tile_fixed(p, tg, g) | (t << p.s_lo);
        uint32_t q = qblk + ql;
        uint64_t addr;
        switch (p.out_mode) {
            case OUT_KSPACE: addr = ((uint64_t)q << p.L) | k; break;
            case OUT_NATURAL: addr = ((uint64_t)bitrev32(k, p.L) << p.beta) | bitrev32(q, p.beta); break;
            default: addr = (k & 1u) ? (half_n + (k >> 1)) : (k >> 1); break;
        }
        out_col[addr] = v;
    }
}


// Cooperative copies with the hot layouts specialised: the per-thread part of
// each address (its run position) is computed once, and k-space tiles are
// streamed with t fastest so every wave access is one contiguous run.  Each
// thread moves its elements in batches of B: all of a batch's global loads
// (or LDS reads) are issued before the first is waited for, so a column's copy
// costs a few memory latencies instead of one per element.  The batch is 8
// values in fft_tile_kernel and 4 in lde_single_kernel, whose registers are
// fuller (measured: larger batches spill there).

// In the fast paths a thread's element i sits at t = t0 + i * tstep (or at
// t = tid + j * nt within tile i >> per_log) with t0 < tstep, so every address
// splits into a part fixed per thread, computed once, and a part that depends
// only on i and is the same for the whole workgroup (scalar registers): the
// index bits of t0 and of i * tstep are disjoint, which also makes
// bitrev_C(t) = bitrev_C(t0) + bitrev_{C - log2 tstep}(i) and
// lidx(t) = lidx(t0) + lidx(i * tstep).
template <int C, uint32_t B = 8>
__device__ __forceinline__ void coop_load(const FftPass &p, const TileGeom &tg, uint32_t *lds, uint32_t col,
                                          uint32_t qblk, uint32_t nq_log, bool replicate_q) {
    constexpr uint32_t T = 1u << C;
    const uint32_t TS = p.TS;
    const uint32_t G = 1u << p.G_log, nt = blockDim.x, tid = threadIdx.x;
    const uint32_t lnt = __builtin_ctz(nt);  // block sizes are powers of two
    const uint32_t *in_col = p.in + (uint64_t)col * p.in_cs;
    if (p.in_mode == IN_NATURAL && p.s_lo == 0 && p.natural_tiles && nq_log == 0 && G <= nt &&
        lnt - p.G_log <= (uint32_t)C) {
        // element (g, t) sits at natural (bitrev_C(t) << (L - C)) | (tau0 + g): runs of G along g
        const uint32_t g = tid & (G - 1), lts = lnt - p.G_log, t0 = tid >> p.G_log, sh = p.L - C;
        const uint32_t *src = in_col + tg.tau0 + g + ((uint64_t)bitrev32(t0, C) << sh);
        uint32_t *dst = lds + g * TS + lidx(t0);
        batched_copy<B>(
            T >> lts, [&](uint32_t i) { return src[(uint64_t)bitrev32(i, C - lts) << sh]; },
            [&](uint32_t i, uint32_t v) { dst[lidx(i << lts)] = v; });
        return;
    }
    if ((p.in_mode == IN_KSPACE || p.in_mode == IN_SHARED) && p.s_lo > 0 && G <= nt &&
        lnt - p.G_log <= (uint32_t)C) {
        // strided tiles: (q, hi, t, lo0 + g) -> runs of G along g
        const uint32_t g = tid & (G - 1), lts = lnt - p.G_log, t0 = tid >> p.G_log, per_log = C - lts;
        const uint32_t nq = (p.in_mode == IN_SHARED) ? 1u : (1u << nq_log);
        const uint32_t *src = in_col + tg.lo0 + g + ((uint64_t)t0 << p.s_lo);
        const uint64_t hi_part = (uint64_t)tg.hi << (p.s_lo + C);
        uint32_t *dst = lds + g * TS + lidx(t0);
        batched_copy<B>(
            nq << per_log,
            [&](uint32_t i) {
                const uint32_t ql = i >> per_log, j = i & ((1u << per_log) - 1);
                const uint64_t q_part = p.in_mode == IN_SHARED ? 0ull : ((uint64_t)(qblk + ql) << p.L);
                return src[q_part + hi_part + ((uint64_t)j << (lts + p.s_lo))];
            },
            [&](uint32_t i, uint32_t v) {
                const uint32_t ql = i >> per_log, j = i & ((1u << per_log) - 1);
                if (replicate_q) {
                    for (uint32_t qq = 0; qq < (1u << p.Q_log); qq++) dst[qq * G * TS + lidx(j << lts)] = v;
                } else {
                    dst[ql * G * TS + lidx(j << lts)] = v;
                }
            });
        return;
    }
    if (p.in_mode == IN_KSPACE && p.s_lo == 0 && !replicate_q && nt <= T) {
        // contiguous tiles, t fastest: element i of the thread is tile i >> per_log, t = tid + j * nt
        const uint32_t per_log = C - lnt;
        const uint32_t *src = in_col + tid;
        uint32_t *dst = lds + lidx(tid);
        batched_copy<B>(
            (G << nq_log) << per_log,
            [&](uint32_t i) {
                const uint32_t tile = i >> per_log, j = i & ((1u << per_log) - 1);
                const uint32_t g = tile & (G - 1), ql = tile >> p.G_log;
                return src[(((uint64_t)(qblk + ql) << p.L) | tile_fixed(p, tg, g)) + (j << lnt)];
            },
            [&](uint32_t i, uint32_t v) {
                const uint32_t tile = i >> per_log, j = i & ((1u << per_log) - 1);
                dst[tile * TS + lidx(j << lnt)] = v;
            });
        return;
    }
    coop_load_generic<C>(p, tg, lds, col, qblk, nq_log, replicate_q);
}

template <int C, uint32_t B = 8>
__device__ __forceinline__ void coop_store(const FftPass &p, const TileGeom &tg, const uint32_t *lds, uint32_t col,
                                           uint32_t qblk, uint32_t nq_log) {
    constexpr uint32_t T = 1u << C;
    const uint32_t TS = p.TS;
    const uint32_t G = 1u << p.G_log, nt = blockDim.x, tid = threadIdx.x;
    const uint32_t lnt = __builtin_ctz(nt);
    uint32_t *out_col = p.out + (uint64_t)col * p.out_cs;
    const uint32_t run_log = p.G_log + nq_log;
    if (p.out_mode == OUT_NATURAL && p.s_lo == 0 && p.natural_tiles && run_log <= lnt && lnt - run_log <= (uint32_t)C &&
        !p.scale) {
        // element (q, g, t) goes to (bitrev_C(t) << (L-C+beta)) | ((tau0 + g) << beta) | bitrev_beta(q):
        // a thread keeps one run position rp = (g, bitrev(q)) and walks t
        const uint32_t rp = tid & ((1u << run_log) - 1), lts = lnt - run_log, t0 = tid >> run_log;
        const uint32_t rq = rp & ((1u << nq_log) - 1), g = rp >> nq_log;
        const uint32_t ql = bitrev32(rq, nq_log), q = qblk + ql;
        const uint32_t sh = p.L - C + p.beta;
        const uint64_t low = ((uint64_t)(tg.tau0 + g) << p.beta) | bitrev32(q, p.beta);
        const uint32_t *src = lds + (ql * G + g) * TS + lidx(t0);
        uint32_t *dst = out_col + low + ((uint64_t)bitrev32(t0, C) << sh);
        batched_copy<B>(
            T >> lts, [&](uint32_t i) { return src[lidx(i << lts)]; },
            [&](uint32_t i, uint32_t v) { dst[(uint64_t)bitrev32(i, C - lts) << sh] = v; });
        return;
    }
    if (p.out_mode == OUT_KSPACE && p.s_lo > 0 && G <= nt && lnt - p.G_log <= (uint32_t)C && !p.scale) {
        const uint32_t g = tid & (G - 1), lts = lnt - p.G_log, t0 = tid >> p.G_log, per_log = C - lts;
        const uint32_t *src = lds + g * TS + lidx(t0);
        uint32_t *dst = out_col + tg.lo0 + g + ((uint64_t)t0 << p.s_lo);
        const uint64_t hi_part = (uint64_t)tg.hi << (p.s_lo + C);
        batched_copy<B>(
            (1u << nq_log) << per_log,
            [&](uint32_t i) {
                const uint32_t ql = i >> per_log, j = i & ((1u << per_log) - 1);
                return src[ql * G * TS + lidx(j << lts)];
            },
            [&](uint32_t i, uint32_t v) {
                const uint32_t ql = i >> per_log, j = i & ((1u << per_log) - 1);
                dst[((uint64_t)(qblk + ql) << p.L) + hi_part + ((uint64_t)j << (lts + p.s_lo))] = v;
            });
        return;
    }
    if (p.out_mode == OUT_KSPACE && p.s_lo == 0 && !p.scale && nt <= T) {
        const uint32_t per_log = C - lnt;
        const uint32_t *src = lds + lidx(tid);
        uint32_t *dst = out_col + tid;
        batched_copy<B>(
            (G << nq_log) << per_log,
            [&](uint32_t i) {
                const uint32_t tile = i >> per_log, j = i & ((1u << per_log) - 1);
                return src[tile * TS + lidx(j << lnt)];
            },
            [&](uint32_t i, uint32_t v) {
                const uint32_t tile = i >> per_log, j = i & ((1u << per_log) - 1);
                const uint32_t g = tile & (G - 1), ql = tile >> p.G_log;
                dst[(((uint64_t)(qblk + ql) << p.L) | tile_fixed(p, tg, g)) + (j << lnt)] = v;
            });
        return;
    }
    coop_store_generic<C>(p, tg, lds, col, qblk, nq_log);
}

// ---------------------------------------------------------------------------
// one pass (same contract as fft_pass_kernel)
// ---------------------------------------------------------------------------
template <int C>
constexpr int tile_block_max() { return C >= 11 ? 512 : 1024; }

template <bool INV, int C>
__device__ __forceinline__ void fft_tile_body(const FftPass &p, const VB vb) {
    extern __shared__ uint32_t lds[];
    const uint32_t TS = p.TS;
    const uint32_t G = 1u << p.G_log, Q = 1u << p.Q_log;
    const uint32_t qblk = vb.z << p.Q_log;
    const TileGeom tg = tile_geom(p, xcd_block(vb.x, vb.gx));
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nwaves = blockDim.x >> 6;
    // tiles of a strided pass share their twiddles per q; contiguous-pass tiles each have their own.
    // The sets are staged once and reused for the workgroup's ncb columns.
    const bool per_tile_tw = p.s_lo == 0;
    const uint32_t nsets = per_tile_tw ? Q * G : Q;
    uint32_t *twl = lds + Q * G * TS;
    for (uint32_t set = 0; set < nsets; set++) {
        uint32_t ql = per_tile_tw ? (set >> p.G_log) : set;
        uint32_t g = per_tile_tw ? (set & (G - 1)) : 0;
        uint32_t Kbase = ((qblk + ql) << p.L) | tile_fixed(p, tg, g);
        stage_twiddles<C>(twl + (set << C), p.tw, p.LT, p.s_lo, Kbase);
    }
    const bool shared_in = p.in_mode == IN_SHARED;
    const uint32_t c0 = vb.y * p.ncb;
    const uint32_t c1 = min(c0 + p.ncb, p.n_cols);
    for (uint32_t col = c0; col < c1; col++) {
        coop_load<C>(p, tg, lds, col, qblk, shared_in ? 0 : p.Q_log, shared_in);
        __syncthreads();
        for (uint32_t tile = wave; tile < Q * G; tile += nwaves) {
            uint32_t set = per_tile_tw ? tile : (tile >> p.G_log);
            tile_run<INV, C, 0>(lds + tile * TS, lane, twl + (set << C));
        }
        __syncthreads();
        coop_store<C>(p, tg, lds, col, qblk, p.Q_log);
        __syncthreads();
    }
}

// The LDE's last evaluation pass (contiguous k-space tiles in, natural order
// out; s_lo == 0, no scale), one wave per tile of the workgroup's G tiles x Q
// cosets.  Each wave loads and transforms its own tile, so only the natural-order
// store, whose runs interleave every tile of the workgroup, needs the workgroup.
// The next column's tile is loaded into registers while the current column is
// transformed, and a column's stores are issued after the next column's tile has
// landed in LDS, so they drain during its butterflies: HBM latency stays off the
// critical path.  A thread's store run position is (g, bitrev(q)) with
// t = t0 + 64 i, as in coop_store's natural-order path.
template <int C>
__device__ __forceinline__ void lde_eval_last_body(const FftPass &p, const VB vb) {
    constexpr uint32_t PER = (1u << C) >> 6;  // tile elements per lane
    using LW = LowTw<C>;
    extern __shared__ uint32_t lds[];
    const uint32_t TS = p.TS, G = 1u << p.G_log, Q = 1u << p.Q_log;
    const uint32_t qblk = vb.z << p.Q_log;
    const uint32_t tau0 = xcd_block(vb.x, vb.gx) << p.G_log;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    uint32_t *twl = lds + Q * G * TS;
    // this wave's tile: coset qblk + (wave >> G_log), k-space tile bitrev(tau0 + (wave & (G - 1)))
    const uint64_t in_off = ((uint64_t)(qblk + (wave >> p.G_log)) << p.L) |
                            ((uint64_t)bitrev32(tau0 + (wave & (G - 1)), p.L - C) << C);
    uint32_t *tl = lds + wave * TS;
    const uint32_t *twt = twl + (wave << LW::SLOG);
    // this thread's natural-order store run position
    const uint32_t run_log = p.G_log + p.Q_log;
    const uint32_t rp = tid & ((1u << run_log) - 1), rq = rp & (Q - 1), g = rp >> p.Q_log;
    const uint32_t ql = bitrev32(rq, p.Q_log), t0 = tid >> run_log;
    const uint32_t sh = p.L - C + p.beta;
    const uint64_t out_off = (((uint64_t)(tau0 + g)) << p.beta) | bitrev32(qblk + ql, p.beta) |
                             ((uint64_t)bitrev32(t0, C) << sh);
    const uint32_t *rd = lds + (ql * G + g) * TS + lidx(t0);
    const uint32_t c0 = vb.y * p.ncb, c1 = min(c0 + p.ncb, p.n_cols);
    uint32_t pf[PER], res[PER];
    if (c0 < c1) {
        const uint32_t *src = p.in + (uint64_t)c0 * p.in_cs + in_off + lane;
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) pf[j] = ldg(src + j * 64);
    }
    // tile set (q, g) covers k-space ((qblk + q) << L) | (bitrev(tau0 + g) << C); the lowest
    // group's stages are this wave's registers, the others its set in LDS
    const auto set_kb = [&](uint32_t set) {
        return ((qblk + (set >> p.G_log)) << p.L) | (bitrev32(tau0 + (set & (G - 1)), p.L - C) << C);
    };
    uint32_t tr[LW::N];
    load_low_twiddles<C>(tr, p.tw, p.LT, set_kb(wave), lane);
    stage_twiddle_sets<C, LW::SLOG>(twl, Q * G, [&](uint32_t set, uint32_t ls, uint32_t i) {
        return ldg(p.tw + (1u << (p.LT - 1 - ls)) + (set_kb(set) >> (ls + 1)) + i);
    });
    __syncthreads();  // twiddle sets staged
    for (uint32_t col = c0; col < c1; col++) {
        // the top group's layout is the load's (element lane + 64 j): it starts from registers
        uint32_t v[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) v[j] = pf[j];
        if (col > c0) {
            uint32_t *dst = p.out + (uint64_t)(col - 1) * p.out_cs + out_off;
#pragma unroll
            for (uint32_t i = 0; i < PER; i++) stg(dst + ((uint64_t)bitrev32(i, C - 6) << sh), res[i]);
        }
        if (col + 1 < c1) {
            const uint32_t *src = p.in + (uint64_t)(col + 1) * p.in_cs + in_off + lane;
#pragma unroll
            for (uint32_t j = 0; j < PER; j++) pf[j] = ldg(src + j * 64);
        }
        tile_regs<false, C, 0, false, true>(v, tl, lane, twt, tr);
        __syncthreads();  // every tile of the workgroup transformed
#pragma unroll
        for (uint32_t i = 0; i < PER; i++) res[i] = rd[lidx(i << 6)];
        __syncthreads();  // every read done: the tiles may take the next column
    }
    if (c0 < c1) {
        uint32_t *dst = p.out + (uint64_t)(c1 - 1) * p.out_cs + out_off;
#pragma unroll
        for (uint32_t i = 0; i < PER; i++) stg(dst + ((uint64_t)bitrev32(i, C - 6) << sh), res[i]);
    }
}

// The LDE's first interpolation pass (natural-order input, contiguous k-space
// tiles out; s_lo == 0, no scale), one wave per tile of the workgroup's G tiles.
// The natural-order load interleaves the G tiles (runs of G words), so it goes
// through the workgroup; each wave then transforms and stores its own tile.  The
// next column is loaded into registers during the butterflies, and a column's
// stores are issued after the next column has landed in LDS.
template <int C>
__device__ __forceinline__ void lde_interp_first_body(const FftPass &p, const VB vb) {
    constexpr uint32_t PER = (1u << C) >> 6;
    using LW = LowTw<C>;
    extern __shared__ uint32_t lds[];
    const uint32_t TS = p.TS, G = 1u << p.G_log;
    const uint32_t tau0 = xcd_block(vb.x, vb.gx) << p.G_log;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    uint32_t *twl = lds + G * TS;
    // load: thread (g, t0) reads elements t = t0 + 64 j of tile g, natural index
    // (bitrev_C(t) << (L - C)) | (tau0 + g), bitrev_C(t) = (bitrev_6(t0) << (C - 6)) | bitrev_{C-6}(j)
    const uint32_t g = tid & (G - 1), t0 = tid >> p.G_log, sh = p.L - C;
    const uint64_t in_off = (uint64_t)(tau0 + g) + ((uint64_t)bitrev32(t0, C) << sh);
    uint32_t *cp = lds + g * TS + lidx(t0);
    // store: this wave's tile, k-space tile bitrev(tau0 + wave), contiguous
    const uint64_t out_off = ((uint64_t)bitrev32(tau0 + wave, p.L - C) << C) + lane;
    uint32_t *tl = lds + wave * TS;
    const uint32_t *twt = twl + (wave << LW::SLOG);
    const uint32_t c0 = vb.y * p.ncb, c1 = min(c0 + p.ncb, p.n_cols);
    uint32_t pf[PER], res[PER];
    if (c0 < c1) {
        const uint32_t *src = p.in + (uint64_t)c0 * p.in_cs + in_off;
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) pf[j] = ldg(src + ((uint64_t)bitrev32(j, C - 6) << sh));
    }
    // tile set g covers k-space bitrev(tau0 + g) << C; the lowest group's stages are this
    // wave's registers, the others its set in LDS
    uint32_t tr[LW::N];
    load_low_twiddles<C>(tr, p.tw, p.LT, bitrev32(tau0 + wave, p.L - C) << C, lane);
    stage_twiddle_sets<C, LW::SLOG>(twl, G, [&](uint32_t set, uint32_t ls, uint32_t i) {
        const uint32_t kb = bitrev32(tau0 + set, p.L - C) << C;
        return ldg(p.tw + (1u << (p.LT - 1 - ls)) + (kb >> (ls + 1)) + i);
    });
    __syncthreads();  // twiddle sets staged
    for (uint32_t col = c0; col < c1; col++) {
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) cp[lidx(j << 6)] = pf[j];
        if (col > c0) {
            uint32_t *dst = p.out + (uint64_t)(col - 1) * p.out_cs + out_off;
#pragma unroll
            for (uint32_t j = 0; j < PER; j++) stg(dst + j * 64, res[j]);
        }
        if (col + 1 < c1) {
            const uint32_t *src = p.in + (uint64_t)(col + 1) * p.in_cs + in_off;
#pragma unroll
            for (uint32_t j = 0; j < PER; j++) pf[j] = ldg(src + ((uint64_t)bitrev32(j, C - 6) << sh));
        }
        __syncthreads();  // every tile loaded
        // the last (top) group's layout is the store's (element lane + 64 j): it stays in registers
        tile_lds<true, C, true, true>(res, tl, lane, twt, tr);
        __syncthreads();  // every read done: the tiles may take the next column
    }
    if (c0 < c1) {
        uint32_t *dst = p.out + (uint64_t)(c1 - 1) * p.out_cs + out_off;
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) stg(dst + j * 64, res[j]);
    }
}
// ---------------------------------------------------------------------------
// fused LDE: interpolation top pass + scale + forward top pass for every q
// ---------------------------------------------------------------------------
// Forward groups 0.. of a tile whose top-group values are in registers.
template <int C>
__device__ __forceinline__ void fwd_from_regs(uint32_t (&v)[TC<C>::R], uint32_t *tl, uint32_t lane,
                                              const uint32_t *twl) {
    using G0 = Grp<false, C, 0>;
    g_stages<false, C, G0::LO, G0::r, 0>(v, lane, twl);
    g_store<C, G0::LO, G0::r>(v, tl, lane);
    wave_lds_sync();
    if constexpr (TC<C>::NG > 1) tile_run<false, C, 1>(tl, lane, twl);
}

struct LdeFused {
    FftPass p;               // geometry: L, C, s_lo, G_log, beta; in/out/strides; in_mode/out_mode
    const uint32_t *tw_inv;  // table of log L (inverse)
    const uint32_t *tw_fwd;  // table of log L + beta (forward)
    uint32_t inv_n;
};

// s_lo > 0: G tiles per workgroup (one wave each).  Twiddle sets in LDS: the
// interpolation's (shared by the G tiles) and one per q for the evaluation.
// Per q the workgroup's G output tiles are stored together (k-space).  The next
// column's tiles are loaded into registers while the current column is
// transformed; the last coset's stores are issued after the next column has
// landed in LDS, so the wait for that load never covers fresh stores.
template <int C>
__device__ __forceinline__ void lde_mid_body(const LdeFused &f, const VB vb) {
    constexpr uint32_t PER = (1u << C) >> 6;  // elements per thread of one column's G tiles
    // Up to 10 stages the next column is prefetched into registers and the last coset's
    // stores are deferred; an 11-stage tile (32 values per lane, held twice: the
    // interpolated tile and the coset being evaluated) has no registers left for either
    // within 128 VGPRs, so it loads and stores each column in place.
    constexpr bool PIPE = true;
    constexpr uint32_t NPF = PIPE ? PER : 1;
    extern __shared__ uint32_t lds[];
    const FftPass &p = f.p;
    const uint32_t TS = p.TS;
    const TileGeom tg = tile_geom(p, xcd_block(vb.x, vb.gx));
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t G = 1u << p.G_log;  // == number of waves
    const uint32_t NQ = 1u << p.beta;
    uint32_t *twl = lds + G * TS;  // set 0: interpolation, set 1 + q: evaluation of q
    const uint32_t hi_bits = tile_fixed(p, tg, 0) & ~((1u << p.s_lo) - 1);
    static_assert(Grp<true, C, TC<C>::NG - 1>::LO == Grp<false, C, 0>::LO, "interpolation ends on the top group");
    const uint32_t no_tr[LowTw<C>::N] = {};  // every group's twiddles are in LDS here
    uint32_t *tl = lds + wave * TS;
    // the strided copies: thread (g, t0) moves elements t = t0 + 64 j of tile g, which sit
    // at (hi, t, lo0 + g) of a column (runs of G words along g)
    const uint32_t g = tid & (G - 1), t0 = tid >> p.G_log;
    const uint64_t off = ((uint64_t)tg.hi << (p.s_lo + C)) + tg.lo0 + g + ((uint64_t)t0 << p.s_lo);
    const uint32_t jsh = 6 + p.s_lo;
    uint32_t *cp = lds + g * TS + lidx(t0);
    const uint32_t c0 = vb.y * p.ncb;
    const uint32_t c1 = min(c0 + p.ncb, p.n_cols);
    uint32_t pf[NPF], res[NPF];
    if (PIPE && c0 < c1) {
        const uint32_t *src = p.in + (uint64_t)c0 * p.in_cs + off;
#pragma unroll
        for (uint32_t j = 0; j < NPF; j++) pf[j] = ldg(src + ((uint64_t)j << jsh));
    }
    // set 0: the interpolation's stages s_lo.. of the size-2^L table; set 1 + q: coset q's of the
    // size-2^(L + beta) table
    stage_twiddle_sets<C>(twl, 1 + NQ, [&](uint32_t set, uint32_t ls, uint32_t i) {
        const uint32_t s = p.s_lo + ls;
        if (set == 0) return ldg(f.tw_inv + (1u << (p.L - 1 - s)) + (hi_bits >> (s + 1)) + i);
        const uint32_t kb = ((set - 1) << p.L) | hi_bits;
        return ldg(f.tw_fwd + (1u << (p.L + p.beta - 1 - s)) + (kb >> (s + 1)) + i);
    });
    __syncthreads();  // twiddle sets staged
    for (uint32_t col = c0; col < c1; col++) {
        if constexpr (PIPE) {
#pragma unroll
            for (uint32_t j = 0; j < PER; j++) cp[lidx(j << 6)] = pf[j];
            if (col > c0) {  // the previous column's last coset, held back until this column's tiles landed
                uint32_t *dst = p.out + (uint64_t)(col - 1) * p.out_cs + ((uint64_t)(NQ - 1) << p.L) + off;
#pragma unroll
                for (uint32_t j = 0; j < PER; j++) stg(dst + ((uint64_t)j << jsh), res[j]);
            }
            if (col + 1 < c1) {
                const uint32_t *src = p.in + (uint64_t)(col + 1) * p.in_cs + off;
#pragma unroll
                for (uint32_t j = 0; j < PER; j++) pf[j] = ldg(src + ((uint64_t)j << jsh));
            }
        } else {
            const uint32_t *src = p.in + (uint64_t)col * p.in_cs + off;
            batched_copy<8>(
                PER, [&](uint32_t j) { return ldg(src + ((uint64_t)j << jsh)); },
                [&](uint32_t j, uint32_t v) { cp[lidx(j << 6)] = v; });
        }
        __syncthreads();
        // the interpolation's last group is the evaluation's first (top bits): it stays in registers
        uint32_t c[TC<C>::R];
        tile_lds<true, C, true, false>(c, tl, lane, twl, no_tr);
#pragma unroll
        for (int j = 0; j < TC<C>::R; j++) c[j] = m31_mul(c[j], f.inv_n);
        uint32_t *out_col = p.out + (uint64_t)col * p.out_cs + off;
        for (uint32_t q = 0; q < NQ; q++) {
            __syncthreads();  // the previous coset's reads (or the interpolation's) are done with the tiles
            uint32_t v[TC<C>::R];
#pragma unroll
            for (int j = 0; j < TC<C>::R; j++) v[j] = c[j];
            fwd_from_regs<C>(v, tl, lane, twl + ((1 + q) << C));
            __syncthreads();
            uint32_t *dst = out_col + ((uint64_t)q << p.L);
            if constexpr (PIPE) {
#pragma unroll
                for (uint32_t j = 0; j < PER; j++) res[j] = cp[lidx(j << 6)];
                if (q + 1 < NQ) {
#pragma unroll
                    for (uint32_t j = 0; j < PER; j++) stg(dst + ((uint64_t)j << jsh), res[j]);
                }
            } else {
                batched_copy<8>(
                    PER, [&](uint32_t j) { return cp[lidx(j << 6)]; },
                    [&](uint32_t j, uint32_t v) { stg(dst + ((uint64_t)j << jsh), v); });
            }
        }
        __syncthreads();  // every read done: the tiles may take the next column
    }
    if (PIPE && c0 < c1) {
        uint32_t *dst = p.out + (uint64_t)(c1 - 1) * p.out_cs + ((uint64_t)(NQ - 1) << p.L) + off;
#pragma unroll
        for (uint32_t j = 0; j < NPF; j++) stg(dst + ((uint64_t)j << jsh), res[j]);
    }
}

// s_lo == 0 and L == C: the whole LDE of a column in one workgroup; wave 0
// interpolates, every wave evaluates some q, all 2^beta outputs leave together
}  // namespace zs
