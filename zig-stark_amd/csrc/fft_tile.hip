// fft_tile.hip -- register-group circle-FFT passes for gfx950.
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
#include "tile_dev.h"

namespace zs {

// the passes as kernels of their own (one workgroup per grid position)
template <bool INV, int C>
__global__ __launch_bounds__(tile_block_max<C>()) void fft_tile_kernel(FftPass p) {
    fft_tile_body<INV, C>(p, VB::hw());
}
// The pipelined LDE passes run two 512-thread workgroups per CU (their LDS allows
// no third), i.e. 4 waves per SIMD: stating that lets the compiler use up to 128
// VGPRs for latency hiding (twiddle loads issued early, butterflies interleaved)
// instead of squeezing registers for an occupancy the LDS never allows.
// An 11-stage pass's LDS (2^11-word tiles and twiddle sets) leaves room for 8 waves
// per CU, 2 per SIMD, so it may use up to 256 VGPRs.
template <int C>
constexpr int lde_pass_waves() { return C >= 11 ? 2 : 4; }
#define ZS_LDE_PASS_BOUNDS \
    __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(lde_pass_waves<C>(), lde_pass_waves<C>())))
// The first and last passes of up to 10 stages may run 16 tiles per workgroup
// (1024 threads, 86 KB of LDS at 10 stages: one workgroup per CU, still 4 waves per
// SIMD), so the natural-order copies move 64-byte runs.
template <int C>
constexpr int lde_pipe_threads() { return C >= 11 ? 512 : 1024; }
#define ZS_LDE_PIPE_BOUNDS \
    __launch_bounds__(lde_pipe_threads<C>()) __attribute__((amdgpu_waves_per_eu(lde_pass_waves<C>(), lde_pass_waves<C>())))
template <int C>
__global__ ZS_LDE_PIPE_BOUNDS void lde_eval_last_kernel(FftPass p) {
    lde_eval_last_body<C>(p, VB::hw());
}
template <int C>
__global__ ZS_LDE_PIPE_BOUNDS void lde_interp_first_kernel(FftPass p) {
    lde_interp_first_body<C>(p, VB::hw());
}
template <int C>
__global__ ZS_LDE_PASS_BOUNDS void lde_mid_kernel(LdeFused f) {
    lde_mid_body<C>(f, VB::hw());
}


template <int C>
static size_t tile_lds_bytes(uint32_t ntiles, uint32_t nsets, uint32_t TS) {
    return (size_t(ntiles) * TS + (size_t(nsets) << C)) * 4;
}

// LDS words from one tile to the next: 2^C plus one pad word per 16 (see lidx),
// plus a pad that spreads the tiles
// of one copy instruction over the 32 LDS banks.  The copies interleave `il` tiles
// across consecutive lanes (tile index fastest, then the position in the tile),
// so a stride of 32 / il banks (mod 32) gives every lane of a half-wave its own bank.
uint32_t tile_stride(uint32_t C, uint32_t il) {
    const uint32_t base = (1u << C) + (1u << (C - 4));
    const uint32_t target = (il >= 2 && il <= 32) ? 32 / il : 1;
    return base + (target + 32 - base % 32) % 32;
}

// tiles one copy instruction of the pass interleaves across lanes (see coop_load, coop_store)
static uint32_t copy_interleave(const FftPass &p, uint32_t q_in_wg) {
    const uint32_t G = 1u << p.G_log;
    uint32_t il = 1;
    if ((p.in_mode == IN_NATURAL && p.s_lo == 0 && p.natural_tiles) ||
        ((p.in_mode == IN_KSPACE || p.in_mode == IN_SHARED) && p.s_lo > 0))
        il = std::max(il, G);
    if (p.out_mode == OUT_NATURAL && p.s_lo == 0 && p.natural_tiles) il = std::max(il, G * q_in_wg);
    if (p.out_mode == OUT_KSPACE && p.s_lo > 0) il = std::max(il, G);
    return il;
}

// Launch-shape defaults, read once from the environment (defaults from measurement):
//   ZS_TILE_ELEMS_LOG: elements of the tiles one workgroup transforms (default 2^13)
//   ZS_COLS_PER_WG:    columns a workgroup runs with one set of staged twiddles (default 4)
static int env_int(const char *name, int dflt, int lo, int hi) {
    const char *e = getenv(name);
    int x = e ? atoi(e) : dflt;
    return std::min(hi, std::max(lo, x));
}
TileConfig default_tile_config() {
    static const TileConfig c = [] {
        TileConfig t;
        t.elems_log = (uint32_t)env_int("ZS_TILE_ELEMS_LOG", 13, 11, 15);
        t.cols_per_wg = (uint32_t)env_int("ZS_COLS_PER_WG", 4, 1, 64);
        return t;
    }();
    return c;
}
// LDE passes without a tuned entry: per kind of pass, the shape the autotuner
// picks for 64 x 2^20 columns (interpolation 2^12 x 4, middle 2^13 x 8,
// evaluation 2^13 x 4); ZS_TILE_ELEMS_LOG / ZS_COLS_PER_WG, when set, apply to all.
LdeTiles default_lde_tiles() {
    static const LdeTiles t = [] {
        LdeTiles d;
        d.role[LDE_INTERP] = {12, 4};
        d.role[LDE_MID] = {13, 8};
        d.role[LDE_EVAL] = {13, 4};
        if (getenv("ZS_TILE_ELEMS_LOG") || getenv("ZS_COLS_PER_WG"))
            for (auto &r : d.role) r = default_tile_config();
        return d;
    }();
    return t;
}

// columns per workgroup, reduced so the grid still has >= 2048 workgroups
static uint32_t pick_ncb(const Ctx *ctx, uint32_t n_cols, uint32_t wgs_per_col) {
    uint32_t ncb = std::min(ctx->tile.cols_per_wg, std::max<uint32_t>(n_cols, 1));
    while (ncb > 1 && uint64_t(wgs_per_col) * ((n_cols + ncb - 1) / ncb) < 2048) ncb >>= 1;
    return ncb;
}

int32_t run_tile_pass(Ctx *ctx, FftPass p, bool inverse, uint32_t n_cols) {
    // the LDE's first interpolation and last evaluation passes have pipelined
    // one-wave-per-tile kernels (8..11 stages, at most 16 tiles per workgroup up to
    // 10 stages, 8 at 11)
    const bool pipe_c = p.C >= 8 && p.C <= 11 && p.s_lo == 0 && p.natural_tiles && !p.scale;
    const bool pipe_interp = pipe_c && inverse && p.in_mode == IN_NATURAL && p.out_mode == OUT_KSPACE &&
                             p.Q_log == 0 && p.beta == 0;
    const bool pipe_eval = pipe_c && !inverse && p.in_mode == IN_KSPACE && p.out_mode == OUT_NATURAL;
    const int32_t pipe_tiles_log = p.C >= 11 ? 3 : 4;
    if (pipe_eval) p.Q_log = std::min<uint32_t>(p.Q_log, (uint32_t)pipe_tiles_log);
    uint32_t avail = p.s_lo > 0 ? p.s_lo : (p.L - p.C);
    int32_t want = (int32_t)ctx->tile.elems_log - (int32_t)p.C - (int32_t)p.Q_log;
    if (pipe_interp || pipe_eval) want = std::min<int32_t>(want, pipe_tiles_log - (int32_t)p.Q_log);
    p.G_log = (uint32_t)std::max<int32_t>(0, std::min<int32_t>((int32_t)avail, want));
    uint32_t tiles = 1u << (p.L - p.C);
    uint32_t ntile = 1u << (p.G_log + p.Q_log);
    uint32_t nsets = p.s_lo == 0 ? ntile : (1u << p.Q_log);
    uint32_t gx = tiles >> p.G_log, gz = (1u << p.beta) >> p.Q_log;
    p.n_cols = n_cols;
    p.ncb = pick_ncb(ctx, n_cols, gx * gz);
    p.TS = tile_stride(p.C, copy_interleave(p, 1u << p.Q_log));
    dim3 grid(gx, (n_cols + p.ncb - 1) / p.ncb, gz);
    uint32_t maxw = p.C >= 11 ? 8 : 16;
    dim3 block(64 * std::min<uint32_t>(ntile, maxw));
    if ((pipe_interp || pipe_eval) && ntile <= (1u << pipe_tiles_log)) {
        // a set keeps only the stages above the lowest register group
        const uint32_t set_log = p.C - low_group_stages((int)p.C);
        const size_t lds = (size_t(ntile) * p.TS + (size_t(ntile) << set_log)) * 4;
        if (lds <= 160 * 1024) {
            const dim3 blk(64 * ntile);
#define ZS_PIPE_CASE(CC)                                                                                        \
    case CC:                                                                                                    \
        if (pipe_interp)                                                                                        \
            ZS_LAUNCH(ctx, "lde_interp_first_kernel<" #CC ">", (lde_interp_first_kernel<CC>), grid, blk, lds,   \
                      ctx->stream, p);                                                                          \
        else                                                                                                    \
            ZS_LAUNCH(ctx, "lde_eval_last_kernel<" #CC ">", (lde_eval_last_kernel<CC>), grid, blk, lds,         \
                      ctx->stream, p);                                                                          \
        break;
            switch (p.C) {
                ZS_PIPE_CASE(8)
                ZS_PIPE_CASE(9)
                ZS_PIPE_CASE(10)
                ZS_PIPE_CASE(11)
            }
#undef ZS_PIPE_CASE
            ZS_HIP(ctx, hipGetLastError());
            return ZS_OK;
        }
    }
#define ZS_TILE_CASE(CC)                                                                                         \
    case CC: {                                                                                                   \
        size_t lds = tile_lds_bytes<CC>(ntile, nsets, p.TS);                                                         \
        if (lds > 160 * 1024) return run_pass(ctx, p, inverse, n_cols);                                          \
        if (inverse)                                                                                             \
            ZS_LAUNCH(ctx, p.s_lo == 0 ? "fft_tile_kernel<inv," #CC ",natural-in>" : "fft_tile_kernel<inv," #CC ">", \
                      (fft_tile_kernel<true, CC>), grid, block, lds, ctx->stream, p);                            \
        else                                                                                                     \
            ZS_LAUNCH(ctx, p.out_mode == OUT_NATURAL ? "fft_tile_kernel<fwd," #CC ",natural-out>"                \
                                                     : "fft_tile_kernel<fwd," #CC ">",                           \
                      (fft_tile_kernel<false, CC>), grid, block, lds, ctx->stream, p);                           \
        break;                                                                                                   \
    }
    switch (p.C) {
        ZS_TILE_CASE(8)
        ZS_TILE_CASE(9)
        ZS_TILE_CASE(10)
        ZS_TILE_CASE(11)
        default: return run_pass(ctx, p, inverse, n_cols);
    }
#undef ZS_TILE_CASE
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

template <int C>
__global__ __launch_bounds__(256) void lde_single_kernel(LdeFused f) {
    extern __shared__ uint32_t lds[];
    const FftPass &p = f.p;
    const uint32_t TS = p.TS;
    const TileGeom tg{0, 0, 0};
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nwaves = blockDim.x >> 6;
    const uint32_t NQ = 1u << p.beta;
    uint32_t *src = lds + NQ * TS;       // the interpolation's tile, after the NQ output tiles
    uint32_t *twl = lds + (NQ + 1) * TS;  // set 0: interpolation, set 1 + q: evaluation of q
    stage_twiddles<C>(twl, f.tw_inv, p.L, 0, 0);
    for (uint32_t q = 0; q < NQ; q++) stage_twiddles<C>(twl + ((1 + q) << C), f.tw_fwd, p.L + p.beta, 0, q << p.L);
    using Top = Grp<false, C, 0>;
    const uint32_t c0 = blockIdx.y * p.ncb;
    const uint32_t c1 = min(c0 + p.ncb, p.n_cols);
    for (uint32_t col = c0; col < c1; col++) {
        coop_load<C, 4>(p, tg, src, col, 0, 0, false);
        __syncthreads();
        if (wave == 0) tile_run<true, C, 0>(src, lane, twl);
        __syncthreads();
        uint32_t c[TC<C>::R];
        g_load<C, Top::LO, Top::r>(c, src, lane);
#pragma unroll
        for (int j = 0; j < TC<C>::R; j++) c[j] = m31_mul(c[j], f.inv_n);
        for (uint32_t q = wave; q < NQ; q += nwaves) {
            uint32_t v[TC<C>::R];
#pragma unroll
            for (int j = 0; j < TC<C>::R; j++) v[j] = c[j];
            fwd_from_regs<C>(v, lds + q * TS, lane, twl + ((1 + q) << C));
        }
        __syncthreads();
        coop_store<C, 4>(p, tg, lds, col, 0, p.beta);
        __syncthreads();
    }
}

bool lde_fused_ok(uint32_t C, uint32_t beta, bool single) {
    if (!tile_pass_ok(C) || beta > 8) return false;
    if (!single) return true;
    // the single-workgroup form keeps 2^beta output tiles, the interpolation tile and
    // 1 + 2^beta twiddle sets in LDS
    size_t ts = tile_stride(C, 1u << beta);
    size_t nq = size_t(1) << beta;
    return ((nq + 1) * ts + ((nq + 1) << C)) * 4 <= 160 * 1024;
}

int32_t run_lde_fused(Ctx *ctx, const uint32_t *d_in, uint64_t in_cs, uint32_t *d_out, uint64_t out_cs,
                      const uint32_t *tw_inv, const uint32_t *tw_fwd, uint32_t L, uint32_t beta, uint32_t s_lo,
                      uint32_t C, uint32_t n_cols) {
    if (!tile_pass_ok(C) || s_lo + C != L) return ZS_ERR_INPUT;
    LdeFused f{};
    FftPass &p = f.p;
    p.in = d_in;
    p.in_cs = in_cs;
    p.out = d_out;
    p.out_cs = out_cs;
    p.L = L;
    p.LT = L;
    p.beta = beta;
    p.s_lo = s_lo;
    p.C = C;
    p.scale = 0;
    f.tw_inv = tw_inv;
    f.tw_fwd = tw_fwd;
    f.inv_n = m31_inv(uint32_t(1) << L);
    dim3 grid;
    dim3 block;
    size_t lds;
    bool single = s_lo == 0;
    if (single) {
        if (!lde_fused_ok(L, beta, true)) return ZS_ERR_INPUT;
        p.G_log = 0;
        p.Q_log = beta;
        p.in_mode = IN_NATURAL;
        p.out_mode = OUT_NATURAL;
        p.natural_tiles = 1;
        p.n_cols = n_cols;
        p.ncb = pick_ncb(ctx, n_cols, 1);
        grid = dim3(1, (n_cols + p.ncb - 1) / p.ncb, 1);
        block = dim3(64 * std::min<uint32_t>(1u << beta, 4));
        p.TS = tile_stride(C, 1u << beta);
        lds = ((size_t(1) << beta) + 1) * p.TS + (((size_t(1) << beta) + 1) << C);
    } else {
        // one wave per tile, at most 8 tiles per workgroup (the kernel's launch bound)
        int32_t want = std::min<int32_t>((int32_t)ctx->tile.elems_log - (int32_t)C, 3);
        p.G_log = (uint32_t)std::max<int32_t>(0, std::min<int32_t>((int32_t)s_lo, want));
        p.Q_log = 0;
        p.in_mode = IN_KSPACE;
        p.out_mode = OUT_KSPACE;
        p.natural_tiles = 0;
        uint32_t gx = (1u << (L - C)) >> p.G_log;
        p.n_cols = n_cols;
        p.ncb = pick_ncb(ctx, n_cols, gx);
        grid = dim3(gx, (n_cols + p.ncb - 1) / p.ncb, 1);
        block = dim3(64u << p.G_log);
        p.TS = tile_stride(C, 1u << p.G_log);
        lds = (size_t(1) << p.G_log) * p.TS + (((size_t(1) << beta) + 1) << C);
        if (lds * 4 > 160 * 1024) return ZS_ERR_INPUT;
    }
#define ZS_FUSED_CASE(CC)                                                                                          \
    case CC:                                                                                                       \
        if (single)                                                                                                \
            ZS_LAUNCH(ctx, "lde_single_kernel<" #CC ">", (lde_single_kernel<CC>), grid, block, lds * 4, ctx->stream, f); \
        else                                                                                                       \
            ZS_LAUNCH(ctx, "lde_mid_kernel<" #CC ">", (lde_mid_kernel<CC>), grid, block, lds * 4, ctx->stream, f);  \
        break;
    switch (C) {
        ZS_FUSED_CASE(8)
        ZS_FUSED_CASE(9)
        ZS_FUSED_CASE(10)
        ZS_FUSED_CASE(11)
        default: return ZS_ERR_INPUT;
    }
#undef ZS_FUSED_CASE
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

}  // namespace zs

