// fft.hip -- circle FFT / IFFT / LDE kernels for gfx950.
//
// The transform is the one of src/m31/ntt/circle.zig:195-346: the A-then-B
// coefficients are interleaved into k-space ([A0, B0, A1, B1, ...]), the
// forward transform runs butterfly stages s = L-1 .. 0 (stage s pairs k and
// k + 2^s with twiddle tw_s[k >> (s+1)]), and the result is read out in
// bit-reversed order; the inverse runs the stages upward with inverted
// twiddles and scales by 1/n.
//
// Twiddles live on the device in one table per half-coset ("heap" layout:
// stage s of a log-L transform at [2^(L-1-s), 2^(L-s))).  Entry values are
// those of slowPrecomputeTwiddles / lineTwiddleSlice / circleTwiddles
// (circle.zig:65-106): stage s >= 1 holds x(2^(s-1) * at(bitrev(h))) of the
// half-coset, stage 0 is the [y, -y, -x, x] expansion of stage 1.
//
// The fast path groups C consecutive stages into one pass: a workgroup loads
// G*Q tiles of 2^C elements into LDS (coalesced: tiles with adjacent addresses
// are loaded together), runs the C stages in LDS, and writes back.  Natural-
// order input (inverse) and output (forward) are handled in the first / last
// pass by choosing tiles whose natural addresses are adjacent, so the
// bit-reversal costs no extra pass.  For the blowup LDE the first log_blowup
// forward stages act on zeros and reduce to copies, so the LDE runs 2^beta
// size-2^log_n transforms per column (q = 0..2^beta-1) on twiddle sub-tables.
#include <vector>

#include "fft_common.h"
#include "pass_dev.h"

namespace zs {

// ---------------------------------------------------------------------------
// twiddle generation
// ---------------------------------------------------------------------------
struct CosetPowers {
    cpoint off[32];   // offset doubled j times
    cpoint step[32];  // step doubled k times
};

__device__ __forceinline__ uint32_t coset_point_x(const CosetPowers &cp, uint32_t j, uint32_t idx) {
    cpoint r = cp.off[j];
    uint32_t b = 0;
    while (idx) {
        if (idx & 1u) r = cp_add(r, cp.step[j + b]);
        idx >>= 1;
        b++;
    }
    return r.x;
}

__global__ void twiddle_gen_kernel(uint32_t *fwd, uint32_t *inv, uint32_t L, CosetPowers cp) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = 1u << L;
    if (e == 0) {
        fwd[0] = 0;
        inv[0] = 0;
        return;
    }
    if (e >= n) return;
    uint32_t lvl = 31 - __builtin_clz(e);
    uint32_t s = L - 1 - lvl;
    uint32_t h = e - (1u << lvl);
    uint32_t v;
    if (s >= 1) {
        v = coset_point_x(cp, s - 1, bitrev32(h, lvl));
    } else {
        uint32_t i = h >> 2, r = h & 3u;
        uint32_t x = coset_point_x(cp, 0, bitrev32(2 * i, L - 2));
        uint32_t y = coset_point_x(cp, 0, bitrev32(2 * i + 1, L - 2));
        v = r == 0 ? y : r == 1 ? m31_neg(y) : r == 2 ? m31_neg(x) : x;
    }
    fwd[e] = v;
    inv[e] = m31_inv(v);
}

hipError_t launch_twiddle_gen(hipStream_t stream, const zs_circle_coset &hc, uint32_t *fwd, uint32_t *inv) {
    const uint32_t L = hc.log_size + 1;
    CosetPowers cp;
    cpoint o{hc.offset_x, hc.offset_y}, st{hc.step_x, hc.step_y};
    for (int j = 0; j < 32; j++) {
        cp.off[j] = o;
        cp.step[j] = st;
        o = cp_double(o);
        st = cp_double(st);
    }
    const uint32_t n = 1u << L;
    hipLaunchKernelGGL(twiddle_gen_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, fwd, inv, L, cp);
    return hipGetLastError();
}

cpoint coset_at_host(const zs_circle_coset &c, uint64_t i) {
    return cp_add({c.offset_x, c.offset_y}, cp_mul({c.step_x, c.step_y}, i));
}

// Twiddle tables are shared by the contexts of a batch or server, so the cache lock
// is held only for lookups and inserts: a missing table is built on the caller's
// stream outside the lock (other workers' lookups proceed meanwhile), and if
// another thread inserted the same table first, that one is kept and this one
// freed.  Tables of arbitrary (non-canonic) cosets are only reachable through the
// synchronous host-buffer transforms, which leave no kernel in flight when they
// return; at most kMaxOtherCosets of them are kept, the oldest evicted first.
constexpr size_t kMaxOtherCosets = 8;

int32_t get_twiddles(Ctx *ctx, const zs_circle_coset &hc, const TwiddleTable **out) {
    auto key = std::make_tuple(hc.offset_x, hc.offset_y, hc.step_x, hc.step_y, hc.log_size);
    DeviceCaches &dc = *ctx->caches;
    {
        std::lock_guard<std::mutex> lk(dc.mu);
        auto it = dc.twiddles.find(key);
        if (it != dc.twiddles.end()) {
            *out = it->second.get();
            return ZS_OK;
        }
    }
    uint32_t L = hc.log_size + 1;
    if (L < 3 || L > 29) return ZS_ERR_INPUT;
    zs_circle_coset canon;
    const bool canonic = zs_circle_coset_canonic_half(L, &canon) == ZS_OK && canon.offset_x == hc.offset_x &&
                         canon.offset_y == hc.offset_y && canon.step_x == hc.step_x && canon.step_y == hc.step_y;
    auto t = std::make_unique<TwiddleTable>();
    t->log_n = L;
    size_t bytes = (size_t(1) << L) * 4;
    ZS_HIP(ctx, hipMalloc(&t->fwd.ptr, bytes));
    t->fwd.bytes = bytes;
    ZS_HIP(ctx, hipMalloc(&t->inv.ptr, bytes));
    t->inv.bytes = bytes;
    ZS_HIP(ctx, launch_twiddle_gen(ctx->stream, hc, t->fwd.as<uint32_t>(), t->inv.as<uint32_t>()));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));  // complete before other streams read it
    std::lock_guard<std::mutex> lk(dc.mu);
    auto it = dc.twiddles.find(key);
    if (it != dc.twiddles.end()) {  // another thread built it meanwhile; ours is freed here
        *out = it->second.get();
        return ZS_OK;
    }
    if (!canonic) {
        dc.other_cosets.push_back(key);
        while (dc.other_cosets.size() > kMaxOtherCosets) {
            dc.twiddles.erase(dc.other_cosets.front());
            dc.other_cosets.pop_front();
        }
    }
    *out = t.get();
    dc.twiddles[key] = std::move(t);
    return ZS_OK;
}

int32_t get_canonic_twiddles(Ctx *ctx, uint32_t log_n, const TwiddleTable **out) {
    zs_circle_coset hc;
    if (zs_circle_coset_canonic_half(log_n, &hc) != ZS_OK) return ZS_ERR_INPUT;
    return get_twiddles(ctx, hc, out);
}

// ---------------------------------------------------------------------------
// multi-stage LDS pass (the body is in pass_dev.h)
// ---------------------------------------------------------------------------
constexpr int PASS_THREADS = 256;

template <bool INV>
__global__ __launch_bounds__(PASS_THREADS) void fft_pass_kernel(FftPass p) {
    fft_pass_body<INV>(p, VB::hw(), threadIdx.x, PASS_THREADS);
}

int32_t run_pass(Ctx *ctx, FftPass p, bool inverse, uint32_t n_cols) {
    uint32_t avail = p.s_lo > 0 ? p.s_lo : (p.L - p.C);
    int32_t want = (int32_t)PASS_ELEMS_LOG - (int32_t)p.C - (int32_t)p.Q_log;
    p.G_log = (uint32_t)std::max<int32_t>(0, std::min<int32_t>((int32_t)avail, want));
    uint32_t tiles = 1u << (p.L - p.C);
    dim3 grid(tiles >> p.G_log, n_cols, (1u << p.beta) >> p.Q_log);
    size_t lds = (size_t(1) << (p.G_log + p.Q_log)) * ((size_t(1) << p.C) + 1) * 4;
    if (inverse)
        hipLaunchKernelGGL(fft_pass_kernel<true>, grid, dim3(PASS_THREADS), lds, ctx->stream, p);
    else
        hipLaunchKernelGGL(fft_pass_kernel<false>, grid, dim3(PASS_THREADS), lds, ctx->stream, p);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

// ---------------------------------------------------------------------------
// n = 2, 4: the reference's special cases (circle.zig:210-220, 290-319)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bfly(uint32_t &v0, uint32_t &v1, uint32_t t) {
    uint32_t m = m31_mul(v1, t);
    v1 = m31_sub(v0, m);
    v0 = m31_add(v0, m);
}
__device__ __forceinline__ void ibfly(uint32_t &v0, uint32_t &v1, uint32_t t) {
    uint32_t a = v0;
    v0 = m31_add(a, v1);
    v1 = m31_mul(m31_sub(a, v1), t);
}

// out_kspace: write the inverse's interleaved coefficients instead of A|B
template <bool INV>
__global__ void fft_small_kernel(const uint32_t *in, uint64_t in_cs, uint32_t *out, uint64_t out_cs, uint32_t n_cols,
                                 uint32_t lg, uint32_t px, uint32_t py, uint32_t out_kspace) {
    uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n_cols) return;
    const uint32_t *a = in + (uint64_t)col * in_cs;
    uint32_t *o = out + (uint64_t)col * out_cs;
    if (!INV) {
        if (lg == 1) {
            uint32_t e0 = a[0], e1 = a[1];
            bfly(e0, e1, py);
            o[0] = e0; o[1] = e1;
        } else {
            uint32_t e0 = a[0], e1 = a[2], e2 = a[1], e3 = a[3];
            bfly(e0, e2, px);
            bfly(e1, e3, px);
            bfly(e0, e1, py);
            bfly(e2, e3, m31_neg(py));
            o[0] = e0; o[1] = e2; o[2] = e1; o[3] = e3;
        }
    } else {
        if (lg == 1) {
            uint32_t yn_inv = m31_inv(m31_add(py, py));
            uint32_t y_inv = m31_add(yn_inv, yn_inv);
            uint32_t n_inv = m31_mul(yn_inv, py);
            uint32_t v0 = a[0], v1 = a[1];
            ibfly(v0, v1, y_inv);
            o[0] = m31_mul(v0, n_inv);
            o[1] = m31_mul(v1, n_inv);
        } else {
            uint32_t xy4 = m31_mul(m31_mul(px, py), 4);
            uint32_t xyn_inv = m31_inv(xy4);
            uint32_t x_inv = m31_mul(m31_mul(xyn_inv, py), 4);
            uint32_t y_inv = m31_mul(m31_mul(xyn_inv, px), 4);
            uint32_t n_inv = m31_mul(m31_mul(xyn_inv, px), py);
            uint32_t v0 = a[0], v1 = a[2], v2 = a[1], v3 = a[3];  // bit-reversed
            ibfly(v0, v1, y_inv);
            ibfly(v2, v3, m31_neg(y_inv));
            ibfly(v0, v2, x_inv);
            ibfly(v1, v3, x_inv);
            v0 = m31_mul(v0, n_inv); v1 = m31_mul(v1, n_inv);
            v2 = m31_mul(v2, n_inv); v3 = m31_mul(v3, n_inv);
            if (out_kspace) {
                o[0] = v0; o[1] = v1; o[2] = v2; o[3] = v3;
            } else {
                o[0] = v0; o[1] = v2; o[2] = v1; o[3] = v3;
            }
        }
    }
}

static int32_t run_small(Ctx *ctx, bool inverse, const zs_circle_coset &hc, const uint32_t *in, uint64_t in_cs,
                         uint32_t *out, uint64_t out_cs, uint32_t n_cols, bool out_kspace) {
    cpoint p0 = coset_at_host(hc, 0);
    uint32_t lg = hc.log_size + 1;
    dim3 grid((n_cols + 127) / 128);
    if (inverse)
        hipLaunchKernelGGL(fft_small_kernel<true>, grid, dim3(128), 0, ctx->stream, in, in_cs, out, out_cs, n_cols, lg,
                           p0.x, p0.y, (uint32_t)out_kspace);
    else
        hipLaunchKernelGGL(fft_small_kernel<false>, grid, dim3(128), 0, ctx->stream, in, in_cs, out, out_cs, n_cols,
                           lg, p0.x, p0.y, 0u);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

// ---------------------------------------------------------------------------
// public launchers
// ---------------------------------------------------------------------------
static int32_t transform(Ctx *ctx, bool inverse, const zs_circle_coset &hc, const uint32_t *d_in, uint64_t in_stride,
                         uint32_t *d_out, uint64_t out_stride, uint32_t n_cols) {
    uint32_t L = hc.log_size + 1;
    if (n_cols == 0) return ZS_OK;
    if (L < 3) return run_small(ctx, inverse, hc, d_in, in_stride, d_out, out_stride, n_cols, false);
    const TwiddleTable *tt;
    ZS_TRY(get_twiddles(ctx, hc, &tt));
    auto plan = plan_passes(L, inverse, PASS_CMAX);
    size_t n = size_t(1) << L;
    uint32_t *tmp = nullptr;
    if (plan.size() > 1) {
        ZS_TRY(ctx->ensure_scratch(0, n * n_cols * 4));
        tmp = ctx->scratch[0].as<uint32_t>();
    }
    for (size_t i = 0; i < plan.size(); i++) {
        FftPass p{};
        bool first = i == 0, last = i + 1 == plan.size();
        p.in = first ? d_in : tmp;
        p.in_cs = first ? in_stride : n;
        p.out = last ? d_out : tmp;
        p.out_cs = last ? out_stride : n;
        p.tw = inverse ? tt->inv.as<uint32_t>() : tt->fwd.as<uint32_t>();
        p.L = L; p.LT = L; p.beta = 0;
        p.s_lo = plan[i].s_lo; p.C = plan[i].C; p.Q_log = 0;
        if (inverse) {
            p.in_mode = first ? IN_NATURAL : IN_KSPACE;
            p.out_mode = last ? OUT_AB : OUT_KSPACE;
            p.natural_tiles = first ? 1 : 0;
            p.scale = last ? m31_inv((uint32_t)(n % P)) : 0;
        } else {
            p.in_mode = first ? IN_AB : IN_KSPACE;
            p.out_mode = last ? OUT_NATURAL : OUT_KSPACE;
            p.natural_tiles = last ? 1 : 0;
            p.scale = 0;
        }
        ZS_TRY(run_tile_pass(ctx, p, inverse, n_cols));
    }
    return ZS_OK;
}

int32_t launch_fft(Ctx *ctx, const zs_circle_coset &hc, const uint32_t *d_in, uint64_t in_stride, uint32_t *d_out,
                   uint64_t out_stride, uint32_t n_cols) {
    return transform(ctx, false, hc, d_in, in_stride, d_out, out_stride, n_cols);
}
int32_t launch_ifft(Ctx *ctx, const zs_circle_coset &hc, const uint32_t *d_in, uint64_t in_stride, uint32_t *d_out,
                    uint64_t out_stride, uint32_t n_cols) {
    return transform(ctx, true, hc, d_in, in_stride, d_out, out_stride, n_cols);
}

size_t lde_scratch_words(uint32_t n_cols, uint32_t log_n, uint32_t log_blowup, size_t *coef_words) {
    *coef_words = (size_t(n_cols) << log_n);
    uint32_t LT = log_n + log_blowup;
    bool multi = log_n > PASS_CMAX || LT < 3;  // multi-pass forward, or the tiny pad path
    return multi ? (size_t(n_cols) << LT) : 0;
}

// zero-pad A|B coefficients of size 2^log_n into A|B of size 2^LT (tiny-size path)
__global__ void pad_ab_kernel(const uint32_t *in, uint32_t *out, uint32_t n_cols, uint32_t log_n, uint32_t LT) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t N = uint64_t(1) << LT;
    if (i >= N * n_cols) return;
    uint64_t col = i >> LT, r = i & (N - 1);
    uint64_t h = uint64_t(1) << (log_n - 1), H = N >> 1;
    uint32_t v = 0;
    if (r < h) v = in[(col << log_n) + r];
    else if (r >= H && r < H + h) v = in[(col << log_n) + h + (r - H)];
    out[i] = v;
}

// LDE with the register-group kernels: interpolation passes below the top one,
// the fused top pass (interpolation top + 1/n + every q's evaluation top), then
// the remaining evaluation passes, the last writing natural order.
// the LDE's two twiddle tables: given by the caller, or from the context's cache
static int32_t lde_tables(Ctx *ctx, const LdeTables *given, const zs_circle_coset &hc_small,
                          const zs_circle_coset &hc_big, const uint32_t **inv_small, const uint32_t **fwd_big) {
    if (given) {
        *inv_small = given->inv_small;
        *fwd_big = given->fwd_big;
        return (*inv_small && *fwd_big) ? ZS_OK : ZS_ERR_INPUT;
    }
    const TwiddleTable *ti, *tf;
    ZS_TRY(get_twiddles(ctx, hc_small, &ti));
    ZS_TRY(get_twiddles(ctx, hc_big, &tf));
    *inv_small = ti->inv.as<uint32_t>();
    *fwd_big = tf->fwd.as<uint32_t>();
    return ZS_OK;
}

static int32_t launch_lde_tiled(Ctx *ctx, const uint32_t *d_in, uint64_t in_stride, uint32_t n_cols, uint32_t log_n,
                                uint32_t log_blowup, uint32_t *d_out, uint64_t out_stride, uint32_t *d_coef,
                                uint32_t *d_big, const zs_circle_coset &hc_small, const zs_circle_coset &hc_big,
                                const std::vector<PassRange> &pi, const std::vector<PassRange> &pf,
                                const LdeTables *given) {
    const uint32_t *tw_inv, *tw_fwd;
    ZS_TRY(lde_tables(ctx, given, hc_small, hc_big, &tw_inv, &tw_fwd));
    const uint32_t LT = log_n + log_blowup;
    const size_t n = size_t(1) << log_n, N = size_t(1) << LT;
    const size_t np = pf.size();
    if (np == 1) {
        ctx->tile = ctx->lde_tiles.role[LDE_MID];
        ctx->lde_role = LDE_MID;
        return run_lde_fused(ctx, d_in, in_stride, d_out, out_stride, tw_inv, tw_fwd,
                             log_n, log_blowup, 0, log_n, n_cols);
    }
    ctx->tile = ctx->lde_tiles.role[LDE_INTERP];
    ctx->lde_role = LDE_INTERP;
    for (size_t i = 0; i + 1 < pi.size(); i++) {
        FftPass p{};
        bool first = i == 0;
        p.in = first ? d_in : d_coef;
        p.in_cs = first ? in_stride : n;
        p.out = d_coef;
        p.out_cs = n;
        p.tw = tw_inv;
        p.L = log_n; p.LT = log_n; p.beta = 0;
        p.s_lo = pi[i].s_lo; p.C = pi[i].C; p.Q_log = 0;
        p.in_mode = first ? IN_NATURAL : IN_KSPACE;
        p.out_mode = OUT_KSPACE;
        p.natural_tiles = first ? 1 : 0;
        p.scale = 0;
        ZS_TRY(run_tile_pass(ctx, p, true, n_cols));
    }
    ctx->tile = ctx->lde_tiles.role[LDE_MID];
    ctx->lde_role = LDE_MID;
    ZS_TRY(run_lde_fused(ctx, d_coef, n, d_big, N, tw_inv, tw_fwd, log_n,
                         log_blowup, pf[0].s_lo, pf[0].C, n_cols));
    ctx->tile = ctx->lde_tiles.role[LDE_EVAL];
    ctx->lde_role = LDE_EVAL;
    for (size_t i = 1; i < np; i++) {
        FftPass p{};
        bool last = i + 1 == np;
        p.in = d_big;
        p.in_cs = N;
        p.out = last ? d_out : d_big;
        p.out_cs = last ? out_stride : N;
        p.tw = tw_fwd;
        p.L = log_n; p.LT = LT; p.beta = log_blowup;
        p.s_lo = pf[i].s_lo; p.C = pf[i].C;
        p.Q_log = last ? std::min<uint32_t>(log_blowup, ctx->tile.elems_log > p.C ? ctx->tile.elems_log - p.C : 0) : 0;
        p.in_mode = IN_KSPACE;
        p.out_mode = last ? OUT_NATURAL : OUT_KSPACE;
        p.natural_tiles = last ? 1 : 0;
        p.scale = 0;
        ZS_TRY(run_tile_pass(ctx, p, false, n_cols));
    }
    return ZS_OK;
}

static int32_t launch_lde_impl(Ctx *ctx, const uint32_t *d_in, uint64_t in_stride, uint32_t n_cols, uint32_t log_n,
                               uint32_t log_blowup, uint32_t *d_out, uint64_t out_stride, uint32_t *d_coef,
                               uint32_t *d_big, const LdeTables *given) {
    uint32_t LT = log_n + log_blowup;
    if (log_n < 1 || LT > 28) return ZS_ERR_INPUT;
    if (n_cols == 0) return ZS_OK;
    size_t n = size_t(1) << log_n;
    zs_circle_coset hc_small, hc_big;
    zs_circle_coset_canonic_half(log_n, &hc_small);
    zs_circle_coset_canonic_half(LT, &hc_big);
    if (LT < 3) {
        // log_n = 1 (and log_blowup <= 1): A|B interpolant, zero-pad, forward transform
        ZS_TRY(run_small(ctx, true, hc_small, d_in, in_stride, d_coef, n, n_cols, false));
        uint64_t tot = uint64_t(n_cols) << LT;
        hipLaunchKernelGGL(pad_ab_kernel, dim3((tot + 255) / 256), dim3(256), 0, ctx->stream, d_coef, d_big, n_cols,
                           log_n, LT);
        ZS_HIP(ctx, hipGetLastError());
        return run_small(ctx, false, hc_big, d_big, size_t(1) << LT, d_out, out_stride, n_cols, false);
    }
    {
        auto pi = plan_passes(log_n, true, PASS_CMAX);
        auto pf = plan_passes(log_n, false, PASS_CMAX);
        if (log_n >= 12 && log_n <= 15) {
            // two even passes would be 6..8 stages each, too few for the register-group
            // kernels: the fused middle pass takes 8..10 stages and a generic LDS pass the
            // 4..5 below it, on each side
            const uint32_t cm = std::min<uint32_t>(10, log_n - 4);
            pi = {{0, log_n - cm}, {log_n - cm, cm}};
            pf = {{log_n - cm, cm}, {0, log_n - cm}};
        }
        // the fused middle pass needs a register-group shape; the other passes run on the
        // register-group kernels where their stage count allows, else on the generic pass
        bool tiled = log_n >= 8 && lde_fused_ok(pf[0].C, log_blowup, pf.size() == 1);
        if (tiled) return launch_lde_tiled(ctx, d_in, in_stride, n_cols, log_n, log_blowup, d_out, out_stride,
                                           d_coef, d_big, hc_small, hc_big, pi, pf, given);
    }
    // 1) interpolate: natural evaluations -> interleaved coefficients scaled by 1/n (k-space)
    ctx->tile = ctx->lde_tiles.role[LDE_INTERP];
    ctx->lde_role = LDE_INTERP;
    if (log_n < 3) {
        ZS_TRY(run_small(ctx, true, hc_small, d_in, in_stride, d_coef, n, n_cols, true));
    } else {
        const uint32_t *tw_inv;
        if (given) {
            if (!(tw_inv = given->inv_small)) return ZS_ERR_INPUT;
        } else {
            const TwiddleTable *ti;
            ZS_TRY(get_twiddles(ctx, hc_small, &ti));
            tw_inv = ti->inv.as<uint32_t>();
        }
        auto plan = plan_passes(log_n, true, PASS_CMAX);
        for (size_t i = 0; i < plan.size(); i++) {
            FftPass p{};
            bool first = i == 0, last = i + 1 == plan.size();
            p.in = first ? d_in : d_coef;
            p.in_cs = first ? in_stride : n;
            p.out = d_coef;
            p.out_cs = n;
            p.tw = tw_inv;
            p.L = log_n; p.LT = log_n; p.beta = 0;
            p.s_lo = plan[i].s_lo; p.C = plan[i].C; p.Q_log = 0;
            p.in_mode = first ? IN_NATURAL : IN_KSPACE;
            p.out_mode = OUT_KSPACE;
            p.natural_tiles = first ? 1 : 0;
            p.scale = last ? m31_inv((uint32_t)n) : 0;
            ZS_TRY(run_tile_pass(ctx, p, true, n_cols));
        }
    }
    // 2) evaluate on the 2^beta quarter transforms of the big canonic domain
    ctx->tile = ctx->lde_tiles.role[LDE_EVAL];
    ctx->lde_role = LDE_EVAL;
    const uint32_t *tw_fwd;
    if (given) {
        if (!(tw_fwd = given->fwd_big)) return ZS_ERR_INPUT;
    } else {
        const TwiddleTable *tf;
        ZS_TRY(get_twiddles(ctx, hc_big, &tf));
        tw_fwd = tf->fwd.as<uint32_t>();
    }
    auto plan = plan_passes(log_n, false, PASS_CMAX);
    size_t N = size_t(1) << LT;
    for (size_t i = 0; i < plan.size(); i++) {
        FftPass p{};
        bool first = i == 0, last = i + 1 == plan.size();
        p.in = first ? d_coef : d_big;
        p.in_cs = first ? n : N;
        p.out = last ? d_out : d_big;
        p.out_cs = last ? out_stride : N;
        p.tw = tw_fwd;
        p.L = log_n; p.LT = LT; p.beta = log_blowup;
        p.s_lo = plan[i].s_lo; p.C = plan[i].C;
        // the last pass keeps 2^Q_log of the 2^beta replicas in one workgroup so its natural-order
        // stores are contiguous; capped so the LDS tile stays at 2^13 elements
        p.Q_log = last ? std::min<uint32_t>(log_blowup, ctx->tile.elems_log > p.C ? ctx->tile.elems_log - p.C : 0) : 0;
        p.in_mode = first ? IN_SHARED : IN_KSPACE;
        p.out_mode = last ? OUT_NATURAL : OUT_KSPACE;
        p.natural_tiles = last ? 1 : 0;
        p.scale = 0;
        ZS_TRY(run_tile_pass(ctx, p, false, n_cols));
    }
    return ZS_OK;
}

// ---------------------------------------------------------------------------
// reference-shaped plain path: one launch per butterfly layer, per column
// (src/cuda/circlefft_gpu.zig:46-175, src/cuda/kernels/circlefft.cu)
// ---------------------------------------------------------------------------
__global__ void plain_interleave_kernel(const uint32_t *ab, uint32_t *out, uint32_t half) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half) return;
    out[2 * i] = ab[i];
    out[2 * i + 1] = ab[half + i];
}
__global__ void plain_layer_kernel(uint32_t *v, const uint32_t *tw, uint32_t n, uint32_t s, uint32_t inverse) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n / 2) return;
    uint32_t b = x >> s, l = x & ((1u << s) - 1);
    uint32_t i0 = (b << (s + 1)) + l, i1 = i0 + (1u << s);
    uint32_t t = tw[b];
    uint32_t v0 = v[i0], v1 = v[i1];
    if (inverse) ibfly(v0, v1, t);
    else bfly(v0, v1, t);
    v[i0] = v0;
    v[i1] = v1;
}
__global__ void plain_bitrev_kernel(const uint32_t *in, uint32_t *out, uint32_t L) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << L)) return;
    out[i] = in[bitrev32(i, L)];
}
__global__ void plain_deinterleave_kernel(const uint32_t *c, uint32_t *ab, uint32_t half, uint32_t inv_n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half) return;
    ab[i] = m31_mul(c[2 * i], inv_n);
    ab[half + i] = m31_mul(c[2 * i + 1], inv_n);
}

// the LDE's passes take their launch shapes from ctx->lde_tiles; the context's
// transform configuration is left as it was
int32_t launch_lde(Ctx *ctx, const uint32_t *d_in, uint64_t in_stride, uint32_t n_cols, uint32_t log_n,
                   uint32_t log_blowup, uint32_t *d_out, uint64_t out_stride, uint32_t *d_coef, uint32_t *d_big,
                   const LdeTables *tables) {
    const TileConfig saved = ctx->tile;
    const int32_t rc = launch_lde_impl(ctx, d_in, in_stride, n_cols, log_n, log_blowup, d_out, out_stride, d_coef,
                                       d_big, tables);
    ctx->tile = saved;
    ctx->lde_role = -1;
    return rc;
}

int32_t launch_lde_plain(Ctx *ctx, const uint32_t *d_in, uint64_t in_stride, uint32_t n_cols, uint32_t log_n,
                         uint32_t log_blowup, uint32_t *d_out, uint64_t out_stride, uint32_t *d_tmp) {
    uint32_t LT = log_n + log_blowup;
    if (log_n < 3) return ZS_ERR_INPUT;
    size_t n = size_t(1) << log_n, N = size_t(1) << LT;
    const TwiddleTable *ti, *tf;
    ZS_TRY(get_canonic_twiddles(ctx, log_n, &ti));
    ZS_TRY(get_canonic_twiddles(ctx, LT, &tf));
    uint32_t *a = d_tmp, *b = d_tmp + N;  // two work vectors of N words
    uint32_t inv_n = m31_inv((uint32_t)n);
    auto grid = [](size_t cnt) { return dim3((uint32_t)((cnt + 255) / 256)); };
    for (uint32_t c = 0; c < n_cols; c++) {
        const uint32_t *src = d_in + (uint64_t)c * in_stride;
        // circleIFFT: bit-reverse, circle layer, line layers 1..L-1, scale + deinterleave
        hipLaunchKernelGGL(plain_bitrev_kernel, grid(n), dim3(256), 0, ctx->stream, src, a, log_n);
        for (uint32_t s = 0; s < log_n; s++)
            hipLaunchKernelGGL(plain_layer_kernel, grid(n / 2), dim3(256), 0, ctx->stream, a,
                               ti->inv.as<uint32_t>() + (size_t(1) << (log_n - 1 - s)), (uint32_t)n, s, 1u);
        hipLaunchKernelGGL(plain_deinterleave_kernel, grid(n / 2), dim3(256), 0, ctx->stream, a, b, (uint32_t)(n / 2),
                           inv_n);
        // zero-pad into A|B of size N
        hipLaunchKernelGGL(pad_ab_kernel, grid(N), dim3(256), 0, ctx->stream, b, a, 1u, log_n, LT);
        // circleFFT: interleave, line layers L-1..1, circle layer, bit-reverse
        hipLaunchKernelGGL(plain_interleave_kernel, grid(N / 2), dim3(256), 0, ctx->stream, a, b, (uint32_t)(N / 2));
        for (int s = (int)LT - 1; s >= 0; s--)
            hipLaunchKernelGGL(plain_layer_kernel, grid(N / 2), dim3(256), 0, ctx->stream, b,
                               tf->fwd.as<uint32_t>() + (size_t(1) << (LT - 1 - s)), (uint32_t)N, (uint32_t)s, 0u);
        hipLaunchKernelGGL(plain_bitrev_kernel, grid(N), dim3(256), 0, ctx->stream, b,
                           d_out + (uint64_t)c * out_stride, LT);
    }
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

// ---------------------------------------------------------------------------
// seeded synthetic input
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__global__ void fill_m31_kernel(uint64_t seed, uint32_t *out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = i + 1;
    uint32_t v;
    for (;;) {
        v = (uint32_t)(splitmix_mix(seed + k * 0x9E3779B97F4A7C15ull) >> 33);
        if (v != P) break;
        k += uint64_t(1) << 40;
    }
    out[i] = v;
}
int32_t launch_fill_m31(Ctx *ctx, uint64_t seed, uint32_t *d_out, uint64_t n) {
    if (n == 0) return ZS_OK;
    hipLaunchKernelGGL(fill_m31_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, seed, d_out, n);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

}  // namespace zs
