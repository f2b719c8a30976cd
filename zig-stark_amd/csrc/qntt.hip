// qntt.hip -- GenericStark's trace commitment on gfx950 (src/m31/stark.zig:343-414).
//
// The reference interpolates each QM31 trace column over H = <w> by O(n^2)
// Lagrange (univariate.zig:79-139) and evaluates the interpolant at every point
// of D = FRI_OFFSET * <w_ev> by Horner, O(n N).  H is a multiplicative group of
// order n = 2^trace_log and D a coset of the group of order N = n 2^log_blowup
// (w_ev^(2^log_blowup) = w), so both steps are QM31 DFTs:
//   c_k = n^-1 sum_i f_i w^(-ik)                 (the unique interpolant)
//   g_i = sum_k (c_k OFFSET^k) w_ev^(ik)         (= interpolant at OFFSET w_ev^i)
// The interpolant of n points is unique, so the coefficients, the codeword and
// every hashQM31 leaf equal the reference's value for value.
//
// Each transform is an iterative radix-2 Cooley-Tukey DFT with a bit-reversed
// gather: the first stages (blocks of up to 2^10 values) run in LDS in one
// launch, the remaining stages one launch each.  The second transform's gather
// scales c_k by n^-1 OFFSET^k and zero-pads k >= n.
#include "internal.h"

namespace zs {

namespace {

constexpr uint32_t QT_MAX_LOG = 10;  // values of one LDS block (16 KB of QM31)

__device__ __forceinline__ qm31 ldq(const uint32_t *p) {
    uint4 v = *reinterpret_cast<const uint4 *>(p);
    return {{v.x, v.y}, {v.z, v.w}};
}
__device__ __forceinline__ void stq(uint32_t *p, qm31 x) {
    *reinterpret_cast<uint4 *>(p) = make_uint4(x.a.c0, x.a.c1, x.b.c0, x.b.c1);
}

// Bit-reversed gather of one column's block [base, base + 2^TL) of a 2^L-point
// transform, then its first TL stages in LDS.  Source index k = bitrev_L(i);
// values with k >= 2^in_log are zero; `scale` (if set) multiplies value k.
// tw[j] = root^j for j < 2^(L-1).
__global__ __launch_bounds__(512) void qntt_first_kernel(const uint32_t *in, uint64_t in_cs, uint32_t in_log,
                                                          const uint32_t *scale, uint32_t *out, uint64_t out_cs,
                                                          uint32_t L, uint32_t TL, const uint32_t *tw) {
    __shared__ qm31 s[1u << QT_MAX_LOG];
    const uint32_t T = 1u << TL, base = blockIdx.x << TL, col = blockIdx.y;
    const uint32_t *src = in + (uint64_t)col * in_cs;
    const uint32_t in_len = 1u << in_log;
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        uint32_t k = bitrev32(base + t, L);
        qm31 v = {{0, 0}, {0, 0}};
        if (k < in_len) {
            v = ldq(src + 4 * (uint64_t)k);
            if (scale) v = qm_mul(v, ldq(scale + 4 * (uint64_t)k));
        }
        s[t] = v;
    }
    __syncthreads();
    for (uint32_t st = 1; st <= TL; st++) {
        const uint32_t half = 1u << (st - 1);
        for (uint32_t j = threadIdx.x; j < T / 2; j += blockDim.x) {
            const uint32_t pos = j & (half - 1), i0 = ((j >> (st - 1)) << st) + pos, i1 = i0 + half;
            const qm31 w = ldq(tw + 4 * (uint64_t)(pos << (L - st)));
            const qm31 u = s[i0], v = qm_mul(s[i1], w);
            s[i0] = qm_add(u, v);
            s[i1] = qm_sub(u, v);
        }
        __syncthreads();
    }
    uint32_t *dst = out + (uint64_t)col * out_cs;
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) stq(dst + 4 * (uint64_t)(base + t), s[t]);
}

// one later stage st (> TL) of every column, in place
__global__ __launch_bounds__(256) void qntt_stage_kernel(uint32_t *data, uint64_t cs, uint32_t L, uint32_t st,
                                                          const uint32_t *tw) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (uint64_t(1) << (L - 1))) return;
    uint32_t *d = data + (uint64_t)blockIdx.y * cs;
    const uint64_t half = uint64_t(1) << (st - 1);
    const uint64_t pos = j & (half - 1), i0 = ((j >> (st - 1)) << st) + pos, i1 = i0 + half;
    const qm31 w = ldq(tw + 4 * (pos << (L - st)));
    const qm31 u = ldq(d + 4 * i0), v = qm_mul(ldq(d + 4 * i1), w);
    stq(d + 4 * i0, qm_add(u, v));
    stq(d + 4 * i1, qm_sub(u, v));
}

// device table of 2^len_log QM31 values first * step^j
// built outside the cache lock (see get_twiddles), inserted unless another thread won
int32_t qm31_table(Ctx *ctx, uint32_t kind, uint32_t log_len, qm31 first, qm31 step, const uint32_t **out) {
    DeviceCaches &dc = *ctx->caches;
    const uint64_t key = (uint64_t(kind) << 32) | log_len;
    {
        std::lock_guard<std::mutex> lk(dc.mu);
        auto it = dc.qtables.find(key);
        if (it != dc.qtables.end()) {
            *out = it->second->as<uint32_t>();
            return ZS_OK;
        }
    }
    const size_t len = size_t(1) << log_len;
    std::vector<uint32_t> host(len * 4);
    qm31 x = first;
    for (size_t j = 0; j < len; j++) {
        qm_to_words(x, &host[4 * j]);
        x = qm_mul(x, step);
    }
    auto buf = std::make_unique<DevBuf>();
    ZS_HIP(ctx, hipMalloc(&buf->ptr, host.size() * 4));
    buf->bytes = host.size() * 4;
    ZS_HIP(ctx, hipMemcpyAsync(buf->ptr, host.data(), host.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> lk(dc.mu);
    auto it = dc.qtables.find(key);
    if (it != dc.qtables.end()) {
        *out = it->second->as<uint32_t>();
        return ZS_OK;
    }
    *out = buf->as<uint32_t>();
    dc.qtables[key] = std::move(buf);
    return ZS_OK;
}

enum { QT_INV_TW = 1, QT_FWD_TW = 2, QT_SCALE = 3 };

// DFT of 2^L points per column (root given by tw), input gathered from `in`
int32_t qntt(Ctx *ctx, const uint32_t *in, uint64_t in_cs, uint32_t in_log, const uint32_t *scale, uint32_t *out,
             uint64_t out_cs, uint32_t L, uint32_t n_cols, const uint32_t *tw) {
    const uint32_t TL = std::min(L, QT_MAX_LOG);
    ZS_LAUNCH(ctx, "qntt_first_kernel", qntt_first_kernel, dim3(1u << (L - TL), n_cols),
              dim3(std::min(512u, std::max(1u, 1u << (TL - 1)))), 0, ctx->stream, in, in_cs, in_log, scale, out,
              out_cs, L, TL, tw);
    ZS_HIP(ctx, hipGetLastError());
    for (uint32_t st = TL + 1; st <= L; st++) {
        ZS_LAUNCH(ctx, "qntt_stage_kernel", qntt_stage_kernel, dim3((uint32_t)(((1u << (L - 1)) + 255) / 256), n_cols),
                  dim3(256), 0, ctx->stream, out, out_cs, L, st, tw);
        ZS_HIP(ctx, hipGetLastError());
    }
    return ZS_OK;
}

}  // namespace

}  // namespace zs

using namespace zs;

extern "C" {

int32_t zs_qm31_trace_commit_dev(zs_hip_ctx *h, const uint32_t *d_cols, uint32_t n_cols, uint32_t trace_log,
                                 uint32_t log_blowup, uint32_t *d_codewords, uint8_t *d_nodes, uint8_t *roots_out,
                                 zs_channel *chan) {
    if (!h || !d_cols || !d_codewords || !d_nodes || !roots_out || n_cols == 0 || trace_log < 1 ||
        log_sum_over(trace_log, log_blowup, 24))
        return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t LT = trace_log + log_blowup;
    const size_t n = size_t(1) << trace_log, N = size_t(1) << LT;
    // tables: (w^-1)^j, w_ev^j (j < half the size) and n^-1 OFFSET^k (k < n)
    const qm31 w = qm31_root_of_unity(trace_log), w_ev = qm31_root_of_unity(LT);
    const qm31 one = {{1, 0}, {0, 0}}, offset = {{1, 0}, {1, 0}};  // FRI_OFFSET = 1 + u (fri.zig:36)
    const qm31 n_inv = {{m31_inv(m31_from_u32((uint32_t)n)), 0}, {0, 0}};
    const uint32_t *tw_inv, *tw_fwd, *scale;
    ZS_TRY(qm31_table(ctx, QT_INV_TW, trace_log - 1, one, qm_inv(w), &tw_inv));
    ZS_TRY(qm31_table(ctx, QT_FWD_TW, LT - 1, one, w_ev, &tw_fwd));
    ZS_TRY(qm31_table(ctx, QT_SCALE, trace_log, n_inv, offset, &scale));
    ZS_TRY(ctx->ensure_scratch(0, n * n_cols * 16));
    uint32_t *coef = ctx->scratch[0].as<uint32_t>();
    ZS_TRY(qntt(ctx, d_cols, 4 * n, trace_log, nullptr, coef, 4 * n, trace_log, n_cols, tw_inv));
    ZS_TRY(qntt(ctx, coef, 4 * n, trace_log, scale, d_codewords, 4 * N, LT, n_cols, tw_fwd));
    // one tree of hashQM31 leaves per column (hashCodeword + MerkleTree.init, stark.zig:396-414)
    uint32_t *nodes = reinterpret_cast<uint32_t *>(d_nodes);
    std::vector<uint64_t> offs(n_cols);
    std::vector<uint32_t> logs(n_cols, LT);
    for (uint32_t c = 0; c < n_cols; c++) {
        offs[c] = c * tree_words(LT);
        ZS_TRY(launch_qm31_leaves(ctx, d_codewords + 4 * N * c, N, nodes + offs[c]));
    }
    ZS_TRY(launch_trees_multi(ctx, nodes, offs.data(), logs.data(), n_cols));
    for (uint32_t c = 0; c < n_cols; c++)
        ZS_HIP(ctx, hipMemcpyAsync(roots_out + 32 * c, nodes + offs[c] + tree_words(LT) - 8, 32,
                                   hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (chan)
        for (uint32_t c = 0; c < n_cols; c++) zs_channel_absorb(chan, roots_out + 32 * c, 32);  // absorbDigest
    return ZS_OK;
}

int32_t zs_fib_trace(uint32_t a0, uint32_t a1, uint32_t log_n, uint32_t *out) {
    if (!out || log_n > 28 || a0 >= P || a1 >= P) return ZS_ERR_INPUT;
    const size_t n = size_t(1) << log_n;
    uint32_t a = a0, b = a1;
    for (size_t i = 0; i < n; i++) {
        uint32_t *p = out + 4 * i, *q = out + 4 * (n + i);
        p[0] = a;
        p[1] = p[2] = p[3] = 0;
        q[0] = b;
        q[1] = q[2] = q[3] = 0;
        const uint32_t nb = m31_add(a, b);
        a = b;
        b = nb;
    }
    return ZS_OK;
}

}  // extern "C"
