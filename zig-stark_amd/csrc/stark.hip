// stark.hip -- GenericStark's prover for the Fibonacci AIR on gfx950
// (src/m31/stark.zig:303-657, FibAir stark.zig:877-904).
//
// The trace commitment (qntt.hip), the quotient and DEEP codewords (the two
// kernels below) and the FRI commit and query phases (fri.hip, capi.cpp) run on
// the device; the transcript, the O(n) DEEP evaluations at z and w z, and the
// proof serialization run on the host.  Every value is the reference's:
//   Q(x_i) = [sum_k alpha_k C_k(x_i) (x_i - w^(n-1))
//             + sum_k beta_k (f_col_k(x_i) - v_k) Z_H(x_i) / (x_i - p_k)] / Z_H(x_i)
// is computed as sum_k alpha_k C_k (x_i - w^(n-1)) Z_H^-1 + sum_k beta_k (f - v_k) / (x_i - p_k),
// the same field element, and Z_H(x_i) = OFFSET^n (w_ev^n)^i - 1 takes only
// 2^log_blowup values (w_ev^n has that order), inverted once on the host.
#include <cstring>
#include <memory>
#include <vector>

#include "internal.h"

namespace zs {

namespace {

constexpr int FIB_M = 2, FIB_NB = 3, FIB_NDEEP = 5;

struct DomainArgs {
    qm31 offset;      // FRI_OFFSET
    qm31 wpow[28];    // w_ev^(2^k)
    uint32_t dlog;    // |D| = 2^dlog
    uint32_t shift;   // 2^log_blowup: index of w x on D is i + shift
};

struct QuotientArgs {
    DomainArgs dom;
    qm31 last;                    // w^(n-1)
    qm31 alpha[2], beta[FIB_NB], p[FIB_NB], v[FIB_NB];
    uint32_t col[FIB_NB];
};

struct DeepArgs {
    DomainArgs dom;
    qm31 z, de[FIB_NDEEP], gamma[FIB_NDEEP];
};

__device__ __forceinline__ qm31 ldq(const uint32_t *p) {
    uint4 v = *reinterpret_cast<const uint4 *>(p);
    return {{v.x, v.y}, {v.z, v.w}};
}
__device__ __forceinline__ void stq(uint32_t *p, qm31 x) {
    *reinterpret_cast<uint4 *>(p) = make_uint4(x.a.c0, x.a.c1, x.b.c0, x.b.c1);
}
// x_i = FRI_OFFSET * w_ev^i
__device__ __forceinline__ qm31 domain_point(const DomainArgs &d, uint32_t i) {
    qm31 x = d.offset;
    for (uint32_t k = 0; k < d.dlog; k++)
        if ((i >> k) & 1u) x = qm_mul(x, d.wpow[k]);
    return x;
}

// cw: the two trace codewords (column c at c * 4N words); zh_inv[r] = Z_H^-1 on x_i with i = r mod shift
__global__ __launch_bounds__(256) void fib_quotient_kernel(const uint32_t *cw, const uint32_t *zh_inv,
                                                            const QuotientArgs a, uint32_t *q_out) {
    const uint32_t N = 1u << a.dom.dlog;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t ni = (i + a.dom.shift) & (N - 1);
    const qm31 x = domain_point(a.dom, i);
    const qm31 cur[2] = {ldq(cw + 4 * (uint64_t)i), ldq(cw + 4 * ((uint64_t)N + i))};
    const qm31 nxt[2] = {ldq(cw + 4 * (uint64_t)ni), ldq(cw + 4 * ((uint64_t)N + ni))};
    // FibAir.evalTransition (stark.zig:886-891)
    const qm31 r0 = qm_sub(nxt[0], cur[1]), r1 = qm_sub(qm_sub(nxt[1], cur[0]), cur[1]);
    const qm31 xl = qm_sub(x, a.last);
    qm31 h = qm_add(qm_mul(a.alpha[0], qm_mul(r0, xl)), qm_mul(a.alpha[1], qm_mul(r1, xl)));
    qm31 q = qm_mul(h, ldq(zh_inv + 4 * (i & (a.dom.shift - 1))));
    for (int k = 0; k < FIB_NB; k++) {
        const qm31 t = qm_mul(qm_sub(cur[a.col[k]], a.v[k]), qm_inv(qm_sub(x, a.p[k])));
        q = qm_add(q, qm_mul(a.beta[k], t));
    }
    stq(q_out + 4 * (uint64_t)i, q);
}

// g(x_i) = (x_i - z)^-1 [sum_j gamma_j (f_j(x_i) - f_j(z)) + gamma_{2+j} (f_j(w x_i) - f_j(w z))
//                        + gamma_4 (Q(x_i) - Q(z))]   (stark.zig:568-580)
__global__ __launch_bounds__(256) void fib_deep_kernel(const uint32_t *cwq, const DeepArgs a, uint32_t *g_out) {
    const uint32_t N = 1u << a.dom.dlog;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t ni = (i + a.dom.shift) & (N - 1);
    const qm31 x = domain_point(a.dom, i);
    qm31 s = {{0, 0}, {0, 0}};
    for (int j = 0; j < FIB_M; j++) {
        s = qm_add(s, qm_mul(a.gamma[j], qm_sub(ldq(cwq + 4 * ((uint64_t)j * N + i)), a.de[j])));
        s = qm_add(s, qm_mul(a.gamma[FIB_M + j], qm_sub(ldq(cwq + 4 * ((uint64_t)j * N + ni)), a.de[FIB_M + j])));
    }
    s = qm_add(s, qm_mul(a.gamma[4], qm_sub(ldq(cwq + 4 * ((uint64_t)2 * N + i)), a.de[4])));
    stq(g_out + 4 * (uint64_t)i, qm_mul(s, qm_inv(qm_sub(x, a.z))));
}

// Hc(x) / Z_H(x) from column values at x and w x (host; computeQuotientAt, stark.zig:793-846)
qm31 fib_quotient_at(qm31 x, const qm31 cur[2], const qm31 nxt[2], qm31 zh, const QuotientArgs &a) {
    const qm31 r0 = qm_sub(nxt[0], cur[1]), r1 = qm_sub(qm_sub(nxt[1], cur[0]), cur[1]);
    const qm31 xl = qm_sub(x, a.last);
    qm31 h = qm_add(qm_mul(a.alpha[0], qm_mul(r0, xl)), qm_mul(a.alpha[1], qm_mul(r1, xl)));
    for (int k = 0; k < FIB_NB; k++)
        h = qm_add(h, qm_mul(a.beta[k], qm_mul(qm_mul(qm_sub(cur[a.col[k]], a.v[k]), zh), qm_inv(qm_sub(x, a.p[k])))));
    return qm_mul(h, qm_inv(zh));
}

qm31 sample(zs_channel *ch) {
    uint32_t w[4];
    zs_channel_sample_qm31(ch, w);
    return qm_from_words(w);
}
void absorb(zs_channel *ch, qm31 x) {
    uint32_t w[4];
    qm_to_words(x, w);
    uint8_t b[16];
    for (int k = 0; k < 4; k++)
        for (int s = 0; s < 4; s++) b[4 * k + s] = (uint8_t)(w[k] >> (8 * s));
    zs_channel_absorb(ch, b, 16);
}
qm31 horner(const std::vector<uint32_t> &co, size_t off, size_t n, qm31 x) {
    qm31 r = {{0, 0}, {0, 0}};
    for (size_t i = n; i > 0; i--) r = qm_add(qm_mul(r, x), qm_from_words(&co[4 * (off + i - 1)]));
    return r;
}

struct Wire {
    std::vector<uint8_t> b;
    void bytes(const void *d, size_t n) {
        const uint8_t *p = static_cast<const uint8_t *>(d);
        b.insert(b.end(), p, p + n);
    }
    void uint(uint64_t v, unsigned n) {
        for (unsigned k = 0; k < n; k++) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void qm(const uint32_t *w) {
        for (int k = 0; k < 4; k++) uint(w[k], 4);
    }
};

}  // namespace

}  // namespace zs

using namespace zs;

extern "C" {

int32_t zs_fib_stark_prove_dev(zs_hip_ctx *h, const uint32_t *d_trace, uint32_t trace_log, uint32_t log_blowup,
                               size_t num_queries, uint32_t remainder_log, const uint32_t a0[4], const uint32_t a1[4],
                               const uint32_t claimed[4], zs_channel *chan, zs_host_allocator host,
                               uint8_t **proof_out, size_t *proof_len) {
    if (!h || !d_trace || !a0 || !a1 || !claimed || !chan || !host.alloc || !proof_out || !proof_len ||
        trace_log < 1 || log_blowup < 1 || log_sum_over(trace_log, log_blowup, 24) || remainder_log < log_blowup ||
        remainder_log >= trace_log + log_blowup || num_queries > 4096)
        return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t dlog = trace_log + log_blowup;
    const size_t n = size_t(1) << trace_log, N = size_t(1) << dlog, shift = size_t(1) << log_blowup;
    const qm31 one = {{1, 0}, {0, 0}}, offset = {{1, 0}, {1, 0}};  // FRI_OFFSET = 1 + u (fri.zig:36)
    const qm31 w = qm31_root_of_unity(trace_log), w_ev = qm31_root_of_unity(dlog);
    // device buffers: [codeword 0 | codeword 1 | quotient], the three trees, g
    DevBuf cwq, trees, g, zhb;
    const size_t tw = tree_words(dlog);
    ZS_HIP(ctx, hipMalloc(&cwq.ptr, 3 * N * 16));
    ZS_HIP(ctx, hipMalloc(&trees.ptr, 3 * tw * 4));
    ZS_HIP(ctx, hipMalloc(&g.ptr, N * 16));
    ZS_HIP(ctx, hipMalloc(&zhb.ptr, shift * 16));
    uint32_t *d_cwq = cwq.as<uint32_t>(), *d_trees = trees.as<uint32_t>();
    // 1. trace commitment; the channel absorbs both roots (stark.zig:343-414)
    uint8_t troots[FIB_M * 32];
    ZS_TRY(zs_qm31_trace_commit_dev(h, d_trace, FIB_M, trace_log, log_blowup, d_cwq,
                                    reinterpret_cast<uint8_t *>(d_trees), troots, chan));
    // n * (interpolation coefficients), left in scratch 0 by the trace commitment (the 1/n is
    // applied in its evaluation gather); scaled below once they are on the host
    std::vector<uint32_t> co(FIB_M * n * 4);
    ZS_HIP(ctx, hipMemcpyAsync(co.data(), ctx->scratch[0].ptr, co.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    // 3. transition and boundary weights (stark.zig:470-483)
    QuotientArgs qa{};
    qa.dom.offset = offset;
    qm31 wp = w_ev;
    for (int k = 0; k < 28; k++) {
        qa.dom.wpow[k] = wp;
        wp = qm_mul(wp, wp);
    }
    qa.dom.dlog = dlog;
    qa.dom.shift = (uint32_t)shift;
    qa.last = qm_pow(w, n - 1);
    for (int k = 0; k < 2; k++) qa.alpha[k] = sample(chan);
    const uint32_t cols[FIB_NB] = {0, 1, 0};
    const uint64_t steps[FIB_NB] = {0, 0, n - 1};
    const uint32_t *vals[FIB_NB] = {a0, a1, claimed};
    for (int k = 0; k < FIB_NB; k++) {
        qa.col[k] = cols[k];
        qa.v[k] = qm_from_words(vals[k]);
        qa.p[k] = qm_pow(w, steps[k]);
    }
    for (int k = 0; k < FIB_NB; k++) qa.beta[k] = sample(chan);
    // Z_H(x_i)^-1 for i mod 2^log_blowup
    std::vector<uint32_t> zh_inv(shift * 4);
    const qm31 off_n = qm_pow(offset, n), wev_n = qm_pow(w_ev, n);
    qm31 t = off_n;
    for (size_t r = 0; r < shift; r++) {
        qm_to_words(qm_inv(qm_sub(t, one)), &zh_inv[4 * r]);
        t = qm_mul(t, wev_n);
    }
    ZS_HIP(ctx, hipMemcpyAsync(zhb.ptr, zh_inv.data(), zh_inv.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    // 4. quotient codeword and its tree (stark.zig:485-538)
    uint32_t *d_q = d_cwq + 2 * N * 4;
    ZS_LAUNCH(ctx, "fib_quotient_kernel", fib_quotient_kernel, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0,
              ctx->stream, d_cwq, zhb.as<uint32_t>(), qa, d_q);
    ZS_HIP(ctx, hipGetLastError());
    ZS_TRY(launch_qm31_leaves(ctx, d_q, N, d_trees + 2 * tw));
    ZS_TRY(launch_tree_levels(ctx, d_trees + 2 * tw, dlog, 0));
    uint8_t qroot[32];
    ZS_HIP(ctx, hipMemcpyAsync(qroot, d_trees + 3 * tw - 8, 32, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    zs_channel_absorb(chan, qroot, 32);
    // 5. z, DEEP evaluations f_j(z), f_j(w z), Q(z), DEEP weights (stark.zig:540-566)
    const uint32_t n_inv = m31_inv(m31_from_u32((uint32_t)n));
    for (uint32_t &v : co) v = m31_mul(v, n_inv);
    DeepArgs da{};
    da.dom = qa.dom;
    da.z = sample(chan);
    const qm31 wz = qm_mul(da.z, w);
    for (int j = 0; j < FIB_M; j++) {
        da.de[j] = horner(co, j * n, n, da.z);
        da.de[FIB_M + j] = horner(co, j * n, n, wz);
    }
    {
        const qm31 cur[2] = {da.de[0], da.de[1]}, nxt[2] = {da.de[2], da.de[3]};
        da.de[4] = fib_quotient_at(da.z, cur, nxt, qm_sub(qm_pow(da.z, n), one), qa);
    }
    for (int k = 0; k < FIB_NDEEP; k++) absorb(chan, da.de[k]);
    for (int k = 0; k < FIB_NDEEP; k++) da.gamma[k] = sample(chan);
    ZS_LAUNCH(ctx, "fib_deep_kernel", fib_deep_kernel, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, ctx->stream,
              d_cwq, da, g.as<uint32_t>());
    ZS_HIP(ctx, hipGetLastError());
    // FRI on g (friParams: log_size = trace_log for FibAir, FRI blowup = log_blowup) and its queries
    const uint32_t L = dlog - remainder_log;
    const size_t rc = size_t(1) << (remainder_log - log_blowup);
    std::vector<uint8_t> froots(size_t(L) * 32);
    std::vector<uint32_t> frem(4 * rc);
    zs_fri_result *res = nullptr;
    ZS_TRY(zs_fri_commit_dev(h, g.as<uint32_t>(), dlog, log_blowup, remainder_log, nullptr, chan, 0, froots.data(),
                             frem.data(), &res));
    std::unique_ptr<zs_fri_result, void (*)(zs_fri_result *)> guard(res, zs_fri_result_free);
    size_t per_q = 0;
    for (uint32_t i = 0; i < L; i++) per_q += 2 * size_t(dlog - i) * 32;
    std::vector<size_t> fidx(num_queries);
    std::vector<uint32_t> fvals(num_queries * L * 8 + 4);
    std::vector<uint8_t> fpaths(num_queries * per_q + 1);
    ZS_TRY(zs_fri_query(res, chan, num_queries, fidx.data(), fvals.data(), fpaths.data()));
    // per-query reveals: f_j(x0), f_j(w x0), Q(x0) and their paths (stark.zig:584-625)
    std::vector<uint64_t> vreq;
    std::vector<PathReq> preq;
    for (size_t q = 0; q < num_queries; q++) {
        const size_t p0 = fidx[q], pn = (p0 + shift) % N;
        const size_t at[FIB_NDEEP] = {p0, p0, pn, pn, p0};
        const uint32_t tr[FIB_NDEEP] = {0, 1, 0, 1, 2};
        for (int k = 0; k < FIB_NDEEP; k++) {
            vreq.push_back((tr[k] * N + at[k]) * 4);
            vreq.push_back((q * FIB_NDEEP + k) * 4);
            preq.push_back({tr[k] * tw, (q * FIB_NDEEP + k) * uint64_t(dlog) * 8, dlog, (uint32_t)at[k]});
        }
    }
    std::vector<uint32_t> rvals(num_queries * FIB_NDEEP * 4 + 4);
    std::vector<uint8_t> rpaths(num_queries * FIB_NDEEP * dlog * 32 + 1);
    if (num_queries) {
        DevBuf dv, dp;
        ZS_HIP(ctx, hipMalloc(&dv.ptr, rvals.size() * 4));
        ZS_HIP(ctx, hipMalloc(&dp.ptr, rpaths.size()));
        ZS_TRY(launch_gather_words(ctx, d_cwq, vreq, 4, dv.as<uint32_t>()));
        ZS_TRY(launch_gather_paths(ctx, d_trees, preq, dp.as<uint32_t>()));
        ZS_HIP(ctx, hipMemcpyAsync(rvals.data(), dv.ptr, rvals.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        ZS_HIP(ctx, hipMemcpyAsync(rpaths.data(), dp.ptr, rpaths.size(), hipMemcpyDeviceToHost, ctx->stream));
        ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    // GenericStark.Proof in the wire format (stark.zig:110-123, serialization.zig)
    Wire wr;
    wr.uint(trace_log, 1);  // StarkParams
    wr.uint(log_blowup, 1);
    wr.uint(num_queries, 8);
    wr.uint(remainder_log, 1);
    wr.uint(FIB_M, 8);  // trace_roots
    wr.bytes(troots, sizeof troots);
    wr.uint(0, 1);  // preprocessed_roots = null
    wr.uint(0, 1);  // accumulator_roots = null
    wr.bytes(qroot, 32);
    wr.uint(FIB_NDEEP, 8);  // deep_evals
    for (int k = 0; k < FIB_NDEEP; k++) {
        uint32_t wd[4];
        qm_to_words(da.de[k], wd);
        wr.qm(wd);
    }
    wr.uint(trace_log, 1);  // fri: Fri.Proof, FriParams {log_size, log_blowup, num_queries, remainder_log}
    wr.uint(log_blowup, 1);
    wr.uint(num_queries, 8);
    wr.uint(remainder_log, 1);
    wr.uint(L, 8);
    wr.bytes(froots.data(), froots.size());
    wr.uint(rc, 8);
    for (size_t k = 0; k < rc; k++) wr.qm(&frem[4 * k]);
    wr.uint(num_queries, 8);
    for (size_t q = 0; q < num_queries; q++) {
        wr.uint(fidx[q], 8);
        wr.uint(L, 8);
        const uint8_t *pp = fpaths.data() + q * per_q;
        for (uint32_t i = 0; i < L; i++) {
            const size_t plen = dlog - i;
            wr.qm(&fvals[8 * (q * L + i)]);
            wr.qm(&fvals[8 * (q * L + i) + 4]);
            wr.uint(plen, 8);
            wr.bytes(pp, plen * 32);
            pp += plen * 32;
            wr.uint(plen, 8);
            wr.bytes(pp, plen * 32);
            pp += plen * 32;
        }
    }
    wr.uint(num_queries, 8);  // queries: QueryReveal {query_index, values, paths}
    for (size_t q = 0; q < num_queries; q++) {
        wr.uint(fidx[q], 8);
        wr.uint(FIB_NDEEP, 8);
        for (int k = 0; k < FIB_NDEEP; k++) wr.qm(&rvals[4 * (q * FIB_NDEEP + k)]);
        wr.uint(FIB_NDEEP, 8);
        for (int k = 0; k < FIB_NDEEP; k++) {
            wr.uint(dlog, 8);
            wr.bytes(&rpaths[(q * FIB_NDEEP + k) * dlog * 32], size_t(dlog) * 32);
        }
    }
    uint8_t *out = static_cast<uint8_t *>(host.alloc(host.ctx, wr.b.size()));
    if (!out) return ZS_ERR_OOM;
    std::memcpy(out, wr.b.data(), wr.b.size());
    *proof_out = out;
    *proof_len = wr.b.size();
    return ZS_OK;
}

}  // extern "C"
