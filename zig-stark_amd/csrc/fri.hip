// fri.hip -- FRI commit-phase folding over QM31 for gfx950 (src/m31/fri.zig).
//
// Layer i lives on D_i = FRI_OFFSET^(2^i) * <w_i>, w_i = w_0^(2^i), and is
// folded with challenge alpha_i (fri.zig:211-218):
//   next[j] = (y0 + y1)/2 + alpha * (y0 - y1) * D_i[j]^{-1} / 2,
//   y0 = cur[j], y1 = cur[j + n/2].
// The reference inverts D_i[j] per element.  Here D_i[j]^{-1} =
// FRI_OFFSET^(-2^i) * w_0^{-(j * 2^i)} is read from two small tables of
// powers of w_0^{-1} (exponent split 10 / rest bits), so a fold costs three
// QM31 products and two halvings (x/2 is a shift and a conditional add of p),
// with no inversion.  Field results are canonical, so the output equals the
// reference's value for value.  The kernel also writes the hashQM31 leaf of
// each new value, fusing the next layer's leaf pass into the fold.
#include "fri_dev.h"
#include "merkle_dev.h"

namespace zs {

static const uint32_t FRI_OFFSET_W[4] = {1, 0, 1, 0};  // fri.zig:36, (1 + 0i) + (1 + 0i)u
static const uint8_t TWO_ADIC_ROOT_BYTES[16] = {0x60, 0x1f, 0x3d, 0x34, 0x32, 0x39, 0xa0, 0x6a,
                                                0xa0, 0xd6, 0x51, 0x61, 0x91, 0xe7, 0x86, 0x3e};

qm31 qm31_root_of_unity(uint32_t log_size) {  // qm31.zig:80-89
    uint32_t w[4];
    for (int i = 0; i < 4; i++) {
        const uint8_t *b = TWO_ADIC_ROOT_BYTES + 4 * i;
        w[i] = m31_from_u32((uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24));
    }
    return qm_pow(qm_from_words(w), uint64_t(1) << (32 - log_size));
}

qm31 fri_offset_pow(uint32_t layer) { return qm_pow(qm_from_words(FRI_OFFSET_W), uint64_t(1) << layer); }

// built outside the cache lock (see get_twiddles), inserted unless another thread won
int32_t get_fri_tables(Ctx *ctx, uint32_t dlog, const uint32_t **out) {
    DeviceCaches &dc = *ctx->caches;
    {
        std::lock_guard<std::mutex> lk(dc.mu);
        auto it = dc.fri_tables.find(dlog);
        if (it != dc.fri_tables.end()) {
            *out = it->second->as<uint32_t>();
            return ZS_OK;
        }
    }
    size_t n_hi = dlog > 11 ? (size_t(1) << (dlog - 11)) : 1;
    std::vector<uint32_t> host((1024 + n_hi) * 4);
    qm31 winv = qm_inv(qm31_root_of_unity(dlog));
    qm31 acc = {{1, 0}, {0, 0}};
    for (size_t a = 0; a < 1024; a++) {
        qm_to_words(acc, &host[4 * a]);
        acc = qm_mul(acc, winv);
    }
    qm31 step = acc;  // w^{-1024}
    acc = {{1, 0}, {0, 0}};
    for (size_t b = 0; b < n_hi; b++) {
        qm_to_words(acc, &host[4 * (1024 + b)]);
        acc = qm_mul(acc, step);
    }
    auto buf = std::make_unique<DevBuf>();
    ZS_HIP(ctx, hipMalloc(&buf->ptr, host.size() * 4));
    buf->bytes = host.size() * 4;
    ZS_HIP(ctx, hipMemcpyAsync(buf->ptr, host.data(), host.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> lk(dc.mu);
    auto it = dc.fri_tables.find(dlog);
    if (it != dc.fri_tables.end()) {
        *out = it->second->as<uint32_t>();
        return ZS_OK;
    }
    *out = buf->as<uint32_t>();
    dc.fri_tables[dlog] = std::move(buf);
    return ZS_OK;
}

__device__ __forceinline__ qm31 ld_qm(const uint32_t *p) {
    uint4 v = *reinterpret_cast<const uint4 *>(p);
    return {{v.x, v.y}, {v.z, v.w}};
}
__device__ __forceinline__ void st_qm(uint32_t *p, qm31 x) {
    *reinterpret_cast<uint4 *>(p) = make_uint4(x.a.c0, x.a.c1, x.b.c0, x.b.c1);
}

__global__ __launch_bounds__(256) void fri_fold_kernel(const uint32_t *cur, uint64_t half, uint32_t layer,
                                                       const uint32_t *tables, qm31 inv_off, const uint32_t *alpha_p,
                                                       uint32_t *next, uint32_t *next_leaves) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= half) return;
    // the factor alpha * inv_off / 2 is the same for every element (uniform, computed once per wave)
    const qm31 a_off = qm_half(qm_mul(ldg_qm(alpha_p), inv_off));
    const qm31 v = fold_value(cur, half, j, layer, tables, a_off);
    stg_qm(next + 4 * j, v);
    if (next_leaves) store_qm31_leaf(next_leaves + j * 8, v);
}

// Two folds in one launch, for challenges known up front: thread k of layer i (n
// values) reads cur[k], cur[k + n/2], cur[k + n/4], cur[k + 3n/4], folds them into
// layer i+1's values at k and k + n/4 (fri.zig:211-219, the same fold_value), and
// those into layer i+2's value at k, writing both layers and their leaves.  Layer
// i+1 is never read back, and one launch replaces two.  With leaves0 the thread also
// writes the hashQM31 leaves of the four layer-i values it read (layer 0's leaves,
// with no launch of their own).
__global__ __launch_bounds__(256) void fri_fold2_kernel(const uint32_t *cur, uint64_t quarter, uint32_t layer,
                                                        const uint32_t *tables, qm31 inv_off0, qm31 inv_off1,
                                                        const uint32_t *alpha_p, uint32_t *leaves0, uint32_t *next1,
                                                        uint32_t *leaves1, uint32_t *next2, uint32_t *leaves2) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= quarter) return;
    const qm31 a0 = qm_half(qm_mul(ldg_qm(alpha_p), inv_off0)), a1 = qm_half(qm_mul(ldg_qm(alpha_p + 4), inv_off1));
    const uint64_t half = 2 * quarter;
    const qm31 y0 = ldg_qm(cur + 4 * k), y1 = ldg_qm(cur + 4 * (k + half));
    const qm31 y2 = ldg_qm(cur + 4 * (k + quarter)), y3 = ldg_qm(cur + 4 * (k + quarter + half));
    if (leaves0) {
        store_qm31_leaf(leaves0 + 8 * k, y0);
        store_qm31_leaf(leaves0 + 8 * (k + half), y1);
        store_qm31_leaf(leaves0 + 8 * (k + quarter), y2);
        store_qm31_leaf(leaves0 + 8 * (k + quarter + half), y3);
    }
    const qm31 lo = fold_values(y0, y1, k, layer, tables, a0);
    const qm31 hi = fold_values(y2, y3, k + quarter, layer, tables, a0);
    stg_qm(next1 + 4 * k, lo);
    stg_qm(next1 + 4 * (k + quarter), hi);
    store_qm31_leaf(leaves1 + 8 * k, lo);
    store_qm31_leaf(leaves1 + 8 * (k + quarter), hi);
    // layer i+1's fold at k from the two values in registers
    const uint64_t e = k << (layer + 1);
    const qm31 t = qm_mul(ldg_qm(tables + 4 * (e & 1023)), ldg_qm(tables + 4 * (1024 + (e >> 10))));
    const qm31 v = qm_add(qm_half(qm_add(lo, hi)), qm_mul(qm_mul(a1, t), qm_sub(lo, hi)));
    stg_qm(next2 + 4 * k, v);
    if (leaves2) store_qm31_leaf(leaves2 + 8 * k, v);
}

int32_t launch_fri_fold2(Ctx *ctx, const uint32_t *d_cur, uint64_t n, uint32_t dlog, uint32_t layer,
                         const uint32_t *d_alpha, uint32_t *d_leaves0, uint32_t *d_next1, uint32_t *d_leaves1,
                         uint32_t *d_next2, uint32_t *d_leaves2) {
    if (n < 4 || !d_leaves1) return ZS_ERR_INPUT;
    const uint32_t *tables;
    ZS_TRY(get_fri_tables(ctx, dlog, &tables));
    const qm31 inv0 = qm_inv(fri_offset_pow(layer)), inv1 = qm_inv(fri_offset_pow(layer + 1));
    const uint64_t quarter = n / 4;
    ZS_LAUNCH(ctx, "fri_fold2_kernel", fri_fold2_kernel, dim3((uint32_t)((quarter + 255) / 256)), dim3(256), 0,
              ctx->stream, d_cur, quarter, layer, tables, inv0, inv1, d_alpha, d_leaves0, d_next1, d_leaves1, d_next2,
              d_leaves2);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

// Channel mode, a layer of fewer than 2^kSubtreeMinLog values: its fold and leaves go
// into the first launch of its tree, each workgroup folding the 2^K values under its
// subtree (no fold launch, and the leaves are hashed from LDS, not read back).
struct FoldLeafSource {
    static constexpr bool on = true;
    FoldLeaves f;
    __device__ void operator()(uint64_t j, uint32_t (&d)[8]) const {
        const qm31 a_off = qm_half(qm_mul(ldg_qm(f.alpha), f.inv_off));
        const qm31 v = fold_value(f.cur, f.half, j, f.layer, f.tables, a_off);
        stg_qm(f.next + 4 * j, v);
        const uint32_t w[4] = {v.a.c0, v.a.c1, v.b.c0, v.b.c1};
        b3_hash_qm31(w, d);
    }
};

__global__ __launch_bounds__(BLOCK_NT) void tree_block_fold_kernel(uint32_t *base, const BlockTrees bt,
                                                                   const FoldLeafSource src) {
    tree_block_body<BLOCK_NT>(base, bt, blockIdx.x, threadIdx.x, src);
}

int32_t launch_tree_block_fold(Ctx *ctx, uint32_t *base, const BlockTrees &bt, uint32_t wgs, size_t lds_bytes,
                               const FoldLeaves &fold) {
    if (bt.n != 1 || (fold.half >> bt.r[0]) != 1) return ZS_ERR_INPUT;
    ZS_LAUNCH(ctx, "tree_block_fold_kernel", tree_block_fold_kernel, dim3(wgs), dim3(BLOCK_NT), lds_bytes,
              ctx->stream, base, bt, FoldLeafSource{fold});
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

int32_t launch_fri_fold(Ctx *ctx, const uint32_t *d_cur, uint64_t n, uint32_t dlog, uint32_t layer,
                        const uint32_t *d_alpha, uint32_t *d_next, uint32_t *d_next_leaves) {
    const uint32_t *tables;
    ZS_TRY(get_fri_tables(ctx, dlog, &tables));
    qm31 inv_off = qm_inv(fri_offset_pow(layer));
    uint64_t half = n / 2;
    ZS_LAUNCH(ctx, "fri_fold_kernel", fri_fold_kernel, dim3((uint32_t)((half + 255) / 256)), dim3(256), 0, ctx->stream,
              d_cur, half, layer, tables, inv_off, d_alpha, d_next, d_next_leaves);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

}  // namespace zs

namespace zs {

// g[i] = sum_c gamma^c * lde_c[i] (QM31 combination of M31 columns); powers: n_cols QM31 values.
// Products of two canonical values are < 2^62, so four of them are summed in 64 bits
// (one multiply-add each) before a reduction.

// with leaves: also the hashQM31 leaf of each value (the FRI's layer-0 leaves)
__global__ __launch_bounds__(256) void combine_kernel(const uint32_t *lde, uint64_t cs, uint32_t n_cols, uint64_t n,
                                                      const uint32_t *powers, uint32_t *out, uint32_t *leaves) {
    extern __shared__ uint32_t pw[];
    for (uint32_t k = threadIdx.x; k < 4 * n_cols; k += blockDim.x) pw[k] = powers[k];
    __syncthreads();
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 v = combine_row(lde, cs, n_cols, i, pw);
    stg4(out + 4 * i, v);
    if (leaves) store_qm31_leaf(leaves + 8 * i, {{v.x, v.y}, {v.z, v.w}});
}

int32_t launch_combine(Ctx *ctx, const uint32_t *d_lde, uint64_t cs, uint32_t n_cols, uint64_t n,
                       const uint32_t *d_powers, uint32_t *d_out, uint32_t *d_leaves) {
    size_t lds = size_t(n_cols) * 16;
    if (lds > 64 * 1024) return ZS_ERR_INPUT;
    ZS_LAUNCH(ctx, "combine_kernel", combine_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), lds, ctx->stream,
              d_lde, cs, n_cols, n, d_powers, d_out, d_leaves);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

__global__ __launch_bounds__(256) void combine_planes_kernel(const uint32_t *cols, uint64_t cs, uint32_t n_cols,
                                                             uint64_t n, const uint32_t *powers, uint32_t *planes) {
    extern __shared__ uint32_t pw[];
    for (uint32_t k = threadIdx.x; k < 4 * n_cols; k += blockDim.x) pw[k] = powers[k];
    __syncthreads();
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 v = combine_row(cols, cs, n_cols, i, pw);
    stg(planes + i, v.x);
    stg(planes + n + i, v.y);
    stg(planes + 2 * n + i, v.z);
    stg(planes + 3 * n + i, v.w);
}

int32_t launch_combine_planes(Ctx *ctx, const uint32_t *d_cols, uint64_t cs, uint32_t n_cols, uint64_t n,
                              const uint32_t *d_powers, uint32_t *d_planes) {
    size_t lds = size_t(n_cols) * 16;
    if (lds > 64 * 1024) return ZS_ERR_INPUT;
    ZS_LAUNCH(ctx, "combine_planes_kernel", combine_planes_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), lds,
              ctx->stream, d_cols, cs, n_cols, n, d_powers, d_planes);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

__global__ __launch_bounds__(256) void interleave4_kernel(const uint32_t *planes, uint64_t n, uint32_t *out,
                                                          uint32_t *leaves) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const qm31 v = {{ldg(planes + i), ldg(planes + n + i)}, {ldg(planes + 2 * n + i), ldg(planes + 3 * n + i)}};
    stg_qm(out + 4 * i, v);
    if (leaves) store_qm31_leaf(leaves + 8 * i, v);
}

int32_t launch_interleave4(Ctx *ctx, const uint32_t *d_planes, uint64_t n, uint32_t *d_out, uint32_t *d_leaves) {
    ZS_LAUNCH(ctx, "interleave4_kernel", interleave4_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0,
              ctx->stream, d_planes, n, d_out, d_leaves);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

}  // namespace zs

namespace zs {


// Fiat-Shamir on the device: absorb the 32-byte root at `digest` and draw a
// QM31 challenge (fri.zig:205-210 with channel.zig:31-43, 49-69 and
// qm31.zig:95-101), so a layer's fold never waits for the host.  The state is
// staged in LDS; the 64 lanes run the hasher's control flow in lockstep
// (identical state, so identical writes), split its copies and run each
// compression as 16 identical quads.
__global__ __launch_bounds__(64) void channel_draw_kernel(B3Hasher *st, const uint32_t *digest, uint32_t *out) {
    __shared__ B3Hasher h, snap;
    __shared__ uint32_t qs[24], work[16];
    channel_draw_body(st, nullptr, digest, out, h, snap, qs, work, threadIdx.x, 64);
}

int32_t launch_channel_draw(Ctx *ctx, B3Hasher *d_state, const uint32_t *d_digest, uint32_t *d_out) {
    ZS_LAUNCH(ctx, "channel_draw_kernel", channel_draw_kernel, dim3(1), dim3(64), 0, ctx->stream, d_state, d_digest,
              d_out);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

// ---------------------------------------------------------------------------
// FRI tail: every layer of at most 2^FRI_TAIL_LOG values in one workgroup.
//
// Small layers are latency-bound: each costs a tree launch or two, a channel
// launch and a fold launch, most of it launch latency.  Here one workgroup keeps
// the layer values and the tree levels in LDS and, per layer, builds the tree
// (quad-lane hash2 per node, each level also written to its place in HBM),
// draws the challenge from the root on wave 0 (channel mode) or reads it
// (given challenges: the trees are then built with the other layers'), folds,
// and hashes the next layer's leaves (fri.zig:199-219).
// ---------------------------------------------------------------------------
constexpr uint32_t FRI_TAIL_NT = 512;

__device__ __forceinline__ void fri_tail_body(const FriTail &t) {
    extern __shared__ uint32_t sm[];
    constexpr uint32_t NV = 1u << FRI_TAIL_LOG;
    uint32_t *const val0 = sm, *const val1 = sm + 4 * NV;       // layer values (QM31 words), ping-pong
    uint32_t *const lv0 = sm + 8 * NV, *const lv1 = sm + 16 * NV;  // tree levels: NV digests, then NV / 2
    auto val = [&](uint32_t k) { return k ? val1 : val0; };
    auto lv = [&](uint32_t k) { return k ? lv1 : lv0; };
    uint32_t *qmb = sm + 20 * NV;                   // 32 words per quad (hash2 blocks)
    B3Hasher *h = reinterpret_cast<B3Hasher *>(qmb + 32 * (FRI_TAIL_NT / 4));
    B3Hasher *snap = h + 1;
    uint32_t *qs = reinterpret_cast<uint32_t *>(snap + 1);  // 24 words: the channel's compressions
    uint32_t *work = qs + 24;                               // 16 words: channel byte staging
    uint32_t *al = work + 16;                               // 4 words: this layer's challenge
    const uint32_t tid = threadIdx.x, j = tid & 3u, quad = tid >> 2;
    const bool chan = t.chan != nullptr;
    uint32_t *qm = qmb + 32 * quad;
    uint32_t mo[28];
#pragma unroll
    for (int k = 0; k < 28; k++) mo[k] = kQuadMsg[j][k];
    const uint32_t iv_lo = kIV[j], iv_hi = kIV[4 + j];
    qm[20 + j] = 0u;
    qm[24 + j] = 0u;
    qm[28 + j] = 0u;
    constexpr uint32_t HW = sizeof(B3Hasher) / 4;
    {
        const uint32_t n = 1u << (t.dlog - t.i0);
        for (uint32_t k = tid; k < 4 * n; k += FRI_TAIL_NT) val0[k] = t.in[k];
        if (chan) {
            const uint32_t *leaves = t.trees + t.tree_off[t.i0];
            for (uint32_t k = tid; k < 8 * n; k += FRI_TAIL_NT) lv0[k] = leaves[k];
            const uint32_t *cw = reinterpret_cast<const uint32_t *>(t.chan);
            uint32_t *hw = reinterpret_cast<uint32_t *>(h);
            for (uint32_t k = tid; k < HW; k += FRI_TAIL_NT) hw[k] = cw[k];
        }
    }
    __syncthreads();
    uint32_t cur = 0;
    for (uint32_t i = t.i0; i < t.L; i++) {
        const uint32_t li = t.dlog - i, n = 1u << li;
        if (chan) {
            // the layer's tree from its leaves in lv[0]
            uint32_t *tg = t.trees + t.tree_off[i];
            uint64_t off = 0;  // words before the child level in tg
            uint32_t P = n, src = 0;
            for (uint32_t l = 1; l <= li; l++) {
                const uint32_t *cs = lv(src);
                uint32_t *ds = lv(src ^ 1u);
                const uint64_t out = off + (uint64_t)P * 8;
                P >>= 1;
                if (P >= FRI_TAIL_NT / 4) {
                    // one or more nodes per quad: one node per lane finishes sooner
                    for (uint32_t p = tid; p < P; p += FRI_TAIL_NT) {
                        uint32_t x[8], y[8], d[8];
                        for (int w = 0; w < 8; w++) {
                            x[w] = cs[16 * p + w];
                            y[w] = cs[16 * p + 8 + w];
                        }
                        b3_hash2(x, y, d);
                        for (int w = 0; w < 8; w++) {
                            ds[8 * p + w] = d[w];
                            tg[out + 8 * p + w] = d[w];
                        }
                    }
                } else {
                    for (uint32_t p = quad; p < P; p += FRI_TAIL_NT / 4) {
                        uint32_t lo, hi;
                        quad_hash2(cs + 16 * p, qm, mo, j, iv_lo, iv_hi, lo, hi);
                        ds[8 * p + j] = lo;
                        ds[8 * p + 4 + j] = hi;
                        tg[out + 8 * p + j] = lo;
                        tg[out + 8 * p + 4 + j] = hi;
                    }
                }
                __syncthreads();
                off = out;
                src ^= 1u;
            }
            // absorb the root (lv(src)) and draw the challenge (fri.zig:205-210)
            if (tid < 64) {  // wave 0
                uint32_t d[8], q[4];
                for (int w = 0; w < 8; w++) d[w] = lv(src)[w];
                if (t.roots && tid < 8) t.roots[8 * i + tid] = lv(src)[tid];
                quad_channel_draw(*h, *snap, qs, work, d, q, j, mo);
                if (tid == 0)
                    for (int k = 0; k < 4; k++) {
                        al[k] = q[k];
                        t.alphas[4 * i + k] = q[k];
                    }
            }
        } else if (tid < 4) {
            al[tid] = t.alphas[4 * i + tid];
        }
        __syncthreads();
        // fold layer i into layer i + 1 (fri.zig:211-219), hashing the new layer's leaves
        const qm31 alpha = {{al[0], al[1]}, {al[2], al[3]}};
        const uint32_t half = n >> 1;
        const uint32_t *vin = val(cur);
        uint32_t *vout = val(cur ^ 1u);
        uint32_t *gout = t.layers + t.lay_off[i + 1];
        uint32_t *gleaves = (i + 1 < t.L) ? t.trees + t.tree_off[i + 1] : nullptr;
        const qm31 a_off = qm_half(qm_mul(alpha, t.inv_off[i]));  // alpha * inv_off / 2, as in fri_fold_kernel
        for (uint32_t k = tid; k < half; k += FRI_TAIL_NT) {
            const qm31 y0 = ld_qm(vin + 4 * k), y1 = ld_qm(vin + 4 * (k + half));
            const uint64_t e = (uint64_t)k << i;
            const qm31 tw = qm_mul(ld_qm(t.tables + 4 * (e & 1023)), ld_qm(t.tables + 4 * (1024 + (e >> 10))));
            const qm31 hs = qm_half(qm_add(y0, y1));
            const qm31 v = qm_add(hs, qm_mul(qm_mul(a_off, tw), qm_sub(y0, y1)));
            st_qm(vout + 4 * k, v);
            st_qm(gout + 4 * k, v);
            if (gleaves) {
                uint32_t w[4] = {v.a.c0, v.a.c1, v.b.c0, v.b.c1};
                uint32_t d[8];
                b3_hash_qm31(w, d);
                uint4 *o = reinterpret_cast<uint4 *>(gleaves + 8 * k);
                o[0] = make_uint4(d[0], d[1], d[2], d[3]);
                o[1] = make_uint4(d[4], d[5], d[6], d[7]);
                if (chan)
                    for (int w8 = 0; w8 < 8; w8++) lv0[8 * k + w8] = d[w8];
            }
        }
        __syncthreads();
        cur ^= 1u;
    }
    if (chan) {
        uint32_t *cw = reinterpret_cast<uint32_t *>(t.chan);
        const uint32_t *hw = reinterpret_cast<const uint32_t *>(h);
        for (uint32_t k = tid; k < HW; k += FRI_TAIL_NT) cw[k] = hw[k];
    }
    // the remainder from layer L in LDS (fri_finish_kernel's map, fri.zig:228-233)
    if (t.rem && tid < t.rc) {
        const uint32_t *last = val(cur);
        qm31 acc = {{0, 0}, {0, 0}};
        for (uint32_t k = 0; k < t.rc; k++)
            acc = qm_add(acc, qm_mul(ld_qm(t.rem_mat + 4 * (tid * t.rc + k)), ld_qm(last + 4 * k)));
        st_qm(t.rem + 4 * tid, acc);
    }
}

__global__ __launch_bounds__(FRI_TAIL_NT) void fri_tail_kernel(FriTail t) { fri_tail_body(t); }

// workgroup b runs the tail of descriptor b
__global__ __launch_bounds__(FRI_TAIL_NT) void fri_tail_multi_kernel(const FriTail *ts) { fri_tail_body(ts[blockIdx.x]); }

constexpr size_t kFriTailLds = (size_t(20) << FRI_TAIL_LOG) * 4 + size_t(32) * (FRI_TAIL_NT / 4) * 4 +
                               2 * sizeof(B3Hasher) + (24 + 16 + 4) * 4;

int32_t launch_fri_tail(Ctx *ctx, const FriTail &t) {
    if (t.rem && (t.rc == 0 || t.rc > 64 || t.rc > (1u << (t.dlog - t.L)))) return ZS_ERR_INPUT;
    ZS_LAUNCH(ctx, "fri_tail_kernel", fri_tail_kernel, dim3(1), dim3(FRI_TAIL_NT), kFriTailLds, ctx->stream, t);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

int32_t launch_fri_tail_multi(Ctx *ctx, const FriTail *d_tails, uint32_t n) {
    if (n == 0) return ZS_OK;
    ZS_LAUNCH(ctx, "fri_tail_multi_kernel", fri_tail_multi_kernel, dim3(n), dim3(FRI_TAIL_NT), kFriTailLds,
              ctx->stream, d_tails);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

// FRI remainder (fri.zig:228-233): the coefficients of the unique interpolant
// (univariate.zig:79-139) of the first rc values of layer L on the points
// x_j = FRI_OFFSET^(2^L) * w^j, w of order 2^remainder_log.  The points depend
// only on (L, remainder_log, rc), so the interpolation is a fixed linear map:
// coeffs = M * ys with M the inverse Vandermonde matrix, built once per shape on
// the host (column j = the interpolant of the j-th unit vector) and cached on
// the device.  Thread k computes coefficient k.
int32_t get_remainder_matrix(Ctx *ctx, uint32_t L, uint32_t remainder_log, uint32_t rc, const uint32_t **out) {
    DeviceCaches &dc = *ctx->caches;
    const uint64_t key = (uint64_t(L) << 32) | (uint64_t(remainder_log) << 16) | rc;
    {
        std::lock_guard<std::mutex> lk(dc.mu);
        auto it = dc.remainder_mats.find(key);
        if (it != dc.remainder_mats.end()) {
            *out = it->second->as<uint32_t>();
            return ZS_OK;
        }
    }
    std::vector<qm31> xs(rc), ys(rc), co;
    const qm31 w = qm31_root_of_unity(remainder_log), zero = {{0, 0}, {0, 0}}, one = {{1, 0}, {0, 0}};
    qm31 acc = fri_offset_pow(L);
    for (uint32_t j = 0; j < rc; j++) {
        xs[j] = acc;
        acc = qm_mul(acc, w);
    }
    std::vector<uint32_t> host(size_t(rc) * rc * 4);
    for (uint32_t j = 0; j < rc; j++) {
        for (uint32_t i = 0; i < rc; i++) ys[i] = i == j ? one : zero;
        interpolate_host(xs, ys, co);
        for (uint32_t k = 0; k < rc; k++) qm_to_words(co[k], &host[4 * (size_t(k) * rc + j)]);
    }
    auto buf = std::make_unique<DevBuf>();
    ZS_HIP(ctx, hipMalloc(&buf->ptr, host.size() * 4));
    buf->bytes = host.size() * 4;
    ZS_HIP(ctx, hipMemcpyAsync(buf->ptr, host.data(), host.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> lk(dc.mu);
    auto it = dc.remainder_mats.find(key);
    if (it != dc.remainder_mats.end()) {
        *out = it->second->as<uint32_t>();
        return ZS_OK;
    }
    *out = buf->as<uint32_t>();
    dc.remainder_mats[key] = std::move(buf);
    return ZS_OK;
}

// The end of a commit phase in one launch: every layer's root gathered into roots
// (after the 8 words of prefix, when given: the commit job's trace root), and the
// remainder coefficients right after the roots.
__global__ __launch_bounds__(256) void fri_finish_kernel(const uint32_t *trees, RootOffsets ro, uint32_t n_roots,
                                                         const uint32_t *prefix, uint32_t *roots, const uint32_t *last,
                                                         const uint32_t *mat, uint32_t rc, uint32_t *rem) {
    const uint32_t t = threadIdx.x;
    if (t < 8 * n_roots) roots[t] = trees[ro.off[t >> 3] + (t & 7u)];
    if (prefix && t < 8) (roots - 8)[t] = prefix[t];  // the 8 words just before the roots
    if (t < rc) {
        qm31 acc = {{0, 0}, {0, 0}};
        for (uint32_t j = 0; j < rc; j++) acc = qm_add(acc, qm_mul(ld_qm(mat + 4 * (t * rc + j)), ld_qm(last + 4 * j)));
        st_qm(rem + 4 * t, acc);
    }
}

int32_t launch_fri_finish(Ctx *ctx, const uint32_t *d_trees, const uint64_t *root_off, uint32_t n_roots,
                          const uint32_t *d_prefix, uint32_t *d_roots, const uint32_t *d_last, const uint32_t *d_mat,
                          uint32_t rc, uint32_t *d_rem) {
    if (n_roots > 28 || rc > 64) return ZS_ERR_INPUT;  // rc == 0: the roots only
    RootOffsets ro{};
    for (uint32_t k = 0; k < n_roots; k++) ro.off[k] = root_off[k];
    ZS_LAUNCH(ctx, "fri_finish_kernel", fri_finish_kernel, dim3(1), dim3(256), 0, ctx->stream, d_trees, ro, n_roots,
              d_prefix, d_roots, d_last, d_mat, rc, d_rem);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

// out[c] = x^c for c < n (square-and-multiply per lane)
__global__ __launch_bounds__(256) void qm31_powers_kernel(const uint32_t *x, uint32_t *out, uint32_t n) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    qm31 b = ld_qm(x), r = {{1, 0}, {0, 0}};
    for (uint32_t e = c; e; e >>= 1) {
        if (e & 1u) r = qm_mul(r, b);
        b = qm_mul(b, b);
    }
    st_qm(out + 4 * c, r);
}

int32_t launch_qm31_powers(Ctx *ctx, const uint32_t *d_x, uint32_t *d_out, uint32_t n) {
    ZS_LAUNCH(ctx, "qm31_powers_kernel", qm31_powers_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_x,
              d_out, n);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

}  // namespace zs
