// fri_dev.h -- device bodies of the commitment's FRI steps (fri.hip runs them as
// kernels; the persistent batch kernel as tasks): the QM31 combination of a row,
// one fold value, and a Fiat-Shamir draw from a root.
#pragma once
#include "b3_quad.h"
#include "internal.h"

namespace zs {

// a QM31 value in global memory
__device__ __forceinline__ qm31 ldg_qm(const uint32_t *p) {
    const uint4 v = ldg4(p);
    return {{v.x, v.y}, {v.z, v.w}};
}
__device__ __forceinline__ void stg_qm(uint32_t *p, qm31 x) { stg4(p, make_uint4(x.a.c0, x.a.c1, x.b.c0, x.b.c1)); }

// g[i] = sum_c gamma^c * lde_c[i] (QM31 combination of M31 columns; pw: the powers, in
// LDS).  Products of two canonical values are < 2^62, so four of them are summed in 64
// bits (one multiply-add each) before a reduction.
__device__ __forceinline__ uint4 combine_row(const uint32_t *lde, uint64_t cs, uint32_t n_cols, uint64_t i,
                                             const uint32_t *pw) {
    uint32_t a[4] = {0, 0, 0, 0};
    uint32_t c = 0;
    for (; c + 4 <= n_cols; c += 4) {
        uint32_t v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = ldg(lde + (uint64_t)(c + k) * cs + i);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint64_t acc = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) acc += (uint64_t)pw[4 * (c + k) + q] * v[k];
            a[q] = m31_add(a[q], m31_reduce64(acc));
        }
    }
    for (; c < n_cols; c++) {
        const uint32_t v = ldg(lde + (uint64_t)c * cs + i);
#pragma unroll
        for (int q = 0; q < 4; q++) a[q] = m31_add(a[q], m31_mul(pw[4 * c + q], v));
    }
    return make_uint4(a[0], a[1], a[2], a[3]);
}

// next[j] of layer `layer` (fri.zig:211-219): (y0 + y1)/2 + alpha (y0 - y1) x^-1 / 2 with
// x^-1 = inv_off * t_lo * t_hi from the two power tables; a_off = alpha * inv_off / 2
__device__ __forceinline__ qm31 fold_values(qm31 y0, qm31 y1, uint64_t j, uint32_t layer, const uint32_t *tables,
                                            qm31 a_off) {
    const uint64_t e = j << layer;
    const qm31 t = qm_mul(ldg_qm(tables + 4 * (e & 1023)), ldg_qm(tables + 4 * (1024 + (e >> 10))));
    const qm31 hs = qm_half(qm_add(y0, y1));
    return qm_add(hs, qm_mul(qm_mul(a_off, t), qm_sub(y0, y1)));
}
__device__ __forceinline__ qm31 fold_value(const uint32_t *cur, uint64_t half, uint64_t j, uint32_t layer,
                                           const uint32_t *tables, qm31 a_off) {
    return fold_values(ldg_qm(cur + 4 * j), ldg_qm(cur + 4 * (j + half)), j, layer, tables, a_off);
}

// hashQM31 leaf of v at leaf
__device__ __forceinline__ void store_qm31_leaf(uint32_t *leaf, qm31 v) {
    uint32_t w[4] = {v.a.c0, v.a.c1, v.b.c0, v.b.c1};
    uint32_t d[8];
    b3_hash_qm31(w, d);
    stg4(leaf, make_uint4(d[0], d[1], d[2], d[3]));
    stg4(leaf + 4, make_uint4(d[4], d[5], d[6], d[7]));
}

// Fiat-Shamir on the device: absorb the 32-byte root at `digest` into the channel at
// `st` (starting from `init` when given: st <- init first) and draw a QM31 challenge
// into out (fri.zig:205-210 with channel.zig:31-43, 49-69 and qm31.zig:95-101).  The
// state is staged in LDS (h, snap: hasher states; qs 24 words, work 16 words); every
// thread of the workgroup (nt threads) helps copy it, the first wave runs the hasher's
// control flow in lockstep (identical state, so identical writes) with each
// compression as 16 identical quads.
__device__ __forceinline__ void channel_draw_body(B3Hasher *st, const B3Hasher *init, const uint32_t *digest,
                                                  uint32_t *out, B3Hasher &h, B3Hasher &snap, uint32_t *qs,
                                                  uint32_t *work, uint32_t tid, uint32_t nt) {
    constexpr uint32_t W = sizeof(B3Hasher) / 4;
    static_assert(sizeof(B3Hasher) % 4 == 0, "hasher state is word-sized");
    uint32_t *hw = reinterpret_cast<uint32_t *>(&h);
    const uint32_t *sw = reinterpret_cast<const uint32_t *>(init ? init : st);
    for (uint32_t k = tid; k < W; k += nt) hw[k] = ldg(sw + k);
    __syncthreads();
    if (tid < 64) {
        const uint32_t j = tid & 3u;
        uint32_t mo[28];
#pragma unroll
        for (int k = 0; k < 28; k++) mo[k] = kQuadMsg[j][k];
        uint32_t d[8], q[4];
        for (int i = 0; i < 8; i++) d[i] = ldg(digest + i);
        quad_channel_draw(h, snap, qs, work, d, q, j, mo);
        if (tid == 0)
            for (int i = 0; i < 4; i++) stg(out + i, q[i]);
    }
    __syncthreads();
    uint32_t *dw = reinterpret_cast<uint32_t *>(st);
    for (uint32_t k = tid; k < W; k += nt) stg(dw + k, hw[k]);
}

}  // namespace zs
