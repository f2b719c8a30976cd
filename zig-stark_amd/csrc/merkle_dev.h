// merkle_dev.h -- device bodies of the BLAKE3 row leaves and Merkle node levels
// (merkle.hip runs them as kernels; the persistent batch kernel as tasks).
#pragma once
#include "b3_quad.h"
#include "internal.h"

namespace zs {

// a digest in global memory
__device__ __forceinline__ void ld_digest_g(const uint32_t *p, uint32_t d[8]) {
    const uint4 a = ldg4(p), b = ldg4(p + 4);
    d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w; d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
}
__device__ __forceinline__ void st_digest_g(uint32_t *p, const uint32_t d[8]) {
    stg4(p, make_uint4(d[0], d[1], d[2], d[3]));
    stg4(p + 4, make_uint4(d[4], d[5], d[6], d[7]));
}
// a digest in global memory or LDS
__device__ __forceinline__ void ld_digest(const uint32_t *p, uint32_t d[8]) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    uint4 a = q[0], b = q[1];
    d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w; d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
}
__device__ __forceinline__ void st_digest(uint32_t *p, const uint32_t d[8]) {
    uint4 *q = reinterpret_cast<uint4 *>(p);
    q[0] = make_uint4(d[0], d[1], d[2], d[3]);
    q[1] = make_uint4(d[4], d[5], d[6], d[7]);
}

// BLAKE3 of one chunk (<= 256 columns = 1024 bytes) of a row.
__device__ __forceinline__ void row_chunk_cv(const uint32_t *cols, uint64_t cs, uint64_t row, uint32_t c0,
                                             uint32_t ncols, uint32_t chunk_idx, bool root, uint32_t cv[8]) {
    b3_iv(cv);
    uint32_t len = ncols * 4;
    uint32_t nblocks = (len + 63) / 64;
    for (uint32_t b = 0; b < nblocks; b++) {
        uint32_t m[16];
#pragma unroll
        for (int w = 0; w < 16; w++) {
            uint32_t c = b * 16 + w;
            m[w] = c < ncols ? ldg(cols + (uint64_t)(c0 + c) * cs + row) : 0u;
        }
        uint32_t blen = (len - b * 64) < 64 ? (len - b * 64) : 64;
        uint32_t flags = (b == 0 ? B3_CHUNK_START : 0) | (b == nblocks - 1 ? (B3_CHUNK_END | (root ? B3_ROOT : 0)) : 0);
        b3_compress_cv(cv, m, chunk_idx, 0, blen, flags);
    }
}

// rows of <= 256 columns: one BLAKE3 chunk per leaf
__device__ __forceinline__ void row_leaf_body(const uint32_t *cols, uint64_t cs, uint32_t n_cols, uint64_t row,
                                              uint32_t *nodes) {
    uint32_t cv[8];
    row_chunk_cv(cols, cs, row, 0, n_cols, 0, true, cv);
    st_digest_g(nodes + row * 8, cv);
}

// rows of more than 256 columns: BLAKE3 chunk tree with a chaining-value stack
__device__ __forceinline__ void row_leaf_multichunk_body(const uint32_t *cols, uint64_t cs, uint32_t n_cols,
                                                         uint64_t row, uint32_t *nodes) {
    uint32_t cv[8];
    uint32_t n_chunks = (n_cols + 255) / 256;
    uint32_t stack[16][8];
    uint32_t sp = 0;
    for (uint32_t c = 0; c + 1 < n_chunks; c++) {
        row_chunk_cv(cols, cs, row, c * 256, 256, c, false, cv);
        uint32_t total = c + 1;
        while ((total & 1u) == 0) {
            sp--;
            uint32_t t[8];
            b3_parent(stack[sp], cv, false, t);
            for (int w = 0; w < 8; w++) cv[w] = t[w];
            total >>= 1;
        }
        for (int w = 0; w < 8; w++) stack[sp][w] = cv[w];
        sp++;
    }
    uint32_t last = n_chunks - 1;
    row_chunk_cv(cols, cs, row, last * 256, n_cols - last * 256, last, false, cv);
    while (sp > 0) {
        sp--;
        uint32_t t[8];
        b3_parent(stack[sp], cv, sp == 0, t);
        for (int w = 0; w < 8; w++) cv[w] = t[w];
    }
    uint4 *o = reinterpret_cast<uint4 *>(nodes + row * 8);
    o[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
    o[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

// NLEV (1..3) levels per launch with no idle lanes: lane i owns the subtree
// over children 2^NLEV * i .. 2^NLEV * (i + 1) - 1 and writes all its nodes.
template <int NLEV>
__device__ __forceinline__ void subtree_lane(const uint32_t *child, uint32_t *lvl1, uint64_t n_lanes, uint64_t i) {
    constexpr int W = 1 << NLEV;
    uint32_t *lvl2 = lvl1 + n_lanes * (W / 2) * 8;
    uint32_t *lvl3 = lvl2 + n_lanes * (W / 4) * 8;
    uint32_t a[8], b[8], n1[W / 2][8];
#pragma unroll
    for (int k = 0; k < W / 2; k++) {
        ld_digest_g(child + (i * W + 2 * k) * 8, a);
        ld_digest_g(child + (i * W + 2 * k + 1) * 8, b);
        b3_hash2(a, b, n1[k]);
        st_digest_g(lvl1 + (i * (W / 2) + k) * 8, n1[k]);
    }
    if constexpr (NLEV >= 2) {
        uint32_t n2[W / 4][8];
#pragma unroll
        for (int k = 0; k < W / 4; k++) {
            b3_hash2(n1[2 * k], n1[2 * k + 1], n2[k]);
            st_digest_g(lvl2 + (i * (W / 4) + k) * 8, n2[k]);
        }
        if constexpr (NLEV >= 3) {
            uint32_t n3[8];
            b3_hash2(n2[0], n2[1], n3);
            st_digest_g(lvl3 + i * 8, n3);
        }
    }
}

// Several independent trees in one launch (the layers of an FRI commitment):
// tree t owns lanes [lane_end[t-1], lane_end[t]) and its levels at word offsets
// child[t], lvl1[t] from `base`.
struct MultiTrees {
    static constexpr int MAX = 32;
    uint32_t n;
    uint64_t child[MAX], lvl1[MAX], lane_end[MAX];
};

// Workgroups that each build K levels of a subtree of 2^K children, keeping
// the levels in LDS.  Levels with at least two nodes per quad hash one node per
// lane; narrower levels use a quad per node.  Several trees can share a launch:
// tree t starts at word offset child[t] (its level with 2^r[t] children) and owns
// workgroups [wg_end[t-1], wg_end[t]), 2^(r[t]-K[t]) of them.
struct BlockTrees {
    static constexpr int MAX = 32;
    uint32_t n;
    uint64_t child[MAX];
    uint32_t r[MAX], K[MAX], wg_end[MAX];
    // set on the launch that computes the root of its one tree: the workgroup then
    // absorbs the root into this device channel and draws a challenge into alpha
    // (starting the channel from chan_init when given); with powers, it also writes
    // the challenge's powers 0 .. n_powers - 1 there
    B3Hasher *chan;
    uint32_t *alpha;
    const B3Hasher *chan_init;
    uint32_t *powers;
    uint32_t n_powers;
};

constexpr uint32_t BLOCK_NT = 256;

// LDS words of a K-level subtree workgroup of NT threads: two level buffers, then 32 words per quad
ZS_HD size_t block_lds_words(uint32_t K, uint32_t NT = BLOCK_NT) { return (size_t(6) << K) + (NT / 4) * 32; }
// the channel state, its snapshot and the draw's staging after the tree words
constexpr size_t kDrawLdsWords = 2 * sizeof(B3Hasher) / 4 + 24 + 16;

// The same description for one tree (a task of the persistent batch kernel).
struct OneTree {
    uint32_t n;
    uint64_t child[1];
    uint32_t r[1], K[1], wg_end[1];
    B3Hasher *chan;
    uint32_t *alpha;
    const B3Hasher *chan_init;
    uint32_t *powers;
    uint32_t n_powers;
};

// The children of a workgroup subtree are read from their level in HBM, or, with a
// leaf source, computed: LEAVES::on, and leaves(j, d) sets d to child j's digest.
struct ReadLeaves {
    static constexpr bool on = false;
    __device__ void operator()(uint64_t, uint32_t (&)[8]) const {}
};
// LDS words after the tree's and the draw's where computed children are kept
ZS_HD size_t block_leaf_lds_words(uint32_t K, uint32_t NT = BLOCK_NT) {
    return block_lds_words(K, NT) + kDrawLdsWords;
}

// workgroup s of tree t (levels from its child level, r child levels left, K of them here);
// BT is BlockTrees or OneTree
template <uint32_t NT, class BT, class LEAVES = ReadLeaves>
__device__ __forceinline__ void tree_block_body(uint32_t *base, const BT &bt, uint32_t b, uint32_t tid,
                                                const LEAVES &leaves = LEAVES{}) {
    extern __shared__ uint32_t sm[];
    uint32_t t = 0;
    while (b >= bt.wg_end[t]) t++;
    const uint32_t s = b - (t ? bt.wg_end[t - 1] : 0u);
    const uint32_t K = bt.K[t], r = bt.r[t];
    const uint32_t j = tid & 3u, quad = tid >> 2;
    constexpr uint32_t NQ = NT / 4;
    uint32_t *bufs[2] = {sm, sm + (4u << K)};
    uint32_t *qm = sm + (6u << K) + quad * 32;
    const uint32_t iv_lo = kIV[j], iv_hi = kIV[4 + j];
    uint32_t mo[28];
#pragma unroll
    for (int k = 0; k < 28; k++) mo[k] = kQuadMsg[j][k];
    qm[20 + j] = 0u;
    qm[24 + j] = 0u;
    qm[28 + j] = 0u;
    uint64_t lvl = bt.child[t];  // word offset of the current child level
    uint64_t width = uint64_t(1) << r;
    const uint32_t *src = base + lvl + (uint64_t)s * (8u << K);
    if constexpr (LEAVES::on) {
        // this subtree's 2^K children: computed, written to their level, kept in LDS
        uint32_t *kids = sm + block_leaf_lds_words(K, NT);
        uint32_t *kids_g = base + lvl + (uint64_t)s * (8u << K);
        for (uint32_t p = tid; p < (1u << K); p += NT) {
            uint32_t d[8];
            leaves(((uint64_t)s << K) + p, d);
            st_digest_g(kids_g + 8 * p, d);
            st_digest(kids + 8 * p, d);
        }
        src = kids;
    }
    __syncthreads();
    for (uint32_t l = 0; l < K; l++) {
        const uint32_t P = 1u << (K - 1 - l);
        uint32_t *out = base + lvl + width * 8 + (uint64_t)s * P * 8;
        uint32_t *dst = bufs[l & 1];
        if (P >= NT / 2) {
            // two or more nodes per quad: one node per lane (a lane's compression has
            // four independent G chains, a quad's one) finishes sooner
            for (uint32_t p = tid; p < P; p += NT) {
                uint32_t x[8], y[8], d[8];
                ld_digest(src + 16 * p, x);
                ld_digest(src + 16 * p + 8, y);
                b3_hash2(x, y, d);
                st_digest_g(out + 8 * p, d);
                st_digest(dst + 8 * p, d);
            }
        } else {
            for (uint32_t p = quad; p < P; p += NQ) {
                uint32_t lo, hi;
                quad_hash2(src + 16 * p, qm, mo, j, iv_lo, iv_hi, lo, hi);
                stg(out + 8 * p + j, lo);
                stg(out + 8 * p + 4 + j, hi);
                dst[8 * p + j] = lo;
                dst[8 * p + 4 + j] = hi;
            }
        }
        __syncthreads();
        src = dst;
        lvl += width * 8;
        width >>= 1;
    }
    if (bt.chan) {
        // src holds the root: absorb it and draw the next challenge
        constexpr uint32_t HW = sizeof(B3Hasher) / 4;
        B3Hasher *hs = reinterpret_cast<B3Hasher *>(sm + block_lds_words(K, NT));
        B3Hasher *snap = hs + 1;
        uint32_t *qs = reinterpret_cast<uint32_t *>(snap + 1), *work = qs + 24;
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(bt.chan_init ? bt.chan_init : bt.chan);
        uint32_t *hw = reinterpret_cast<uint32_t *>(hs);
        for (uint32_t k = tid; k < HW; k += NT) hw[k] = cw[k];
        __syncthreads();
        if (tid < 64) {  // wave 0
            uint32_t d[8], q[4];
            for (int w = 0; w < 8; w++) d[w] = src[w];
            quad_channel_draw(*hs, *snap, qs, work, d, q, j, mo);
            if (tid == 0) {
                for (int k = 0; k < 4; k++) bt.alpha[k] = q[k];
                for (int k = 0; k < 4; k++) work[k] = q[k];  // for the powers below
            }
        }
        __syncthreads();
        uint32_t *ow = reinterpret_cast<uint32_t *>(bt.chan);
        for (uint32_t k = tid; k < HW; k += NT) ow[k] = hw[k];
        if (bt.powers) {
            // powers of the challenge, square-and-multiply per power (as qm31_powers_kernel)
            const qm31 g = {{work[0], work[1]}, {work[2], work[3]}};
            for (uint32_t c = tid; c < bt.n_powers; c += NT) {
                qm31 b = g, r = {{1u, 0u}, {0u, 0u}};
                for (uint32_t e = c; e; e >>= 1) {
                    if (e & 1u) r = qm_mul(r, b);
                    b = qm_mul(b, b);
                }
                stg4(bt.powers + 4 * c, make_uint4(r.a.c0, r.a.c1, r.b.c0, r.b.c1));
            }
        }
    }
}

}  // namespace zs
