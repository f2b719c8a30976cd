// b3_quad.h -- BLAKE3 compression shared by the four lanes of a quad (device only).
//
// Used where a compression's latency matters more than its throughput: the
// narrow levels of Merkle trees near the root and the Fiat-Shamir channel.
#pragma once
#include "b3_hasher.h"

namespace zs {

// LDS writes of a wave must land before its lanes read each other's values
__device__ __forceinline__ void wave_lds_sync_m() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---------------------------------------------------------------------------
// Four lanes per compression.
//
// One compression on one lane is about 700 dependent instructions; where only
// a few compressions are in flight (tree levels near the root, the channel),
// that latency is the cost.  Here the four lanes of a quad share a compression: lane j
// holds column j of the 4x4 state (s[j], s[4+j], s[8+j], s[12+j]) and runs
// the column G on it; for the diagonal G, rows 1-3 are rotated across the quad
// by 1, 2 and 3 lanes (DPP quad_perm) and rotated back afterwards.  The
// message words a lane needs in round r are m[PERM^r(2j)], m[PERM^r(2j+1)]
// (column) and m[PERM^r(8+2j)], m[PERM^r(9+2j)] (diagonal): the quad writes
// the block's 16 words to LDS and each lane reads its 28 through a fixed
// per-lane index table.
// ---------------------------------------------------------------------------

// kQuadMsg[j][4r + s]: message word of lane j, round r, slot s (column x, y, diagonal x, y)
static __constant__ uint8_t kQuadMsg[4][28] = {
    {0, 1, 8, 9, 2, 6, 1, 11, 3, 4, 6, 5, 10, 7, 4, 0, 12, 13, 7, 2, 9, 14, 13, 3, 11, 15, 14, 10},
    {2, 3, 10, 11, 3, 10, 12, 5, 10, 12, 9, 0, 12, 9, 11, 2, 9, 11, 5, 3, 11, 5, 0, 10, 5, 0, 2, 12},
    {4, 5, 12, 13, 7, 0, 9, 14, 13, 2, 11, 15, 14, 3, 5, 8, 15, 10, 0, 1, 8, 12, 2, 6, 1, 9, 3, 4},
    {6, 7, 14, 15, 4, 13, 15, 8, 7, 14, 8, 1, 13, 15, 1, 6, 14, 8, 6, 4, 15, 1, 4, 7, 8, 6, 7, 13},
};

static __constant__ uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                       0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

// lane j of the quad reads lane p[j]: rotations by 1, 2, 3 lanes
constexpr int QROT1 = 0x39, QROT2 = 0x4E, QROT3 = 0x93;
template <int CTRL>
__device__ __forceinline__ uint32_t qperm(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

__device__ __forceinline__ void quad_compress(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, const uint32_t *msg,
                                              const uint32_t (&mo)[28]) {
    // all 28 message words are read before the first round, so the rounds never
    // wait on LDS
    uint32_t mm[28];
#pragma unroll
    for (int k = 0; k < 28; k++) mm[k] = msg[mo[k]];
#pragma unroll
    for (int r = 0; r < 7; r++) {
        const uint32_t m0 = mm[4 * r], m1 = mm[4 * r + 1];
        const uint32_t m2 = mm[4 * r + 2], m3 = mm[4 * r + 3];
        ZS_B3_G(a, b, c, d, m0, m1);
        b = qperm<QROT1>(b);
        c = qperm<QROT2>(c);
        d = qperm<QROT3>(d);
        ZS_B3_G(a, b, c, d, m2, m3);
        b = qperm<QROT3>(b);
        c = qperm<QROT2>(c);
        d = qperm<QROT1>(d);
    }
}

// hash2 of the 16 words at xp (left || right digest) by the quad; lane j gets
// digest words j and 4 + j.  qm: the quad's 32 LDS words (block 0, then block 1
// whose words 4..15 are zero).  iv_lo / iv_hi: IV[j], IV[4 + j].
__device__ __forceinline__ void quad_hash2(const uint32_t *xp, uint32_t *qm, const uint32_t (&mo)[28], uint32_t j,
                                           uint32_t iv_lo, uint32_t iv_hi, uint32_t &o_lo, uint32_t &o_hi) {
    if (j == 0) {
        qm[0] = 0x2d67697au;  // "zig-"
        qm[1] = 0x72617473u;  // "star"
        qm[2] = 0x61703a6bu;  // "k:pa"
        qm[3] = 0x7269u | (xp[0] << 16);
    } else {
        const uint32_t xa = xp[4 * j - 4], xb = xp[4 * j - 3], xc = xp[4 * j - 2], xd = xp[4 * j - 1], xe = xp[4 * j];
        qm[4 * j] = __builtin_amdgcn_alignbit(xb, xa, 16);
        qm[4 * j + 1] = __builtin_amdgcn_alignbit(xc, xb, 16);
        qm[4 * j + 2] = __builtin_amdgcn_alignbit(xd, xc, 16);
        qm[4 * j + 3] = __builtin_amdgcn_alignbit(xe, xd, 16);
    }
    const uint32_t y0 = xp[12 + j], y1 = xp[j < 3 ? 13 + j : 15u];
    qm[16 + j] = __builtin_amdgcn_alignbit(j < 3 ? y1 : 0u, y0, 16);
    wave_lds_sync_m();
    uint32_t a = iv_lo, b = iv_hi, c = iv_lo, d = j == 2 ? 64u : (j == 3 ? B3_CHUNK_START : 0u);
    quad_compress(a, b, c, d, qm, mo);
    const uint32_t h0 = a ^ c, h1 = b ^ d;
    a = h0;
    b = h1;
    c = iv_lo;
    d = j == 2 ? 14u : (j == 3 ? (B3_CHUNK_END | B3_ROOT) : 0u);
    quad_compress(a, b, c, d, qm + 16, mo);
    o_lo = a ^ c;
    o_hi = b ^ d;
    wave_lds_sync_m();  // every lane has read qm before the quad's next hash2 rewrites it
}

// A hasher's compression (B3Hasher's CF) run by the 16 quads of one whole wave:
// every lane holds the same chaining value and message; lane j of each quad
// computes column j of the state, and the quads trade the result through the
// same LDS words (identical values).  The hasher's copies are split over the
// 64 lanes.
struct B3QuadCF {
    uint32_t *qs;               // 24 LDS words: message, then chaining value
    uint32_t j;                 // lane in the quad
    const uint32_t (&mo)[28];   // the lane's message-word table (kQuadMsg[j])
    __device__ uint32_t lane() const { return __lane_id(); }
    __device__ uint32_t lanes() const { return 64; }
    __device__ void sync() const { wave_lds_sync_m(); }
    __device__ void operator()(uint32_t cv[8], const uint32_t m[16], uint32_t counter_lo, uint32_t counter_hi,
                               uint32_t block_len, uint32_t flags) {
        for (int i = 0; i < 16; i++) qs[i] = m[i];
        for (int i = 0; i < 8; i++) qs[16 + i] = cv[i];
        wave_lds_sync_m();
        uint32_t a = qs[16 + j], b = qs[20 + j], c = kIV[j];
        uint32_t d = j == 0 ? counter_lo : (j == 1 ? counter_hi : (j == 2 ? block_len : flags));
        quad_compress(a, b, c, d, qs, mo);
        wave_lds_sync_m();  // every lane has read the message and the chaining value
        qs[16 + j] = a ^ c;
        qs[20 + j] = b ^ d;
        wave_lds_sync_m();
        for (int i = 0; i < 8; i++) cv[i] = qs[16 + i];
        wave_lds_sync_m();  // read before the next compression rewrites qs
    }
};

// Fiat-Shamir step on the device (fri.zig:205-210 with channel.zig:31-43,
// 49-69, qm31.zig:95-101): the 64 lanes of one wave (all active) absorb the
// digest d into the LDS-resident hasher h and draw a QM31 challenge into q
// (every lane gets it).  qs: 24 LDS words, work: 16 LDS words, snap: LDS
// working storage.
__device__ __forceinline__ void quad_channel_draw(B3Hasher &h, B3Hasher &snap, uint32_t *qs, uint32_t *work,
                                                  const uint32_t d[8], uint32_t q[4], uint32_t j,
                                                  const uint32_t (&mo)[28]) {
    B3QuadCF cf{qs, j, mo};
    channel_absorb_digest(h, d, work, cf);
    channel_sample_qm31(h, snap, q, work, cf);
}

}  // namespace zs
