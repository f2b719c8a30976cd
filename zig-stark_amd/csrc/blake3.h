// blake3.h -- BLAKE3 compression and the reference's message layouts, for host
// and device.
//
// The reference hashes with Zig std's std.crypto.hash.Blake3 (unkeyed,
// 32-byte output).  The messages on this path are:
//   hashQM31(v)     = BLAKE3(16 bytes)           src/m31/hash.zig:33-37
//   hashM31s(row)   = BLAKE3(4 * c bytes)        src/m31/hash.zig:15-25
//   hash2(l, r)     = BLAKE3("zig-stark:pair" || l || r), 78 bytes
//                                                src/core/hash/hash.zig:17-25
// Digests are kept as 8 little-endian u32 words, which is their byte order.
#pragma once
#include "field.h"

namespace zs {

constexpr uint32_t B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8;

ZS_HD uint32_t rotr(uint32_t x, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(x, x, n);
#else
    return (x >> n) | (x << (32 - n));
#endif
}

#define ZS_B3_G(a, b, c, d, mx, my)        \
    do {                                   \
        a = a + b + (mx);                  \
        d = rotr(d ^ a, 16);               \
        c = c + d;                         \
        b = rotr(b ^ c, 12);               \
        a = a + b + (my);                  \
        d = rotr(d ^ a, 8);                \
        c = c + d;                         \
        b = rotr(b ^ c, 7);                \
    } while (0)

// One round with the message words already in schedule order.
#define ZS_B3_ROUND(m0, m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11, m12, m13, m14, m15) \
    do {                                                                                   \
        ZS_B3_G(s0, s4, s8, s12, m0, m1);                                                  \
        ZS_B3_G(s1, s5, s9, s13, m2, m3);                                                  \
        ZS_B3_G(s2, s6, s10, s14, m4, m5);                                                 \
        ZS_B3_G(s3, s7, s11, s15, m6, m7);                                                 \
        ZS_B3_G(s0, s5, s10, s15, m8, m9);                                                 \
        ZS_B3_G(s1, s6, s11, s12, m10, m11);                                               \
        ZS_B3_G(s2, s7, s8, s13, m12, m13);                                                \
        ZS_B3_G(s3, s4, s9, s14, m14, m15);                                                \
    } while (0)

// Chaining-value output of one compression (the first 8 output words).
// The 7 rounds are written out with the message permutation applied at
// compile time: round r reads m[PERM^r[k]].
ZS_HD void b3_compress_cv(uint32_t cv[8], const uint32_t m[16], uint32_t counter_lo, uint32_t counter_hi,
                          uint32_t block_len, uint32_t flags) {
    uint32_t s0 = cv[0], s1 = cv[1], s2 = cv[2], s3 = cv[3], s4 = cv[4], s5 = cv[5], s6 = cv[6], s7 = cv[7];
    uint32_t s8 = 0x6A09E667u, s9 = 0xBB67AE85u, s10 = 0x3C6EF372u, s11 = 0xA54FF53Au;
    uint32_t s12 = counter_lo, s13 = counter_hi, s14 = block_len, s15 = flags;
    ZS_B3_ROUND(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11], m[12], m[13], m[14], m[15]);
    ZS_B3_ROUND(m[2], m[6], m[3], m[10], m[7], m[0], m[4], m[13], m[1], m[11], m[12], m[5], m[9], m[14], m[15], m[8]);
    ZS_B3_ROUND(m[3], m[4], m[10], m[12], m[13], m[2], m[7], m[14], m[6], m[5], m[9], m[0], m[11], m[15], m[8], m[1]);
    ZS_B3_ROUND(m[10], m[7], m[12], m[9], m[14], m[3], m[13], m[15], m[4], m[0], m[11], m[2], m[5], m[8], m[1], m[6]);
    ZS_B3_ROUND(m[12], m[13], m[9], m[11], m[15], m[10], m[14], m[8], m[7], m[2], m[5], m[3], m[0], m[1], m[6], m[4]);
    ZS_B3_ROUND(m[9], m[14], m[11], m[5], m[8], m[12], m[15], m[1], m[13], m[3], m[0], m[10], m[2], m[6], m[4], m[7]);
    ZS_B3_ROUND(m[11], m[15], m[5], m[0], m[1], m[9], m[8], m[6], m[14], m[10], m[2], m[12], m[3], m[4], m[7], m[13]);
    cv[0] = s0 ^ s8; cv[1] = s1 ^ s9; cv[2] = s2 ^ s10; cv[3] = s3 ^ s11;
    cv[4] = s4 ^ s12; cv[5] = s5 ^ s13; cv[6] = s6 ^ s14; cv[7] = s7 ^ s15;
}

ZS_HD void b3_iv(uint32_t cv[8]) {
    cv[0] = 0x6A09E667u; cv[1] = 0xBB67AE85u; cv[2] = 0x3C6EF372u; cv[3] = 0xA54FF53Au;
    cv[4] = 0x510E527Fu; cv[5] = 0x9B05688Cu; cv[6] = 0x1F83D9ABu; cv[7] = 0x5BE0CD19u;
}

// hashQM31: one 16-byte block, flags CHUNK_START | CHUNK_END | ROOT.
ZS_HD void b3_hash_qm31(const uint32_t v[4], uint32_t out[8]) {
    uint32_t m[16] = {v[0], v[1], v[2], v[3], 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    b3_iv(out);
    b3_compress_cv(out, m, 0, 0, 16, B3_CHUNK_START | B3_CHUNK_END | B3_ROOT);
}

// hash2: "zig-stark:pair" (14 bytes) then l and r, so the digest words sit
// 2 bytes off word alignment: message word j >= 4 is (X[j-4] >> 16) | (X[j-3] << 16)
// over X = l || r.
ZS_HD void b3_hash2(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
    const uint32_t W0 = 0x2d67697au;  // "zig-"
    const uint32_t W1 = 0x72617473u;  // "star"
    const uint32_t W2 = 0x61703a6bu;  // "k:pa"
    uint32_t x[16];
    for (int i = 0; i < 8; i++) { x[i] = l[i]; x[8 + i] = r[i]; }
    uint32_t m[16];
    m[0] = W0; m[1] = W1; m[2] = W2;
    m[3] = 0x7269u | (x[0] << 16);  // "ir" + first two digest bytes
    for (int j = 4; j < 16; j++) m[j] = (x[j - 4] >> 16) | (x[j - 3] << 16);
    b3_iv(out);
    b3_compress_cv(out, m, 0, 0, 64, B3_CHUNK_START);
    uint32_t m2[16] = {(x[12] >> 16) | (x[13] << 16), (x[13] >> 16) | (x[14] << 16), (x[14] >> 16) | (x[15] << 16),
                       x[15] >> 16, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    b3_compress_cv(out, m2, 0, 0, 14, B3_CHUNK_END | B3_ROOT);
}

// parent node of BLAKE3's own chunk tree (used for messages over 1024 bytes)
ZS_HD void b3_parent(const uint32_t l[8], const uint32_t r[8], bool root, uint32_t out[8]) {
    uint32_t m[16];
    for (int i = 0; i < 8; i++) { m[i] = l[i]; m[8 + i] = r[i]; }
    b3_iv(out);
    b3_compress_cv(out, m, 0, 0, 64, B3_PARENT | (root ? B3_ROOT : 0));
}

}  // namespace zs
