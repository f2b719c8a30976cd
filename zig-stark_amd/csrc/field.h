// field.h -- M31 / CM31 / QM31 and circle-group arithmetic for host and device.
//
// Semantics follow src/m31/field/m31.zig, cm31.zig, qm31.zig and
// src/m31/circle/point.zig: every value is canonical in [0, p), p = 2^31 - 1,
// so any exact formula for a field operation yields the same bits as the
// reference's.  The device forms use the unsigned-min reduction
// (r = min(r, r - p)), which keeps each add/sub/mul to one 32-bit compare-free
// select; there is no floating point and no matrix-core use anywhere.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ZS_HD __host__ __device__ __forceinline__

namespace zs {

constexpr uint32_t P = 0x7fffffffu;

// r in [0, 2p) -> [0, p)
ZS_HD uint32_t m31_reduce_once(uint32_t r) {
    uint32_t s = r - P;
    return s < r ? s : r;
}
ZS_HD uint32_t m31_add(uint32_t a, uint32_t b) { return m31_reduce_once(a + b); }
ZS_HD uint32_t m31_sub(uint32_t a, uint32_t b) {
    uint32_t d = a - b;
    uint32_t e = d + P;
    return e < d ? e : d;
}
ZS_HD uint32_t m31_neg(uint32_t a) { return a == 0 ? 0 : P - a; }

ZS_HD uint32_t m31_mul(uint32_t a, uint32_t b) {
    uint64_t prod = (uint64_t)a * (uint64_t)b;
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t hi = __builtin_amdgcn_alignbit((uint32_t)(prod >> 32), (uint32_t)prod, 31);
#else
    uint32_t hi = (uint32_t)(prod >> 31);
#endif
    return m31_reduce_once(((uint32_t)prod & P) + hi);
}

// a * b mod p with the second factor given doubled (w2 = 2b, b < 2^31): the
// 64-bit product 2ab has ab >> 31 as its high word and (ab mod 2^31) << 1 as
// its low word, so the Mersenne fold needs no bit-align or mask.
ZS_HD uint32_t m31_mul_dbl(uint32_t a, uint32_t w2) {
    const uint64_t prod = (uint64_t)a * (uint64_t)w2;
    return m31_reduce_once((uint32_t)(prod >> 32) + ((uint32_t)prod >> 1));
}

// x mod p for any 64-bit x (2^31 = 1 mod p)
ZS_HD uint32_t m31_reduce64(uint64_t x) {
    const uint64_t y = (x & P) + (x >> 31);                      // < 2^31 + 2^33
    const uint32_t z = (uint32_t)(y & P) + (uint32_t)(y >> 31);  // < 2^31 + 8
    return m31_reduce_once(z);
}

// x / 2 mod p without a multiply (p is odd)
ZS_HD uint32_t m31_half(uint32_t x) { return (x & 1u) ? (uint32_t)((x + P) >> 1) : (x >> 1); }

ZS_HD uint32_t m31_pow(uint32_t a, uint32_t e) {
    uint32_t r = 1;
    while (e) {
        if (e & 1u) r = m31_mul(r, a);
        a = m31_mul(a, a);
        e >>= 1;
    }
    return r;
}
ZS_HD uint32_t m31_inv(uint32_t a) { return m31_pow(a, P - 2); }
ZS_HD uint32_t m31_from_u32(uint32_t x) { return x >= P ? x % P : x; }

// ---- CM31 = M31[i]/(i^2 + 1) ----
struct cm31 {
    uint32_t c0, c1;
};
ZS_HD cm31 cm_add(cm31 a, cm31 b) { return {m31_add(a.c0, b.c0), m31_add(a.c1, b.c1)}; }
ZS_HD cm31 cm_sub(cm31 a, cm31 b) { return {m31_sub(a.c0, b.c0), m31_sub(a.c1, b.c1)}; }
ZS_HD cm31 cm_mul(cm31 a, cm31 b) {
    return {m31_sub(m31_mul(a.c0, b.c0), m31_mul(a.c1, b.c1)), m31_add(m31_mul(a.c0, b.c1), m31_mul(a.c1, b.c0))};
}
ZS_HD cm31 cm_mul_by_i(cm31 a) { return {m31_neg(a.c1), a.c0}; }
ZS_HD cm31 cm_neg(cm31 a) { return {m31_neg(a.c0), m31_neg(a.c1)}; }
ZS_HD cm31 cm_inv(cm31 a) {
    uint32_t ni = m31_inv(m31_add(m31_mul(a.c0, a.c0), m31_mul(a.c1, a.c1)));
    return {m31_mul(a.c0, ni), m31_mul(m31_neg(a.c1), ni)};
}

// ---- QM31 = CM31[u]/(u^2 + i); words a.c0, a.c1, b.c0, b.c1 ----
struct qm31 {
    cm31 a, b;
};
ZS_HD qm31 qm_add(qm31 x, qm31 y) { return {cm_add(x.a, y.a), cm_add(x.b, y.b)}; }
ZS_HD qm31 qm_sub(qm31 x, qm31 y) { return {cm_sub(x.a, y.a), cm_sub(x.b, y.b)}; }
// x = (a0 + a1 i) + (b0 + b1 i) u and y = (c0 + c1 i) + (d0 + d1 i) u with
// i^2 = -1, u^2 = -i (qm31.zig:25-33, expanded over M31):
//   r0 = a0 c0 - a1 c1 + b0 d1 + b1 d0      r1 = a0 c1 + a1 c0 - b0 d0 + b1 d1
//   r2 = a0 d0 - a1 d1 + b0 c0 - b1 c1      r3 = a0 d1 + a1 d0 + b0 c1 + b1 c0
// Each word is a sum of four products of canonical values (a subtracted one
// takes p - operand), below 2^64 in total, so it is accumulated in 64 bits and
// reduced once: 16 multiply-adds and 4 reductions instead of 12 separately
// reduced products and the CM31 sums around them.
ZS_HD qm31 qm_mul(qm31 x, qm31 y) {
    const uint64_t a0 = x.a.c0, a1 = x.a.c1, b0 = x.b.c0, b1 = x.b.c1;
    const uint64_t c0 = y.a.c0, c1 = y.a.c1, d0 = y.b.c0, d1 = y.b.c1;
    const uint64_t nc1 = P - c1, nd0 = P - d0, nd1 = P - d1;
    const uint32_t r0 = m31_reduce64(a0 * c0 + a1 * nc1 + b0 * d1 + b1 * d0);
    const uint32_t r1 = m31_reduce64(a0 * c1 + a1 * c0 + b0 * nd0 + b1 * d1);
    const uint32_t r2 = m31_reduce64(a0 * d0 + a1 * nd1 + b0 * c0 + b1 * nc1);
    const uint32_t r3 = m31_reduce64(a0 * d1 + a1 * d0 + b0 * c1 + b1 * c0);
    return {{r0, r1}, {r2, r3}};
}
ZS_HD qm31 qm_half(qm31 x) { return {{m31_half(x.a.c0), m31_half(x.a.c1)}, {m31_half(x.b.c0), m31_half(x.b.c1)}}; }
ZS_HD qm31 qm_inv(qm31 x) {
    cm31 norm = cm_add(cm_mul(x.a, x.a), cm_mul_by_i(cm_mul(x.b, x.b)));
    cm31 ni = cm_inv(norm);
    return {cm_mul(x.a, ni), cm_mul(cm_neg(x.b), ni)};
}
ZS_HD qm31 qm_pow(qm31 x, uint64_t e) {
    qm31 r = {{1, 0}, {0, 0}};
    while (e) {
        if (e & 1u) r = qm_mul(r, x);
        x = qm_mul(x, x);
        e >>= 1;
    }
    return r;
}
ZS_HD qm31 qm_from_words(const uint32_t *w) { return {{w[0], w[1]}, {w[2], w[3]}}; }
ZS_HD void qm_to_words(qm31 x, uint32_t *w) {
    w[0] = x.a.c0; w[1] = x.a.c1; w[2] = x.b.c0; w[3] = x.b.c1;
}
ZS_HD bool qm_eq(qm31 x, qm31 y) {
    return x.a.c0 == y.a.c0 && x.a.c1 == y.a.c1 && x.b.c0 == y.b.c0 && x.b.c1 == y.b.c1;
}

// ---- circle group: x^2 + y^2 = 1, (x1,y1)(x2,y2) = (x1x2 - y1y2, x1y2 + y1x2) ----
struct cpoint {
    uint32_t x, y;
};
ZS_HD cpoint cp_add(cpoint a, cpoint b) {
    return {m31_sub(m31_mul(a.x, b.x), m31_mul(a.y, b.y)), m31_add(m31_mul(a.x, b.y), m31_mul(a.y, b.x))};
}
ZS_HD cpoint cp_double(cpoint a) {
    return {m31_sub(m31_mul(a.x, a.x), m31_mul(a.y, a.y)), m31_add(m31_mul(a.x, a.y), m31_mul(a.x, a.y))};
}
ZS_HD cpoint cp_neg(cpoint a) { return {a.x, m31_neg(a.y)}; }
ZS_HD cpoint cp_mul(cpoint a, uint64_t s) {
    cpoint r = {1, 0};
    while (s) {
        if (s & 1u) r = cp_add(r, a);
        a = cp_double(a);
        s >>= 1;
    }
    return r;
}

ZS_HD uint32_t bitrev32(uint32_t x, uint32_t width) {
#if defined(__HIP_DEVICE_COMPILE__)
    return width ? (__builtin_bitreverse32(x) >> (32 - width)) : 0;
#else
    uint32_t r = 0;
    for (uint32_t i = 0; i < width; i++) { r = (r << 1) | (x & 1u); x >>= 1; }
    return r;
#endif
}

}  // namespace zs
