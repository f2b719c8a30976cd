/* zs_oracle.c -- CPU restatement of zig-stark's M31 trace-commitment path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Nothing in the product library links or calls this file.
 *
 * Each function restates the reference function cited above it (paths are
 * relative to the zig-stark source tree).  The restatement keeps the
 * reference's algorithm (per-call slow twiddles, scalar BLAKE3 per node,
 * O(n^2) Lagrange, one QM31 inversion per FRI fold) because it doubles as the
 * reference CPU path that the GPU figures are compared against.
 *
 * BLAKE3 is not in the reference tree: it is Zig std's std.crypto.hash.Blake3
 * (Zig 0.16.0, pinned by build.zig.zon:5).  zo_blake3 restates the published
 * BLAKE3 specification (compression function, chunk chaining, binary tree).
 */
#include "zs_oracle.h"

#include <stdlib.h>
#include <string.h>

#define P ZO_P

/* ======================================================================== */
/* M31: src/m31/field/m31.zig:9-100                                          */
/* ======================================================================== */

/* m31.zig:19-22 */
uint32_t zo_m31_from_u32(uint32_t x) { return x >= P ? x % P : x; }

/* m31.zig:28-35 */
uint32_t zo_m31_add(uint32_t a, uint32_t b) {
    uint64_t r = (uint64_t)a + (uint64_t)b;
    if (r >= P) r -= P;
    r = (r & P) + (r >> 31);
    if (r >= P) r -= P;
    return (uint32_t)r;
}

/* m31.zig:37-42 */
uint32_t zo_m31_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : P - (b - a); }

/* m31.zig:44-52 */
uint32_t zo_m31_mul(uint32_t a, uint32_t b) {
    uint64_t prod = (uint64_t)a * (uint64_t)b;
    uint64_t r = (prod & P) + (prod >> 31);
    if (r >= P) r -= P;
    if (r >= P) r -= P;
    return (uint32_t)r;
}

/* m31.zig:54-57 */
uint32_t zo_m31_neg(uint32_t a) { return a == 0 ? 0 : P - a; }

/* m31.zig:64-72 */
uint32_t zo_m31_pow(uint32_t a, uint32_t e) {
    uint32_t r = 1, b = a;
    while (e) {
        if (e & 1) r = zo_m31_mul(r, b);
        b = zo_m31_mul(b, b);
        e >>= 1;
    }
    return r;
}

/* m31.zig:59-62 */
uint32_t zo_m31_inv(uint32_t a) { return zo_m31_pow(a, P - 2); }

/* ======================================================================== */
/* CM31: src/m31/field/cm31.zig:4-84 (i^2 = -1)                              */
/* ======================================================================== */
typedef struct { uint32_t c0, c1; } cm31;

static cm31 cm_add(cm31 a, cm31 b) { return (cm31){zo_m31_add(a.c0, b.c0), zo_m31_add(a.c1, b.c1)}; }
static cm31 cm_sub(cm31 a, cm31 b) { return (cm31){zo_m31_sub(a.c0, b.c0), zo_m31_sub(a.c1, b.c1)}; }
/* cm31.zig:30-39 */
static cm31 cm_mul(cm31 a, cm31 b) {
    uint32_t ac = zo_m31_mul(a.c0, b.c0), bd = zo_m31_mul(a.c1, b.c1);
    uint32_t ad = zo_m31_mul(a.c0, b.c1), bc = zo_m31_mul(a.c1, b.c0);
    return (cm31){zo_m31_sub(ac, bd), zo_m31_add(ad, bc)};
}
static cm31 cm_neg(cm31 a) { return (cm31){zo_m31_neg(a.c0), zo_m31_neg(a.c1)}; }
/* cm31.zig:45-52 */
static cm31 cm_inv(cm31 a) {
    uint32_t norm = zo_m31_add(zo_m31_mul(a.c0, a.c0), zo_m31_mul(a.c1, a.c1));
    uint32_t ni = zo_m31_inv(norm);
    return (cm31){zo_m31_mul(a.c0, ni), zo_m31_mul(zo_m31_neg(a.c1), ni)};
}
/* cm31.zig:67-69 */
static cm31 cm_mul_by_i(cm31 a) { return (cm31){zo_m31_neg(a.c1), a.c0}; }

/* ======================================================================== */
/* QM31: src/m31/field/qm31.zig:5-101 (QM31 = CM31[u]/(u^2 + i))             */
/* ======================================================================== */
static cm31 qa(zo_qm31 x) { return (cm31){x.v[0], x.v[1]}; }
static cm31 qb(zo_qm31 x) { return (cm31){x.v[2], x.v[3]}; }
static zo_qm31 qm(cm31 a, cm31 b) { return (zo_qm31){{a.c0, a.c1, b.c0, b.c1}}; }

zo_qm31 zo_qm31_add(zo_qm31 x, zo_qm31 y) { return qm(cm_add(qa(x), qa(y)), cm_add(qb(x), qb(y))); }
zo_qm31 zo_qm31_sub(zo_qm31 x, zo_qm31 y) { return qm(cm_sub(qa(x), qa(y)), cm_sub(qb(x), qb(y))); }

/* qm31.zig:25-33 */
zo_qm31 zo_qm31_mul(zo_qm31 x, zo_qm31 y) {
    cm31 ab0 = cm_mul(qa(x), qa(y));
    cm31 ab1 = cm_mul(qb(x), qb(y));
    cm31 cross = cm_mul(cm_add(qa(x), qb(x)), cm_add(qa(y), qb(y)));
    return qm(cm_sub(ab0, cm_mul_by_i(ab1)), cm_sub(cm_sub(cross, ab0), ab1));
}

/* qm31.zig:39-48 */
zo_qm31 zo_qm31_inv(zo_qm31 x) {
    cm31 a2 = cm_mul(qa(x), qa(x));
    cm31 b2i = cm_mul_by_i(cm_mul(qb(x), qb(x)));
    cm31 ni = cm_inv(cm_add(a2, b2i));
    return qm(cm_mul(qa(x), ni), cm_mul(cm_neg(qb(x)), ni));
}

/* qm31.zig:57-67 */
zo_qm31 zo_qm31_pow(zo_qm31 x, uint64_t e) {
    zo_qm31 r = {{1, 0, 0, 0}}, b = x;
    while (e) {
        if (e & 1) r = zo_qm31_mul(r, b);
        b = zo_qm31_mul(b, b);
        e >>= 1;
    }
    return r;
}

/* qm31.zig:80-89: TWO_ADIC_ROOT from its LE bytes, primitiveRootOfUnity(k) = ROOT^(2^(32-k)) */
static const uint8_t TWO_ADIC_ROOT_BYTES[16] = {0x60, 0x1f, 0x3d, 0x34, 0x32, 0x39, 0xa0, 0x6a,
                                                0xa0, 0xd6, 0x51, 0x61, 0x91, 0xe7, 0x86, 0x3e};
static uint32_t rd32(const uint8_t *b) {
    return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
}
static void wr32(uint8_t *b, uint32_t v) {
    b[0] = (uint8_t)v; b[1] = (uint8_t)(v >> 8); b[2] = (uint8_t)(v >> 16); b[3] = (uint8_t)(v >> 24);
}
/* qm31.zig:95-101 fromBytes: each word reduced by M31.fromInt */
static zo_qm31 qm31_from_bytes(const uint8_t *b) {
    zo_qm31 r;
    for (int i = 0; i < 4; i++) r.v[i] = zo_m31_from_u32(rd32(b + 4 * i));
    return r;
}
static void qm31_to_bytes(zo_qm31 x, uint8_t *b) {
    for (int i = 0; i < 4; i++) wr32(b + 4 * i, x.v[i]);
}
zo_qm31 zo_qm31_root_of_unity(uint32_t log_size) {
    zo_qm31 root = qm31_from_bytes(TWO_ADIC_ROOT_BYTES);
    return zo_qm31_pow(root, (uint64_t)1 << (32 - log_size));
}

void zo_qm31_mul_p(const uint32_t *x, const uint32_t *y, uint32_t *out) {
    zo_qm31 a, b;
    memcpy(a.v, x, 16); memcpy(b.v, y, 16);
    zo_qm31 r = zo_qm31_mul(a, b);
    memcpy(out, r.v, 16);
}
void zo_qm31_inv_p(const uint32_t *x, uint32_t *out) {
    zo_qm31 a;
    memcpy(a.v, x, 16);
    zo_qm31 r = zo_qm31_inv(a);
    memcpy(out, r.v, 16);
}
void zo_qm31_pow_p(const uint32_t *x, uint64_t e, uint32_t *out) {
    zo_qm31 a;
    memcpy(a.v, x, 16);
    zo_qm31 r = zo_qm31_pow(a, e);
    memcpy(out, r.v, 16);
}
void zo_qm31_root_of_unity_p(uint32_t log_size, uint32_t *out) {
    zo_qm31 r = zo_qm31_root_of_unity(log_size);
    memcpy(out, r.v, 16);
}

/* ======================================================================== */
/* Circle group: src/m31/circle/point.zig:8-87, coset.zig:10-66             */
/* ======================================================================== */

/* point.zig:16-21 */
zo_point zo_point_add(zo_point a, zo_point b) {
    return (zo_point){zo_m31_sub(zo_m31_mul(a.x, b.x), zo_m31_mul(a.y, b.y)),
                      zo_m31_add(zo_m31_mul(a.x, b.y), zo_m31_mul(a.y, b.x))};
}
/* point.zig:23-28 */
zo_point zo_point_double(zo_point a) {
    return (zo_point){zo_m31_sub(zo_m31_mul(a.x, a.x), zo_m31_mul(a.y, a.y)),
                      zo_m31_mul(zo_m31_mul(a.x, a.y), 2)};
}
/* point.zig:30-32 */
zo_point zo_point_neg(zo_point a) { return (zo_point){a.x, zo_m31_neg(a.y)}; }
/* point.zig:49-51 */
zo_point zo_point_generator(void) { return (zo_point){1268011823u, 2u}; }

/* point.zig:64-86: fixed 4-bit window over all 64 bits of the scalar */
zo_point zo_point_mul_scalar(zo_point a, uint64_t s) {
    zo_point table[16];
    table[0] = (zo_point){1, 0};
    table[1] = a;
    table[2] = zo_point_add(a, a);
    for (int i = 3; i < 16; i++) table[i] = zo_point_add(table[i - 1], a);
    zo_point r = (zo_point){1, 0};
    for (int d = 15; d >= 0; d--) {
        r = zo_point_double(zo_point_double(zo_point_double(zo_point_double(r))));
        r = zo_point_add(r, table[(s >> (4 * d)) & 0xf]);
    }
    return r;
}

/* coset.zig:55-65 */
zo_coset zo_coset_canonic_half(uint32_t log_size) {
    zo_point g = zo_point_neg(zo_point_generator());
    zo_coset c;
    c.offset = zo_point_mul_scalar(g, (uint64_t)1 << (30 - log_size));
    c.step = zo_point_mul_scalar(g, (uint64_t)1 << (32 - log_size));
    c.log_size = log_size - 1;
    return c;
}
/* coset.zig:23-26 */
zo_point zo_coset_at(zo_coset c, size_t index) {
    return zo_point_add(c.offset, zo_point_mul_scalar(c.step, (uint64_t)index));
}
/* coset.zig:38-41 */
zo_coset zo_coset_half(zo_coset c) {
    return (zo_coset){zo_point_double(c.offset), zo_point_double(c.step), c.log_size - 1};
}
/* coset.zig:44-46 */
zo_coset zo_coset_conjugate(zo_coset c) {
    return (zo_coset){zo_point_neg(c.offset), zo_point_neg(c.step), c.log_size};
}

static zo_coset coset_from5(const uint32_t *c5) {
    return (zo_coset){{c5[0], c5[1]}, {c5[2], c5[3]}, c5[4]};
}
void zo_coset_canonic_half_p(uint32_t log_size, uint32_t *out5) {
    zo_coset c = zo_coset_canonic_half(log_size);
    out5[0] = c.offset.x; out5[1] = c.offset.y; out5[2] = c.step.x; out5[3] = c.step.y;
    out5[4] = c.log_size;
}
void zo_coset_at_p(const uint32_t *coset5, size_t index, uint32_t *out2) {
    zo_point p = zo_coset_at(coset_from5(coset5), index);
    out2[0] = p.x; out2[1] = p.y;
}
/* circle.zig:186-190: index into half_coset ++ conjugate */
static zo_point domain_point(zo_coset hc, size_t i) {
    size_t half = (size_t)1 << hc.log_size;
    if (i < half) return zo_coset_at(hc, i);
    return zo_coset_at(zo_coset_conjugate(hc), i - half);
}
void zo_domain_point_p(const uint32_t *coset5, size_t index, uint32_t *out2) {
    zo_point p = domain_point(coset_from5(coset5), index);
    out2[0] = p.x; out2[1] = p.y;
}
void zo_point_mul_scalar_p(const uint32_t *p2, uint64_t s, uint32_t *out2) {
    zo_point p = zo_point_mul_scalar((zo_point){p2[0], p2[1]}, s);
    out2[0] = p.x; out2[1] = p.y;
}

/* ======================================================================== */
/* Circle FFT: src/m31/ntt/circle.zig                                         */
/* ======================================================================== */

/* circle.zig:27-31 */
static inline void butterfly(uint32_t *v0, uint32_t *v1, uint32_t t) {
    uint32_t tmp = zo_m31_mul(*v1, t);
    *v1 = zo_m31_sub(*v0, tmp);
    *v0 = zo_m31_add(*v0, tmp);
}
/* circle.zig:33-37 */
static inline void ibutterfly(uint32_t *v0, uint32_t *v1, uint32_t t) {
    uint32_t tmp = *v0;
    *v0 = zo_m31_add(tmp, *v1);
    *v1 = zo_m31_mul(zo_m31_sub(tmp, *v1), t);
}
/* circle.zig:39-55 */
static void fft_layer_loop(uint32_t *values, unsigned i, size_t h, uint32_t t, int inverse) {
    size_t step = (size_t)1 << i;
    size_t base = h << (i + 1);
    for (size_t l = 0; l < step; l++) {
        if (inverse) ibutterfly(&values[base + l], &values[base + l + step], t);
        else butterfly(&values[base + l], &values[base + l + step], t);
    }
}
/* circle.zig:149-158 */
static size_t bit_reverse_index(size_t k, unsigned width) {
    size_t r = 0;
    for (unsigned i = 0; i < width; i++) { r = (r << 1) | (k & 1); k >>= 1; }
    return r;
}
static unsigned log2_exact(size_t n) {
    unsigned l = 0;
    while (((size_t)1 << l) < n) l++;
    return l;
}
/* circle.zig:161-167 */
static void bit_reverse_slice(uint32_t *v, size_t n) {
    unsigned w = log2_exact(n);
    for (size_t i = 0; i < n; i++) {
        size_t j = bit_reverse_index(i, w);
        if (i < j) { uint32_t t = v[i]; v[i] = v[j]; v[j] = t; }
    }
}

/* circle.zig:65-82 (slowPrecomputeTwiddles): level k holds the bit-reversed
 * x-coordinates of the first 2^(k-1) points of the k-times-halved coset, each
 * point computed with CircleCoset.at (one mulScalar per twiddle). */
void zo_precompute_twiddles(const uint32_t *coset5, uint32_t *out) {
    zo_coset cur = coset_from5(coset5);
    size_t pos = 0;
    for (uint32_t k = cur.log_size; k > 0; k--) {
        size_t m = (size_t)1 << (k - 1);
        for (size_t i = 0; i < m; i++) out[pos + i] = zo_coset_at(cur, i).x;
        bit_reverse_slice(out + pos, m);
        pos += m;
        cur = zo_coset_half(cur);
    }
    out[pos] = 1;
}
/* circle.zig:86-91 */
static const uint32_t *line_twiddle_slice(const uint32_t *tw, size_t tw_len, unsigned j, unsigned lg,
                                          size_t *len_out) {
    size_t len = (size_t)1 << (lg - 2 - j);
    *len_out = len;
    return tw + (tw_len - 2 * len);
}
/* circle.zig:96-106 */
static void circle_twiddles(const uint32_t *first, size_t first_len, uint32_t *out) {
    for (size_t i = 0; i < first_len / 2; i++) {
        uint32_t x = first[2 * i], y = first[2 * i + 1];
        out[4 * i] = y;
        out[4 * i + 1] = zo_m31_neg(y);
        out[4 * i + 2] = zo_m31_neg(x);
        out[4 * i + 3] = x;
    }
}

/* circle.zig:195-246 */
int zo_circle_fft(const uint32_t *coeffs, size_t n, const uint32_t *coset5, uint32_t *evals) {
    zo_coset hc = coset_from5(coset5);
    if (n < 2 || (n & (n - 1)) || ((size_t)1 << hc.log_size) != n / 2) return -2;
    size_t half = n / 2;
    unsigned lg = log2_exact(n);
    /* interleave, circle.zig:130-137 */
    for (size_t i = 0; i < half; i++) { evals[2 * i] = coeffs[i]; evals[2 * i + 1] = coeffs[half + i]; }
    if (lg == 1) {
        uint32_t y = zo_coset_at(hc, 0).y;
        butterfly(&evals[0], &evals[1], y);
    } else if (lg == 2) {
        zo_point p = zo_coset_at(hc, 0);
        butterfly(&evals[0], &evals[2], p.x);
        butterfly(&evals[1], &evals[3], p.x);
        butterfly(&evals[0], &evals[1], p.y);
        butterfly(&evals[2], &evals[3], zo_m31_neg(p.y));
    } else {
        uint32_t *tw = (uint32_t *)malloc(half * sizeof(uint32_t));
        uint32_t *ctw = (uint32_t *)malloc(half * sizeof(uint32_t));
        if (!tw || !ctw) { free(tw); free(ctw); return -3; }
        zo_precompute_twiddles(coset5, tw);
        for (unsigned jj = 0; jj < lg - 1; jj++) {
            unsigned j = (lg - 2) - jj;
            size_t len;
            const uint32_t *s = line_twiddle_slice(tw, half, j, lg, &len);
            for (size_t h = 0; h < len; h++) fft_layer_loop(evals, j + 1, h, s[h], 0);
        }
        size_t len0;
        const uint32_t *s0 = line_twiddle_slice(tw, half, 0, lg, &len0);
        circle_twiddles(s0, len0, ctw);
        for (size_t h = 0; h < half; h++) fft_layer_loop(evals, 0, h, ctw[h], 0);
        free(tw);
        free(ctw);
    }
    bit_reverse_slice(evals, n);
    return 0;
}

/* circle.zig:272-346 */
int zo_circle_ifft(const uint32_t *evals, size_t n, const uint32_t *coset5, uint32_t *coeffs) {
    zo_coset hc = coset_from5(coset5);
    if (n < 2 || (n & (n - 1)) || ((size_t)1 << hc.log_size) != n / 2) return -2;
    size_t half = n / 2;
    unsigned lg = log2_exact(n);
    uint32_t *v = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!v) return -3;
    memcpy(v, evals, n * sizeof(uint32_t));
    bit_reverse_slice(v, n);
    if (lg == 1) {
        uint32_t y = zo_coset_at(hc, 0).y;
        uint32_t yn_inv = zo_m31_inv(zo_m31_mul(y, 2));
        uint32_t y_inv = zo_m31_mul(yn_inv, 2);
        uint32_t n_inv = zo_m31_mul(yn_inv, y);
        ibutterfly(&v[0], &v[1], y_inv);
        v[0] = zo_m31_mul(v[0], n_inv);
        v[1] = zo_m31_mul(v[1], n_inv);
    } else if (lg == 2) {
        zo_point p = zo_coset_at(hc, 0);
        uint32_t xyn_inv = zo_m31_inv(zo_m31_mul(zo_m31_mul(p.x, p.y), 4));
        uint32_t x_inv = zo_m31_mul(zo_m31_mul(xyn_inv, p.y), 4);
        uint32_t y_inv = zo_m31_mul(zo_m31_mul(xyn_inv, p.x), 4);
        uint32_t n_inv = zo_m31_mul(zo_m31_mul(xyn_inv, p.x), p.y);
        ibutterfly(&v[0], &v[1], y_inv);
        ibutterfly(&v[2], &v[3], zo_m31_neg(y_inv));
        ibutterfly(&v[0], &v[2], x_inv);
        ibutterfly(&v[1], &v[3], x_inv);
        for (int i = 0; i < 4; i++) v[i] = zo_m31_mul(v[i], n_inv);
    } else {
        uint32_t *tw = (uint32_t *)malloc(half * sizeof(uint32_t));
        uint32_t *ctw = (uint32_t *)malloc(half * sizeof(uint32_t));
        if (!tw || !ctw) { free(tw); free(ctw); free(v); return -3; }
        zo_precompute_twiddles(coset5, tw);
        for (size_t i = 0; i < half; i++) tw[i] = zo_m31_inv(tw[i]);
        size_t len0;
        const uint32_t *s0 = line_twiddle_slice(tw, half, 0, lg, &len0);
        circle_twiddles(s0, len0, ctw);
        for (size_t h = 0; h < half; h++) fft_layer_loop(v, 0, h, ctw[h], 1);
        for (unsigned j = 0; j < lg - 1; j++) {
            size_t len;
            const uint32_t *s = line_twiddle_slice(tw, half, j, lg, &len);
            for (size_t h = 0; h < len; h++) fft_layer_loop(v, j + 1, h, s[h], 1);
        }
        uint32_t inv_n = zo_m31_inv(zo_m31_from_u32((uint32_t)n));
        for (size_t i = 0; i < n; i++) v[i] = zo_m31_mul(v[i], inv_n);
        free(tw);
        free(ctw);
    }
    /* deinterleave, circle.zig:140-147 */
    for (size_t i = 0; i < half; i++) { coeffs[i] = v[2 * i]; coeffs[half + i] = v[2 * i + 1]; }
    free(v);
    return 0;
}

/* circle.zig:108-110 */
static uint32_t pi_map(uint32_t x) { return zo_m31_sub(zo_m31_mul(zo_m31_mul(x, x), 2), 1); }

/* circle.zig:170-173 */
static uint32_t interleaved_at(const uint32_t *c, size_t half, size_t k) {
    return (k % 2 == 0) ? c[k / 2] : c[half + (k - 1) / 2];
}
/* circle.zig:177-183 */
static uint32_t fold_range(const uint32_t *c, size_t half, size_t start, size_t len, const uint32_t *facs) {
    if (len == 1) return interleaved_at(c, half, start);
    size_t h = len / 2;
    uint32_t lhs = fold_range(c, half, start, h, facs + 1);
    uint32_t rhs = fold_range(c, half, start + h, h, facs + 1);
    return zo_m31_add(lhs, zo_m31_mul(rhs, facs[0]));
}
/* circle.zig:250-258 with getFoldingAlphas (circle.zig:114-127) */
uint32_t zo_eval_at_point(const uint32_t *coeffs, size_t n, uint32_t px, uint32_t py) {
    if (n == 1) return coeffs[0];
    unsigned lg = log2_exact(n);
    uint32_t alphas[64];
    alphas[lg - 1] = py;
    uint32_t x = px;
    for (int i = (int)lg - 2; i >= 0; i--) { alphas[i] = x; x = pi_map(x); }
    return fold_range(coeffs, n / 2, 0, n, alphas);
}

/* LDE composition (no single reference function): circleIFFT on
 * canonicHalf(log_n), zero-pad A and B to 2^(log_n+log_blowup-1) each, then
 * circleFFT on canonicHalf(log_n + log_blowup). */
int zo_lde_padded_coeffs(const uint32_t *col, uint32_t log_n, uint32_t log_blowup, uint32_t *out) {
    if (log_n < 1 || log_n + log_blowup > 28) return -2;
    size_t n = (size_t)1 << log_n, h = n / 2;
    size_t N = (size_t)1 << (log_n + log_blowup), H = N / 2;
    uint32_t c5[5];
    zo_coset_canonic_half_p(log_n, c5);
    uint32_t *c = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!c) return -3;
    int rc = zo_circle_ifft(col, n, c5, c);
    if (rc) { free(c); return rc; }
    memset(out, 0, N * sizeof(uint32_t));
    memcpy(out, c, h * sizeof(uint32_t));
    memcpy(out + H, c + h, h * sizeof(uint32_t));
    free(c);
    return 0;
}
int zo_lde(const uint32_t *col, uint32_t log_n, uint32_t log_blowup, uint32_t *out) {
    size_t N = (size_t)1 << (log_n + log_blowup);
    uint32_t *pad = (uint32_t *)malloc(N * sizeof(uint32_t));
    if (!pad) return -3;
    int rc = zo_lde_padded_coeffs(col, log_n, log_blowup, pad);
    if (!rc) {
        uint32_t c5[5];
        zo_coset_canonic_half_p(log_n + log_blowup, c5);
        rc = zo_circle_fft(pad, N, c5, out);
    }
    free(pad);
    return rc;
}

/* ======================================================================== */
/* BLAKE3 (published specification; Zig std.crypto.hash.Blake3)             */
/* ======================================================================== */
static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

static inline uint32_t rotr32(uint32_t x, unsigned n) { return (x >> n) | (x << (32 - n)); }

/* the quarter-round shared by BLAKE2s and BLAKE3 */
static inline void g_mix(uint32_t *s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx;
    s[d] = rotr32(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr32(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my;
    s[d] = rotr32(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr32(s[b] ^ s[c], 7);
}
/* one round: columns then diagonals, message words taken through `sched` */
static void round_mix(uint32_t *s, const uint32_t *m, const uint8_t *sched) {
    g_mix(s, 0, 4, 8, 12, m[sched[0]], m[sched[1]]);
    g_mix(s, 1, 5, 9, 13, m[sched[2]], m[sched[3]]);
    g_mix(s, 2, 6, 10, 14, m[sched[4]], m[sched[5]]);
    g_mix(s, 3, 7, 11, 15, m[sched[6]], m[sched[7]]);
    g_mix(s, 0, 5, 10, 15, m[sched[8]], m[sched[9]]);
    g_mix(s, 1, 6, 11, 12, m[sched[10]], m[sched[11]]);
    g_mix(s, 2, 7, 8, 13, m[sched[12]], m[sched[13]]);
    g_mix(s, 3, 4, 9, 14, m[sched[14]], m[sched[15]]);
}
static const uint8_t IDENT16[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};

void zo_blake3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out16[16]) {
    uint32_t s[16], m[16], t[16];
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    for (int i = 0; i < 4; i++) s[8 + i] = B3_IV[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    memcpy(m, block, 64);
    for (int r = 0; r < 7; r++) {
        round_mix(s, m, IDENT16);
        if (r < 6) {
            for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]];
            memcpy(m, t, 64);
        }
    }
    for (int i = 0; i < 8; i++) {
        out16[i] = s[i] ^ s[i + 8];
        out16[i + 8] = s[i + 8] ^ cv[i];
    }
}

static void load_block(const uint8_t *p, size_t len, uint32_t m[16]) {
    uint8_t buf[64];
    memset(buf, 0, 64);
    memcpy(buf, p, len);
    for (int i = 0; i < 16; i++) m[i] = rd32(buf + 4 * i);
}

/* chaining value (or root output) of one chunk of <= 1024 bytes */
static void chunk_cv(const uint8_t *in, size_t len, uint64_t chunk_counter, int is_root, uint32_t cv_out[8]) {
    uint32_t cv[8], m[16], out[16];
    memcpy(cv, B3_IV, 32);
    size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    for (size_t b = 0; b < nblocks; b++) {
        size_t off = b * 64;
        size_t blen = len - off < 64 ? len - off : 64;
        load_block(in + off, blen, m);
        uint32_t flags = 0;
        if (b == 0) flags |= CHUNK_START;
        if (b == nblocks - 1) {
            flags |= CHUNK_END;
            if (is_root) flags |= ROOT;
        }
        zo_blake3_compress(cv, m, chunk_counter, (uint32_t)blen, flags, out);
        memcpy(cv, out, 32);
    }
    memcpy(cv_out, cv, 32);
}

/* binary tree: the left subtree takes the largest power-of-two number of
 * whole chunks that leaves at least one byte for the right subtree */
static void subtree_cv(const uint8_t *in, size_t len, uint64_t chunk_counter, int is_root, uint32_t cv_out[8]) {
    if (len <= 1024) { chunk_cv(in, len, chunk_counter, is_root, cv_out); return; }
    size_t full_chunks = (len - 1) / 1024;
    size_t left_chunks = 1;
    while (left_chunks * 2 <= full_chunks) left_chunks *= 2;
    size_t left_len = left_chunks * 1024;
    uint32_t block[16], out[16];
    subtree_cv(in, left_len, chunk_counter, 0, block);
    subtree_cv(in + left_len, len - left_len, chunk_counter + left_chunks, 0, block + 8);
    zo_blake3_compress(B3_IV, block, 0, 64, PARENT | (is_root ? ROOT : 0), out);
    memcpy(cv_out, out, 32);
}

void zo_blake3(const uint8_t *msg, size_t len, uint8_t out[32]) {
    uint32_t cv[8];
    subtree_cv(msg, len, 0, 1, cv);
    for (int i = 0; i < 8; i++) wr32(out + 4 * i, cv[i]);
}

/* BLAKE2s-256 (RFC 7693) assembled from g_mix/round_mix above; used only to
 * pin the shared quarter-round and round layout against an independent
 * implementation (Python's hashlib.blake2s). */
static const uint8_t B2S_SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
    {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4},
    {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13},
    {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11},
    {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5},
    {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
};
static void b2s_compress(uint32_t h[8], const uint32_t m[16], uint64_t t, int last) {
    uint32_t v[16];
    for (int i = 0; i < 8; i++) { v[i] = h[i]; v[i + 8] = B3_IV[i]; }
    v[12] ^= (uint32_t)t;
    v[13] ^= (uint32_t)(t >> 32);
    if (last) v[14] = ~v[14];
    for (int r = 0; r < 10; r++) round_mix(v, m, B2S_SIGMA[r]);
    for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}
void zo_blake2s_from_g(const uint8_t *msg, size_t len, uint8_t out[32]) {
    uint32_t h[8], m[16];
    memcpy(h, B3_IV, 32);
    h[0] ^= 0x01010000u ^ 32u;
    size_t off = 0;
    while (len - off > 64) {
        load_block(msg + off, 64, m);
        off += 64;
        b2s_compress(h, m, off, 0);
    }
    load_block(msg + off, len - off, m);
    b2s_compress(h, m, len, 1);
    for (int i = 0; i < 8; i++) wr32(out + 4 * i, h[i]);
}

/* src/core/hash/hash.zig:17-25: BLAKE3("zig-stark:pair" || a || b) */
void zo_hash2(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]) {
    uint8_t msg[78];
    memcpy(msg, "zig-stark:pair", 14);
    memcpy(msg + 14, a, 32);
    memcpy(msg + 46, b, 32);
    zo_blake3(msg, 78, out);
}
/* src/m31/hash.zig:15-25: BLAKE3 over the u32 LE bytes of each value */
void zo_hash_m31s(const uint32_t *vals, size_t n, uint8_t out[32]) {
    uint8_t stackbuf[1024];
    uint8_t *buf = n * 4 <= sizeof stackbuf ? stackbuf : (uint8_t *)malloc(n * 4);
    for (size_t i = 0; i < n; i++) wr32(buf + 4 * i, vals[i]);
    zo_blake3(buf, n * 4, out);
    if (buf != stackbuf) free(buf);
}
/* src/m31/hash.zig:33-37 */
void zo_hash_qm31(const uint32_t v[4], uint8_t out[32]) {
    uint8_t buf[16];
    for (int i = 0; i < 4; i++) wr32(buf + 4 * i, v[i]);
    zo_blake3(buf, 16, out);
}

/* ======================================================================== */
/* Merkle tree: src/core/merkle/merkle.zig:17-89                            */
/* ======================================================================== */

/* merkle.zig:17-45 */
int zo_merkle_build(const uint8_t *leaves, size_t n, uint8_t *nodes) {
    if (n == 0 || (n & (n - 1))) return -2;
    memcpy(nodes, leaves, n * 32);
    uint8_t *lvl = nodes;
    size_t len = n;
    while (len > 1) {
        uint8_t *next = lvl + len * 32;
        for (size_t i = 0; i < len / 2; i++) zo_hash2(lvl + 64 * i, lvl + 64 * i + 32, next + 32 * i);
        lvl = next;
        len /= 2;
    }
    return 0;
}
/* merkle.zig:57-68: siblings bottom-up */
int zo_merkle_open(const uint8_t *nodes, size_t n, size_t index, uint8_t *path) {
    if (n == 0 || (n & (n - 1)) || index >= n) return -2;
    const uint8_t *lvl = nodes;
    size_t len = n, idx = index, level = 0;
    while (len > 1) {
        memcpy(path + 32 * level, lvl + 32 * (idx ^ 1), 32);
        lvl += len * 32;
        len /= 2;
        idx >>= 1;
        level++;
    }
    return 0;
}
/* merkle.zig:72-89 */
int zo_merkle_verify(const uint8_t root[32], size_t index, const uint8_t leaf[32], const uint8_t *path,
                     size_t path_len) {
    uint8_t d[32], t[32];
    memcpy(d, leaf, 32);
    size_t idx = index;
    for (size_t l = 0; l < path_len; l++) {
        if ((idx & 1) == 0) zo_hash2(d, path + 32 * l, t);
        else zo_hash2(path + 32 * l, d, t);
        memcpy(d, t, 32);
        idx >>= 1;
    }
    return memcmp(d, root, 32) == 0;
}

/* row leaves: hashM31s over (col_0[i], ..., col_{c-1}[i]), then MerkleTree.init */
int zo_row_commit(const uint32_t *cols, size_t n_cols, size_t n_rows, uint8_t *nodes) {
    if (n_rows == 0 || (n_rows & (n_rows - 1)) || n_cols == 0) return -2;
    uint32_t *row = (uint32_t *)malloc(n_cols * sizeof(uint32_t));
    uint8_t *leaves = (uint8_t *)malloc(n_rows * 32);
    if (!row || !leaves) { free(row); free(leaves); return -3; }
    for (size_t i = 0; i < n_rows; i++) {
        for (size_t c = 0; c < n_cols; c++) row[c] = cols[c * n_rows + i];
        zo_hash_m31s(row, n_cols, leaves + 32 * i);
    }
    int rc = zo_merkle_build(leaves, n_rows, nodes);
    free(row);
    free(leaves);
    return rc;
}

/* ======================================================================== */
/* Channel: src/core/channel/channel.zig:11-101                              */
/* The hasher state is kept as the full transcript; BLAKE3 of the transcript */
/* equals the incremental hasher's output.                                   */
/* ======================================================================== */
struct zo_channel {
    uint8_t *buf;
    size_t len, cap;
};
static void ch_append(zo_channel *c, const uint8_t *d, size_t n) {
    if (c->len + n > c->cap) {
        size_t cap = c->cap ? c->cap : 256;
        while (cap < c->len + n) cap *= 2;
        c->buf = (uint8_t *)realloc(c->buf, cap);
        c->cap = cap;
    }
    memcpy(c->buf + c->len, d, n);
    c->len += n;
}
/* channel.zig:16-21 */
zo_channel *zo_channel_new(const uint8_t *sep, size_t sep_len) {
    zo_channel *c = (zo_channel *)calloc(1, sizeof *c);
    ch_append(c, (const uint8_t *)"zig-stark:channel", 17);
    if (sep_len) ch_append(c, sep, sep_len);
    return c;
}
zo_channel *zo_channel_clone(const zo_channel *c) {
    zo_channel *d = (zo_channel *)calloc(1, sizeof *d);
    if (c->len) ch_append(d, c->buf, c->len);
    return d;
}
void zo_channel_free(zo_channel *c) {
    if (!c) return;
    free(c->buf);
    free(c);
}
/* channel.zig:23-43 */
void zo_channel_absorb(zo_channel *c, const uint8_t *data, size_t len) { ch_append(c, data, len); }
size_t zo_channel_transcript_len(const zo_channel *c) { return c->len; }
/* channel.zig:49-69 */
void zo_channel_sample_bytes(zo_channel *c, uint8_t *out, size_t len) {
    uint64_t counter = 0;
    while (len > 0) {
        size_t base = c->len;
        uint8_t cbuf[8];
        for (int i = 0; i < 8; i++) cbuf[i] = (uint8_t)(counter >> (8 * i));
        ch_append(c, (const uint8_t *)"zig-stark:sample", 16);
        ch_append(c, cbuf, 8);
        uint8_t block[32];
        zo_blake3(c->buf, c->len, block);
        c->len = base; /* drop the snapshot's suffix */
        ch_append(c, block, 32);
        size_t k = len < 32 ? len : 32;
        memcpy(out, block, k);
        out += k;
        len -= k;
        counter++;
    }
}
/* channel.zig:72-76 with QM31.fromBytes */
void zo_channel_sample_qm31(zo_channel *c, uint32_t out[4]) {
    uint8_t b[16];
    zo_channel_sample_bytes(c, b, 16);
    zo_qm31 q = qm31_from_bytes(b);
    memcpy(out, q.v, 16);
}
/* channel.zig:79-101 */
size_t zo_channel_sample_index(zo_channel *c, size_t n) {
    unsigned bits = 0;
    while (((size_t)1 << (bits + 1)) <= n && bits < 63) bits++;
    bits += 1; /* floor(log2 n) + 1 */
    size_t bytes = (bits + 7) / 8;
    uint8_t buf[8];
    size_t mask = bits >= 64 ? ~(size_t)0 : (((size_t)1 << bits) - 1);
    for (;;) {
        zo_channel_sample_bytes(c, buf, bytes);
        size_t v = 0;
        for (size_t i = 0; i < bytes; i++) v = (v << 8) | buf[i];
        v &= mask;
        if (v < n) return v;
    }
}

/* ======================================================================== */
/* Univariate QM31 polynomials: src/m31/poly/univariate.zig                   */
/* ======================================================================== */
static zo_qm31 ld4(const uint32_t *p) { zo_qm31 r; memcpy(r.v, p, 16); return r; }
static void st4(uint32_t *p, zo_qm31 x) { memcpy(p, x.v, 16); }
static const zo_qm31 Q_ZERO = {{0, 0, 0, 0}}, Q_ONE = {{1, 0, 0, 0}};
static zo_qm31 q_neg(zo_qm31 x) { return zo_qm31_sub(Q_ZERO, x); }

/* univariate.zig:11-19 (Horner) */
void zo_poly_eval(const uint32_t *coeffs, size_t n, const uint32_t *x, uint32_t *out) {
    zo_qm31 r = Q_ZERO, xx = ld4(x);
    for (size_t i = n; i > 0; i--) r = zo_qm31_add(zo_qm31_mul(r, xx), ld4(coeffs + 4 * (i - 1)));
    st4(out, r);
}
/* univariate.zig:79-139 (O(n^2) Lagrange) */
int zo_poly_interpolate(const uint32_t *xs, const uint32_t *ys, size_t n, uint32_t *coeffs) {
    zo_qm31 *p = (zo_qm31 *)calloc(n + 1, sizeof(zo_qm31));
    zo_qm31 *w = (zo_qm31 *)calloc(n, sizeof(zo_qm31));
    zo_qm31 *acc = (zo_qm31 *)calloc(n, sizeof(zo_qm31));
    zo_qm31 *tmp = (zo_qm31 *)calloc(n, sizeof(zo_qm31));
    if (!p || !w || !acc || !tmp) { free(p); free(w); free(acc); free(tmp); return -3; }
    p[0] = Q_ONE;
    size_t plen = 1;
    for (size_t i = 0; i < n; i++) {
        zo_qm31 xi = ld4(xs + 4 * i);
        for (size_t j = plen; j > 0; j--) p[j] = zo_qm31_sub(p[j - 1], zo_qm31_mul(p[j], xi));
        p[0] = zo_qm31_mul(q_neg(p[0]), xi);
        plen++;
    }
    for (size_t i = 0; i < n; i++) {
        zo_qm31 xi = ld4(xs + 4 * i), wi = Q_ONE;
        for (size_t j = 0; j < n; j++)
            if (i != j) wi = zo_qm31_mul(wi, zo_qm31_sub(xi, ld4(xs + 4 * j)));
        w[i] = wi;
    }
    for (size_t i = 0; i < n; i++) {
        zo_qm31 xi = ld4(xs + 4 * i);
        zo_qm31 sc = zo_qm31_mul(ld4(ys + 4 * i), zo_qm31_inv(w[i]));
        tmp[n - 1] = p[n];
        for (size_t j = n - 1; j > 0; j--) tmp[j - 1] = zo_qm31_add(p[j], zo_qm31_mul(xi, tmp[j]));
        for (size_t k = 0; k < n; k++) acc[k] = zo_qm31_add(acc[k], zo_qm31_mul(tmp[k], sc));
    }
    for (size_t k = 0; k < n; k++) st4(coeffs + 4 * k, acc[k]);
    free(p); free(w); free(acc); free(tmp);
    return 0;
}

/* ======================================================================== */
/* FRI: src/m31/fri.zig                                                        */
/* ======================================================================== */
/* fri.zig:36: FRI_OFFSET = (1 + 0i) + (1 + 0i) u */
const uint32_t ZO_FRI_OFFSET[4] = {1, 0, 1, 0};

/* fri.zig:96-101 */
void zo_fri_domain_point(uint32_t dlog, uint32_t layer, size_t j, uint32_t out[4]) {
    zo_qm31 w = zo_qm31_root_of_unity(dlog - layer);
    zo_qm31 off = zo_qm31_pow(ld4(ZO_FRI_OFFSET), (uint64_t)1 << layer);
    st4(out, zo_qm31_mul(off, zo_qm31_pow(w, (uint64_t)j)));
}
/* fri.zig:104-117 */
static zo_qm31 *domain_points(uint32_t dlog, uint32_t layer) {
    size_t n = (size_t)1 << (dlog - layer);
    zo_qm31 *pts = (zo_qm31 *)malloc(n * sizeof(zo_qm31));
    zo_qm31 w = zo_qm31_root_of_unity(dlog - layer);
    zo_qm31 acc = zo_qm31_pow(ld4(ZO_FRI_OFFSET), (uint64_t)1 << layer);
    for (size_t j = 0; j < n; j++) { pts[j] = acc; acc = zo_qm31_mul(acc, w); }
    return pts;
}
/* fri.zig:119-121 */
static zo_qm31 inv_two(void) { zo_qm31 r = {{zo_m31_inv(2), 0, 0, 0}}; return r; }

/* fri.zig:211-218 */
static void fold_layer(const zo_qm31 *cur, size_t n, const zo_qm31 *dom, zo_qm31 alpha, zo_qm31 *next) {
    zo_qm31 i2 = inv_two();
    size_t nn = n / 2;
    for (size_t j = 0; j < nn; j++) {
        zo_qm31 y0 = cur[j], y1 = cur[j + nn];
        zo_qm31 hs = zo_qm31_mul(zo_qm31_add(y0, y1), i2);
        zo_qm31 hd = zo_qm31_mul(zo_qm31_mul(zo_qm31_sub(y0, y1), zo_qm31_inv(dom[j])), i2);
        next[j] = zo_qm31_add(hs, zo_qm31_mul(alpha, hd));
    }
}
int zo_fri_fold(const uint32_t *layer_vals, size_t n, uint32_t dlog, uint32_t layer, const uint32_t alpha[4],
                uint32_t *next) {
    if (n != ((size_t)1 << (dlog - layer)) || n < 2) return -2;
    zo_qm31 *dom = domain_points(dlog, layer);
    fold_layer((const zo_qm31 *)layer_vals, n, dom, ld4(alpha), (zo_qm31 *)next);
    free(dom);
    return 0;
}

typedef struct {
    uint32_t L;
    zo_qm31 **layers;  /* L + 1 */
    uint8_t **trees;   /* L */
} fri_state;

static void fri_state_free(fri_state *s) {
    if (s->layers) for (uint32_t i = 0; i <= s->L; i++) free(s->layers[i]);
    if (s->trees) for (uint32_t i = 0; i < s->L; i++) free(s->trees[i]);
    free(s->layers);
    free(s->trees);
}

/* fri.zig:157-233: commit loop + remainder */
static int fri_commit_impl(const uint32_t *codeword, uint32_t dlog, uint32_t log_blowup, uint32_t remainder_log,
                           const uint32_t *alphas, zo_channel *chan, uint8_t *roots, uint32_t *remainder,
                           uint32_t *alphas_out, fri_state *st) {
    if (remainder_log > dlog || remainder_log < log_blowup) return -2;
    uint32_t L = dlog - remainder_log;
    if (L < 1) return -2;
    if (!alphas && !chan) return -2;
    size_t n0 = (size_t)1 << dlog;
    st->L = L;
    st->layers = (zo_qm31 **)calloc(L + 1, sizeof(zo_qm31 *));
    st->trees = (uint8_t **)calloc(L, sizeof(uint8_t *));
    st->layers[0] = (zo_qm31 *)malloc(n0 * sizeof(zo_qm31));
    memcpy(st->layers[0], codeword, n0 * sizeof(zo_qm31));
    for (uint32_t i = 0; i < L; i++) {
        size_t ni = n0 >> i;
        uint8_t *leaves = (uint8_t *)malloc(ni * 32);
        for (size_t j = 0; j < ni; j++) zo_hash_qm31(st->layers[i][j].v, leaves + 32 * j);
        st->trees[i] = (uint8_t *)malloc((2 * ni - 1) * 32);
        zo_merkle_build(leaves, ni, st->trees[i]);
        free(leaves);
        const uint8_t *root = st->trees[i] + (2 * ni - 2) * 32;
        memcpy(roots + 32 * i, root, 32);
        zo_qm31 alpha;
        if (alphas) {
            alpha = ld4(alphas + 4 * i);
        } else {
            zo_channel_absorb(chan, root, 32);
            zo_channel_sample_qm31(chan, alpha.v);
        }
        if (alphas_out) st4(alphas_out + 4 * i, alpha);
        zo_qm31 *dom = domain_points(dlog, i);
        st->layers[i + 1] = (zo_qm31 *)malloc((ni / 2) * sizeof(zo_qm31));
        fold_layer(st->layers[i], ni, dom, alpha, st->layers[i + 1]);
        free(dom);
    }
    size_t rc = (size_t)1 << (remainder_log - log_blowup);
    zo_qm31 *domL = domain_points(dlog, L);
    zo_poly_interpolate((const uint32_t *)domL, (const uint32_t *)st->layers[L], rc, remainder);
    free(domL);
    if (!alphas && chan) {
        uint8_t b[16];
        for (size_t k = 0; k < rc; k++) {
            qm31_to_bytes(ld4(remainder + 4 * k), b);
            zo_channel_absorb(chan, b, 16);
        }
    }
    return 0;
}

int zo_fri_commit(const uint32_t *codeword, uint32_t dlog, uint32_t log_blowup, uint32_t remainder_log,
                  const uint32_t *alphas, zo_channel *chan, uint8_t *roots, uint32_t *layers, uint32_t *remainder,
                  uint32_t *alphas_out) {
    fri_state st = {0};
    int rc = fri_commit_impl(codeword, dlog, log_blowup, remainder_log, alphas, chan, roots, remainder,
                             alphas_out, &st);
    if (!rc && layers) {
        size_t off = 0;
        for (uint32_t i = 0; i <= st.L; i++) {
            size_t ni = (size_t)1 << (dlog - i);
            memcpy(layers + 4 * off, st.layers[i], ni * 16);
            off += ni;
        }
    }
    fri_state_free(&st);
    return rc;
}

size_t zo_fri_query_path_bytes(uint32_t dlog, uint32_t remainder_log) {
    size_t s = 0;
    for (uint32_t i = 0; i < dlog - remainder_log; i++) s += 2 * (size_t)(dlog - i) * 32;
    return s;
}

/* fri.zig:157-273: commit phase, then query phase (fri.zig:236-256) */
int zo_fri_prove(const uint32_t *codeword, uint32_t dlog, uint32_t log_blowup, uint32_t remainder_log,
                 size_t num_queries, zo_channel *chan, uint8_t *roots, uint32_t *remainder, size_t *query_idx,
                 uint32_t *query_vals, uint8_t *query_paths) {
    if (!chan) return -2;
    fri_state st = {0};
    int rc = fri_commit_impl(codeword, dlog, log_blowup, remainder_log, NULL, chan, roots, remainder, NULL, &st);
    if (rc) { fri_state_free(&st); return rc; }
    uint32_t L = st.L;
    size_t n0 = (size_t)1 << dlog;
    size_t per_q = zo_fri_query_path_bytes(dlog, remainder_log);
    for (size_t q = 0; q < num_queries; q++) {
        size_t p = zo_channel_sample_index(chan, n0);
        query_idx[q] = p;
        uint8_t *pp = query_paths + q * per_q;
        for (uint32_t i = 0; i < L; i++) {
            size_t ni = n0 >> i, nn = ni / 2, pn = p % nn;
            st4(query_vals + 4 * ((q * L + i) * 2), st.layers[i][pn]);
            st4(query_vals + 4 * ((q * L + i) * 2 + 1), st.layers[i][pn + nn]);
            size_t plen = dlog - i;
            zo_merkle_open(st.trees[i], ni, pn, pp);
            pp += plen * 32;
            zo_merkle_open(st.trees[i], ni, pn + nn, pp);
            pp += plen * 32;
            p = pn;
        }
    }
    fri_state_free(&st);
    return 0;
}

/* fri.zig:277-336 */
int zo_fri_verify(uint32_t dlog, uint32_t log_blowup, uint32_t remainder_log, size_t num_queries, zo_channel *chan,
                  const uint8_t *roots, const uint32_t *remainder, const size_t *query_idx, const uint32_t *query_vals,
                  const uint8_t *query_paths) {
    if (remainder_log > dlog || remainder_log < log_blowup) return 0;
    uint32_t L = dlog - remainder_log;
    size_t n0 = (size_t)1 << dlog;
    size_t rc = (size_t)1 << (remainder_log - log_blowup);
    zo_qm31 *alphas = (zo_qm31 *)malloc(L * sizeof(zo_qm31));
    for (uint32_t i = 0; i < L; i++) {
        zo_channel_absorb(chan, roots + 32 * i, 32);
        zo_channel_sample_qm31(chan, alphas[i].v);
    }
    uint8_t b[16];
    for (size_t k = 0; k < rc; k++) {
        qm31_to_bytes(ld4(remainder + 4 * k), b);
        zo_channel_absorb(chan, b, 16);
    }
    size_t per_q = zo_fri_query_path_bytes(dlog, remainder_log);
    zo_qm31 i2 = inv_two();
    int ok = 1;
    for (size_t q = 0; q < num_queries && ok; q++) {
        if (zo_channel_sample_index(chan, n0) != query_idx[q]) { ok = 0; break; }
        size_t p = query_idx[q];
        const uint8_t *pp = query_paths + q * per_q;
        zo_qm31 prev = Q_ZERO;
        for (uint32_t i = 0; i < L && ok; i++) {
            size_t ni = n0 >> i, nn = ni / 2, pn = p % nn, plen = dlog - i;
            zo_qm31 v0 = ld4(query_vals + 4 * ((q * L + i) * 2));
            zo_qm31 v1 = ld4(query_vals + 4 * ((q * L + i) * 2 + 1));
            uint8_t h0[32], h1[32];
            zo_hash_qm31(v0.v, h0);
            zo_hash_qm31(v1.v, h1);
            if (!zo_merkle_verify(roots + 32 * i, pn, h0, pp, plen)) ok = 0;
            if (!zo_merkle_verify(roots + 32 * i, pn + nn, h1, pp + plen * 32, plen)) ok = 0;
            pp += 2 * plen * 32;
            if (!ok) break;
            if (i > 0) {
                zo_qm31 m = (p == pn) ? v0 : v1;
                if (memcmp(prev.v, m.v, 16) != 0) { ok = 0; break; }
            }
            uint32_t g[4];
            zo_fri_domain_point(dlog, i, pn, g);
            zo_qm31 hs = zo_qm31_mul(zo_qm31_add(v0, v1), i2);
            zo_qm31 hd = zo_qm31_mul(zo_qm31_mul(zo_qm31_sub(v0, v1), zo_qm31_inv(ld4(g))), i2);
            zo_qm31 fv = zo_qm31_add(hs, zo_qm31_mul(alphas[i], hd));
            if (i == L - 1) {
                uint32_t h[4], e[4];
                zo_fri_domain_point(dlog, i + 1, pn, h);
                zo_poly_eval(remainder, rc, h, e);
                if (memcmp(e, fv.v, 16) != 0) { ok = 0; break; }
            }
            prev = fv;
            p = pn;
        }
    }
    free(alphas);
    return ok;
}

/* ======================================================================== */
/* Seeded synthetic inputs                                                   */
/* ======================================================================== */
static uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* element i = top 31 bits of the (i+1)-th SplitMix64 output for `seed`; the
 * value 2^31 - 1 is rejected by moving to the candidate 2^40 positions later
 * (keeps the stream index-addressable so the device generator matches). */
void zo_fill_m31(uint64_t seed, uint32_t *out, size_t n) {
    const uint64_t gamma = 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i < n; i++) {
        uint64_t k = (uint64_t)i + 1;
        uint32_t v;
        for (;;) {
            v = (uint32_t)(splitmix_mix(seed + k * gamma) >> 33);
            if (v != P) break;
            k += (uint64_t)1 << 40;
        }
        out[i] = v;
    }
}

/* ======================================================================== */
/* GenericStark trace commitment: src/m31/stark.zig:343-414 (no preprocessed  */
/* columns), with FibAir's trace (stark.zig:877-931)                         */
/* ======================================================================== */
/* stark.zig:343-414: per column, interpolate over H = {w^i} (O(n^2) Lagrange,
 * univariate.zig:79-139), evaluate at D = {FRI_OFFSET * w_ev^i} by Horner,
 * hashQM31 leaves (hashCodeword, stark.zig:786-790), MerkleTree.init, and
 * absorbDigest(root) in column order. */
int zo_trace_commit(const uint32_t *cols, size_t n_cols, uint32_t trace_log, uint32_t log_blowup, zo_channel *chan,
                    uint32_t *codewords, uint8_t *roots) {
    const size_t n = (size_t)1 << trace_log, N = (size_t)1 << (trace_log + log_blowup);
    zo_qm31 w = zo_qm31_root_of_unity(trace_log), w_ev = zo_qm31_root_of_unity(trace_log + log_blowup);
    uint32_t *h = (uint32_t *)malloc(n * 16), *d = (uint32_t *)malloc(N * 16), *co = (uint32_t *)malloc(n * 16);
    uint32_t *cw = (uint32_t *)malloc(N * 16);
    uint8_t *leaves = (uint8_t *)malloc(N * 32), *nodes = (uint8_t *)malloc((2 * N - 1) * 32);
    if (!h || !d || !co || !cw || !leaves || !nodes) {
        free(h); free(d); free(co); free(cw); free(leaves); free(nodes);
        return -3;
    }
    zo_qm31 x = Q_ONE;
    for (size_t i = 0; i < n; i++) { st4(h + 4 * i, x); x = zo_qm31_mul(x, w); }
    x = ld4(ZO_FRI_OFFSET);
    for (size_t i = 0; i < N; i++) { st4(d + 4 * i, x); x = zo_qm31_mul(x, w_ev); }
    int rc = 0;
    for (size_t c = 0; c < n_cols && !rc; c++) {
        rc = zo_poly_interpolate(h, cols + 4 * n * c, n, co);
        if (rc) break;
        for (size_t i = 0; i < N; i++) {
            zo_poly_eval(co, n, d + 4 * i, cw + 4 * i);
            zo_hash_qm31(cw + 4 * i, leaves + 32 * i);
        }
        if (codewords) memcpy(codewords + 4 * N * c, cw, N * 16);
        zo_merkle_build(leaves, N, nodes);
        memcpy(roots + 32 * c, nodes + (2 * N - 2) * 32, 32);
        if (chan) zo_channel_absorb(chan, roots + 32 * c, 32);
    }
    free(h); free(d); free(co); free(cw); free(leaves); free(nodes);
    return rc;
}

/* FibAir.generateTrace (stark.zig:918-935) with seeds a0, a1 instead of 0, 1:
 * column 0 holds a_i, column 1 holds a_{i+1}, a_{i+2} = a_i + a_{i+1}; values
 * are QM31 with the M31 seed in the first coordinate. */
void zo_fib_trace(uint32_t a0, uint32_t a1, size_t n, uint32_t *cols) {
    uint32_t a = a0, b = a1;
    for (size_t i = 0; i < n; i++) {
        uint32_t *p = cols + 4 * i, *q = cols + 4 * (n + i);
        p[0] = a; p[1] = p[2] = p[3] = 0;
        q[0] = b; q[1] = q[2] = q[3] = 0;
        uint32_t nb = zo_m31_add(a, b);
        a = b;
        b = nb;
    }
}

/* ======================================================================== */
/* Wire format: src/core/serialization.zig:37-140, docs/wire.md              */
/* Little-endian unsigned ints (usize as 8 bytes), [N]u8 verbatim, slices as */
/* a u64 length then the elements, ?T as a presence byte then T, structs as  */
/* their fields in declaration order (allocator fields skipped), field       */
/* elements as their toBytes (QM31: a.c0, a.c1, b.c0, b.c1 as u32 LE).       */
/* ======================================================================== */
typedef struct { uint8_t *p; size_t n; } wire_buf;
static void w_bytes(wire_buf *w, const void *d, size_t n) {
    if (w->p) memcpy(w->p + w->n, d, n);
    w->n += n;
}
static void w_uint(wire_buf *w, uint64_t v, unsigned bytes) {
    uint8_t t[8];
    for (unsigned b = 0; b < bytes; b++) t[b] = (uint8_t)(v >> (8 * b));
    w_bytes(w, t, bytes);
}
static void w_qm31(wire_buf *w, const uint32_t v[4]) {
    for (int k = 0; k < 4; k++) w_uint(w, v[k], 4);
}

/* serialization.zig:219-247: the golden layout of
 * struct { name: [3]u8, count: u32, rows: []const u16, flag: bool, maybe: ?[2]u8 }
 * with ("abc", 0x01020304, {0x0102, 0x0304}, true, {0xde, 0xad}), written with the
 * writers above (pins them) */
size_t zo_wire_golden_example(uint8_t *out) {
    wire_buf w = {out, 0};
    w_bytes(&w, "abc", 3);
    w_uint(&w, 0x01020304, 4);
    w_uint(&w, 2, 8);
    w_uint(&w, 0x0102, 2);
    w_uint(&w, 0x0304, 2);
    w_uint(&w, 1, 1);
    w_uint(&w, 1, 1);
    const uint8_t maybe[2] = {0xde, 0xad};
    w_bytes(&w, maybe, 2);
    return w.n;
}

/* Fri.Proof (fri.zig:38-80): params {log_size u8, log_blowup u8, num_queries usize,
 * remainder_log u8}, roots []Digest, remainder []QM31, queries []Query{index usize,
 * pairs []Pair{value0 QM31, value1 QM31, path0 []Digest, path1 []Digest}}.
 * Inputs in zo_fri_prove's layouts; out may be NULL (returns the length). */
size_t zo_fri_proof_serialize(uint32_t log_size, uint32_t log_blowup, uint32_t remainder_log, size_t num_queries,
                              const uint8_t *roots, const uint32_t *remainder, const size_t *query_idx,
                              const uint32_t *query_vals, const uint8_t *query_paths, uint8_t *out) {
    wire_buf w = {out, 0};
    const uint32_t dlog = log_size + log_blowup, L = dlog - remainder_log;
    const size_t rc = (size_t)1 << (remainder_log - log_blowup);
    w_uint(&w, log_size, 1);
    w_uint(&w, log_blowup, 1);
    w_uint(&w, num_queries, 8);
    w_uint(&w, remainder_log, 1);
    w_uint(&w, L, 8);
    w_bytes(&w, roots, (size_t)L * 32);
    w_uint(&w, rc, 8);
    for (size_t k = 0; k < rc; k++) w_qm31(&w, remainder + 4 * k);
    w_uint(&w, num_queries, 8);
    const size_t per_q = zo_fri_query_path_bytes(dlog, remainder_log);
    for (size_t q = 0; q < num_queries; q++) {
        w_uint(&w, query_idx[q], 8);
        w_uint(&w, L, 8);
        const uint8_t *pp = query_paths + q * per_q;
        for (uint32_t i = 0; i < L; i++) {
            const size_t plen = dlog - i;
            w_qm31(&w, query_vals + 4 * ((q * L + i) * 2));
            w_qm31(&w, query_vals + 4 * ((q * L + i) * 2 + 1));
            w_uint(&w, plen, 8);
            w_bytes(&w, pp, plen * 32);
            pp += plen * 32;
            w_uint(&w, plen, 8);
            w_bytes(&w, pp, plen * 32);
            pp += plen * 32;
        }
    }
    return w.n;
}

/* ======================================================================== */
/* GenericStark(FibSeededAir): src/m31/stark.zig:303-657 (prove) and        */
/* :659-786 (verify), without preprocessed columns or lookups, for FibAir    */
/* (stark.zig:877-904) with seeded boundary assertions (0,0) = a0,          */
/* (1,0) = a1, (0,n-1) = claimed.  Proofs are GenericStark.Proof in the     */
/* wire format (serialization.zig).                                          */
/* ======================================================================== */
#define FIB_M 2
#define FIB_NDEEP 5
static void fib_boundary(const uint32_t *a0, const uint32_t *a1, const uint32_t *claimed, size_t n, size_t col[3],
                         size_t step[3], zo_qm31 val[3]) {
    col[0] = 0; step[0] = 0; val[0] = ld4(a0);
    col[1] = 1; step[1] = 0; val[1] = ld4(a1);
    col[2] = 0; step[2] = n - 1; val[2] = ld4(claimed);
}
/* Hc(x) of stark.zig:504-528 (without lookups) for FibAir's transitions:
 * next[0] - cur[1] and next[1] - cur[0] - cur[1] (stark.zig:886-891) */
static zo_qm31 fib_composition(zo_qm31 x, const zo_qm31 cur[2], const zo_qm31 nxt[2], zo_qm31 last, zo_qm31 zh,
                               const zo_qm31 al[2], const zo_qm31 be[3], const size_t col[3], const zo_qm31 pk[3],
                               const zo_qm31 val[3]) {
    zo_qm31 res[2] = {zo_qm31_sub(nxt[0], cur[1]), zo_qm31_sub(zo_qm31_sub(nxt[1], cur[0]), cur[1])};
    zo_qm31 h = Q_ZERO, xl = zo_qm31_sub(x, last);
    for (int k = 0; k < 2; k++) h = zo_qm31_add(h, zo_qm31_mul(al[k], zo_qm31_mul(res[k], xl)));
    for (int k = 0; k < 3; k++) {
        zo_qm31 term = zo_qm31_mul(zo_qm31_mul(zo_qm31_sub(cur[col[k]], val[k]), zh), zo_qm31_inv(zo_qm31_sub(x, pk[k])));
        h = zo_qm31_add(h, zo_qm31_mul(be[k], term));
    }
    return h;
}
static void absorb_qm31(zo_channel *c, zo_qm31 x) {
    uint8_t b[16];
    qm31_to_bytes(x, b);
    zo_channel_absorb(c, b, 16);
}
static zo_qm31 sample_q(zo_channel *c) {
    zo_qm31 r;
    zo_channel_sample_qm31(c, r.v);
    return r;
}
static void w_stark_params(wire_buf *w, uint32_t tl, uint32_t lb, size_t nq, uint32_t rl) {
    w_uint(w, tl, 1);
    w_uint(w, lb, 1);
    w_uint(w, nq, 8);
    w_uint(w, rl, 1);
}

uint8_t *zo_fib_stark_prove(uint32_t tl, uint32_t lb, size_t nq, uint32_t rl, const uint32_t *trace,
                            const uint32_t a0[4], const uint32_t a1[4], const uint32_t claimed[4], zo_channel *chan,
                            size_t *out_len) {
    const size_t n = (size_t)1 << tl, N = (size_t)1 << (tl + lb), shift = (size_t)1 << lb;
    const uint32_t dlog = tl + lb;
    if (!chan || tl < 1 || rl < lb || rl >= dlog) return NULL;
    zo_qm31 w = zo_qm31_root_of_unity(tl), w_ev = zo_qm31_root_of_unity(dlog);
    uint32_t *h = malloc(n * 16), *co = malloc(FIB_M * n * 16), *cw = malloc(FIB_M * N * 16), *qc = malloc(N * 16);
    uint32_t *g = malloc(N * 16);
    zo_qm31 *d = malloc(N * sizeof(zo_qm31));
    uint8_t *leaves = malloc(N * 32), *trees[FIB_M + 1];
    for (int t = 0; t <= FIB_M; t++) trees[t] = malloc((2 * N - 1) * 32);
    /* 1. trace: interpolate over H, evaluate on D, commit (stark.zig:343-414) */
    zo_qm31 x = Q_ONE;
    for (size_t i = 0; i < n; i++) { st4(h + 4 * i, x); x = zo_qm31_mul(x, w); }
    d[0] = ld4(ZO_FRI_OFFSET);
    for (size_t i = 1; i < N; i++) d[i] = zo_qm31_mul(d[i - 1], w_ev);
    uint8_t roots[FIB_M][32], qroot[32];
    for (int j = 0; j < FIB_M; j++) {
        zo_poly_interpolate(h, trace + 4 * n * j, n, co + 4 * n * j);
        for (size_t i = 0; i < N; i++) {
            zo_poly_eval(co + 4 * n * j, n, d[i].v, cw + 4 * (N * j + i));
            zo_hash_qm31(cw + 4 * (N * j + i), leaves + 32 * i);
        }
        zo_merkle_build(leaves, N, trees[j]);
        memcpy(roots[j], trees[j] + (2 * N - 2) * 32, 32);
        zo_channel_absorb(chan, roots[j], 32);
    }
    /* 3. transition weights, boundary assertions, boundary weights (stark.zig:470-483) */
    zo_qm31 al[2], be[3], val[3], pk[3];
    for (int k = 0; k < 2; k++) al[k] = sample_q(chan);
    size_t col[3], step[3];
    fib_boundary(a0, a1, claimed, n, col, step, val);
    for (int k = 0; k < 3; k++) be[k] = sample_q(chan);
    for (int k = 0; k < 3; k++) pk[k] = zo_qm31_pow(w, step[k]);
    const zo_qm31 last = ld4(h + 4 * (n - 1));
    /* 4. quotient on D (stark.zig:485-531) */
    for (size_t i = 0; i < N; i++) {
        zo_qm31 zh = zo_qm31_sub(zo_qm31_pow(d[i], n), Q_ONE);
        size_t ni = (i + shift) % N;
        zo_qm31 cur[2] = {ld4(cw + 4 * i), ld4(cw + 4 * (N + i))}, nxt[2] = {ld4(cw + 4 * ni), ld4(cw + 4 * (N + ni))};
        zo_qm31 hv = fib_composition(d[i], cur, nxt, last, zh, al, be, col, pk, val);
        st4(qc + 4 * i, zo_qm31_mul(hv, zo_qm31_inv(zh)));
        zo_hash_qm31(qc + 4 * i, leaves + 32 * i);
    }
    zo_merkle_build(leaves, N, trees[FIB_M]);
    memcpy(qroot, trees[FIB_M] + (2 * N - 2) * 32, 32);
    zo_channel_absorb(chan, qroot, 32);
    /* 5. DEEP evaluations and weights (stark.zig:540-566, computeQuotientAt :793-846) */
    zo_qm31 z = sample_q(chan), wz = zo_qm31_mul(z, w), de[FIB_NDEEP], ga[FIB_NDEEP];
    for (int j = 0; j < FIB_M; j++) {
        zo_poly_eval(co + 4 * n * j, n, z.v, de[j].v);
        zo_poly_eval(co + 4 * n * j, n, wz.v, de[FIB_M + j].v);
    }
    {
        zo_qm31 zh = zo_qm31_sub(zo_qm31_pow(z, n), Q_ONE);
        zo_qm31 cur[2] = {de[0], de[1]}, nxt[2] = {de[2], de[3]};
        de[4] = zo_qm31_mul(fib_composition(z, cur, nxt, last, zh, al, be, col, pk, val), zo_qm31_inv(zh));
    }
    for (int k = 0; k < FIB_NDEEP; k++) absorb_qm31(chan, de[k]);
    for (int k = 0; k < FIB_NDEEP; k++) ga[k] = sample_q(chan);
    /* DEEP combined codeword (stark.zig:568-580) */
    for (size_t i = 0; i < N; i++) {
        zo_qm31 inv_dz = zo_qm31_inv(zo_qm31_sub(d[i], z)), gv = Q_ZERO;
        size_t ni = (i + shift) % N;
        for (int j = 0; j < FIB_M; j++) {
            gv = zo_qm31_add(gv, zo_qm31_mul(ga[j], zo_qm31_mul(zo_qm31_sub(ld4(cw + 4 * (N * j + i)), de[j]), inv_dz)));
            gv = zo_qm31_add(gv, zo_qm31_mul(ga[FIB_M + j],
                                             zo_qm31_mul(zo_qm31_sub(ld4(cw + 4 * (N * j + ni)), de[FIB_M + j]), inv_dz)));
        }
        gv = zo_qm31_add(gv, zo_qm31_mul(ga[4], zo_qm31_mul(zo_qm31_sub(ld4(qc + 4 * i), de[4]), inv_dz)));
        st4(g + 4 * i, gv);
    }
    /* FRI on g: friParams = {log_size = trace_log, log_blowup, num_queries, remainder_log} for FibAir */
    const uint32_t L = dlog - rl;
    const size_t rc = (size_t)1 << (rl - lb), per_q = zo_fri_query_path_bytes(dlog, rl);
    uint8_t *froots = malloc(32 * L), *fpaths = malloc(per_q * nq + 1);
    uint32_t *frem = malloc(16 * rc), *fvals = malloc(16 * 2 * L * nq + 16);
    size_t *fidx = malloc(sizeof(size_t) * (nq + 1));
    int rcode = zo_fri_prove(g, dlog, lb, rl, nq, chan, froots, frem, fidx, fvals, fpaths);
    uint8_t *out = NULL;
    if (rcode == 0) {
        /* serialize GenericStark.Proof (stark.zig:110-123) */
        for (int pass = 0; pass < 2; pass++) {
            wire_buf wb = {pass ? out : NULL, 0};
            w_stark_params(&wb, tl, lb, nq, rl);
            w_uint(&wb, FIB_M, 8);
            for (int j = 0; j < FIB_M; j++) w_bytes(&wb, roots[j], 32);
            w_uint(&wb, 0, 1); /* preprocessed_roots = null */
            w_uint(&wb, 0, 1); /* accumulator_roots = null */
            w_bytes(&wb, qroot, 32);
            w_uint(&wb, FIB_NDEEP, 8);
            for (int k = 0; k < FIB_NDEEP; k++) w_qm31(&wb, de[k].v);
            size_t fl = zo_fri_proof_serialize(tl, lb, rl, nq, froots, frem, fidx, fvals, fpaths,
                                               pass ? out + wb.n : NULL);
            wb.n += fl;
            w_uint(&wb, nq, 8);
            for (size_t q = 0; q < nq; q++) {
                size_t p0 = fidx[q], pn = (p0 + shift) % N;
                const size_t at[FIB_NDEEP] = {p0, p0, pn, pn, p0};
                const int tr[FIB_NDEEP] = {0, 1, 0, 1, FIB_M};
                w_uint(&wb, p0, 8);
                w_uint(&wb, FIB_NDEEP, 8);
                for (int k = 0; k < FIB_NDEEP; k++) {
                    const uint32_t *src = tr[k] == FIB_M ? qc + 4 * at[k] : cw + 4 * (N * tr[k] + at[k]);
                    w_qm31(&wb, src);
                }
                w_uint(&wb, FIB_NDEEP, 8);
                uint8_t path[32 * 32];
                for (int k = 0; k < FIB_NDEEP; k++) {
                    zo_merkle_open(trees[tr[k]], N, at[k], path);
                    w_uint(&wb, dlog, 8);
                    w_bytes(&wb, path, (size_t)dlog * 32);
                }
            }
            if (!pass) { out = malloc(wb.n); *out_len = wb.n; }
        }
    }
    free(h); free(co); free(cw); free(qc); free(g); free(d); free(leaves);
    for (int t = 0; t <= FIB_M; t++) free(trees[t]);
    free(froots); free(fpaths); free(frem); free(fvals); free(fidx);
    return out;
}

void zo_free(void *p) { free(p); }

/* ---- verify (stark.zig:659-786) over the wire bytes ---- */
typedef struct { const uint8_t *b; size_t n, pos; int bad; } wire_rd;
static const uint8_t *r_take(wire_rd *r, size_t k) {
    if (r->bad || r->pos + k > r->n) { r->bad = 1; return NULL; }
    const uint8_t *p = r->b + r->pos;
    r->pos += k;
    return p;
}
static uint64_t r_uint(wire_rd *r, unsigned k) {
    const uint8_t *p = r_take(r, k);
    uint64_t v = 0;
    if (p) for (unsigned i = 0; i < k; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}
static zo_qm31 r_qm31(wire_rd *r) {
    zo_qm31 x;
    for (int k = 0; k < 4; k++) x.v[k] = (uint32_t)r_uint(r, 4);
    return x;
}

int zo_fib_stark_verify(const uint8_t *proof, size_t len, const uint32_t a0[4], const uint32_t a1[4],
                        const uint32_t claimed[4], zo_channel *chan) {
    wire_rd r = {proof, len, 0, 0};
    const uint32_t tl = (uint32_t)r_uint(&r, 1), lb = (uint32_t)r_uint(&r, 1);
    const size_t nq = (size_t)r_uint(&r, 8);
    const uint32_t rl = (uint32_t)r_uint(&r, 1);
    if (r.bad || tl < 1 || tl + lb > 24 || rl < lb || rl >= tl + lb || nq > 4096) return 0;
    const size_t n = (size_t)1 << tl, N = (size_t)1 << (tl + lb), shift = (size_t)1 << lb;
    const uint32_t dlog = tl + lb, L = dlog - rl;
    const size_t rc = (size_t)1 << (rl - lb);
    if (r_uint(&r, 8) != FIB_M) return 0;
    const uint8_t *roots = r_take(&r, 32 * FIB_M);
    if (r_uint(&r, 1) != 0 || r_uint(&r, 1) != 0) return 0; /* no preprocessed / accumulator roots */
    const uint8_t *qroot = r_take(&r, 32);
    if (r_uint(&r, 8) != FIB_NDEEP) return 0;
    zo_qm31 de[FIB_NDEEP];
    for (int k = 0; k < FIB_NDEEP; k++) de[k] = r_qm31(&r);
    /* nested Fri.Proof */
    if (r_uint(&r, 1) != tl || r_uint(&r, 1) != lb || r_uint(&r, 8) != nq || r_uint(&r, 1) != rl) return 0;
    if (r_uint(&r, 8) != L) return 0;
    const uint8_t *froots = r_take(&r, 32 * (size_t)L);
    if (r_uint(&r, 8) != rc) return 0;
    uint32_t *frem = malloc(16 * rc), *fvals = malloc(16 * 2 * (size_t)L * nq + 16);
    const size_t per_q = zo_fri_query_path_bytes(dlog, rl);
    uint8_t *fpaths = malloc(per_q * nq + 1);
    size_t *fidx = malloc(sizeof(size_t) * (nq + 1));
    for (size_t k = 0; k < rc; k++) st4(frem + 4 * k, r_qm31(&r));
    if (r_uint(&r, 8) != nq) r.bad = 1;
    for (size_t q = 0; q < nq && !r.bad; q++) {
        fidx[q] = (size_t)r_uint(&r, 8);
        if (r_uint(&r, 8) != L) { r.bad = 1; break; }
        uint8_t *pp = fpaths + q * per_q;
        for (uint32_t i = 0; i < L && !r.bad; i++) {
            st4(fvals + 4 * ((q * L + i) * 2), r_qm31(&r));
            st4(fvals + 4 * ((q * L + i) * 2 + 1), r_qm31(&r));
            for (int s = 0; s < 2 && !r.bad; s++) {
                if (r_uint(&r, 8) != dlog - i) { r.bad = 1; break; }
                const uint8_t *p = r_take(&r, 32 * (size_t)(dlog - i));
                if (p) memcpy(pp, p, 32 * (size_t)(dlog - i));
                pp += 32 * (size_t)(dlog - i);
            }
        }
    }
    int ok = !r.bad;
    /* replay the transcript (stark.zig:676-705) */
    zo_qm31 al[2], be[3], val[3], pk[3], ga[FIB_NDEEP];
    size_t col[3], step[3];
    const zo_qm31 w = zo_qm31_root_of_unity(tl), w_ev = zo_qm31_root_of_unity(dlog);
    if (ok) {
        for (int j = 0; j < FIB_M; j++) zo_channel_absorb(chan, roots + 32 * j, 32);
        for (int k = 0; k < 2; k++) al[k] = sample_q(chan);
        fib_boundary(a0, a1, claimed, n, col, step, val);
        for (int k = 0; k < 3; k++) be[k] = sample_q(chan);
        for (int k = 0; k < 3; k++) pk[k] = zo_qm31_pow(w, step[k]);
        zo_channel_absorb(chan, qroot, 32);
        zo_qm31 z = sample_q(chan);
        for (int k = 0; k < FIB_NDEEP; k++) absorb_qm31(chan, de[k]);
        for (int k = 0; k < FIB_NDEEP; k++) ga[k] = sample_q(chan);
        ok = zo_fri_verify(dlog, lb, rl, nq, chan, froots, frem, fidx, fvals, fpaths);
        const zo_qm31 last = zo_qm31_pow(w, n - 1);
        if (ok && r_uint(&r, 8) != nq) ok = 0;
        for (size_t q = 0; q < nq && ok; q++) {
            size_t p0 = (size_t)r_uint(&r, 8);
            if (p0 != fidx[q] || r_uint(&r, 8) != FIB_NDEEP) { ok = 0; break; }
            zo_qm31 v[FIB_NDEEP];
            for (int k = 0; k < FIB_NDEEP; k++) v[k] = r_qm31(&r);
            if (r_uint(&r, 8) != FIB_NDEEP) { ok = 0; break; }
            const size_t pn = (p0 + shift) % N, at[FIB_NDEEP] = {p0, p0, pn, pn, p0};
            for (int k = 0; k < FIB_NDEEP && ok; k++) {
                if (r_uint(&r, 8) != dlog) { ok = 0; break; }
                const uint8_t *path = r_take(&r, 32 * (size_t)dlog);
                uint8_t leaf[32];
                zo_hash_qm31(v[k].v, leaf);
                const uint8_t *root = k == 4 ? qroot : roots + 32 * (k % 2);
                if (!path || !zo_merkle_verify(root, at[k], leaf, path, dlog)) ok = 0;
            }
            if (!ok) break;
            zo_qm31 x0 = zo_qm31_mul(ld4(ZO_FRI_OFFSET), zo_qm31_pow(w_ev, p0));
            /* DEEP identity against the FRI layer-0 leaf (stark.zig:736-751) */
            zo_qm31 inv_dz = zo_qm31_inv(zo_qm31_sub(x0, z)), gv = Q_ZERO;
            for (int j = 0; j < FIB_M; j++) {
                gv = zo_qm31_add(gv, zo_qm31_mul(ga[j], zo_qm31_mul(zo_qm31_sub(v[j], de[j]), inv_dz)));
                gv = zo_qm31_add(gv, zo_qm31_mul(ga[FIB_M + j], zo_qm31_mul(zo_qm31_sub(v[FIB_M + j], de[FIB_M + j]), inv_dz)));
            }
            gv = zo_qm31_add(gv, zo_qm31_mul(ga[4], zo_qm31_mul(zo_qm31_sub(v[4], de[4]), inv_dz)));
            const size_t p1 = p0 % (N / 2);
            zo_qm31 leafv = ld4(fvals + 4 * ((q * L) * 2 + (p0 == p1 ? 0 : 1)));
            if (memcmp(gv.v, leafv.v, 16) != 0) { ok = 0; break; }
            /* Hc(x0) == Z_H(x0) * Q(x0) (stark.zig:753-779) */
            zo_qm31 zh = zo_qm31_sub(zo_qm31_pow(x0, n), Q_ONE);
            zo_qm31 cur[2] = {v[0], v[1]}, nxt[2] = {v[2], v[3]};
            zo_qm31 hv = fib_composition(x0, cur, nxt, last, zh, al, be, col, pk, val);
            zo_qm31 rhs = zo_qm31_mul(zh, v[4]);
            if (memcmp(hv.v, rhs.v, 16) != 0) ok = 0;
        }
        if (ok && (r.bad || r.pos != r.n)) ok = 0; /* truncated or trailing bytes */
    }
    free(frem); free(fvals); free(fpaths); free(fidx);
    return ok;
}

/* ======================================================================== */
/* GenericStark(Air): src/m31/stark.zig:311-657 (proveWithPreprocessed) and */
/* :660-780 (verify), with preprocessed columns and the synthesized LogUp    */
/* machinery, for an AIR given as data (zo_air, zs_oracle.h).                */
/* ======================================================================== */
typedef struct {
    const zo_air *air;
    uint32_t m, n_pre, n_rel, n_lc, T, B, m_total, n_trans_total, total_bound, n_deep, reveal_len;
} stark_dims;
static stark_dims stark_dims_of(const zo_air *a) {
    stark_dims d;
    d.air = a;
    d.m = a->num_columns; d.n_pre = a->num_preprocessed; d.n_rel = a->num_lookup_relations;
    d.n_lc = a->num_lookup_columns; d.T = a->num_transition_constraints; d.B = a->num_boundary;
    d.m_total = d.m + d.n_rel;                 /* stark.zig:184-188 */
    d.n_trans_total = d.T + 2 * d.n_rel;
    d.total_bound = d.B + d.n_rel;
    d.n_deep = 2 * d.m_total + 1;
    d.reveal_len = 2 * d.m_total + 1 + d.n_pre;
    return d;
}
/* the AIR's maxConstraintDegree(n), then GenericStark's (stark.zig:212-216) */
static uint64_t air_degree(const zo_air *a, uint64_t n) {
    if (a->degree_mul || a->degree_add) return (uint64_t)a->degree_mul * (n - 1) + a->degree_add;
    uint64_t best = 0;
    for (uint32_t t = 0; t < a->num_terms; t++) {
        uint64_t dg = 0;
        for (uint32_t f = 0; f < a->terms[t].n_factors; f++) dg += a->terms[t].factor[f].kind == ZO_AIR_X ? 1 : n - 1;
        if (dg > best) best = dg;
    }
    return best;
}
static uint64_t max_constraint_degree(const zo_air *a, uint64_t n) {
    uint64_t air = air_degree(a, n), lookup = a->num_lookup_relations ? 3 * (n - 1) : 0;
    return air > lookup ? air : lookup;
}
/* compositionLogSize (stark.zig:197-206) and friParams (:208-222) */
int zo_stark_fri_params(const zo_air *air, uint32_t tl, uint32_t lb, uint32_t *fri_log_size, uint32_t *fri_lb) {
    const uint64_t n = (uint64_t)1 << tl, dc = max_constraint_degree(air, n);
    const uint64_t dg = (n - 1) > (dc > n ? dc - n : 0) ? (n - 1) : dc - n;
    uint32_t log = 0;
    for (uint64_t bound = 1; bound <= dg; bound <<= 1) log++;
    if (log > tl + lb) return -2;
    *fri_log_size = log;
    *fri_lb = tl + lb - log;
    return 0;
}
/* combinedKey (stark.zig:225-229): sum_j alpha_j * row[cols[j]] */
static zo_qm31 combined_key(const uint32_t *cols, uint32_t k0, uint32_t k1, const zo_qm31 *alpha_cols,
                            const zo_qm31 *row) {
    zo_qm31 k = Q_ZERO;
    for (uint32_t j = k0; j < k1; j++) k = zo_qm31_add(k, zo_qm31_mul(alpha_cols[j - k0], row[cols[j]]));
    return k;
}
/* rowContribution (stark.zig:260-275) */
static zo_qm31 row_contribution(const zo_air *a, uint32_t r, const zo_qm31 *ac, const zo_qm31 *ar,
                                const zo_qm31 *row, const zo_qm31 *row_pre) {
    const zo_qm31 sel = row[a->selector_column[r]], mult = row[a->multiplicity_column[r]];
    const zo_qm31 key = combined_key(a->key_columns, a->key_start[r], a->key_start[r + 1], ac, row);
    const zo_qm31 tkey = combined_key(a->table_columns, a->table_start[r], a->table_start[r + 1], ac, row_pre);
    const zo_qm31 lt = zo_qm31_mul(sel, zo_qm31_mul(mult, zo_qm31_inv(zo_qm31_sub(ar[r], key))));
    const zo_qm31 tt = zo_qm31_mul(zo_qm31_sub(Q_ONE, sel), zo_qm31_inv(zo_qm31_sub(ar[r], tkey)));
    return zo_qm31_sub(lt, tt);
}
/* lookupConstraintAt (stark.zig:231-253) */
static zo_qm31 lookup_constraint(const zo_air *a, uint32_t r, const zo_qm31 *ac, const zo_qm31 *ar,
                                 const zo_qm31 *cur, const zo_qm31 *row_pre, zo_qm31 acc_cur, zo_qm31 acc_next) {
    const zo_qm31 sel = cur[a->selector_column[r]], mult = cur[a->multiplicity_column[r]];
    const zo_qm31 key = combined_key(a->key_columns, a->key_start[r], a->key_start[r + 1], ac, cur);
    const zo_qm31 tkey = combined_key(a->table_columns, a->table_start[r], a->table_start[r + 1], ac, row_pre);
    const zo_qm31 dacc = zo_qm31_sub(acc_next, acc_cur);
    const zo_qm31 lb = zo_qm31_sub(zo_qm31_mul(dacc, zo_qm31_sub(ar[r], key)), mult);
    const zo_qm31 tb = zo_qm31_add(zo_qm31_mul(dacc, zo_qm31_sub(ar[r], tkey)), Q_ONE);
    return zo_qm31_add(zo_qm31_mul(sel, lb), zo_qm31_mul(zo_qm31_sub(Q_ONE, sel), tb));
}
/* Air.evalTransition for the AIR's terms */
static void eval_transition(const zo_air *a, zo_qm31 x, const zo_qm31 *cur, const zo_qm31 *nxt, zo_qm31 *res) {
    for (uint32_t k = 0; k < a->num_transition_constraints; k++) res[k] = Q_ZERO;
    for (uint32_t t = 0; t < a->num_terms; t++) {
        const zo_air_term *tm = &a->terms[t];
        zo_qm31 p = ld4(tm->coeff);
        for (uint32_t f = 0; f < tm->n_factors; f++) {
            const zo_air_factor fc = tm->factor[f];
            p = zo_qm31_mul(p, fc.kind == ZO_AIR_X ? x : fc.kind == ZO_AIR_NEXT ? nxt[fc.column] : cur[fc.column]);
        }
        res[tm->constraint] = zo_qm31_add(res[tm->constraint], p);
    }
}
typedef struct {
    const zo_qm31 *alphas, *betas, *alpha_cols, *alpha_rels, *pk, *bval;
    const uint32_t *bcol;
    zo_qm31 last;
} stark_weights;
/* Hc(x) of stark.zig:504-530 (also :753-778 and :819-845): cur / nxt hold the
 * m_total trace and accumulator values at x and w x, row_pre the preprocessed ones */
static zo_qm31 composition(const stark_dims *d, const stark_weights *wt, zo_qm31 x, const zo_qm31 *cur,
                           const zo_qm31 *nxt, const zo_qm31 *row_pre, zo_qm31 zh, zo_qm31 *res) {
    const zo_air *a = d->air;
    eval_transition(a, x, cur, nxt, res);
    zo_qm31 h = Q_ZERO, xl = zo_qm31_sub(x, wt->last);
    for (uint32_t k = 0; k < d->T; k++) h = zo_qm31_add(h, zo_qm31_mul(wt->alphas[k], zo_qm31_mul(res[k], xl)));
    for (uint32_t r = 0; r < d->n_rel; r++) {
        const zo_qm31 c = lookup_constraint(a, r, wt->alpha_cols, wt->alpha_rels, cur, row_pre, cur[d->m + r],
                                            nxt[d->m + r]);
        h = zo_qm31_add(h, zo_qm31_mul(wt->alphas[d->T + r], c));
    }
    for (uint32_t r = 0; r < d->n_rel; r++) {
        const zo_qm31 sel = cur[a->selector_column[r]];
        h = zo_qm31_add(h, zo_qm31_mul(wt->alphas[d->T + d->n_rel + r], zo_qm31_mul(sel, zo_qm31_sub(Q_ONE, sel))));
    }
    for (uint32_t k = 0; k < d->total_bound; k++) {
        const zo_qm31 term = zo_qm31_mul(zo_qm31_mul(zo_qm31_sub(cur[wt->bcol[k]], wt->bval[k]), zh),
                                         zo_qm31_inv(zo_qm31_sub(x, wt->pk[k])));
        h = zo_qm31_add(h, zo_qm31_mul(wt->betas[k], term));
    }
    return h;
}
/* stark.zig:430-449 */
void zo_stark_accumulators(const zo_air *a, uint32_t tl, const uint32_t *pre, const uint32_t *trace,
                           const uint32_t *alpha_cols, const uint32_t *alpha_rels, uint32_t *out) {
    const size_t n = (size_t)1 << tl;
    const uint32_t m = a->num_columns, np = a->num_preprocessed, nr = a->num_lookup_relations;
    zo_qm31 *row = malloc(sizeof(zo_qm31) * (m + 1)), *rp = malloc(sizeof(zo_qm31) * (np + 1));
    zo_qm31 *ac = malloc(sizeof(zo_qm31) * (a->num_lookup_columns + 1)), *ar = malloc(sizeof(zo_qm31) * (nr + 1));
    for (uint32_t j = 0; j < a->num_lookup_columns; j++) ac[j] = ld4(alpha_cols + 4 * j);
    for (uint32_t r = 0; r < nr; r++) ar[r] = ld4(alpha_rels + 4 * r);
    for (uint32_t r = 0; r < nr; r++) st4(out + 4 * n * r, Q_ZERO);
    for (size_t i = 1; i < n; i++) {
        for (uint32_t j = 0; j < m; j++) row[j] = ld4(trace + 4 * (n * j + i - 1));
        for (uint32_t j = 0; j < np; j++) rp[j] = ld4(pre + 4 * (n * j + i - 1));
        for (uint32_t r = 0; r < nr; r++)
            st4(out + 4 * (n * r + i),
                zo_qm31_add(ld4(out + 4 * (n * r + i - 1)), row_contribution(a, r, ac, ar, row, rp)));
    }
    free(row); free(rp); free(ac); free(ar);
}

/* interpolate columns over H, evaluate on D, build hashQM31 trees; roots absorbed when chan is set */
static void commit_columns(const uint32_t *cols, uint32_t ncols, size_t n, size_t N, const uint32_t *h,
                           const zo_qm31 *dp, uint32_t *coeffs, uint32_t *codewords, uint8_t *trees, uint8_t *roots,
                           uint8_t *leaves, zo_channel *chan) {
    for (uint32_t j = 0; j < ncols; j++) {
        zo_poly_interpolate(h, cols + 4 * n * j, n, coeffs + 4 * n * j);
        for (size_t i = 0; i < N; i++) {
            zo_poly_eval(coeffs + 4 * n * j, n, dp[i].v, codewords + 4 * (N * j + i));
            zo_hash_qm31(codewords + 4 * (N * j + i), leaves + 32 * i);
        }
        uint8_t *t = trees + (2 * N - 1) * 32 * j;
        zo_merkle_build(leaves, N, t);
        memcpy(roots + 32 * j, t + (2 * N - 2) * 32, 32);
        if (chan) zo_channel_absorb(chan, roots + 32 * j, 32);
    }
}

uint8_t *zo_stark_prove(const zo_air *air, uint32_t tl, uint32_t lb, size_t nq, uint32_t rl, const uint32_t *pre,
                        const uint32_t *trace, zo_channel *chan, size_t *out_len) {
    const stark_dims D = stark_dims_of(air);
    const uint32_t dlog = tl + lb;
    uint32_t fls, flb;
    if (!chan || tl < 1 || zo_stark_fri_params(air, tl, lb, &fls, &flb) || rl >= dlog || rl < flb) return NULL;
    const size_t n = (size_t)1 << tl, N = (size_t)1 << dlog, shift = (size_t)1 << lb;
    const zo_qm31 w = zo_qm31_root_of_unity(tl), w_ev = zo_qm31_root_of_unity(dlog);
    const size_t tsz = (2 * N - 1) * 32;
    uint32_t *h = malloc(n * 16), *co = malloc((size_t)(D.m_total + 1) * n * 16), *pco = malloc((D.n_pre + 1) * n * 16);
    uint32_t *cw = malloc((size_t)(D.m_total + 1) * N * 16), *pcw = malloc((D.n_pre + 1) * N * 16);
    uint32_t *qc = malloc(N * 16), *g = malloc(N * 16), *acc = malloc((D.n_rel + 1) * n * 16);
    zo_qm31 *dp = malloc(N * sizeof(zo_qm31));
    uint8_t *leaves = malloc(N * 32), *trees = malloc((D.m_total + 1) * tsz), *ptrees = malloc((D.n_pre + 1) * tsz);
    uint8_t *roots = malloc(32 * (D.m_total + 1)), *proots = malloc(32 * (D.n_pre + 1)), qroot[32];
    zo_qm31 *ac = malloc(sizeof(zo_qm31) * (D.n_lc + 1)), *ar = malloc(sizeof(zo_qm31) * (D.n_rel + 1));
    zo_qm31 *al = malloc(sizeof(zo_qm31) * (D.n_trans_total + 1)), *be = malloc(sizeof(zo_qm31) * (D.total_bound + 1));
    zo_qm31 *pk = malloc(sizeof(zo_qm31) * (D.total_bound + 1)), *bv = malloc(sizeof(zo_qm31) * (D.total_bound + 1));
    uint32_t *bc = malloc(sizeof(uint32_t) * (D.total_bound + 1));
    zo_qm31 *cur = malloc(sizeof(zo_qm31) * (D.m_total + 1)), *nxt = malloc(sizeof(zo_qm31) * (D.m_total + 1));
    zo_qm31 *rp = malloc(sizeof(zo_qm31) * (D.n_pre + 1)), *res = malloc(sizeof(zo_qm31) * (D.T + 1));
    zo_qm31 *de = malloc(sizeof(zo_qm31) * D.n_deep), *ga = malloc(sizeof(zo_qm31) * D.n_deep);
    /* H and D points (stark.zig:343-372) */
    zo_qm31 x = Q_ONE;
    for (size_t i = 0; i < n; i++) { st4(h + 4 * i, x); x = zo_qm31_mul(x, w); }
    dp[0] = ld4(ZO_FRI_OFFSET);
    for (size_t i = 1; i < N; i++) dp[i] = zo_qm31_mul(dp[i - 1], w_ev);
    /* preprocessed commitment, absorbed first; then the trace (stark.zig:374-414) */
    commit_columns(pre, D.n_pre, n, N, h, dp, pco, pcw, ptrees, proots, leaves, chan);
    commit_columns(trace, D.m, n, N, h, dp, co, cw, trees, roots, leaves, chan);
    /* lookup challenges, accumulators on H, their commitment (stark.zig:416-468) */
    for (uint32_t j = 0; j < D.n_lc; j++) ac[j] = sample_q(chan);
    for (uint32_t r = 0; r < D.n_rel; r++) ar[r] = sample_q(chan);
    if (D.n_rel) {
        uint32_t *acw = malloc(4 * D.n_lc * 4 + 16), *arw = malloc(4 * D.n_rel * 4 + 16);
        for (uint32_t j = 0; j < D.n_lc; j++) st4(acw + 4 * j, ac[j]);
        for (uint32_t r = 0; r < D.n_rel; r++) st4(arw + 4 * r, ar[r]);
        zo_stark_accumulators(air, tl, pre, trace, acw, arw, acc);
        free(acw); free(arw);
        commit_columns(acc, D.n_rel, n, N, h, dp, co + 4 * n * D.m, cw + 4 * N * D.m, trees + tsz * D.m,
                       roots + 32 * D.m, leaves, chan);
    }
    /* transition and boundary weights (stark.zig:470-483) */
    for (uint32_t k = 0; k < D.n_trans_total; k++) al[k] = sample_q(chan);
    for (uint32_t k = 0; k < D.B; k++) { bc[k] = air->boundary[k].column; bv[k] = ld4(air->boundary[k].value);
                                         pk[k] = zo_qm31_pow(w, air->boundary[k].step); }
    for (uint32_t r = 0; r < D.n_rel; r++) { bc[D.B + r] = D.m + r; bv[D.B + r] = Q_ZERO; pk[D.B + r] = Q_ONE; }
    for (uint32_t k = 0; k < D.total_bound; k++) be[k] = sample_q(chan);
    stark_weights wt = {al, be, ac, ar, pk, bv, bc, ld4(h + 4 * (n - 1))};
    /* quotient on D (stark.zig:485-531) and its commitment */
    for (size_t i = 0; i < N; i++) {
        const zo_qm31 zh = zo_qm31_sub(zo_qm31_pow(dp[i], n), Q_ONE);
        const size_t ni = (i + shift) % N;
        for (uint32_t j = 0; j < D.m_total; j++) { cur[j] = ld4(cw + 4 * (N * j + i)); nxt[j] = ld4(cw + 4 * (N * j + ni)); }
        for (uint32_t j = 0; j < D.n_pre; j++) rp[j] = ld4(pcw + 4 * (N * j + i));
        st4(qc + 4 * i, zo_qm31_mul(composition(&D, &wt, dp[i], cur, nxt, rp, zh, res), zo_qm31_inv(zh)));
        zo_hash_qm31(qc + 4 * i, leaves + 32 * i);
    }
    uint8_t *qtree = malloc(tsz);
    zo_merkle_build(leaves, N, qtree);
    memcpy(qroot, qtree + (2 * N - 2) * 32, 32);
    zo_channel_absorb(chan, qroot, 32);
    /* z, DEEP evaluations, Q(z) (computeQuotientAt, stark.zig:793-846), DEEP weights */
    const zo_qm31 z = sample_q(chan), wz = zo_qm31_mul(z, w);
    for (uint32_t j = 0; j < D.m_total; j++) {
        zo_poly_eval(co + 4 * n * j, n, z.v, de[j].v);
        zo_poly_eval(co + 4 * n * j, n, wz.v, de[D.m_total + j].v);
    }
    {
        for (uint32_t j = 0; j < D.n_pre; j++) zo_poly_eval(pco + 4 * n * j, n, z.v, rp[j].v);
        const zo_qm31 zh = zo_qm31_sub(zo_qm31_pow(z, n), Q_ONE);
        de[2 * D.m_total] = zo_qm31_mul(composition(&D, &wt, z, de, de + D.m_total, rp, zh, res), zo_qm31_inv(zh));
    }
    for (uint32_t k = 0; k < D.n_deep; k++) absorb_qm31(chan, de[k]);
    for (uint32_t k = 0; k < D.n_deep; k++) ga[k] = sample_q(chan);
    /* DEEP combined codeword (stark.zig:568-580) */
    for (size_t i = 0; i < N; i++) {
        const zo_qm31 inv_dz = zo_qm31_inv(zo_qm31_sub(dp[i], z));
        const size_t ni = (i + shift) % N;
        zo_qm31 gv = Q_ZERO;
        for (uint32_t j = 0; j < D.m_total; j++) {
            gv = zo_qm31_add(gv, zo_qm31_mul(ga[j], zo_qm31_mul(zo_qm31_sub(ld4(cw + 4 * (N * j + i)), de[j]), inv_dz)));
            gv = zo_qm31_add(gv, zo_qm31_mul(ga[D.m_total + j], zo_qm31_mul(zo_qm31_sub(ld4(cw + 4 * (N * j + ni)),
                                                                                      de[D.m_total + j]), inv_dz)));
        }
        gv = zo_qm31_add(gv, zo_qm31_mul(ga[2 * D.m_total],
                                         zo_qm31_mul(zo_qm31_sub(ld4(qc + 4 * i), de[2 * D.m_total]), inv_dz)));
        st4(g + 4 * i, gv);
    }
    /* FRI on g under friParams (stark.zig:582-583) */
    const uint32_t L = dlog - rl;
    const size_t rc = (size_t)1 << (rl - flb), per_q = zo_fri_query_path_bytes(dlog, rl);
    uint8_t *froots = malloc(32 * L + 1), *fpaths = malloc(per_q * nq + 1);
    uint32_t *frem = malloc(16 * rc), *fvals = malloc(16 * 2 * L * nq + 16);
    size_t *fidx = malloc(sizeof(size_t) * (nq + 1));
    int rcode = zo_fri_prove(g, dlog, flb, rl, nq, chan, froots, frem, fidx, fvals, fpaths);
    uint8_t *out = NULL;
    if (rcode == 0) {
        for (int pass = 0; pass < 2; pass++) { /* GenericStark.Proof (stark.zig:107-123) */
            wire_buf wb = {pass ? out : NULL, 0};
            w_stark_params(&wb, tl, lb, nq, rl);
            w_uint(&wb, D.m, 8);
            w_bytes(&wb, roots, 32 * (size_t)D.m);
            w_uint(&wb, D.n_pre ? 1 : 0, 1); /* ?[]Digest */
            if (D.n_pre) { w_uint(&wb, D.n_pre, 8); w_bytes(&wb, proots, 32 * (size_t)D.n_pre); }
            w_uint(&wb, D.n_rel ? 1 : 0, 1);
            if (D.n_rel) { w_uint(&wb, D.n_rel, 8); w_bytes(&wb, roots + 32 * D.m, 32 * (size_t)D.n_rel); }
            w_bytes(&wb, qroot, 32);
            w_uint(&wb, D.n_deep, 8);
            for (uint32_t k = 0; k < D.n_deep; k++) w_qm31(&wb, de[k].v);
            wb.n += zo_fri_proof_serialize(fls, flb, rl, nq, froots, frem, fidx, fvals, fpaths,
                                           pass ? out + wb.n : NULL);
            w_uint(&wb, nq, 8);
            uint8_t path[32 * 32];
            for (size_t q = 0; q < nq; q++) {
                const size_t p0 = fidx[q], pn = (p0 + shift) % N;
                w_uint(&wb, p0, 8);
                w_uint(&wb, D.reveal_len, 8);
                /* values: columns at p0, at pn, Q at p0, preprocessed at p0 (stark.zig:593-625) */
                for (int half = 0; half < 2; half++)
                    for (uint32_t j = 0; j < D.m_total; j++) w_qm31(&wb, cw + 4 * (N * j + (half ? pn : p0)));
                w_qm31(&wb, qc + 4 * p0);
                for (uint32_t k = 0; k < D.n_pre; k++) w_qm31(&wb, pcw + 4 * (N * k + p0));
                w_uint(&wb, D.reveal_len, 8);
                for (int half = 0; half < 2; half++)
                    for (uint32_t j = 0; j < D.m_total; j++) {
                        zo_merkle_open(trees + tsz * j, N, half ? pn : p0, path);
                        w_uint(&wb, dlog, 8);
                        w_bytes(&wb, path, (size_t)dlog * 32);
                    }
                zo_merkle_open(qtree, N, p0, path);
                w_uint(&wb, dlog, 8);
                w_bytes(&wb, path, (size_t)dlog * 32);
                for (uint32_t k = 0; k < D.n_pre; k++) {
                    zo_merkle_open(ptrees + tsz * k, N, p0, path);
                    w_uint(&wb, dlog, 8);
                    w_bytes(&wb, path, (size_t)dlog * 32);
                }
            }
            if (!pass) { out = malloc(wb.n); *out_len = wb.n; }
        }
    }
    free(h); free(co); free(pco); free(cw); free(pcw); free(qc); free(g); free(acc); free(dp); free(leaves);
    free(trees); free(ptrees); free(roots); free(proots); free(ac); free(ar); free(al); free(be); free(pk); free(bv);
    free(bc); free(cur); free(nxt); free(rp); free(res); free(de); free(ga); free(qtree);
    free(froots); free(fpaths); free(frem); free(fvals); free(fidx);
    return out;
}

/* read a ?[]Digest of exactly `want` digests when `want` > 0, absent otherwise */
static const uint8_t *r_opt_digests(wire_rd *r, uint32_t want) {
    const uint64_t present = r_uint(r, 1);
    if (present > 1 || (present != (want > 0))) { r->bad = 1; return NULL; }
    if (!present) return NULL;
    if (r_uint(r, 8) != want) { r->bad = 1; return NULL; }
    return r_take(r, 32 * (size_t)want);
}

int zo_stark_verify(const zo_air *air, const uint8_t *proof, size_t len, zo_channel *chan) {
    const stark_dims D = stark_dims_of(air);
    wire_rd r = {proof, len, 0, 0};
    const uint32_t tl = (uint32_t)r_uint(&r, 1), lb = (uint32_t)r_uint(&r, 1);
    const size_t nq = (size_t)r_uint(&r, 8);
    const uint32_t rl = (uint32_t)r_uint(&r, 1);
    uint32_t fls, flb;
    if (r.bad || tl < 1 || tl + lb > 24 || nq > 4096 || zo_stark_fri_params(air, tl, lb, &fls, &flb) ||
        rl >= tl + lb || rl < flb)
        return 0;
    const size_t n = (size_t)1 << tl, N = (size_t)1 << (tl + lb), shift = (size_t)1 << lb;
    const uint32_t dlog = tl + lb, L = dlog - rl;
    const size_t rc = (size_t)1 << (rl - flb);
    if (r_uint(&r, 8) != D.m) return 0;
    const uint8_t *troots = r_take(&r, 32 * (size_t)D.m);
    const uint8_t *proots = r_opt_digests(&r, D.n_pre), *aroots = r_opt_digests(&r, D.n_rel);
    const uint8_t *qroot = r_take(&r, 32);
    if (r.bad || r_uint(&r, 8) != D.n_deep) return 0;
    zo_qm31 *de = malloc(sizeof(zo_qm31) * D.n_deep), *ga = malloc(sizeof(zo_qm31) * D.n_deep);
    for (uint32_t k = 0; k < D.n_deep; k++) de[k] = r_qm31(&r);
    /* nested Fri.Proof; its parameters must be friParams(params) */
    if (r_uint(&r, 1) != fls || r_uint(&r, 1) != flb || r_uint(&r, 8) != nq || r_uint(&r, 1) != rl ||
        r_uint(&r, 8) != L) { free(de); free(ga); return 0; }
    const uint8_t *froots = r_take(&r, 32 * (size_t)L);
    if (r_uint(&r, 8) != rc) { free(de); free(ga); return 0; }
    uint32_t *frem = malloc(16 * rc), *fvals = malloc(16 * 2 * (size_t)L * nq + 16);
    const size_t per_q = zo_fri_query_path_bytes(dlog, rl);
    uint8_t *fpaths = malloc(per_q * nq + 1);
    size_t *fidx = malloc(sizeof(size_t) * (nq + 1));
    for (size_t k = 0; k < rc; k++) st4(frem + 4 * k, r_qm31(&r));
    if (r_uint(&r, 8) != nq) r.bad = 1;
    for (size_t q = 0; q < nq && !r.bad; q++) {
        fidx[q] = (size_t)r_uint(&r, 8);
        if (r_uint(&r, 8) != L) { r.bad = 1; break; }
        uint8_t *pp = fpaths + q * per_q;
        for (uint32_t i = 0; i < L && !r.bad; i++) {
            st4(fvals + 4 * ((q * L + i) * 2), r_qm31(&r));
            st4(fvals + 4 * ((q * L + i) * 2 + 1), r_qm31(&r));
            for (int s = 0; s < 2 && !r.bad; s++) {
                if (r_uint(&r, 8) != dlog - i) { r.bad = 1; break; }
                const uint8_t *p = r_take(&r, 32 * (size_t)(dlog - i));
                if (p) memcpy(pp, p, 32 * (size_t)(dlog - i));
                pp += 32 * (size_t)(dlog - i);
            }
        }
    }
    int ok = !r.bad;
    zo_qm31 *ac = malloc(sizeof(zo_qm31) * (D.n_lc + 1)), *ar = malloc(sizeof(zo_qm31) * (D.n_rel + 1));
    zo_qm31 *al = malloc(sizeof(zo_qm31) * (D.n_trans_total + 1)), *be = malloc(sizeof(zo_qm31) * (D.total_bound + 1));
    zo_qm31 *pk = malloc(sizeof(zo_qm31) * (D.total_bound + 1)), *bv = malloc(sizeof(zo_qm31) * (D.total_bound + 1));
    uint32_t *bc = malloc(sizeof(uint32_t) * (D.total_bound + 1));
    zo_qm31 *v = malloc(sizeof(zo_qm31) * D.reveal_len), *res = malloc(sizeof(zo_qm31) * (D.T + 1));
    if (ok) {
        const zo_qm31 w = zo_qm31_root_of_unity(tl), w_ev = zo_qm31_root_of_unity(dlog);
        /* replay the transcript (stark.zig:676-705) */
        for (uint32_t k = 0; k < D.n_pre; k++) zo_channel_absorb(chan, proots + 32 * k, 32);
        for (uint32_t j = 0; j < D.m; j++) zo_channel_absorb(chan, troots + 32 * j, 32);
        for (uint32_t j = 0; j < D.n_lc; j++) ac[j] = sample_q(chan);
        for (uint32_t rr = 0; rr < D.n_rel; rr++) ar[rr] = sample_q(chan);
        for (uint32_t rr = 0; rr < D.n_rel; rr++) zo_channel_absorb(chan, aroots + 32 * rr, 32);
        for (uint32_t k = 0; k < D.n_trans_total; k++) al[k] = sample_q(chan);
        for (uint32_t k = 0; k < D.B; k++) { bc[k] = air->boundary[k].column; bv[k] = ld4(air->boundary[k].value);
                                             pk[k] = zo_qm31_pow(w, air->boundary[k].step); }
        for (uint32_t rr = 0; rr < D.n_rel; rr++) { bc[D.B + rr] = D.m + rr; bv[D.B + rr] = Q_ZERO; pk[D.B + rr] = Q_ONE; }
        for (uint32_t k = 0; k < D.total_bound; k++) be[k] = sample_q(chan);
        zo_channel_absorb(chan, qroot, 32);
        const zo_qm31 z = sample_q(chan);
        for (uint32_t k = 0; k < D.n_deep; k++) absorb_qm31(chan, de[k]);
        for (uint32_t k = 0; k < D.n_deep; k++) ga[k] = sample_q(chan);
        ok = zo_fri_verify(dlog, flb, rl, nq, chan, froots, frem, fidx, fvals, fpaths);
        stark_weights wt = {al, be, ac, ar, pk, bv, bc, zo_qm31_pow(w, n - 1)};
        if (ok && r_uint(&r, 8) != nq) ok = 0;
        for (size_t q = 0; q < nq && ok; q++) {
            const size_t p0 = (size_t)r_uint(&r, 8);
            if (p0 != fidx[q] || r_uint(&r, 8) != D.reveal_len) { ok = 0; break; }
            for (uint32_t k = 0; k < D.reveal_len; k++) v[k] = r_qm31(&r);
            if (r_uint(&r, 8) != D.reveal_len) { ok = 0; break; }
            const size_t pn = (p0 + shift) % N;
            /* Merkle checks (stark.zig:716-731) */
            for (uint32_t k = 0; k < D.reveal_len && ok; k++) {
                if (r_uint(&r, 8) != dlog) { ok = 0; break; }
                const uint8_t *path = r_take(&r, 32 * (size_t)dlog), *root;
                size_t at;
                if (k < 2 * D.m_total) {
                    const uint32_t j = k % D.m_total;
                    root = j < D.m ? troots + 32 * j : aroots + 32 * (j - D.m);
                    at = k < D.m_total ? p0 : pn;
                } else if (k == 2 * D.m_total) {
                    root = qroot; at = p0;
                } else {
                    root = proots + 32 * (k - 2 * D.m_total - 1); at = p0;
                }
                uint8_t leaf[32];
                zo_hash_qm31(v[k].v, leaf);
                if (!path || !zo_merkle_verify(root, at, leaf, path, dlog)) ok = 0;
            }
            if (!ok) break;
            const zo_qm31 x0 = zo_qm31_mul(ld4(ZO_FRI_OFFSET), zo_qm31_pow(w_ev, p0));
            /* DEEP identity against the FRI layer-0 leaf (stark.zig:733-751) */
            const zo_qm31 inv_dz = zo_qm31_inv(zo_qm31_sub(x0, z));
            zo_qm31 gv = Q_ZERO;
            for (uint32_t j = 0; j < D.m_total; j++) {
                gv = zo_qm31_add(gv, zo_qm31_mul(ga[j], zo_qm31_mul(zo_qm31_sub(v[j], de[j]), inv_dz)));
                gv = zo_qm31_add(gv, zo_qm31_mul(ga[D.m_total + j],
                                                 zo_qm31_mul(zo_qm31_sub(v[D.m_total + j], de[D.m_total + j]), inv_dz)));
            }
            gv = zo_qm31_add(gv, zo_qm31_mul(ga[2 * D.m_total],
                                             zo_qm31_mul(zo_qm31_sub(v[2 * D.m_total], de[2 * D.m_total]), inv_dz)));
            const size_t p1 = p0 % (N / 2);
            const zo_qm31 leafv = ld4(fvals + 4 * ((q * L) * 2 + (p0 == p1 ? 0 : 1)));
            if (memcmp(gv.v, leafv.v, 16) != 0) { ok = 0; break; }
            /* Hc(x0) == Z_H(x0) * Q(x0) (stark.zig:753-779) */
            const zo_qm31 zh = zo_qm31_sub(zo_qm31_pow(x0, n), Q_ONE);
            const zo_qm31 hv = composition(&D, &wt, x0, v, v + D.m_total, v + 2 * D.m_total + 1, zh, res);
            const zo_qm31 rhs = zo_qm31_mul(zh, v[2 * D.m_total]);
            if (memcmp(hv.v, rhs.v, 16) != 0) ok = 0;
        }
        if (ok && (r.bad || r.pos != r.n)) ok = 0; /* truncated or trailing bytes */
    }
    free(de); free(ga); free(frem); free(fvals); free(fpaths); free(fidx);
    free(ac); free(ar); free(al); free(be); free(pk); free(bv); free(bc); free(v); free(res);
    return ok;
}
