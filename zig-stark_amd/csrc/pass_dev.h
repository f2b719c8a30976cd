// pass_dev.h -- the generic multi-stage LDS pass of the circle FFT (any C), as a
// device body: a workgroup of any size (nthreads) loads its G x Q tiles of 2^C
// elements into LDS, runs the C butterfly stages, and stores them.  fft.hip runs
// it as fft_pass_kernel; the persistent batch kernel runs it as a task.
#pragma once
#include "tile_dev.h"

namespace zs {

template <bool INV>
__device__ __forceinline__ void fft_pass_body(const FftPass &p, const VB vb, uint32_t tid, uint32_t nthreads) {
    extern __shared__ uint32_t lds[];
    const uint32_t C = p.C, T = 1u << C, TP = T + 1;
    const uint32_t G = 1u << p.G_log, Q = 1u << p.Q_log;
    const uint32_t col = vb.y;
    const uint32_t qblk = vb.z << p.Q_log;

    uint32_t hi = 0, lo0 = 0, tau0 = 0;
    if (p.s_lo > 0) {
        uint32_t lo_groups = (1u << p.s_lo) >> p.G_log;
        hi = vb.x / lo_groups;
        lo0 = (vb.x % lo_groups) << p.G_log;
    } else {
        tau0 = vb.x << p.G_log;
    }
    auto fixed = [&](uint32_t g) -> uint32_t {
        if (p.s_lo > 0) return (hi << (p.s_lo + C)) | (lo0 + g);
        uint32_t h = tau0 + g;
        if (p.natural_tiles) h = bitrev32(h, p.L - C);
        return h << C;
    };
    const bool g_fast = (p.s_lo > 0) || p.natural_tiles;
    const uint32_t half_n = 1u << (p.L - 1);

    // ---- load ----
    const uint32_t *in_col = p.in + (uint64_t)col * p.in_cs;
    const uint32_t qload_log = (p.in_mode == IN_SHARED) ? 0 : p.Q_log;
    const uint32_t nload = (G << qload_log) << C;
    for (uint32_t e = tid; e < nload; e += nthreads) {
        uint32_t g, t, ql;
        if (g_fast) {
            g = e & (G - 1);
            uint32_t r = e >> p.G_log;
            ql = r & ((1u << qload_log) - 1);
            t = r >> qload_log;
        } else {
            t = e & (T - 1);
            uint32_t r = e >> C;
            g = r & (G - 1);
            ql = r >> p.G_log;
        }
        uint32_t k = fixed(g) | (t << p.s_lo);
        uint64_t addr;
        switch (p.in_mode) {
            case IN_KSPACE: addr = ((uint64_t)(qblk + ql) << p.L) | k; break;
            case IN_SHARED: addr = k; break;
            case IN_NATURAL: addr = bitrev32(k, p.L); break;
            default: addr = (k & 1u) ? (half_n + (k >> 1)) : (k >> 1); break;
        }
        uint32_t v = ldg(in_col + addr);
        if (p.in_mode == IN_SHARED) {
            for (uint32_t qq = 0; qq < Q; qq++) lds[(qq * G + g) * TP + t] = v;
        } else {
            lds[(ql * G + g) * TP + t] = v;
        }
    }
    __syncthreads();

    // ---- stages ----
    const uint32_t half_pairs = T >> 1;
    const uint32_t npairs = (Q * G) * half_pairs;
    for (uint32_t si = 0; si < C; si++) {
        const uint32_t ls = INV ? si : (C - 1 - si);
        const uint32_t s = p.s_lo + ls;
        const uint32_t *tws = p.tw + (1u << (p.LT - 1 - s));
        for (uint32_t b = tid; b < npairs; b += nthreads) {
            uint32_t j = b & (half_pairs - 1);
            uint32_t tile = b >> (C - 1);
            uint32_t g = tile & (G - 1), ql = tile >> p.G_log;
            uint32_t t0 = ((j >> ls) << (ls + 1)) | (j & ((1u << ls) - 1));
            uint32_t t1 = t0 | (1u << ls);
            uint32_t K = ((qblk + ql) << p.L) | fixed(g) | (t0 << p.s_lo);
            uint32_t w = ldg(tws + (K >> (s + 1)));
            uint32_t *base = lds + tile * TP;
            uint32_t v0 = base[t0], v1 = base[t1];
            if (INV) {
                base[t0] = m31_add(v0, v1);
                base[t1] = m31_mul(m31_sub(v0, v1), w);
            } else {
                uint32_t m = m31_mul(v1, w);
                base[t0] = m31_add(v0, m);
                base[t1] = m31_sub(v0, m);
            }
        }
        __syncthreads();
    }

    // ---- store ----
    uint32_t *out_col = p.out + (uint64_t)col * p.out_cs;
    const uint32_t nstore = (G * Q) << C;
    for (uint32_t e = tid; e < nstore; e += nthreads) {
        uint32_t g, t, ql;
        if (p.out_mode == OUT_NATURAL) {
            uint32_t rq = e & (Q - 1);
            g = (e >> p.Q_log) & (G - 1);
            t = e >> (p.Q_log + p.G_log);
            ql = bitrev32(rq, p.Q_log);
        } else if (g_fast) {
            g = e & (G - 1);
            uint32_t r = e >> p.G_log;
            ql = r & (Q - 1);
            t = r >> p.Q_log;
        } else {
            t = e & (T - 1);
            uint32_t r = e >> C;
            g = r & (G - 1);
            ql = r >> p.G_log;
        }
        uint32_t v = lds[(ql * G + g) * TP + t];
        if (p.scale) v = m31_mul(v, p.scale);
        uint32_t k = fixed(g) | (t << p.s_lo);
        uint32_t q = qblk + ql;
        uint64_t addr;
        switch (p.out_mode) {
            case OUT_KSPACE: addr = ((uint64_t)q << p.L) | k; break;
            case OUT_NATURAL: addr = ((uint64_t)bitrev32(k, p.L) << p.beta) | bitrev32(q, p.beta); break;
            default: addr = (k & 1u) ? (half_n + (k >> 1)) : (k >> 1); break;
        }
        stg(out_col + addr, v);
    }
}

}  // namespace zs
