// air.hip -- GenericStark(Air).proveWithPreprocessed for an AIR given as data, on
// gfx950 (src/m31/stark.zig:311-657): preprocessed columns, the synthesized LogUp
// accumulators (stark.zig:416-468), the quotient with every transition,
// lookup, selector-binarity and boundary term (:485-531) and the DEEP codeword
// (:568-580) on the device, and so are the evaluations of every committed
// column at z and w z (from interpolation coefficients kept on the device); the
// transcript, the composition at z and the proof bytes on the host.
//
// Inversions: the reference inverts per element (alpha - key, alpha - tkey per row
// and relation; x - p_k per boundary assertion and x - z per point of D).  Here
// the accumulator contributions and the DEEP codeword batch their inversions over
// a run of consecutive rows or points per thread (one QM31 inversion and three
// products per value, an input of zero mapped to zero as QM31.inv does), and the
// quotient sums its boundary terms as one fraction num/den, inverted in the same
// batches.  That fraction equals the reference's sum because D and H are disjoint
// (Z_H, checked nonzero on D on the host, vanishes on H), so no x - p_k is zero.
// Q = c Z_H^-1 + num/den is Hc Z_H^-1 expanded: the same field element.
#include <cstring>
#include <memory>
#include <vector>

#include "air.h"

namespace zs {

int32_t air_from_abi(const zs_air *a, AirDesc *d) {
    if (!a) return ZS_ERR_INPUT;
    std::memset(d, 0, sizeof *d);
    d->m = a->num_columns;
    d->n_pre = a->num_preprocessed;
    d->n_rel = a->num_lookup_relations;
    d->n_lc = a->num_lookup_columns;
    d->T = a->num_transition_constraints;
    d->B = a->num_boundary;
    d->n_terms = a->num_terms;
    d->degree_mul = a->degree_mul;
    d->degree_add = a->degree_add;
    d->m_total = d->m + d->n_rel;
    d->n_trans_total = d->T + 2 * d->n_rel;
    d->total_bound = d->B + d->n_rel;
    d->n_deep = 2 * d->m_total + 1;
    d->reveal_len = 2 * d->m_total + 1 + d->n_pre;
    if (d->m == 0 || d->m_total + d->n_pre > AIR_MAX_COLS || d->n_terms > AIR_MAX_TERMS || d->T > AIR_MAX_TRANS ||
        d->n_rel > AIR_MAX_REL || d->B > AIR_MAX_BOUND || d->n_lc > AIR_MAX_KEYS)
        return ZS_ERR_INPUT;
    if ((d->n_terms && !a->terms) || (d->B && !a->boundary)) return ZS_ERR_INPUT;
    for (uint32_t t = 0; t < d->n_terms; t++) {
        const zs_air_term &tm = a->terms[t];
        if (tm.constraint >= d->T || tm.n_factors > 4) return ZS_ERR_INPUT;
        d->term_con[t] = tm.constraint;
        d->term_nf[t] = tm.n_factors;
        uint32_t c[4];
        for (int k = 0; k < 4; k++) c[k] = m31_from_u32(tm.coeff[k]);
        d->term_coeff[t] = qm_from_words(c);
        for (uint32_t f = 0; f < tm.n_factors; f++) {
            const zs_air_factor fc = tm.factor[f];
            if (fc.kind > ZS_AIR_X || (fc.kind != ZS_AIR_X && fc.column >= d->m)) return ZS_ERR_INPUT;
            d->term_kind[t][f] = fc.kind;
            d->term_col[t][f] = fc.kind == ZS_AIR_X ? 0 : fc.column;
        }
    }
    for (uint32_t k = 0; k < d->B; k++) {
        const zs_boundary &b = a->boundary[k];
        if (b.column >= d->m) return ZS_ERR_INPUT;
        d->bcol[k] = b.column;
        d->bstep[k] = b.step;
        uint32_t v[4];
        for (int i = 0; i < 4; i++) v[i] = m31_from_u32(b.value[i]);
        d->bval[k] = qm_from_words(v);
    }
    for (uint32_t r = 0; r < d->n_rel; r++) {  // acc_r[0] = 0 (stark.zig:479-481)
        d->bcol[d->B + r] = d->m + r;
        d->bstep[d->B + r] = 0;
        d->bval[d->B + r] = {{0, 0}, {0, 0}};
    }
    if (d->n_rel) {
        if (d->n_pre == 0 || d->n_lc == 0 || !a->selector_column || !a->multiplicity_column || !a->key_start ||
            !a->table_start || !a->key_columns || !a->table_columns)
            return ZS_ERR_INPUT;  // the reference's comptime checks (stark.zig:157-172)
        uint32_t kp = 0, tp = 0;
        for (uint32_t r = 0; r <= d->n_rel; r++) {
            d->key_s[r] = a->key_start[r];
            d->tab_s[r] = a->table_start[r];
        }
        if (d->key_s[0] != 0 || d->tab_s[0] != 0) return ZS_ERR_INPUT;
        for (uint32_t r = 0; r < d->n_rel; r++) {
            if (a->selector_column[r] >= d->m || a->multiplicity_column[r] >= d->m) return ZS_ERR_INPUT;
            d->sel[r] = a->selector_column[r];
            d->mult[r] = a->multiplicity_column[r];
            const uint32_t nk = d->key_s[r + 1] - d->key_s[r], nt = d->tab_s[r + 1] - d->tab_s[r];
            if (d->key_s[r + 1] < d->key_s[r] || d->tab_s[r + 1] < d->tab_s[r] || nk > d->n_lc || nt > d->n_lc)
                return ZS_ERR_INPUT;
            for (uint32_t j = d->key_s[r]; j < d->key_s[r + 1]; j++) {
                if (kp >= AIR_MAX_REL * AIR_MAX_KEYS || a->key_columns[j] >= d->m) return ZS_ERR_INPUT;
                d->key_c[kp++] = a->key_columns[j];
            }
            for (uint32_t j = d->tab_s[r]; j < d->tab_s[r + 1]; j++) {
                if (tp >= AIR_MAX_REL * AIR_MAX_KEYS || a->table_columns[j] >= d->n_pre) return ZS_ERR_INPUT;
                d->tab_c[tp++] = a->table_columns[j];
            }
        }
    }
    return ZS_OK;
}

// maxConstraintDegree of the AIR and of GenericStark (stark.zig:212-216)
static uint64_t air_max_degree(const AirDesc &d, uint64_t n) {
    uint64_t air = 0;
    if (d.degree_mul || d.degree_add) {
        air = uint64_t(d.degree_mul) * (n - 1) + d.degree_add;
    } else {
        for (uint32_t t = 0; t < d.n_terms; t++) {
            uint64_t dg = 0;
            for (uint32_t f = 0; f < d.term_nf[t]; f++) dg += d.term_kind[t][f] == ZS_AIR_X ? 1 : n - 1;
            air = std::max(air, dg);
        }
    }
    return std::max<uint64_t>(air, d.n_rel ? 3 * (n - 1) : 0);
}

bool air_fri_params(const AirDesc &d, uint32_t trace_log, uint32_t log_blowup, uint32_t *fls, uint32_t *flb) {
    const uint64_t n = uint64_t(1) << trace_log, dc = air_max_degree(d, n);
    const uint64_t dg = std::max<uint64_t>(n - 1, dc > n ? dc - n : 0);  // degree_g (stark.zig:199-200)
    uint32_t log = 0;
    for (uint64_t bound = 1; bound <= dg; bound <<= 1) log++;
    if (log > trace_log + log_blowup) return false;
    *fls = log;
    *flb = trace_log + log_blowup - log;
    return true;
}

namespace {

constexpr uint32_t kRowsPerThread = 8;   // accumulator contributions
constexpr uint32_t kPointsPerThread = 4; // quotient and DEEP values

__device__ __forceinline__ qm31 ldq(const uint32_t *p) {
    uint4 v = *reinterpret_cast<const uint4 *>(p);
    return {{v.x, v.y}, {v.z, v.w}};
}
__device__ __forceinline__ void stq(uint32_t *p, qm31 x) {
    *reinterpret_cast<uint4 *>(p) = make_uint4(x.a.c0, x.a.c1, x.b.c0, x.b.c1);
}
__device__ __forceinline__ bool qm_is_zero(qm31 x) { return (x.a.c0 | x.a.c1 | x.b.c0 | x.b.c1) == 0; }

// v[k] <- v[k]^-1 for k < cnt with one inversion (Montgomery's trick); a zero stays
// zero, as QM31.inv maps zero to zero
template <int K>
__device__ __forceinline__ void batch_inv(qm31 (&v)[K], uint32_t cnt) {
    const qm31 one = {{1, 0}, {0, 0}};
    qm31 pre[K];
    qm31 acc = one;
#pragma unroll
    for (int k = 0; k < K; k++) {
        pre[k] = acc;
        if (k < (int)cnt && !qm_is_zero(v[k])) acc = qm_mul(acc, v[k]);
    }
    qm31 inv = qm_inv(acc);
#pragma unroll
    for (int k = K - 1; k >= 0; k--) {
        if (k >= (int)cnt || qm_is_zero(v[k])) continue;
        const qm31 vi = qm_mul(inv, pre[k]);
        inv = qm_mul(inv, v[k]);
        v[k] = vi;
    }
}

// rowContribution (stark.zig:260-275) of rows [i0, i0 + kRowsPerThread) for every relation
__global__ __launch_bounds__(256) void logup_contrib_kernel(const AirDesc *__restrict__ dp,
                                                            const AirWeights *__restrict__ wp, const uint32_t *trace,
                                                            const uint32_t *pre, uint32_t n, uint32_t *contrib) {
    const AirDesc &d = *dp;
    const AirWeights &w = *wp;
    const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * kRowsPerThread;
    if (i0 >= n) return;
    const uint32_t cnt = min(kRowsPerThread, n - i0);
    const qm31 one = {{1, 0}, {0, 0}}, zero = {{0, 0}, {0, 0}};
    for (uint32_t r = 0; r < d.n_rel; r++) {
        qm31 den[2 * kRowsPerThread];
        for (uint32_t k = 0; k < kRowsPerThread; k++) {
            qm31 key = zero, tkey = zero;
            if (k < cnt) {
                const uint32_t i = i0 + k;
                for (uint32_t j = d.key_s[r]; j < d.key_s[r + 1]; j++)
                    key = qm_add(key, qm_mul(w.alpha_cols[j - d.key_s[r]], ldq(trace + 4 * ((uint64_t)d.key_c[j] * n + i))));
                for (uint32_t j = d.tab_s[r]; j < d.tab_s[r + 1]; j++)
                    tkey = qm_add(tkey, qm_mul(w.alpha_cols[j - d.tab_s[r]], ldq(pre + 4 * ((uint64_t)d.tab_c[j] * n + i))));
            }
            den[2 * k] = qm_sub(w.alpha_rels[r], key);
            den[2 * k + 1] = qm_sub(w.alpha_rels[r], tkey);
        }
        batch_inv(den, 2 * cnt);
        for (uint32_t k = 0; k < cnt; k++) {
            const uint32_t i = i0 + k;
            const qm31 sel = ldq(trace + 4 * ((uint64_t)d.sel[r] * n + i));
            const qm31 mult = ldq(trace + 4 * ((uint64_t)d.mult[r] * n + i));
            const qm31 lt = qm_mul(sel, qm_mul(mult, den[2 * k]));
            const qm31 tt = qm_mul(qm_sub(one, sel), den[2 * k + 1]);
            stq(contrib + 4 * ((uint64_t)r * n + i), qm_sub(lt, tt));
        }
    }
}

// acc_r[0] = 0, acc_r[i] = acc_r[i-1] + contribution_r(i-1) (stark.zig:436-449): an
// exclusive prefix sum per relation in three launches -- tile sums, an exclusive
// scan of the tile sums, each tile rescanned from its offset.  Field addition is
// exact, so any grouping gives the reference's values.
constexpr uint32_t kScanThreads = 256, kScanPer = 8, kScanTile = kScanThreads * kScanPer;
constexpr uint32_t kScanSumsThreads = 1024;

// the sum of tile b of relation r: thread t adds elements t, t + 256, ... of the tile
__global__ __launch_bounds__(kScanThreads) void acc_tile_sums_kernel(const uint32_t *contrib, uint32_t n,
                                                                     uint32_t nb, uint32_t *sums) {
    __shared__ uint32_t part[4 * kScanThreads];
    const uint32_t r = blockIdx.y, b = blockIdx.x, t = threadIdx.x;
    const uint32_t *c = contrib + 4 * (uint64_t)r * n;
    qm31 s = {{0, 0}, {0, 0}};
    for (uint32_t k = 0; k < kScanPer; k++) {
        const uint64_t i = (uint64_t)b * kScanTile + k * kScanThreads + t;
        if (i < n) s = qm_add(s, ldq(c + 4 * i));
    }
    stq(part + 4 * t, s);
    __syncthreads();
    for (uint32_t h = kScanThreads / 2; h > 0; h >>= 1) {
        if (t < h) stq(part + 4 * t, qm_add(ldq(part + 4 * t), ldq(part + 4 * (t + h))));
        __syncthreads();
    }
    if (t == 0) stq(sums + 4 * ((uint64_t)r * nb + b), ldq(part));
}

// sums[r][b] <- the sum of tiles 0 .. b-1 of relation r, one workgroup per relation
__global__ __launch_bounds__(kScanSumsThreads) void acc_scan_sums_kernel(uint32_t *sums, uint32_t nb) {
    __shared__ uint32_t run[4 * kScanSumsThreads];
    const uint32_t r = blockIdx.x, t = threadIdx.x;
    const uint32_t per = (nb + kScanSumsThreads - 1) / kScanSumsThreads;
    const uint32_t b0 = min(nb, t * per), b1 = min(nb, b0 + per);
    uint32_t *s = sums + 4 * (uint64_t)r * nb;
    qm31 v = {{0, 0}, {0, 0}};
    for (uint32_t i = b0; i < b1; i++) v = qm_add(v, ldq(s + 4 * i));
    stq(run + 4 * t, v);
    __syncthreads();
    for (uint32_t off = 1; off < kScanSumsThreads; off <<= 1) {  // inclusive scan of the thread sums
        qm31 u = ldq(run + 4 * t);
        if (t >= off) u = qm_add(u, ldq(run + 4 * (t - off)));
        __syncthreads();
        stq(run + 4 * t, u);
        __syncthreads();
    }
    qm31 acc = t ? ldq(run + 4 * (t - 1)) : qm31{{0, 0}, {0, 0}};
    for (uint32_t i = b0; i < b1; i++) {
        const qm31 x = ldq(s + 4 * i);
        stq(s + 4 * i, acc);
        acc = qm_add(acc, x);
    }
}

// tile b of relation r: thread t holds elements [8t, 8t + 8) of the tile; their
// exclusive prefix within the tile plus the tile's offset
__global__ __launch_bounds__(kScanThreads) void acc_apply_kernel(const uint32_t *contrib, uint32_t n,
                                                                 const uint32_t *offs, uint32_t nb, uint32_t *acc) {
    __shared__ uint32_t part[4 * kScanThreads];
    const uint32_t r = blockIdx.y, b = blockIdx.x, t = threadIdx.x;
    const uint64_t i0 = (uint64_t)b * kScanTile + (uint64_t)t * kScanPer;
    const uint32_t *c = contrib + 4 * (uint64_t)r * n;
    qm31 v[kScanPer];
    qm31 s = {{0, 0}, {0, 0}};
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; k++) {
        v[k] = i0 + k < n ? ldq(c + 4 * (i0 + k)) : qm31{{0, 0}, {0, 0}};
        s = qm_add(s, v[k]);
    }
    stq(part + 4 * t, s);
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        qm31 u = ldq(part + 4 * t);
        if (t >= off) u = qm_add(u, ldq(part + 4 * (t - off)));
        __syncthreads();
        stq(part + 4 * t, u);
        __syncthreads();
    }
    qm31 run = ldq(offs + 4 * ((uint64_t)r * nb + b));
    if (t) run = qm_add(run, ldq(part + 4 * (t - 1)));
    uint32_t *a = acc + 4 * (uint64_t)r * n;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; k++) {
        if (i0 + k < n) stq(a + 4 * (i0 + k), run);
        run = qm_add(run, v[k]);
    }
}

// UnivariateQM31.eval (sum_i c_i x^i) of column cols[e] at point pts[e], for every
// pair e = blockIdx.y: thread t of tile b adds c_i x^i for i = tile start + t + 256 k;
// one partial per tile, summed on the host (field addition: exact in any order)
constexpr uint32_t kEvalThreads = 256, kEvalPer = 16, kEvalTile = kEvalThreads * kEvalPer;
struct EvalJob {
    uint32_t col[2 * AIR_MAX_COLS + 1], pt[2 * AIR_MAX_COLS + 1];
    qm31 x[2], x_step[2];  // the points and their 256th powers
};
__global__ __launch_bounds__(kEvalThreads) void poly_eval_kernel(const uint32_t *co, uint32_t n, const EvalJob job,
                                                                 uint32_t nb, uint32_t *partial) {
    __shared__ uint32_t part[4 * kEvalThreads];
    const uint32_t e = blockIdx.y, b = blockIdx.x, t = threadIdx.x;
    const uint32_t *c = co + 4 * (uint64_t)job.col[e] * n;
    const uint64_t i0 = (uint64_t)b * kEvalTile + t;
    const qm31 step = job.x_step[job.pt[e]];
    qm31 p = qm_pow(job.x[job.pt[e]], i0), s = {{0, 0}, {0, 0}};
    for (uint32_t k = 0; k < kEvalPer; k++) {
        const uint64_t i = i0 + (uint64_t)k * kEvalThreads;
        if (i >= n) break;
        s = qm_add(s, qm_mul(ldq(c + 4 * i), p));
        p = qm_mul(p, step);
    }
    stq(part + 4 * t, s);
    __syncthreads();
    for (uint32_t h = kEvalThreads / 2; h > 0; h >>= 1) {
        if (t < h) stq(part + 4 * t, qm_add(ldq(part + 4 * t), ldq(part + 4 * (t + h))));
        __syncthreads();
    }
    if (t == 0) stq(partial + 4 * ((uint64_t)e * nb + b), ldq(part));
}

struct AirDomain {
    qm31 offset, w_ev;
    qm31 wpow[28];   // w_ev^(2^k)
    uint32_t dlog;   // |D| = 2^dlog
    uint32_t shift;  // 2^log_blowup: w x_i = x_(i + shift)
};
__device__ __forceinline__ qm31 domain_x(const AirDomain &dm, uint32_t i) {
    qm31 x = dm.offset;
    for (uint32_t k = 0; k < dm.dlog; k++)
        if ((i >> k) & 1u) x = qm_mul(x, dm.wpow[k]);
    return x;
}

// Q on D (stark.zig:485-531): cw holds the m_total trace and accumulator codewords
// and pcw the preprocessed ones (column c at 4 N c); zh_inv[r] = Z_H^-1 at x_i, i = r mod shift
__global__ __launch_bounds__(256) void air_quotient_kernel(const AirDesc *__restrict__ dp,
                                                           const AirWeights *__restrict__ wp, const uint32_t *cw,
                                                           const uint32_t *pcw, const uint32_t *zh_inv,
                                                           const AirDomain dm, uint32_t *q_out) {
    const AirDesc &d = *dp;
    const AirWeights &w = *wp;
    const uint32_t N = 1u << dm.dlog;
    const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * kPointsPerThread;
    if (i0 >= N) return;
    const uint32_t cnt = min(kPointsPerThread, N - i0);
    qm31 x = domain_x(dm, i0);
    qm31 base[kPointsPerThread], num[kPointsPerThread], den[kPointsPerThread];
    for (uint32_t k = 0; k < kPointsPerThread; k++) {
        if (k >= cnt) {
            den[k] = {{1, 0}, {0, 0}};
            continue;
        }
        const uint32_t i = i0 + k, ni = (i + dm.shift) & (N - 1);
        auto val = [&](uint32_t kind, uint32_t col) {
            if (kind == AIR_PRE) return ldq(pcw + 4 * ((uint64_t)col * N + i));
            return ldq(cw + 4 * ((uint64_t)col * N + (kind == ZS_AIR_NEXT ? ni : i)));
        };
        base[k] = qm_mul(air_constraints(d, w, x, val), ldq(zh_inv + 4 * (i & (dm.shift - 1))));
        // sum_b beta_b (f_b(x) - v_b) / (x - p_b) as one fraction
        qm31 nu = {{0, 0}, {0, 0}}, de = {{1, 0}, {0, 0}};
        for (uint32_t b = 0; b < d.total_bound; b++) {
            const qm31 t = qm_mul(w.beta[b], qm_sub(val(ZS_AIR_CUR, d.bcol[b]), d.bval[b]));
            const qm31 xp = qm_sub(x, w.pk[b]);
            nu = qm_add(qm_mul(nu, xp), qm_mul(t, de));
            de = qm_mul(de, xp);
        }
        num[k] = nu;
        den[k] = de;
        x = qm_mul(x, dm.w_ev);
    }
    batch_inv(den, cnt);
    for (uint32_t k = 0; k < cnt; k++) stq(q_out + 4 * (uint64_t)(i0 + k), qm_add(base[k], qm_mul(num[k], den[k])));
}

struct DeepConst {
    qm31 z;
    qm31 de[2 * AIR_MAX_COLS + 1], gamma[2 * AIR_MAX_COLS + 1];
};

// g on D (stark.zig:568-580): cwq holds the m_total codewords and then Q
__global__ __launch_bounds__(256) void air_deep_kernel(const DeepConst *__restrict__ cp, const uint32_t *cwq,
                                                       uint32_t m_total, const AirDomain dm, uint32_t *g_out) {
    const DeepConst &c = *cp;
    const uint32_t N = 1u << dm.dlog;
    const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * kPointsPerThread;
    if (i0 >= N) return;
    const uint32_t cnt = min(kPointsPerThread, N - i0);
    qm31 x = domain_x(dm, i0);
    qm31 s[kPointsPerThread], dz[kPointsPerThread];
    for (uint32_t k = 0; k < kPointsPerThread; k++) {
        if (k >= cnt) {
            dz[k] = {{1, 0}, {0, 0}};
            continue;
        }
        const uint32_t i = i0 + k, ni = (i + dm.shift) & (N - 1);
        qm31 a = {{0, 0}, {0, 0}};
        for (uint32_t j = 0; j < m_total; j++) {
            a = qm_add(a, qm_mul(c.gamma[j], qm_sub(ldq(cwq + 4 * ((uint64_t)j * N + i)), c.de[j])));
            a = qm_add(a, qm_mul(c.gamma[m_total + j], qm_sub(ldq(cwq + 4 * ((uint64_t)j * N + ni)), c.de[m_total + j])));
        }
        a = qm_add(a, qm_mul(c.gamma[2 * m_total], qm_sub(ldq(cwq + 4 * ((uint64_t)m_total * N + i)), c.de[2 * m_total])));
        s[k] = a;
        dz[k] = qm_sub(x, c.z);
        x = qm_mul(x, dm.w_ev);
    }
    batch_inv(dz, cnt);
    for (uint32_t k = 0; k < cnt; k++) stq(g_out + 4 * (uint64_t)(i0 + k), qm_mul(s[k], dz[k]));
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
struct Wire {
    std::vector<uint8_t> b;
    void bytes(const void *d, size_t n) {
        const uint8_t *p = static_cast<const uint8_t *>(d);
        b.insert(b.end(), p, p + n);
    }
    void uint(uint64_t v, unsigned n) {
        for (unsigned k = 0; k < n; k++) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void qm(qm31 x) {
        uint32_t w[4];
        qm_to_words(x, w);
        for (int k = 0; k < 4; k++) uint(w[k], 4);
    }
};

}  // namespace
}  // namespace zs

using namespace zs;

extern "C" int32_t zs_stark_prove_dev(zs_hip_ctx *h, const zs_air *air, const uint32_t *d_pre,
                                      const uint32_t *d_trace, uint32_t trace_log, uint32_t log_blowup,
                                      size_t num_queries, uint32_t remainder_log, zs_channel *chan,
                                      zs_host_allocator host, uint8_t **proof_out, size_t *proof_len) {
    if (!h || !air || !d_trace || !chan || !host.alloc || !proof_out || !proof_len || trace_log < 1 ||
        log_sum_over(trace_log, log_blowup, 24) || num_queries > 4096)
        return ZS_ERR_INPUT;
    auto dsc = std::make_unique<AirDesc>();
    AirDesc &d = *dsc;
    ZS_TRY(air_from_abi(air, &d));
    if (d.n_pre && !d_pre) return ZS_ERR_INPUT;
    uint32_t fls, flb;
    const uint32_t dlog = trace_log + log_blowup;
    if (!air_fri_params(d, trace_log, log_blowup, &fls, &flb) || remainder_log >= dlog || remainder_log < flb ||
        remainder_log - flb > kMaxRemainderLog)
        return ZS_ERR_INPUT;  // friParams' and proveCodeword's assertions (stark.zig:208-222, fri.zig:221-227)
    Ctx *ctx = &h->c;
    ZS_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = size_t(1) << trace_log, N = size_t(1) << dlog, shift = size_t(1) << log_blowup;
    const qm31 one = {{1, 0}, {0, 0}}, offset = {{1, 0}, {1, 0}};  // FRI_OFFSET = 1 + u (fri.zig:36)
    const qm31 w = qm31_root_of_unity(trace_log), w_ev = qm31_root_of_unity(dlog);
    const uint32_t mt = d.m_total, ncw = mt + 1 + d.n_pre;  // codewords: trace | acc | Q | pre
    const size_t tw = tree_words(dlog);
    DevBuf cwb, trees, g, zhb, contrib, acc, dconst, wconst, deep, coef, sums;
    ZS_HIP(ctx, hipMalloc(&cwb.ptr, ncw * N * 16));
    ZS_HIP(ctx, hipMalloc(&coef.ptr, size_t(mt + d.n_pre) * n * 16));
    ZS_HIP(ctx, hipMalloc(&trees.ptr, ncw * tw * 4));
    ZS_HIP(ctx, hipMalloc(&g.ptr, N * 16));
    ZS_HIP(ctx, hipMalloc(&zhb.ptr, shift * 16));
    ZS_HIP(ctx, hipMalloc(&dconst.ptr, sizeof(AirDesc)));
    ZS_HIP(ctx, hipMalloc(&wconst.ptr, sizeof(AirWeights)));
    ZS_HIP(ctx, hipMalloc(&deep.ptr, sizeof(DeepConst)));
    const uint32_t scan_nb = (uint32_t)((n + kScanTile - 1) / kScanTile);
    if (d.n_rel) {
        ZS_HIP(ctx, hipMalloc(&contrib.ptr, d.n_rel * n * 16));
        ZS_HIP(ctx, hipMalloc(&acc.ptr, d.n_rel * n * 16));
        ZS_HIP(ctx, hipMalloc(&sums.ptr, size_t(d.n_rel) * scan_nb * 16));
    }
    uint32_t *cw = cwb.as<uint32_t>(), *tr = trees.as<uint32_t>();
    uint32_t *q_cw = cw + 4 * N * mt, *pre_cw = cw + 4 * N * (mt + 1);
    uint32_t *q_tree = tr + tw * mt, *pre_tree = tr + tw * (mt + 1);
    ZS_HIP(ctx, hipMemcpyAsync(dconst.ptr, &d, sizeof(AirDesc), hipMemcpyHostToDevice, ctx->stream));
    // n * (the interpolation coefficients) of every committed column, kept on the
    // device for the evaluations at z: [trace | acc | pre], column c at 4 n c.  The
    // trace commitment leaves them in scratch 0; copied out before the next one.
    auto take_coeffs = [&](size_t first_col, uint32_t cols) -> int32_t {
        ZS_HIP(ctx, hipMemcpyAsync(coef.as<uint32_t>() + 4 * first_col * n, ctx->scratch[0].ptr,
                                   size_t(cols) * n * 16, hipMemcpyDeviceToDevice, ctx->stream));
        return ZS_OK;
    };
    // 1. preprocessed commitment, absorbed first, then the trace (stark.zig:374-414)
    std::vector<uint8_t> proots(32 * d.n_pre + 1), roots(32 * mt + 1);
    if (d.n_pre) {
        ZS_TRY(zs_qm31_trace_commit_dev(h, d_pre, d.n_pre, trace_log, log_blowup, pre_cw,
                                        reinterpret_cast<uint8_t *>(pre_tree), proots.data(), chan));
        ZS_TRY(take_coeffs(mt, d.n_pre));
    }
    ZS_TRY(zs_qm31_trace_commit_dev(h, d_trace, d.m, trace_log, log_blowup, cw, reinterpret_cast<uint8_t *>(tr),
                                    roots.data(), chan));
    ZS_TRY(take_coeffs(0, d.m));
    // 2. lookup challenges, accumulators on H and their commitment (stark.zig:416-468)
    auto wt = std::make_unique<AirWeights>();
    for (uint32_t j = 0; j < d.n_lc; j++) wt->alpha_cols[j] = sample(chan);
    for (uint32_t r = 0; r < d.n_rel; r++) wt->alpha_rels[r] = sample(chan);
    if (d.n_rel) {
        ZS_HIP(ctx, hipMemcpyAsync(wconst.ptr, wt.get(), sizeof(AirWeights), hipMemcpyHostToDevice, ctx->stream));
        const uint32_t threads = (uint32_t)((n + kRowsPerThread - 1) / kRowsPerThread);
        ZS_LAUNCH(ctx, "logup_contrib_kernel", logup_contrib_kernel, dim3((threads + 255) / 256), dim3(256), 0,
                  ctx->stream, dconst.as<AirDesc>(), wconst.as<AirWeights>(), d_trace, d_pre, (uint32_t)n,
                  contrib.as<uint32_t>());
        ZS_HIP(ctx, hipGetLastError());
        ZS_LAUNCH(ctx, "acc_tile_sums_kernel", acc_tile_sums_kernel, dim3(scan_nb, d.n_rel), dim3(kScanThreads), 0,
                  ctx->stream, contrib.as<uint32_t>(), (uint32_t)n, scan_nb, sums.as<uint32_t>());
        ZS_HIP(ctx, hipGetLastError());
        ZS_LAUNCH(ctx, "acc_scan_sums_kernel", acc_scan_sums_kernel, dim3(d.n_rel), dim3(kScanSumsThreads), 0,
                  ctx->stream, sums.as<uint32_t>(), scan_nb);
        ZS_HIP(ctx, hipGetLastError());
        ZS_LAUNCH(ctx, "acc_apply_kernel", acc_apply_kernel, dim3(scan_nb, d.n_rel), dim3(kScanThreads), 0,
                  ctx->stream, contrib.as<uint32_t>(), (uint32_t)n, sums.as<uint32_t>(), scan_nb,
                  acc.as<uint32_t>());
        ZS_HIP(ctx, hipGetLastError());
        ZS_TRY(zs_qm31_trace_commit_dev(h, acc.as<uint32_t>(), d.n_rel, trace_log, log_blowup, cw + 4 * N * d.m,
                                        reinterpret_cast<uint8_t *>(tr + tw * d.m), roots.data() + 32 * d.m, chan));
        ZS_TRY(take_coeffs(d.m, d.n_rel));
    }
    // 3. transition weights, boundary assertions, boundary weights (stark.zig:470-483)
    for (uint32_t k = 0; k < d.n_trans_total; k++) wt->alpha[k] = sample(chan);
    for (uint32_t k = 0; k < d.total_bound; k++) wt->pk[k] = qm_pow(w, d.bstep[k]);
    for (uint32_t k = 0; k < d.total_bound; k++) wt->beta[k] = sample(chan);
    wt->last = qm_pow(w, n - 1);
    ZS_HIP(ctx, hipMemcpyAsync(wconst.ptr, wt.get(), sizeof(AirWeights), hipMemcpyHostToDevice, ctx->stream));
    // Z_H(x_i)^-1: x_i^n = OFFSET^n (w_ev^n)^i takes 2^log_blowup values
    std::vector<uint32_t> zh_inv(shift * 4);
    {
        const qm31 wev_n = qm_pow(w_ev, n);
        qm31 t = qm_pow(offset, n);
        for (size_t r = 0; r < shift; r++) {
            const qm31 zh = qm_sub(t, one);
            if (zh.a.c0 == 0 && zh.a.c1 == 0 && zh.b.c0 == 0 && zh.b.c1 == 0) {
                ctx->set_error("Z_H vanishes on the evaluation domain");
                return ZS_ERR_PROTOCOL;
            }
            qm_to_words(qm_inv(zh), &zh_inv[4 * r]);
            t = qm_mul(t, wev_n);
        }
    }
    ZS_HIP(ctx, hipMemcpyAsync(zhb.ptr, zh_inv.data(), zh_inv.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    AirDomain dm{};
    dm.offset = offset;
    dm.w_ev = w_ev;
    qm31 wp = w_ev;
    for (int k = 0; k < 28; k++) {
        dm.wpow[k] = wp;
        wp = qm_mul(wp, wp);
    }
    dm.dlog = dlog;
    dm.shift = (uint32_t)shift;
    // 4. quotient codeword and its tree (stark.zig:485-538)
    const uint32_t pthreads = (uint32_t)((N + kPointsPerThread - 1) / kPointsPerThread);
    ZS_LAUNCH(ctx, "air_quotient_kernel", air_quotient_kernel, dim3((pthreads + 255) / 256), dim3(256), 0,
              ctx->stream, dconst.as<AirDesc>(), wconst.as<AirWeights>(), cw, pre_cw, zhb.as<uint32_t>(), dm, q_cw);
    ZS_HIP(ctx, hipGetLastError());
    ZS_TRY(launch_qm31_leaves(ctx, q_cw, N, q_tree));
    ZS_TRY(launch_tree_levels(ctx, q_tree, dlog, 0));
    uint8_t qroot[32];
    ZS_HIP(ctx, hipMemcpyAsync(qroot, q_tree + tw - 8, 32, hipMemcpyDeviceToHost, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    zs_channel_absorb(chan, qroot, 32);
    // 5. z, DEEP evaluations, Q(z) (computeQuotientAt, stark.zig:793-846), DEEP weights
    auto dc = std::make_unique<DeepConst>();
    dc->z = sample(chan);
    const qm31 wz = qm_mul(dc->z, w);
    std::vector<qm31> pz(d.n_pre + 1);
    {
        // every trace and accumulator column at z and w z, the preprocessed ones at z
        const uint32_t ne = 2 * mt + d.n_pre, eval_nb = (uint32_t)((n + kEvalTile - 1) / kEvalTile);
        EvalJob job{};
        for (uint32_t j = 0; j < mt; j++) {
            job.col[j] = job.col[mt + j] = j;
            job.pt[mt + j] = 1;
        }
        for (uint32_t j = 0; j < d.n_pre; j++) job.col[2 * mt + j] = mt + j;
        job.x[0] = dc->z;
        job.x[1] = wz;
        for (int k = 0; k < 2; k++) job.x_step[k] = qm_pow(job.x[k], kEvalThreads);
        DevBuf dpart;
        std::vector<uint32_t> part(size_t(ne) * eval_nb * 4);
        ZS_HIP(ctx, hipMalloc(&dpart.ptr, part.size() * 4));
        ZS_LAUNCH(ctx, "poly_eval_kernel", poly_eval_kernel, dim3(eval_nb, ne), dim3(kEvalThreads), 0, ctx->stream,
                  coef.as<uint32_t>(), (uint32_t)n, job, eval_nb, dpart.as<uint32_t>());
        ZS_HIP(ctx, hipGetLastError());
        ZS_HIP(ctx, hipMemcpyAsync(part.data(), dpart.ptr, part.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const qm31 n_inv = {{m31_inv(m31_from_u32((uint32_t)n)), 0}, {0, 0}};
        for (uint32_t e = 0; e < ne; e++) {
            qm31 v = {{0, 0}, {0, 0}};
            for (uint32_t b = 0; b < eval_nb; b++) v = qm_add(v, qm_from_words(&part[4 * (size_t(e) * eval_nb + b)]));
            v = qm_mul(v, n_inv);
            if (e < 2 * mt) dc->de[e] = v;
            else pz[e - 2 * mt] = v;
        }
    }
    {
        const qm31 zh = qm_sub(qm_pow(dc->z, n), one);
        dc->de[2 * mt] = qm_mul(air_composition_host(d, *wt, dc->z, dc->de, dc->de + mt, pz.data(), zh), qm_inv(zh));
    }
    for (uint32_t k = 0; k < d.n_deep; k++) absorb(chan, dc->de[k]);
    for (uint32_t k = 0; k < d.n_deep; k++) dc->gamma[k] = sample(chan);
    ZS_HIP(ctx, hipMemcpyAsync(deep.ptr, dc.get(), sizeof(DeepConst), hipMemcpyHostToDevice, ctx->stream));
    ZS_LAUNCH(ctx, "air_deep_kernel", air_deep_kernel, dim3((pthreads + 255) / 256), dim3(256), 0, ctx->stream,
              deep.as<DeepConst>(), cw, mt, dm, g.as<uint32_t>());
    ZS_HIP(ctx, hipGetLastError());
    // 6. FRI on g under friParams (stark.zig:582-583), then its queries
    const uint32_t L = dlog - remainder_log;
    const size_t rc = size_t(1) << (remainder_log - flb);
    std::vector<uint8_t> froots(size_t(L) * 32);
    std::vector<uint32_t> frem(4 * rc);
    zs_fri_result *res = nullptr;
    ZS_TRY(zs_fri_commit_dev(h, g.as<uint32_t>(), dlog, flb, remainder_log, nullptr, chan, 0, froots.data(),
                             frem.data(), &res));
    std::unique_ptr<zs_fri_result, void (*)(zs_fri_result *)> guard(res, zs_fri_result_free);
    size_t per_q = 0;
    for (uint32_t i = 0; i < L; i++) per_q += 2 * size_t(dlog - i) * 32;
    std::vector<size_t> fidx(num_queries + 1);
    std::vector<uint32_t> fvals(num_queries * L * 8 + 4);
    std::vector<uint8_t> fpaths(num_queries * per_q + 1);
    ZS_TRY(zs_fri_query(res, chan, num_queries, fidx.data(), fvals.data(), fpaths.data()));
    // 7. per-query reveals (stark.zig:584-625): trace and accumulators at x0 and w x0, Q at
    // x0, preprocessed at x0; codeword k of the gather is [trace | acc | Q | pre] column k
    const uint32_t R = d.reveal_len;
    std::vector<uint64_t> vreq;
    std::vector<PathReq> preq;
    for (size_t q = 0; q < num_queries; q++) {
        const size_t p0 = fidx[q], pn = (p0 + shift) % N;
        for (uint32_t k = 0; k < R; k++) {
            uint32_t col;
            size_t at;
            if (k < 2 * mt) {
                col = k % mt;
                at = k < mt ? p0 : pn;
            } else {
                col = k - mt;  // Q is codeword mt, preprocessed column j is mt + 1 + j
                at = p0;
            }
            vreq.push_back((col * N + at) * 4);
            vreq.push_back((q * R + k) * 4);
            preq.push_back({col * tw, (q * R + k) * uint64_t(dlog) * 8, dlog, (uint32_t)at});
        }
    }
    std::vector<uint32_t> rvals(num_queries * R * 4 + 4);
    std::vector<uint8_t> rpaths(num_queries * R * dlog * 32 + 1);
    if (num_queries) {
        DevBuf dv, dpth;
        ZS_HIP(ctx, hipMalloc(&dv.ptr, rvals.size() * 4));
        ZS_HIP(ctx, hipMalloc(&dpth.ptr, rpaths.size()));
        ZS_TRY(launch_gather_words(ctx, cw, vreq, 4, dv.as<uint32_t>()));
        ZS_TRY(launch_gather_paths(ctx, tr, preq, dpth.as<uint32_t>()));
        ZS_HIP(ctx, hipMemcpyAsync(rvals.data(), dv.ptr, rvals.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        ZS_HIP(ctx, hipMemcpyAsync(rpaths.data(), dpth.ptr, rpaths.size(), hipMemcpyDeviceToHost, ctx->stream));
        ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    // 8. GenericStark.Proof in the wire format (stark.zig:107-123, serialization.zig)
    Wire wr;
    wr.uint(trace_log, 1);  // StarkParams
    wr.uint(log_blowup, 1);
    wr.uint(num_queries, 8);
    wr.uint(remainder_log, 1);
    wr.uint(d.m, 8);  // trace_roots
    wr.bytes(roots.data(), 32 * size_t(d.m));
    wr.uint(d.n_pre ? 1 : 0, 1);  // preprocessed_roots: ?[]Digest
    if (d.n_pre) {
        wr.uint(d.n_pre, 8);
        wr.bytes(proots.data(), 32 * size_t(d.n_pre));
    }
    wr.uint(d.n_rel ? 1 : 0, 1);  // accumulator_roots
    if (d.n_rel) {
        wr.uint(d.n_rel, 8);
        wr.bytes(roots.data() + 32 * d.m, 32 * size_t(d.n_rel));
    }
    wr.bytes(qroot, 32);
    wr.uint(d.n_deep, 8);
    for (uint32_t k = 0; k < d.n_deep; k++) wr.qm(dc->de[k]);
    wr.uint(fls, 1);  // fri: Fri.Proof under friParams
    wr.uint(flb, 1);
    wr.uint(num_queries, 8);
    wr.uint(remainder_log, 1);
    wr.uint(L, 8);
    wr.bytes(froots.data(), froots.size());
    wr.uint(rc, 8);
    for (size_t k = 0; k < rc; k++) wr.qm(qm_from_words(&frem[4 * k]));
    wr.uint(num_queries, 8);
    for (size_t q = 0; q < num_queries; q++) {
        wr.uint(fidx[q], 8);
        wr.uint(L, 8);
        const uint8_t *pp = fpaths.data() + q * per_q;
        for (uint32_t i = 0; i < L; i++) {
            const size_t plen = dlog - i;
            wr.qm(qm_from_words(&fvals[8 * (q * L + i)]));
            wr.qm(qm_from_words(&fvals[8 * (q * L + i) + 4]));
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
        wr.uint(R, 8);
        for (uint32_t k = 0; k < R; k++) wr.qm(qm_from_words(&rvals[4 * (q * R + k)]));
        wr.uint(R, 8);
        for (uint32_t k = 0; k < R; k++) {
            wr.uint(dlog, 8);
            wr.bytes(&rpaths[(q * R + k) * size_t(dlog) * 32], size_t(dlog) * 32);
        }
    }
    uint8_t *out = static_cast<uint8_t *>(host.alloc(host.ctx, wr.b.size()));
    if (!out) return ZS_ERR_OOM;
    std::memcpy(out, wr.b.data(), wr.b.size());
    *proof_out = out;
    *proof_len = wr.b.size();
    return ZS_OK;
}
