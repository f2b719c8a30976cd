// air.h -- an AIR given as data (zs_air) in fixed-capacity form for the device
// prover and the host verifier, and the parts of GenericStark's composition
// (src/m31/stark.zig:19-69, 180-275, 484-531) that both evaluate.
#pragma once
#include "internal.h"

namespace zs {

constexpr uint32_t AIR_MAX_COLS = 32;    // trace + accumulator + preprocessed columns
constexpr uint32_t AIR_MAX_TERMS = 64;
constexpr uint32_t AIR_MAX_TRANS = 32;   // the AIR's transition constraints
constexpr uint32_t AIR_MAX_REL = 4;      // LogUp relations
constexpr uint32_t AIR_MAX_KEYS = 8;     // key (and table) columns per relation
constexpr uint32_t AIR_MAX_BOUND = 16;   // the AIR's boundary assertions (plus one acc[0] = 0 per relation)

struct AirDesc {
    uint32_t m, n_pre, n_rel, n_lc, T, B;                  // stark.zig:140-147
    uint32_t m_total, n_trans_total, total_bound, n_deep, reveal_len;  // stark.zig:184-189
    uint32_t n_terms, degree_mul, degree_add;
    uint32_t term_con[AIR_MAX_TERMS], term_nf[AIR_MAX_TERMS];
    uint32_t term_kind[AIR_MAX_TERMS][4], term_col[AIR_MAX_TERMS][4];
    qm31 term_coeff[AIR_MAX_TERMS];
    uint32_t sel[AIR_MAX_REL], mult[AIR_MAX_REL];
    uint32_t key_s[AIR_MAX_REL + 1], key_c[AIR_MAX_REL * AIR_MAX_KEYS];
    uint32_t tab_s[AIR_MAX_REL + 1], tab_c[AIR_MAX_REL * AIR_MAX_KEYS];
    // boundary assertions, the AIR's then acc_r[0] = 0 (stark.zig:476-481)
    uint32_t bcol[AIR_MAX_BOUND + AIR_MAX_REL];
    uint64_t bstep[AIR_MAX_BOUND + AIR_MAX_REL];
    qm31 bval[AIR_MAX_BOUND + AIR_MAX_REL];
};

// the channel's draws a composition needs
struct AirWeights {
    qm31 alpha[AIR_MAX_TRANS + 2 * AIR_MAX_REL];     // transition weights (incl. the synthesized ones)
    qm31 beta[AIR_MAX_BOUND + AIR_MAX_REL];          // boundary weights
    qm31 pk[AIR_MAX_BOUND + AIR_MAX_REL];            // w^step of each boundary assertion
    qm31 alpha_cols[AIR_MAX_KEYS], alpha_rels[AIR_MAX_REL];
    qm31 last;                                       // w^(n-1)
};

// Validated copy of the caller's description: ZS_ERR_INPUT when it is malformed
// or beyond the capacities above.
int32_t air_from_abi(const zs_air *a, AirDesc *d);
// compositionLogSize / friParams (stark.zig:197-222): false when log_size > domain
bool air_fri_params(const AirDesc &d, uint32_t trace_log, uint32_t log_blowup, uint32_t *fri_log_size,
                    uint32_t *fri_log_blowup);

// Hc(x) without its boundary part (stark.zig:504-527): the AIR's transitions times
// (x - w^(n-1)), the cyclic LogUp constraints (lookupConstraintAt, :231-253) and the
// selector binarity, each with its weight.  val(kind, col) gives the trace or
// accumulator column col at x (ZS_AIR_CUR) or w x (ZS_AIR_NEXT), or the preprocessed
// column col at x (AIR_PRE).
constexpr uint32_t AIR_PRE = 3;
template <class VAL>
ZS_HD qm31 air_constraints(const AirDesc &d, const AirWeights &w, qm31 x, const VAL &val) {
    const qm31 one = {{1, 0}, {0, 0}}, zero = {{0, 0}, {0, 0}};
    const qm31 xl = qm_sub(x, w.last);
    qm31 h = zero;
    // Air.evalTransition as terms; constraint k's value is summed, then weighted
    for (uint32_t k = 0; k < d.T; k++) {
        qm31 c = zero;
        for (uint32_t t = 0; t < d.n_terms; t++) {
            if (d.term_con[t] != k) continue;
            qm31 p = d.term_coeff[t];
            for (uint32_t f = 0; f < d.term_nf[t]; f++) {
                const uint32_t kind = d.term_kind[t][f];
                p = qm_mul(p, kind == ZS_AIR_X ? x : val(kind, d.term_col[t][f]));
            }
            c = qm_add(c, p);
        }
        h = qm_add(h, qm_mul(w.alpha[k], qm_mul(c, xl)));
    }
    for (uint32_t r = 0; r < d.n_rel; r++) {
        const qm31 sel = val(ZS_AIR_CUR, d.sel[r]), mult = val(ZS_AIR_CUR, d.mult[r]);
        qm31 key = zero, tkey = zero;  // combinedKey (stark.zig:225-229)
        for (uint32_t j = d.key_s[r]; j < d.key_s[r + 1]; j++)
            key = qm_add(key, qm_mul(w.alpha_cols[j - d.key_s[r]], val(ZS_AIR_CUR, d.key_c[j])));
        for (uint32_t j = d.tab_s[r]; j < d.tab_s[r + 1]; j++)
            tkey = qm_add(tkey, qm_mul(w.alpha_cols[j - d.tab_s[r]], val(AIR_PRE, d.tab_c[j])));
        const qm31 dacc = qm_sub(val(ZS_AIR_NEXT, d.m + r), val(ZS_AIR_CUR, d.m + r));
        const qm31 lb = qm_sub(qm_mul(dacc, qm_sub(w.alpha_rels[r], key)), mult);
        const qm31 tb = qm_add(qm_mul(dacc, qm_sub(w.alpha_rels[r], tkey)), one);
        const qm31 c = qm_add(qm_mul(sel, lb), qm_mul(qm_sub(one, sel), tb));
        h = qm_add(h, qm_mul(w.alpha[d.T + r], c));
    }
    for (uint32_t r = 0; r < d.n_rel; r++) {
        const qm31 sel = val(ZS_AIR_CUR, d.sel[r]);
        h = qm_add(h, qm_mul(w.alpha[d.T + d.n_rel + r], qm_mul(sel, qm_sub(one, sel))));
    }
    return h;
}

// the full Hc(x) (stark.zig:504-530) on the host, from value arrays (cur / nxt:
// m_total values at x and w x, pre: the preprocessed values at x), one inversion
// per boundary term
inline qm31 air_composition_host(const AirDesc &d, const AirWeights &w, qm31 x, const qm31 *cur, const qm31 *nxt,
                                 const qm31 *pre, qm31 zh) {
    auto val = [&](uint32_t kind, uint32_t col) {
        return kind == ZS_AIR_CUR ? cur[col] : (kind == ZS_AIR_NEXT ? nxt[col] : pre[col]);
    };
    qm31 h = air_constraints(d, w, x, val);
    for (uint32_t k = 0; k < d.total_bound; k++) {
        const qm31 term = qm_mul(qm_mul(qm_sub(cur[d.bcol[k]], d.bval[k]), zh), qm_inv(qm_sub(x, w.pk[k])));
        h = qm_add(h, qm_mul(w.beta[k], term));
    }
    return h;
}

}  // namespace zs
