// stark.hip -- GenericStark(FibAir).prove (src/m31/stark.zig:303-657, FibAir
// stark.zig:877-904) with seeded boundary assertions: the Fibonacci AIR written as
// data and proved by the generic device prover (air.hip: trace commitment,
// quotient and DEEP codewords with batch inversion, FRI, query reveals).
#include <cstring>
#include <memory>
#include <vector>

#include "internal.h"

using namespace zs;

extern "C" {

int32_t zs_fib_stark_prove_dev(zs_hip_ctx *h, const uint32_t *d_trace, uint32_t trace_log, uint32_t log_blowup,
                               size_t num_queries, uint32_t remainder_log, const uint32_t a0[4], const uint32_t a1[4],
                               const uint32_t claimed[4], zs_channel *chan, zs_host_allocator host,
                               uint8_t **proof_out, size_t *proof_len) {
    if (!h || !d_trace || !a0 || !a1 || !claimed || !chan || !host.alloc || !proof_out || !proof_len ||
        trace_log < 1 || log_blowup < 1 || log_sum_over(trace_log, log_blowup, 24) || remainder_log < log_blowup ||
        remainder_log >= trace_log + log_blowup || num_queries > 4096)
        return ZS_ERR_INPUT;
    // FibAir (stark.zig:877-904) as data: next[0] - cur[1], next[1] - cur[0] - cur[1];
    // boundary (0,0) = a0, (1,0) = a1, (0,n-1) = claimed; maxConstraintDegree n - 1
    const uint32_t minus_one = P - 1;
    zs_air_term terms[5] = {};
    const uint32_t con[5] = {0, 0, 1, 1, 1}, kind[5] = {ZS_AIR_NEXT, ZS_AIR_CUR, ZS_AIR_NEXT, ZS_AIR_CUR, ZS_AIR_CUR};
    const uint32_t col[5] = {0, 1, 1, 0, 1}, coef[5] = {1, minus_one, 1, minus_one, minus_one};
    for (int t = 0; t < 5; t++) {
        terms[t].constraint = con[t];
        terms[t].coeff[0] = coef[t];
        terms[t].n_factors = 1;
        terms[t].factor[0] = {kind[t], col[t]};
    }
    zs_boundary bnd[3] = {};
    const uint32_t *vals[3] = {a0, a1, claimed};
    const uint32_t bcol[3] = {0, 1, 0};
    const uint64_t bstep[3] = {0, 0, (uint64_t(1) << trace_log) - 1};
    for (int k = 0; k < 3; k++) {
        bnd[k].column = bcol[k];
        bnd[k].step = bstep[k];
        std::memcpy(bnd[k].value, vals[k], 16);
    }
    zs_air air{};
    air.num_columns = 2;
    air.num_transition_constraints = 2;
    air.num_terms = 5;
    air.terms = terms;
    air.degree_mul = 1;
    air.num_boundary = 3;
    air.boundary = bnd;
    return zs_stark_prove_dev(h, &air, nullptr, d_trace, trace_log, log_blowup, num_queries, remainder_log, chan,
                              host, proof_out, proof_len);
}

}  // extern "C"
