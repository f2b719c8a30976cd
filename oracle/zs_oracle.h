/* zs_oracle.h -- CPU restatement of zig-stark's M31 trace-commitment path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library (libzs_hip.so) never
 * links, loads or calls it.
 *
 * Every function restates one reference function; the citation is on the
 * definition in zs_oracle.c.  All arithmetic is exact integer arithmetic.
 *
 * Conventions
 *   - M31 values are uint32_t in [0, p), p = 2^31 - 1.
 *   - QM31 values are 4 x uint32_t in the order a.c0, a.c1, b.c0, b.c1
 *     (src/m31/field/qm31.zig:91-101).
 *   - Digests are 32 bytes.
 *   - Merkle node arrays hold every level, leaves first: level 0 (n digests),
 *     level 1 (n/2), ..., root (1): (2n - 1) digests in total.
 */
#ifndef ZS_ORACLE_H
#define ZS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZO_P 2147483647u

/* ---- M31 / CM31 / QM31 ---------------------------------------------------- */
uint32_t zo_m31_from_u32(uint32_t x);
uint32_t zo_m31_add(uint32_t a, uint32_t b);
uint32_t zo_m31_sub(uint32_t a, uint32_t b);
uint32_t zo_m31_mul(uint32_t a, uint32_t b);
uint32_t zo_m31_neg(uint32_t a);
uint32_t zo_m31_pow(uint32_t a, uint32_t e);
uint32_t zo_m31_inv(uint32_t a);

typedef struct { uint32_t v[4]; } zo_qm31;
zo_qm31 zo_qm31_add(zo_qm31 x, zo_qm31 y);
zo_qm31 zo_qm31_sub(zo_qm31 x, zo_qm31 y);
zo_qm31 zo_qm31_mul(zo_qm31 x, zo_qm31 y);
zo_qm31 zo_qm31_inv(zo_qm31 x);
zo_qm31 zo_qm31_pow(zo_qm31 x, uint64_t e);
zo_qm31 zo_qm31_root_of_unity(uint32_t log_size);
/* ptr-based wrappers for FFI callers: out may alias inputs */
void zo_qm31_mul_p(const uint32_t *x, const uint32_t *y, uint32_t *out);
void zo_qm31_inv_p(const uint32_t *x, uint32_t *out);
void zo_qm31_pow_p(const uint32_t *x, uint64_t e, uint32_t *out);
void zo_qm31_root_of_unity_p(uint32_t log_size, uint32_t *out);

/* ---- circle group ------------------------------------------------------------ */
typedef struct { uint32_t x, y; } zo_point;
typedef struct { zo_point offset, step; uint32_t log_size; } zo_coset;

zo_point zo_point_add(zo_point a, zo_point b);
zo_point zo_point_double(zo_point a);
zo_point zo_point_neg(zo_point a);
zo_point zo_point_generator(void);
zo_point zo_point_mul_scalar(zo_point a, uint64_t s);
zo_coset zo_coset_canonic_half(uint32_t log_size);
zo_point zo_coset_at(zo_coset c, size_t index);
zo_coset zo_coset_half(zo_coset c);
zo_coset zo_coset_conjugate(zo_coset c);
/* ptr-based wrappers: point = 2 x u32, coset = 5 x u32 (ox, oy, sx, sy, log) */
void zo_coset_canonic_half_p(uint32_t log_size, uint32_t *out5);
void zo_coset_at_p(const uint32_t *coset5, size_t index, uint32_t *out2);
void zo_domain_point_p(const uint32_t *coset5, size_t index, uint32_t *out2);
void zo_point_mul_scalar_p(const uint32_t *p2, uint64_t s, uint32_t *out2);

/* ---- circle FFT (src/m31/ntt/circle.zig) ------------------------------------ */
/* twiddle tree, length 2^coset.log_size; last slot is 1 */
void zo_precompute_twiddles(const uint32_t *coset5, uint32_t *out);
int zo_circle_fft(const uint32_t *coeffs, size_t n, const uint32_t *coset5, uint32_t *evals);
int zo_circle_ifft(const uint32_t *evals, size_t n, const uint32_t *coset5, uint32_t *coeffs);
uint32_t zo_eval_at_point(const uint32_t *coeffs, size_t n, uint32_t px, uint32_t py);
/* blowup LDE of one column on canonic cosets: out has 2^(log_n + log_blowup) values */
int zo_lde(const uint32_t *col, uint32_t log_n, uint32_t log_blowup, uint32_t *out);
/* the zero-padded A-then-B coefficient vector the LDE evaluates */
int zo_lde_padded_coeffs(const uint32_t *col, uint32_t log_n, uint32_t log_blowup, uint32_t *out);

/* ---- BLAKE3 and the reference's hash wrappers ----------------------------- */
void zo_blake3(const uint8_t *msg, size_t len, uint8_t out[32]);
void zo_blake3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out16[16]);
/* BLAKE2s-256 built from the same G function and round schedule (cross-check) */
void zo_blake2s_from_g(const uint8_t *msg, size_t len, uint8_t out[32]);
void zo_hash2(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]);
void zo_hash_m31s(const uint32_t *vals, size_t n, uint8_t out[32]);
void zo_hash_qm31(const uint32_t v[4], uint8_t out[32]);

/* ---- Merkle tree (src/core/merkle/merkle.zig) ------------------------------ */
/* nodes: (2n - 1) * 32 bytes, levels concatenated leaves first */
int zo_merkle_build(const uint8_t *leaves, size_t n, uint8_t *nodes);
int zo_merkle_open(const uint8_t *nodes, size_t n, size_t index, uint8_t *path);
int zo_merkle_verify(const uint8_t root[32], size_t index, const uint8_t leaf[32],
                     const uint8_t *path, size_t path_len);
/* row-leaf tree over column-major columns: leaf i = hashM31s(col_0[i], ..., col_{c-1}[i]) */
int zo_row_commit(const uint32_t *cols, size_t n_cols, size_t n_rows, uint8_t *nodes);

/* ---- Fiat-Shamir channel (src/core/channel/channel.zig) -------------------- */
typedef struct zo_channel zo_channel;
zo_channel *zo_channel_new(const uint8_t *sep, size_t sep_len);
zo_channel *zo_channel_clone(const zo_channel *c);
void zo_channel_free(zo_channel *c);
void zo_channel_absorb(zo_channel *c, const uint8_t *data, size_t len);
void zo_channel_sample_bytes(zo_channel *c, uint8_t *out, size_t len);
void zo_channel_sample_qm31(zo_channel *c, uint32_t out[4]);
size_t zo_channel_sample_index(zo_channel *c, size_t n);
size_t zo_channel_transcript_len(const zo_channel *c);

/* ---- univariate QM31 polynomials (src/m31/poly/univariate.zig) ----------- */
void zo_poly_eval(const uint32_t *coeffs, size_t n, const uint32_t *x, uint32_t *out);
int zo_poly_interpolate(const uint32_t *xs, const uint32_t *ys, size_t n, uint32_t *coeffs);

/* ---- GenericStark trace commitment (src/m31/stark.zig:343-414) ------------- */
/* cols: n_cols columns of 2^trace_log QM31 values; codewords (may be NULL): out,
 * n_cols * 2^(trace_log+log_blowup) QM31 values on D = FRI_OFFSET * <w_ev>;
 * roots: out, n_cols digests, each absorbed into chan (if not NULL) in order. */
int zo_trace_commit(const uint32_t *cols, size_t n_cols, uint32_t trace_log, uint32_t log_blowup, zo_channel *chan,
                    uint32_t *codewords, uint8_t *roots);
/* FibAir trace with seeds: cols = [a_0..a_{n-1}], [a_1..a_n] as QM31 (M31 in the first word) */
void zo_fib_trace(uint32_t a0, uint32_t a1, size_t n, uint32_t *cols);

/* ---- FRI commit phase (src/m31/fri.zig) ------------------------------------ */
extern const uint32_t ZO_FRI_OFFSET[4];
void zo_fri_domain_point(uint32_t dlog, uint32_t layer, size_t j, uint32_t out[4]);
/* fold one layer: next[j] = fold(x_j, y[j], y[j + n/2], alpha), x_j = D_layer[j] */
int zo_fri_fold(const uint32_t *layer_vals, size_t n, uint32_t dlog, uint32_t layer,
                const uint32_t alpha[4], uint32_t *next);

/* Commit phase of proveCodeword (fri.zig:157-233).
 *   codeword: 2^dlog QM31 values (4 u32 each)
 *   alphas:   NULL => sampled from `chan` (absorb root, sample QM31 per layer);
 *             else L fixed fold challenges (the channel is then not touched
 *             for alphas, but the remainder is still absorbed if chan != NULL)
 *   roots:    out, L digests
 *   layers:   out (may be NULL), every layer 0..L concatenated (2^(dlog+1) - 2^(dlog-L) values)
 *   remainder: out, 2^(remainder_log - log_blowup) QM31 coefficients
 *   alphas_out: out (may be NULL), the L challenges used
 * L = dlog - remainder_log. */
int zo_fri_commit(const uint32_t *codeword, uint32_t dlog, uint32_t log_blowup,
                  uint32_t remainder_log, const uint32_t *alphas, zo_channel *chan,
                  uint8_t *roots, uint32_t *layers, uint32_t *remainder, uint32_t *alphas_out);

/* Full proveCodeword: commit phase + query phase (fri.zig:157-273).
 * query_idx: out, num_queries p0 values.
 * query_vals: out, num_queries * L * 2 QM31 (value0, value1 per layer).
 * query_paths: out, num_queries * L * 2 paths, path of layer i has (dlog - i) digests;
 *              laid out query-major, then layer, then (path0, path1). */
int zo_fri_prove(const uint32_t *codeword, uint32_t dlog, uint32_t log_blowup,
                 uint32_t remainder_log, size_t num_queries, zo_channel *chan,
                 uint8_t *roots, uint32_t *remainder, size_t *query_idx,
                 uint32_t *query_vals, uint8_t *query_paths);
size_t zo_fri_query_path_bytes(uint32_t dlog, uint32_t remainder_log);
/* verify (fri.zig:277-336); same layouts as zo_fri_prove. returns 1 accept, 0 reject */
int zo_fri_verify(uint32_t dlog, uint32_t log_blowup, uint32_t remainder_log, size_t num_queries,
                  zo_channel *chan, const uint8_t *roots, const uint32_t *remainder,
                  const size_t *query_idx, const uint32_t *query_vals, const uint8_t *query_paths);

/* ---- wire format (src/core/serialization.zig, docs/wire.md) ------------------ */
/* the golden example of serialization.zig:219-247 written with the oracle's writers */
size_t zo_wire_golden_example(uint8_t *out);
/* Fri.Proof bytes from zo_fri_prove's outputs; out may be NULL (length only) */
size_t zo_fri_proof_serialize(uint32_t log_size, uint32_t log_blowup, uint32_t remainder_log, size_t num_queries,
                              const uint8_t *roots, const uint32_t *remainder, const size_t *query_idx,
                              const uint32_t *query_vals, const uint8_t *query_paths, uint8_t *out);

/* ---- GenericStark(FibSeededAir) prove / verify (src/m31/stark.zig) ------------ */
/* trace: 2 QM31 columns of 2^trace_log values; a0, a1, claimed: boundary values of
 * (0,0), (1,0), (0,n-1).  Returns malloc'd GenericStark.Proof wire bytes (free with
 * zo_free), NULL on error.  Channel state advances as the prover's. */
uint8_t *zo_fib_stark_prove(uint32_t trace_log, uint32_t log_blowup, size_t num_queries, uint32_t remainder_log,
                            const uint32_t *trace, const uint32_t a0[4], const uint32_t a1[4],
                            const uint32_t claimed[4], zo_channel *chan, size_t *out_len);
/* 1 accept, 0 reject (including malformed, truncated or trailing bytes) */
int zo_fib_stark_verify(const uint8_t *proof, size_t len, const uint32_t a0[4], const uint32_t a1[4],
                        const uint32_t claimed[4], zo_channel *chan);
void zo_free(void *p);

/* ---- GenericStark(Air).proveWithPreprocessed / verify (src/m31/stark.zig:311-780) ----
 * The AIR as data.  Transition constraint k at a point is the sum of its terms,
 * a term being coeff * (product of its factors): current[column], next[column]
 * or the evaluation point x (evalTransition, stark.zig:26-35).  Boundary
 * assertions (column, step, value) are what Air.boundaryAssertions returns for
 * the public inputs.  The LogUp metadata per relation: the selector and
 * multiplicity trace columns and the key (trace) and table (preprocessed)
 * column lists, relation r's at key_columns[key_start[r] .. key_start[r+1]).
 * maxConstraintDegree(n) = degree_mul * (n - 1) + degree_add, or, with both 0,
 * the largest term degree (n - 1 per column factor, 1 per x factor). */
#define ZO_AIR_CUR 0u
#define ZO_AIR_NEXT 1u
#define ZO_AIR_X 2u
typedef struct { uint32_t kind, column; } zo_air_factor;
typedef struct {
    uint32_t constraint;
    uint32_t coeff[4];
    uint32_t n_factors; /* 0..4 */
    zo_air_factor factor[4];
} zo_air_term;
typedef struct { uint32_t column, reserved; uint64_t step; uint32_t value[4]; } zo_boundary;
typedef struct {
    uint32_t num_columns, num_preprocessed, num_transition_constraints, num_terms;
    const zo_air_term *terms;
    uint32_t degree_mul, degree_add;
    uint32_t num_boundary, reserved;
    const zo_boundary *boundary;
    uint32_t num_lookup_columns, num_lookup_relations;
    const uint32_t *selector_column, *multiplicity_column;
    const uint32_t *key_start, *key_columns, *table_start, *table_columns;
} zo_air;
/* pre: num_preprocessed QM31 columns, trace: num_columns QM31 columns, each of
 * 2^trace_log values (column-major).  Returns malloc'd GenericStark.Proof wire
 * bytes (free with zo_free) or NULL when the parameters are not ones the
 * reference prover accepts (friParams' assertions). */
uint8_t *zo_stark_prove(const zo_air *air, uint32_t trace_log, uint32_t log_blowup, size_t num_queries,
                        uint32_t remainder_log, const uint32_t *pre, const uint32_t *trace, zo_channel *chan,
                        size_t *out_len);
int zo_stark_verify(const zo_air *air, const uint8_t *proof, size_t len, zo_channel *chan);
/* FRI parameters of the DEEP polynomial (friParams, stark.zig:208-222): log_size and FRI log_blowup */
int zo_stark_fri_params(const zo_air *air, uint32_t trace_log, uint32_t log_blowup, uint32_t *fri_log_size,
                        uint32_t *fri_log_blowup);
/* the accumulator columns on H (stark.zig:430-449) for the given lookup challenges:
 * out = num_lookup_relations QM31 columns of 2^trace_log values */
void zo_stark_accumulators(const zo_air *air, uint32_t trace_log, const uint32_t *pre, const uint32_t *trace,
                           const uint32_t *alpha_cols, const uint32_t *alpha_rels, uint32_t *out);

/* ---- seeded synthetic inputs ------------------------------------------------ */
/* SplitMix64 stream; M31 = 31 random bits with 0x7FFFFFFF rejected */
void zo_fill_m31(uint64_t seed, uint32_t *out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
