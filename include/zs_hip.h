/* zs_hip.h -- C ABI of the MI355X (gfx950) M31 trace-commitment path.
 *
 * Replaces zig-stark's src/cuda layer (src/cuda/circlefft_gpu.zig,
 * src/cuda/cuda.zig) and follows the conventions of zig-capi.h:
 *   - return codes: 0 ok, -1 generic, -2 malformed input, -3 out of memory,
 *     -4 protocol error (zig-capi.h:20-21), plus -5 for "no usable HIP device
 *     or a HIP runtime failure" (the reference's error.CudaFailure /
 *     error.NoCudaDevice, src/cuda/cuda.zig:54-106).  There is no CPU
 *     fallback: without a device every compute entry point returns -5;
 *   - buffers passed in are borrowed for the duration of the call;
 *   - buffers the library returns are allocated with the caller's
 *     zs_host_allocator and released with zs_free (zig-capi.h:33-39, 73-74).
 *
 * Layouts: M31 values are uint32_t in [0, 2^31 - 1).  QM31 values are four
 * uint32_t (a.c0, a.c1, b.c0, b.c1; src/m31/field/qm31.zig:91-101).  Digests
 * are 32 bytes.  Columns are column-major: column c starts at c * col_stride.
 * A Merkle node array holds every level, leaves first (level 0 = 2^k
 * digests, ..., the root last), as MerkleTree.nodes does
 * (src/core/merkle/merkle.zig:17-45).
 *
 * Functions whose name ends in _dev take device pointers and run
 * asynchronously on the context's stream; the others take host pointers and
 * return when the result is in host memory (as circlefft_gpu.zig does).
 * Calls on one context must not overlap from several threads; use one
 * context per thread (contexts on one device share nothing).
 */
#ifndef ZS_HIP_H
#define ZS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZS_OK 0
#define ZS_ERR_GENERIC -1
#define ZS_ERR_INPUT -2
#define ZS_ERR_OOM -3
#define ZS_ERR_PROTOCOL -4
#define ZS_ERR_DEVICE -5

/* flags for zs_m31_lde_commit_dev / zs_fri_commit_dev */
#define ZS_FLAG_PLAIN 1u /* reference-shaped per-column, per-layer launches (comparison path) */
#define ZS_FLAG_NO_GRAPH 2u /* FRI commit: launch each kernel (the default; kept for callers that pass it) */
#define ZS_FLAG_PERSISTENT 4u /* batch: LDE, row leaves and node levels of all traces in one persistent launch */
#define ZS_FLAG_GRAPH 8u /* FRI commit (and a commit job's FRI): capture the launches once per buffer set as a HIP graph and replay it */

typedef struct zs_hip_ctx zs_hip_ctx;
typedef struct zs_tree zs_tree;
typedef struct zs_channel zs_channel;
typedef struct zs_fri_result zs_fri_result;

/* zig-capi.h:35-39 */
typedef struct {
    void *ctx;
    void *(*alloc)(void *ctx, size_t size);
    void (*free)(void *ctx, void *ptr, size_t size);
} zs_host_allocator;

/* CircleCoset (src/m31/circle/coset.zig:10-14): offset + <step>, 2^log_size points */
typedef struct {
    uint32_t offset_x, offset_y, step_x, step_y, log_size;
} zs_circle_coset;

/* ---- library and context --------------------------------------------------- */
/* zig-capi.h:76-77 (zs_version) */
const char *zs_hip_version(void);
/* Lazily-initialised context on `device` (src/cuda/circlefft_gpu.zig:20-39,
 * src/cuda/cuda.zig:81-106).  flags: reserved, pass 0. */
int32_t zs_hip_init(int32_t device, uint32_t flags, zs_hip_ctx **out_ctx);
void zs_hip_destroy(zs_hip_ctx *ctx);
const char *zs_hip_last_error(const zs_hip_ctx *ctx);
/* Run _dev calls on this hipStream_t (NULL = the null stream). */
int32_t zs_hip_set_stream(zs_hip_ctx *ctx, void *hip_stream);
int32_t zs_hip_synchronize(zs_hip_ctx *ctx);
/* Per-kernel timing (no reference counterpart; the reference times whole calls
 * with clock_gettime, benchmarks/ntt/src/main.zig:15-47).  enable = 1 clears the
 * records and starts bracketing every hot kernel with HIP events on the
 * context's stream; 0 stops. */
int32_t zs_hip_profile(zs_hip_ctx *ctx, int32_t enable);
/* Synchronise and aggregate the recorded launches by kernel name: names are
 * written NUL-separated into `names` (names_cap bytes), ms[i] is the total
 * event time and counts[i] the launches of entry i; *n_out entries (<= max). */
int32_t zs_hip_profile_read(zs_hip_ctx *ctx, char *names, size_t names_cap, double *ms, uint32_t *counts,
                            uint32_t max, uint32_t *n_out);
/* zig-capi.h:73-74 */
void zs_free(zs_host_allocator host, uint8_t *ptr, size_t len);

/* CircleCoset.canonicHalf(log_size), 1 <= log_size <= 28 (coset.zig:55-65). */
int32_t zs_circle_coset_canonic_half(uint32_t log_size, zs_circle_coset *out);

/* ---- circle FFT ---------------------------------------------------------------- */
/* circleFFT / circleIFFT (src/cuda/circlefft_gpu.zig:46 and :111, the GPU
 * replacements of src/m31/ntt/circle.zig:195 and :272).  n = 2 * 2^half_coset.log_size,
 * n >= 2.  Host buffers; every size runs on the device (no CPU fallback). */
int32_t zs_m31_circle_fft(zs_hip_ctx *ctx, const uint32_t *coeffs, size_t n, const zs_circle_coset *half_coset,
                          uint32_t *evals);
int32_t zs_m31_circle_ifft(zs_hip_ctx *ctx, const uint32_t *evals, size_t n, const zs_circle_coset *half_coset,
                           uint32_t *coeffs);
/* Batched device variants on canonicHalf(log_n): n_cols columns, column c at c * col_stride. */
int32_t zs_m31_circle_fft_batch_dev(zs_hip_ctx *ctx, const uint32_t *d_coeffs, uint32_t *d_evals, uint32_t log_n,
                                    uint32_t n_cols, size_t col_stride);
int32_t zs_m31_circle_ifft_batch_dev(zs_hip_ctx *ctx, const uint32_t *d_evals, uint32_t *d_coeffs, uint32_t log_n,
                                     uint32_t n_cols, size_t col_stride);

/* ---- low-degree extension and row-leaf commitment -------------------------- */
/* Blowup LDE of each column (circleIFFT on canonicHalf(log_n), zero-pad the A and
 * B halves, circleFFT on canonicHalf(log_n + log_blowup)); then leaf i =
 * hashM31s(lde_0[i], ..., lde_{c-1}[i]) (src/m31/hash.zig:15-25) and
 * MerkleTree.init over the 2^(log_n + log_blowup) leaves (merkle.zig:17-45).
 *   d_cols:  n_cols * 2^log_n words, column-major, contiguous
 *   d_lde:   n_cols * 2^(log_n+log_blowup) words, column-major, contiguous
 *   d_nodes: (2^(log_n+log_blowup+1) - 1) * 32 bytes
 * 1 <= log_n, log_n + log_blowup <= 28, 1 <= n_cols. */
int32_t zs_m31_lde_commit_dev(zs_hip_ctx *ctx, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_n,
                              uint32_t log_blowup, uint32_t *d_lde, uint8_t *d_nodes, uint32_t flags);
/* The LDE alone (no commitment); same layouts as zs_m31_lde_commit_dev. */
int32_t zs_m31_lde_dev(zs_hip_ctx *ctx, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_n, uint32_t log_blowup,
                       uint32_t *d_lde);
/* No counterpart in the reference: zs_m31_lde_commit_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments (no flags), layouts and
 * results; it calls no hipMalloc/hipFree.  The LDE's intermediate buffers and
 * twiddle tables live in one fixed device array of 256 MiB per GPU, reserved for
 * the life of the process and set up by zs_hip_init.  Limits: log_n + log_blowup
 * <= 22, and n_cols * 2^log_n words of coefficients plus, when log_n > 11, n_cols *
 * 2^(log_n + log_blowup) words of intermediate values within 192 MiB (8 columns of
 * 2^20 at blowup 4 fit); larger shapes return ZS_ERR_INPUT.  Every call on a GPU
 * uses the same array: calls that overlap on one GPU (other streams or threads) race. */
int32_t zs_m31_lde_commit_dev_noalloc(zs_hip_ctx *ctx, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_n,
                                      uint32_t log_blowup, uint32_t *d_lde, uint8_t *d_nodes);
/* Launch-configuration autotuner (no reference counterpart; the reference has
 * no GPU LDE).  Runs zs_m31_lde_dev for one shape under each candidate
 * configuration of the register-group passes (2^11..2^14 elements per
 * workgroup x 1, 2, 4, 8 columns per workgroup) on synthetic device data, checks
 * that every candidate yields the same row-tree root, times each pass kernel
 * with events, and keeps for each kind of pass (interpolation passes, the fused
 * middle kernel, evaluation passes) the candidate it ran fastest under, for that
 * shape (log_n, log_blowup, n_cols rounded up to a power of two): later LDE
 * calls of the shape on the same device use it, from every context of the
 * process (batch workers and server workers included); a shape not tuned takes
 * the pick of the nearest tuned column count with the same log sizes.  Picks are
 * appended to a per-device file ($ZS_AUTOTUNE_CACHE, else the user's cache
 * directory; "off" disables it) that the next process reads instead of tuning.
 * out (may be NULL): tile elements log2 and columns per workgroup of the
 * interpolation, middle and evaluation passes, then the microseconds of one
 * LDE under the chosen combination. */
int32_t zs_autotune_lde(zs_hip_ctx *ctx, uint32_t n_cols, uint32_t log_n, uint32_t log_blowup, uint32_t out[7]);
/* Row leaves alone: d_leaves[i] = hashM31s(row i), 2^log_rows digests (level 0 only). */
int32_t zs_m31_row_leaves_dev(zs_hip_ctx *ctx, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_rows,
                              size_t col_stride, uint8_t *d_leaves);
/* Host-buffer form: returns a device-resident tree; lde_out may be NULL. */
int32_t zs_m31_lde_commit(zs_hip_ctx *ctx, const uint32_t *cols, uint32_t n_cols, uint32_t log_n,
                          uint32_t log_blowup, zs_tree **out_tree, uint8_t root[32], uint32_t *lde_out);
/* Row-leaf tree over existing device columns of 2^log_rows values. */
int32_t zs_m31_row_commit_dev(zs_hip_ctx *ctx, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_rows,
                              size_t col_stride, uint8_t *d_nodes, uint32_t flags);

/* ---- Merkle trees (src/core/merkle/merkle.zig) ---------------------------- */
uint32_t zs_tree_log_leaves(const zs_tree *tree);
/* MerkleTree.root (merkle.zig:47-49) */
int32_t zs_tree_root(const zs_tree *tree, uint8_t out[32]);
/* MerkleTree.open (merkle.zig:57-68): log_leaves sibling digests, bottom-up. */
int32_t zs_merkle_open(const zs_tree *tree, size_t index, uint8_t *path_out);
/* Copy one whole level (2^(log_leaves - level) digests) to host memory. */
int32_t zs_tree_read_level(const zs_tree *tree, uint32_t level, uint8_t *out);
void zs_tree_free(zs_tree *tree);

/* ---- Fiat-Shamir channel (src/core/channel/channel.zig), host side ------ */
zs_channel *zs_channel_new(const uint8_t *domain_separator, size_t len);
zs_channel *zs_channel_clone(const zs_channel *ch);
void zs_channel_free(zs_channel *ch);
void zs_channel_absorb(zs_channel *ch, const uint8_t *data, size_t len);      /* channel.zig:23-29 */
void zs_channel_sample_bytes(zs_channel *ch, uint8_t *out, size_t len);       /* channel.zig:49-69 */
void zs_channel_sample_qm31(zs_channel *ch, uint32_t out[4]);                 /* channel.zig:72-76 */
size_t zs_channel_sample_index(zs_channel *ch, size_t n);                     /* channel.zig:79-101 */

/* ---- FRI commit phase (src/m31/fri.zig:157-233) ----------------------------- */
/* Commit a QM31 codeword of 2^dlog values on D_0 = FRI_OFFSET * <w_0>:
 * L = dlog - remainder_log layers are Merkle-committed (hashQM31 leaves) and
 * folded; the remainder is the 2^(remainder_log - log_blowup) interpolation
 * coefficients of the last layer.
 *   alphas: L fold challenges (QM31), or NULL to draw them from `chan`
 *           (absorb root_i, sample QM31) exactly as proveCodeword does; with
 *           NULL the remainder is absorbed into `chan` afterwards.
 *   roots_out: L * 32 bytes; remainder_out: 4 * 2^(remainder_log - log_blowup) words.
 *   out_result: optional; keeps every layer and tree on the device for queries.
 * d_codeword is a device pointer; it is not modified.  The leaf, node-level and
 * fold kernels of every layer are launched one by one; with ZS_FLAG_GRAPH (and no
 * out_result) they are captured once per buffer set into a HIP graph and replayed
 * on later calls (on ROCm 7 the replay is 0.5-3 % slower than the launches it
 * replaces, so it is opt-in).  ZS_FLAG_NO_GRAPH overrides ZS_FLAG_GRAPH. */
int32_t zs_fri_commit_dev(zs_hip_ctx *ctx, const uint32_t *d_codeword, uint32_t dlog, uint32_t log_blowup,
                          uint32_t remainder_log, const uint32_t *alphas, zs_channel *chan, uint32_t flags,
                          uint8_t *roots_out, uint32_t *remainder_out, zs_fri_result **out_result);
/* Host-buffer form of zs_fri_commit_dev. */
int32_t zs_fri_commit(zs_hip_ctx *ctx, const uint32_t *codeword, uint32_t dlog, uint32_t log_blowup,
                      uint32_t remainder_log, const uint32_t *alphas, zs_channel *chan, uint8_t *roots_out,
                      uint32_t *remainder_out, zs_fri_result **out_result);
/* Query phase (fri.zig:236-256): num_queries indices from `chan`, then per
 * layer the pair (p mod n/2, p mod n/2 + n/2) values and both Merkle paths.
 *   idx_out: num_queries; vals_out: num_queries * L * 2 * 4 words;
 *   paths_out: per query, per layer i, path0 then path1 of (dlog - i) digests. */
int32_t zs_fri_query(zs_fri_result *res, zs_channel *chan, size_t num_queries, size_t *idx_out, uint32_t *vals_out,
                     uint8_t *paths_out);
/* Fri.proveCodeword (src/m31/fri.zig:157-273) on a device codeword of
 * 2^(log_size + log_blowup) QM31 values, with every challenge and query index
 * drawn from `chan`, returned as the reference's wire bytes
 * (src/core/serialization.zig, docs/wire.md) of its Fri.Proof: FriParams
 * {log_size u8, log_blowup u8, num_queries u64, remainder_log u8}, roots
 * []Digest, remainder []QM31, queries []Query{index u64, pairs []Pair{value0,
 * value1 QM31, path0, path1 []Digest}}, every slice after its u64 length.
 * The bytes are allocated with `host` and released with zs_free. */
int32_t zs_fri_prove_dev(zs_hip_ctx *ctx, const uint32_t *d_codeword, uint32_t log_size, uint32_t log_blowup,
                         uint32_t remainder_log, size_t num_queries, zs_channel *chan, zs_host_allocator host,
                         uint8_t **proof_out, size_t *proof_len);
/* Copy FRI layer `layer` (2^(dlog - layer) QM31 values) of a kept commitment to host
 * memory (the layers proveCodeword holds between its commit and query phases,
 * fri.zig:185-219). */
int32_t zs_fri_read_layer(const zs_fri_result *res, uint32_t layer, uint32_t *out);
void zs_fri_result_free(zs_fri_result *res);
/* Fri.verify (src/m31/fri.zig:277-336) of a proof in the wire format above, on
 * the host (no device needed): `chan` in the transcript state the prover's
 * channel had before the commit phase; it is advanced as the reference's
 * verify advances it (absorb roots, resample alphas, absorb the remainder,
 * resample the query indices).  Returns ZS_OK with *out_ok 1 (accepted) or 0
 * (a root count, query count, index, Merkle path, fold or remainder mismatch),
 * or ZS_ERR_INPUT when the bytes do not deserialize (truncated, trailing bytes),
 * as the capi's verify entry points return -2 (src/capi.zig:175-204). */
int32_t zs_fri_verify(const uint8_t *proof, size_t proof_len, zs_channel *chan, uint8_t *out_ok);
/* GenericStark(FibSeededAir).verify (src/m31/stark.zig:660-780) of the
 * GenericStark.Proof wire bytes zs_fib_stark_prove_dev writes, on the host:
 * public inputs a0, a1, claimed (QM31 words) and `chan` in the prover's starting
 * state, advanced as the reference's verify advances it (trace roots, alphas,
 * betas, quotient root, z, DEEP values, gammas, then Fri.verify under the
 * STARK's parameters, then every query's reveals: Merkle paths, the DEEP
 * identity against the FRI layer-0 value, and Hc(x0) = Z_H(x0) Q(x0)).  One
 * reveal per FRI query is required (the prover writes exactly that many), and
 * the FRI proof's own copy of the parameters must equal the STARK's (the
 * reference reads but does not compare them; an honest prover writes the same).
 * Returns ZS_OK with *out_ok 1 or 0, or ZS_ERR_INPUT for bytes that do not
 * deserialize. */
int32_t zs_fib_stark_verify(const uint8_t *proof, size_t proof_len, const uint32_t a0[4], const uint32_t a1[4],
                            const uint32_t claimed[4], zs_channel *chan, uint8_t *out_ok);

/* ---- one trace's full commitment, batches, and the serving loop --------------- */
/* The commitment of one trace (the composition the reference leaves open; the
 * channel, FRI and tree steps each follow the reference):
 *   1. LDE of the columns (blowup 2^log_blowup) and the row-leaf Merkle tree -> trace root;
 *   2. channel = Channel.init("zig-stark:hip:commit"), absorbDigest(trace root),
 *      gamma = sample(QM31)                     (src/core/channel/channel.zig)
 *   3. g[i] = sum_c gamma^c * lde_c[i] over the LDE rows (QM31);
 *   4. the FRI commit phase of g with challenges drawn from the same channel
 *      (src/m31/fri.zig:157-233), dlog = log_n + log_blowup.
 * The tree stays in d_nodes; fri_roots gets (dlog - remainder_log) digests and
 * remainder 4 * 2^(remainder_log - log_blowup) words.  Requires
 * log_blowup <= remainder_log < log_n + log_blowup and n_cols <= 4096. */
int32_t zs_commit_job_dev(zs_hip_ctx *ctx, const uint32_t *d_cols, uint32_t n_cols, uint32_t log_n,
                          uint32_t log_blowup, uint32_t remainder_log, uint32_t *d_lde, uint8_t *d_nodes,
                          uint8_t trace_root[32], uint8_t *fri_roots, uint32_t *remainder, uint32_t flags);

/* One trace of a batch: device input, device outputs (LDE, tree), host results. */
typedef struct {
    const uint32_t *d_cols;  /* n_cols * 2^log_n words, column-major */
    uint32_t n_cols, log_n;
    uint32_t *d_lde;         /* n_cols * 2^(log_n+log_blowup) words */
    uint8_t *d_nodes;        /* (2^(log_n+log_blowup+1) - 1) * 32 bytes */
    uint8_t trace_root[32];  /* out */
    uint8_t *fri_roots;      /* out, (log_n + log_blowup - remainder_log) * 32 bytes */
    uint32_t *remainder;     /* out, 4 * 2^(remainder_log - log_blowup) words */
    int32_t status;          /* out, ZS_OK or an error code */
} zs_job;
/* Commit many traces of uneven sizes.  n_workers > 1 (at most 64) runs them on that many
 * worker contexts with their own streams, taking jobs from a shared queue so
 * their kernels overlap on the device; ZS_FLAG_PLAIN runs the reference-shaped
 * per-trace loop (one trace after another, per-column per-layer launches).
 * ZS_FLAG_PERSISTENT first runs the LDE, row leaves and node levels of every trace
 * with log_n >= 3 and at most 256 columns (log_blowup <= 8, log_n + log_blowup <= 26) in one persistent
 * launch (several when their scratch -- n_cols * 2^(log_n+log_blowup) words per trace --
 * exceeds the free device memory), then each trace's gamma draw, combination and FRI
 * on the workers; other traces take the worker path.  Results are the same bytes on
 * every path.  Returns ZS_OK when every job's status is ZS_OK. */
int32_t zs_commit_batch_dev(zs_hip_ctx *ctx, zs_job *jobs, size_t n_jobs, uint32_t log_blowup,
                            uint32_t remainder_log, uint32_t n_workers, uint32_t flags);

/* Serving loop (continuous batching): the HBM budget is allocated once when the
 * server starts; each request's buffers (input, LDE, row tree, FRI layers and
 * trees, scratch) are carved from it first-fit.  A queued request is admitted as
 * soon as a worker is idle and its memory is free, and runs on that worker's
 * stream concurrently with the requests already on the device.  Twiddle and
 * FRI tables are shared by the workers and stay resident.  A request's result
 * does not depend on what else is running. */
typedef struct zs_server zs_server;
typedef struct {
    uint32_t n_cols, log_n, log_blowup, remainder_log;
    uint64_t seed;            /* synthetic input: zs_fill_m31_dev stream of `seed` */
    const uint32_t *host_cols;/* or NULL; else n_cols * 2^log_n host words (copied at submit) */
} zs_request;
typedef struct {
    int32_t status;
    uint32_t fri_layers;
    uint8_t trace_root[32];
    uint8_t fri_roots[28 * 32];
    uint32_t remainder[4 * 64];
    double t_submit, t_admit, t_done;  /* seconds, steady clock */
    uint64_t device_bytes;             /* the request's admitted footprint */
} zs_response;
/* NULL if the device, the budget allocation or a worker stream fails. */
zs_server *zs_server_new(int32_t device, uint32_t n_workers, uint64_t hbm_budget_bytes);
/* Enqueue; returns immediately with a ticket. ZS_ERR_OOM if the request alone exceeds the budget. */
int32_t zs_server_submit(zs_server *srv, const zs_request *req, uint64_t *ticket);
/* Block until the request is answered (the response is then released). */
int32_t zs_server_wait(zs_server *srv, uint64_t ticket, zs_response *out);
/* Non-blocking: *ready = 1 and the response (then released) if answered, else *ready = 0. */
int32_t zs_server_poll(zs_server *srv, uint64_t ticket, zs_response *out, int32_t *ready);
/* Device bytes admitted right now, and the largest value seen. */
int32_t zs_server_memory(zs_server *srv, uint64_t *in_use, uint64_t *peak);
void zs_server_free(zs_server *srv);

/* ---- GenericStark trace commitment (src/m31/stark.zig:343-414) -------------- */
/* The trace half of proveWithPreprocessed for AIRs without preprocessed
 * columns: each QM31 column of 2^trace_log values is interpolated over
 * H = <w> (w = primitiveRootOfUnity(trace_log)), evaluated on
 * D = FRI_OFFSET * <w_ev> (w_ev = primitiveRootOfUnity(trace_log + log_blowup)),
 * committed with hashQM31 leaves (hashCodeword, stark.zig:786-790) and
 * MerkleTree.init, and each root is absorbed into `chan` in column order
 * (absorbDigest).  The interpolant is unique, so the DFTs used here give the
 * reference's Lagrange/Horner values exactly.
 *   d_cols:      n_cols * 2^trace_log QM31 values, column-major
 *   d_codewords: out, n_cols * 2^(trace_log+log_blowup) QM31 values
 *   d_nodes:     out, n_cols trees; tree c at c * (2^(trace_log+log_blowup+1) - 1) * 32 bytes
 *   roots_out:   n_cols * 32 bytes;  chan: may be NULL
 * 1 <= trace_log, trace_log + log_blowup <= 24. */
int32_t zs_qm31_trace_commit_dev(zs_hip_ctx *ctx, const uint32_t *d_cols, uint32_t n_cols, uint32_t trace_log,
                                 uint32_t log_blowup, uint32_t *d_codewords, uint8_t *d_nodes, uint8_t *roots_out,
                                 zs_channel *chan);
/* FibAir.generateTrace (stark.zig:918-935) with seeds a0, a1 in place of 0, 1:
 * out = two QM31 columns of 2^log_n values, [a_0 .. a_{n-1}] and [a_1 .. a_n],
 * a_{i+2} = a_i + a_{i+1}, each M31 value in the first QM31 coordinate.  Host buffers. */
int32_t zs_fib_trace(uint32_t a0, uint32_t a1, uint32_t log_n, uint32_t *out);

/* GenericStark(FibAir).prove (src/m31/stark.zig:303-657) with the Fibonacci AIR
 * of stark.zig:877-904 and the boundary assertions (0,0) = a0, (1,0) = a1,
 * (0,n-1) = claimed, on a device trace of two QM31 columns of 2^trace_log
 * values: trace commitment, quotient and DEEP codewords and FRI on the device,
 * transcript on the host.  The proof is returned as the reference's wire bytes
 * of GenericStark.Proof (serialization.zig), allocated with `host`; the
 * channel ends where the reference prover's does.
 * 1 <= log_blowup <= remainder_log < trace_log + log_blowup <= 24. */
int32_t zs_fib_stark_prove_dev(zs_hip_ctx *ctx, const uint32_t *d_trace, uint32_t trace_log, uint32_t log_blowup,
                               size_t num_queries, uint32_t remainder_log, const uint32_t a0[4], const uint32_t a1[4],
                               const uint32_t claimed[4], zs_channel *chan, zs_host_allocator host,
                               uint8_t **proof_out, size_t *proof_len);

/* ---- GenericStark(Air) with preprocessed columns and LogUp lookups ---------------- */
/* An AIR as data (the reference's comptime Air type, src/m31/stark.zig:19-69):
 *  - transition constraint k (Air.evalTransition) is the sum of the terms whose
 *    `constraint` is k, a term being coeff * (product of its factors), a factor
 *    being current[column], next[column] or the evaluation point x;
 *  - boundary assertions (Air.boundaryAssertions for the public inputs):
 *    (column, step, value);
 *  - LogUp metadata per relation r (lookup_selector_columns, ..._multiplicity_columns,
 *    lookup_key_columns, lookup_table_columns): relation r's key (trace) columns are
 *    key_columns[key_start[r] .. key_start[r+1]), its table (preprocessed) columns
 *    table_columns[table_start[r] .. table_start[r+1]);
 *  - Air.maxConstraintDegree(n) = degree_mul * (n - 1) + degree_add, or, with both
 *    0, the largest term degree (n - 1 per column factor, 1 per x factor).
 * Limits of the device prover: 32 committed columns (trace + accumulators +
 * preprocessed), 64 terms, 16 boundary assertions, 4 relations, 8 key columns
 * per relation. */
#define ZS_AIR_CUR 0u
#define ZS_AIR_NEXT 1u
#define ZS_AIR_X 2u
typedef struct { uint32_t kind, column; } zs_air_factor;
typedef struct {
    uint32_t constraint;
    uint32_t coeff[4];   /* QM31 */
    uint32_t n_factors;  /* 0..4 */
    zs_air_factor factor[4];
} zs_air_term;
typedef struct { uint32_t column, reserved; uint64_t step; uint32_t value[4]; } zs_boundary;
typedef struct {
    uint32_t num_columns, num_preprocessed, num_transition_constraints, num_terms;
    const zs_air_term *terms;
    uint32_t degree_mul, degree_add;
    uint32_t num_boundary, reserved;
    const zs_boundary *boundary;
    uint32_t num_lookup_columns, num_lookup_relations;
    const uint32_t *selector_column, *multiplicity_column;
    const uint32_t *key_start, *key_columns, *table_start, *table_columns;
} zs_air;
/* GenericStark(Air).proveWithPreprocessed (src/m31/stark.zig:311-657) on device
 * columns (d_pre: num_preprocessed, d_trace: num_columns QM31 columns of
 * 2^trace_log values, column-major; d_pre may be NULL without preprocessed
 * columns): preprocessed, trace and accumulator commitments, the quotient (every
 * synthesized LogUp, selector-binarity and boundary term) and the DEEP codeword
 * with batch inversion, FRI under friParams (stark.zig:197-222) and the query
 * reveals on the device; the transcript and the O(n) evaluations at z on the host.
 * The proof is returned as GenericStark.Proof wire bytes allocated with `host`.
 * ZS_ERR_INPUT for parameters friParams rejects (remainder_log below the FRI
 * blowup or not below trace_log + log_blowup) and for AIRs beyond the limits. */
int32_t zs_stark_prove_dev(zs_hip_ctx *ctx, const zs_air *air, const uint32_t *d_pre, const uint32_t *d_trace,
                           uint32_t trace_log, uint32_t log_blowup, size_t num_queries, uint32_t remainder_log,
                           zs_channel *chan, zs_host_allocator host, uint8_t **proof_out, size_t *proof_len);
/* GenericStark(Air).verify (stark.zig:660-780) of the wire bytes on the host, with
 * the checks of zs_fib_stark_verify (one reveal per FRI query, the nested FRI
 * parameters equal to friParams).  ZS_OK with *out_ok 1 or 0; ZS_ERR_INPUT for
 * bytes that do not deserialize or an AIR beyond the limits. */
int32_t zs_stark_verify(const uint8_t *proof, size_t proof_len, const zs_air *air, zs_channel *chan,
                        uint8_t *out_ok);

/* ---- synthetic input ---------------------------------------------------------- */
/* Seeded uniform M31 values: element i is the top 31 bits of the (i+1)-th
 * SplitMix64 output from `seed` (2^31 - 1 rejected). */
int32_t zs_fill_m31_dev(zs_hip_ctx *ctx, uint64_t seed, uint32_t *d_out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
