// merkle.hip -- BLAKE3 leaf hashing and Merkle node levels for gfx950.
//
// Leaves: hashM31s(row) (src/m31/hash.zig:15-25) over the values of one row
// across all columns, and hashQM31 (src/m31/hash.zig:33-37).  Nodes:
// hash2(left, right) (src/core/hash/hash.zig:17-25), level l+1 node i over
// level l nodes 2i and 2i+1 (src/core/merkle/merkle.zig:36-41).  Every level
// is stored, leaves first, as 8 little-endian u32 words per digest.
//
// The row-leaf kernel reads columns at the row's index (lanes take adjacent
// rows, so every column load is a coalesced 256-byte wave access) and
// compresses 16 columns per BLAKE3 block; rows of more than 256 columns span
// several 1024-byte chunks and are merged with BLAKE3's chunk tree.  Node
// levels are built bottom-up; the lowest levels are hashed inside a wave
// (a lane keeps its node, the partner's digest arrives by lane shuffle) so
// that a launch writes several levels.
#include "blake3.h"
#include "internal.h"

namespace zs {

// BLAKE3 of one chunk (<= 256 columns = 1024 bytes) of a row.
__device__ __forceinline__ void row_chunk_cv(const uint32_t *cols, uint64_t cs, uint64_t row, uint32_t c0,
                                             uint32_t ncols, uint32_t chunk_idx, bool root, uint32_t cv[8]) {
    b3_iv(cv);
    uint32_t len = ncols * 4;
    uint32_t nblocks = (len + 63) / 64;
    for (uint32_t b = 0; b < nblocks; b++) {
        uint32_t m[16];
#pragma unroll
        for (int w = 0; w < 16; w++) {
            uint32_t c = b * 16 + w;
            m[w] = c < ncols ? cols[(uint64_t)(c0 + c) * cs + row] : 0u;
        }
        uint32_t blen = (len - b * 64) < 64 ? (len - b * 64) : 64;
        uint32_t flags = (b == 0 ? B3_CHUNK_START : 0) | (b == nblocks - 1 ? (B3_CHUNK_END | (root ? B3_ROOT : 0)) : 0);
        b3_compress_cv(cv, m, chunk_idx, 0, blen, flags);
    }
}

// rows of <= 256 columns: one BLAKE3 chunk per leaf
__global__ __launch_bounds__(256) void row_leaf_kernel(const uint32_t *cols, uint64_t cs, uint32_t n_cols,
                                                       uint64_t n_rows, uint32_t *nodes) {
    uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n_rows) return;
    uint32_t cv[8];
    row_chunk_cv(cols, cs, row, 0, n_cols, 0, true, cv);
    uint4 *o = reinterpret_cast<uint4 *>(nodes + row * 8);
    o[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
    o[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

// rows of more than 256 columns: BLAKE3 chunk tree with a chaining-value stack
__global__ __launch_bounds__(256) void row_leaf_multichunk_kernel(const uint32_t *cols, uint64_t cs, uint32_t n_cols,
                                                                  uint64_t n_rows, uint32_t *nodes) {
    uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n_rows) return;
    uint32_t cv[8];
    uint32_t n_chunks = (n_cols + 255) / 256;
    uint32_t stack[16][8];
    uint32_t sp = 0;
    for (uint32_t c = 0; c + 1 < n_chunks; c++) {
        row_chunk_cv(cols, cs, row, c * 256, 256, c, false, cv);
        uint32_t total = c + 1;
        while ((total & 1u) == 0) {
            sp--;
            uint32_t t[8];
            b3_parent(stack[sp], cv, false, t);
            for (int w = 0; w < 8; w++) cv[w] = t[w];
            total >>= 1;
        }
        for (int w = 0; w < 8; w++) stack[sp][w] = cv[w];
        sp++;
    }
    uint32_t last = n_chunks - 1;
    row_chunk_cv(cols, cs, row, last * 256, n_cols - last * 256, last, false, cv);
    while (sp > 0) {
        sp--;
        uint32_t t[8];
        b3_parent(stack[sp], cv, sp == 0, t);
        for (int w = 0; w < 8; w++) cv[w] = t[w];
    }
    uint4 *o = reinterpret_cast<uint4 *>(nodes + row * 8);
    o[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
    o[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

int32_t launch_row_leaves(Ctx *ctx, const uint32_t *d_cols, uint64_t col_stride, uint32_t n_cols,
                                 uint64_t n, uint32_t *d_nodes) {
    dim3 grid((uint32_t)((n + 255) / 256));
    if (n_cols <= 256)
        hipLaunchKernelGGL(row_leaf_kernel, grid, dim3(256), 0, ctx->stream, d_cols, col_stride, n_cols, n, d_nodes);
    else
        hipLaunchKernelGGL(row_leaf_multichunk_kernel, grid, dim3(256), 0, ctx->stream, d_cols, col_stride, n_cols,
                           n, d_nodes);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

__global__ __launch_bounds__(256) void qm31_leaf_kernel(const uint32_t *vals, uint64_t n, uint32_t *nodes) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 v = reinterpret_cast<const uint4 *>(vals)[i];
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t d[8];
    b3_hash_qm31(w, d);
    uint4 *o = reinterpret_cast<uint4 *>(nodes + i * 8);
    o[0] = make_uint4(d[0], d[1], d[2], d[3]);
    o[1] = make_uint4(d[4], d[5], d[6], d[7]);
}

// Up to `nlev` levels per launch: lane i hashes node i of the first new level
// from two children in memory, then pairs of lanes combine by shuffle.
__global__ __launch_bounds__(256) void tree_levels_kernel(uint32_t *child_level, uint32_t *parent_level,
                                                          uint64_t n_parents, uint32_t nlev, uint64_t n_parents_total) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool active = i < n_parents;
    uint32_t d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (active) {
        const uint4 *c = reinterpret_cast<const uint4 *>(child_level + i * 16);
        uint4 a0 = c[0], a1 = c[1], b0 = c[2], b1 = c[3];
        uint32_t l[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        uint32_t r[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
        b3_hash2(l, r, d);
        uint4 *o = reinterpret_cast<uint4 *>(parent_level + i * 8);
        o[0] = make_uint4(d[0], d[1], d[2], d[3]);
        o[1] = make_uint4(d[4], d[5], d[6], d[7]);
    }
    // further levels inside the wave: node i at this level pairs with i ^ 1
    uint32_t *lvl = parent_level;
    uint64_t cnt = n_parents;
    uint64_t idx = i;
    for (uint32_t k = 1; k < nlev; k++) {
        uint32_t partner[8];
        uint32_t lane_off = 1u << (k - 1);  // partner's lane distance at level k-1
#pragma unroll
        for (int w = 0; w < 8; w++) partner[w] = __shfl_xor(d[w], lane_off);
        uint64_t next_cnt = cnt >> 1;
        uint32_t *next_lvl = lvl + cnt * 8;
        bool left = ((idx >> (k - 1)) & 1u) == 0 && ((i & ((1u << (k - 1)) - 1)) == 0);
        if (active && left && (idx >> k) < next_cnt) {
            uint32_t t[8];
            b3_hash2(d, partner, t);
            for (int w = 0; w < 8; w++) d[w] = t[w];
            uint4 *o = reinterpret_cast<uint4 *>(next_lvl + (idx >> k) * 8);
            o[0] = make_uint4(d[0], d[1], d[2], d[3]);
            o[1] = make_uint4(d[4], d[5], d[6], d[7]);
        }
        lvl = next_lvl;
        cnt = next_cnt;
    }
    (void)n_parents_total;
}

// Build levels from_level+1 .. top over level from_level (already written).
int32_t launch_tree_levels(Ctx *ctx, uint32_t *d_nodes, uint32_t log_leaves, uint32_t from_level) {
    uint32_t level = from_level;
    while (level < log_leaves) {
        uint32_t remaining = log_leaves - level;
        uint32_t nlev = remaining < 6 ? remaining : 6;  // 6 levels fit a 64-lane wave
        uint64_t n_parents = (uint64_t(1) << log_leaves) >> (level + 1);
        uint32_t *child = d_nodes + level_offset_words(log_leaves, level);
        uint32_t *parent = d_nodes + level_offset_words(log_leaves, level + 1);
        // shuffle levels only combine lanes of one wave
        if ((uint64_t(1) << (nlev - 1)) > n_parents) nlev = 1;
        uint32_t blocks = (uint32_t)((n_parents + 255) / 256);
        hipLaunchKernelGGL(tree_levels_kernel, dim3(blocks), dim3(256), 0, ctx->stream, child, parent, n_parents, nlev,
                           n_parents);
        ZS_HIP(ctx, hipGetLastError());
        level += nlev;
    }
    return ZS_OK;
}

int32_t launch_row_commit(Ctx *ctx, const uint32_t *d_cols, uint64_t col_stride, uint32_t n_cols, uint32_t log_rows,
                          uint32_t *d_nodes) {
    uint64_t n = uint64_t(1) << log_rows;
    ZS_TRY(launch_row_leaves(ctx, d_cols, col_stride, n_cols, n, d_nodes));
    return launch_tree_levels(ctx, d_nodes, log_rows, 0);
}

// plain: the same leaves, then one launch per node level
__global__ __launch_bounds__(256) void tree_level_plain_kernel(const uint32_t *child, uint32_t *parent, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t d[8];
    b3_hash2(child + i * 16, child + i * 16 + 8, d);
    for (int w = 0; w < 8; w++) parent[i * 8 + w] = d[w];
}

int32_t launch_row_commit_plain(Ctx *ctx, const uint32_t *d_cols, uint64_t col_stride, uint32_t n_cols,
                                uint32_t log_rows, uint32_t *d_nodes) {
    uint64_t n = uint64_t(1) << log_rows;
    ZS_TRY(launch_row_leaves(ctx, d_cols, col_stride, n_cols, n, d_nodes));
    for (uint32_t l = 0; l < log_rows; l++) {
        uint64_t np = n >> (l + 1);
        hipLaunchKernelGGL(tree_level_plain_kernel, dim3((uint32_t)((np + 255) / 256)), dim3(256), 0, ctx->stream,
                           d_nodes + level_offset_words(log_rows, l), d_nodes + level_offset_words(log_rows, l + 1),
                           np);
    }
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

int32_t launch_qm31_leaves(Ctx *ctx, const uint32_t *d_vals, uint64_t n, uint32_t *d_nodes) {
    hipLaunchKernelGGL(qm31_leaf_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_vals, n,
                       d_nodes);
    ZS_HIP(ctx, hipGetLastError());
    return ZS_OK;
}

}  // namespace zs

namespace zs {

// ---- opening gathers: Merkle paths and stored values --------------------------------
__global__ void gather_paths_kernel(const uint32_t *trees, const PathReq *reqs, uint32_t n, uint32_t *out) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    PathReq q = reqs[r];
    uint64_t lvl_off = 0, cnt = uint64_t(1) << q.log_leaves, idx = q.index;
    for (uint32_t l = 0; l < q.log_leaves; l++) {
        const uint32_t *src = trees + q.tree_off_words + (lvl_off + (idx ^ 1)) * 8;
        uint32_t *dst = out + q.out_off_words + uint64_t(l) * 8;
        for (int w = 0; w < 8; w++) dst[w] = src[w];
        lvl_off += cnt;
        cnt >>= 1;
        idx >>= 1;
    }
}

__global__ void gather_words_kernel(const uint32_t *src, const uint64_t *pairs, uint32_t n, uint32_t words,
                                    uint32_t *out) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint64_t s = pairs[2 * r], d = pairs[2 * r + 1];
    for (uint32_t w = 0; w < words; w++) out[d + w] = src[s + w];
}

int32_t launch_gather_paths(Ctx *ctx, const uint32_t *d_trees, const std::vector<PathReq> &reqs, uint32_t *d_out) {
    if (reqs.empty()) return ZS_OK;
    DevBuf r;
    ZS_HIP(ctx, hipMalloc(&r.ptr, reqs.size() * sizeof(PathReq)));
    ZS_HIP(ctx, hipMemcpy(r.ptr, reqs.data(), reqs.size() * sizeof(PathReq), hipMemcpyHostToDevice));
    uint32_t n = (uint32_t)reqs.size();
    hipLaunchKernelGGL(gather_paths_kernel, dim3((n + 127) / 128), dim3(128), 0, ctx->stream, d_trees,
                       r.as<PathReq>(), n, d_out);
    ZS_HIP(ctx, hipGetLastError());
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

int32_t launch_gather_words(Ctx *ctx, const uint32_t *d_src, const std::vector<uint64_t> &pairs, uint32_t words,
                            uint32_t *d_out) {
    if (pairs.empty()) return ZS_OK;
    DevBuf r;
    ZS_HIP(ctx, hipMalloc(&r.ptr, pairs.size() * 8));
    ZS_HIP(ctx, hipMemcpy(r.ptr, pairs.data(), pairs.size() * 8, hipMemcpyHostToDevice));
    uint32_t n = (uint32_t)(pairs.size() / 2);
    hipLaunchKernelGGL(gather_words_kernel, dim3((n + 127) / 128), dim3(128), 0, ctx->stream, d_src,
                       r.as<uint64_t>(), n, words, d_out);
    ZS_HIP(ctx, hipGetLastError());
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

}  // namespace zs
