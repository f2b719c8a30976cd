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
// levels are built bottom-up, three per launch: a lane owns an 8-child
// subtree and hashes all 7 of its nodes, so no lane idles while the level
// narrows; from 2^16 nodes down, workgroups keep 8-level subtrees in LDS and
// hash the narrow levels with four lanes per compression.
#include "merkle_dev.h"

namespace zs {

// rows of <= 256 columns: one BLAKE3 chunk per leaf
__global__ __launch_bounds__(256) void row_leaf_kernel(const uint32_t *cols, uint64_t cs, uint32_t n_cols,
                                                       uint64_t n_rows, uint32_t *nodes) {
    const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row < n_rows) row_leaf_body(cols, cs, n_cols, row, nodes);
}
// rows of more than 256 columns
__global__ __launch_bounds__(256) void row_leaf_multichunk_kernel(const uint32_t *cols, uint64_t cs, uint32_t n_cols,
                                                                  uint64_t n_rows, uint32_t *nodes) {
    const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row < n_rows) row_leaf_multichunk_body(cols, cs, n_cols, row, nodes);
}

int32_t launch_row_leaves(Ctx *ctx, const uint32_t *d_cols, uint64_t col_stride, uint32_t n_cols,
                                 uint64_t n, uint32_t *d_nodes) {
    dim3 grid((uint32_t)((n + 255) / 256));
    if (n_cols <= 256)
        ZS_LAUNCH(ctx, "row_leaf_kernel", row_leaf_kernel, grid, dim3(256), 0, ctx->stream, d_cols, col_stride, n_cols, n,
                  d_nodes);
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


__global__ __launch_bounds__(256, 8) void tree_subtree3_multi_kernel(uint32_t *base, const MultiTrees mt) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= mt.lane_end[mt.n - 1]) return;
    uint32_t t = 0;
    while (g >= mt.lane_end[t]) t++;
    uint64_t start = t ? mt.lane_end[t - 1] : 0;
    subtree_lane<3>(base + mt.child[t], base + mt.lvl1[t], mt.lane_end[t] - start, g - start);
}


__global__ __launch_bounds__(BLOCK_NT) void tree_block_kernel(uint32_t *base, const BlockTrees bt) {
    tree_block_body<BLOCK_NT>(base, bt, blockIdx.x, threadIdx.x);
}

// levels with at least 2^kSubtreeMinLog children are hashed 3 at a time by one lane
// per 8-child subtree (throughput); below that, by workgroup subtrees whose quads
// share each hash (a lane's chain of 14 compressions would set the latency)
// once a tree that wide makes the 3-level pass worth a launch, narrower trees of the
// same launch join it down to 2^11 children: their lanes add little to a pass whose
// time one lane's 7 compressions set, and it takes 3 levels (7/8 of the workgroups)
// off the workgroup launch that follows
constexpr uint32_t kSubtreeJoinLog = 11;

// Every node level of several trees whose level `level[k]` is written: tree k
// has 2^log_leaves[k] leaves at word offset off_words[k] from `base`.  While some
// tree has a level of 2^18 children or more, every tree with at least 2^11
// children goes 3 levels at a time (one lane per 8-child subtree, all trees in one
// launch); then workgroup subtrees of 8 levels, then one workgroup per tree for
// the rest.
int32_t launch_trees_multi(Ctx *ctx, uint32_t *base, const uint64_t *off_words, const uint32_t *log_leaves,
                           uint32_t n_trees, B3Hasher *draw_chan, uint32_t *draw_alpha, const B3Hasher *chan_init,
                           uint32_t *powers, uint32_t n_powers, const FoldLeaves *fold) {
    // a fused draw needs the root computed by one workgroup: one tree with at least one level
    if (draw_chan && (n_trees != 1 || log_leaves[0] == 0)) return ZS_ERR_INPUT;
    // computed leaves go into the first launch: one tree whose first launch is a workgroup-subtree one
    if (fold && (n_trees != 1 || log_leaves[0] == 0 || log_leaves[0] >= kSubtreeMinLog)) return ZS_ERR_INPUT;
    std::vector<uint32_t> level(n_trees, 0);
    for (uint32_t start = 0; start < n_trees; start += MultiTrees::MAX) {
        const uint32_t cnt = std::min<uint32_t>(MultiTrees::MAX, n_trees - start);
        for (;;) {
            MultiTrees mt{};
            uint64_t lanes = 0;
            std::vector<uint32_t> in;
            bool wide = false;
            for (uint32_t k = start; k < start + cnt; k++) wide |= log_leaves[k] - level[k] >= kSubtreeMinLog;
            if (!wide) break;
            for (uint32_t k = start; k < start + cnt; k++) {
                uint32_t rem = log_leaves[k] - level[k];
                if (rem < kSubtreeJoinLog) continue;
                mt.child[mt.n] = off_words[k] + level_offset_words(log_leaves[k], level[k]);
                mt.lvl1[mt.n] = off_words[k] + level_offset_words(log_leaves[k], level[k] + 1);
                lanes += (uint64_t(1) << rem) >> 3;
                mt.lane_end[mt.n] = lanes;
                mt.n++;
                in.push_back(k);
            }
            if (mt.n == 0) break;
            ZS_LAUNCH(ctx, "tree_subtree3_multi_kernel", tree_subtree3_multi_kernel,
                      dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, ctx->stream, base, mt);
            ZS_HIP(ctx, hipGetLastError());
            for (uint32_t k : in) level[k] += 3;
        }
        for (;;) {
            BlockTrees bt{};
            uint32_t wgs = 0, kmax = 0;
            std::vector<uint32_t> in;
            for (uint32_t k = start; k < start + cnt; k++) {
                uint32_t rem = log_leaves[k] - level[k];
                if (rem == 0) continue;
                uint32_t K = rem > 8 ? 8 : rem;
                bt.child[bt.n] = off_words[k] + level_offset_words(log_leaves[k], level[k]);
                bt.r[bt.n] = rem;
                bt.K[bt.n] = K;
                wgs += 1u << (rem - K);
                bt.wg_end[bt.n] = wgs;
                bt.n++;
                kmax = std::max(kmax, K);
                in.push_back(k);
            }
            if (bt.n == 0) break;
            const bool last = draw_chan && bt.K[0] == bt.r[0];  // this launch computes the root
            bt.chan = last ? draw_chan : nullptr;
            bt.alpha = last ? draw_alpha : nullptr;
            bt.chan_init = last ? chan_init : nullptr;
            bt.powers = last ? powers : nullptr;
            bt.n_powers = last ? n_powers : 0u;
            if (fold && level[in[0]] == 0) {
                ZS_TRY(launch_tree_block_fold(ctx, base, bt, wgs, (block_leaf_lds_words(kmax) + (8u << kmax)) * 4,
                                              *fold));
            } else {
                ZS_LAUNCH(ctx, "tree_block_kernel", tree_block_kernel, dim3(wgs), dim3(BLOCK_NT),
                          (block_lds_words(kmax) + (last ? kDrawLdsWords : 0)) * 4, ctx->stream, base, bt);
                ZS_HIP(ctx, hipGetLastError());
            }
            for (uint32_t i = 0; i < bt.n; i++) level[in[i]] += bt.K[i];
        }
    }
    return ZS_OK;
}

// Build levels from_level+1 .. top over level from_level (already written): the
// levels from there up are laid out like a tree of 2^(log_leaves - from_level) leaves.
int32_t launch_tree_levels(Ctx *ctx, uint32_t *d_nodes, uint32_t log_leaves, uint32_t from_level,
                           B3Hasher *draw_chan, uint32_t *draw_alpha, const B3Hasher *chan_init, uint32_t *powers,
                           uint32_t n_powers) {
    const uint64_t off = level_offset_words(log_leaves, from_level);
    const uint32_t lg = log_leaves - from_level;
    return launch_trees_multi(ctx, d_nodes, &off, &lg, 1, draw_chan, draw_alpha, chan_init, powers, n_powers);
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
    ZS_LAUNCH(ctx, "qm31_leaf_kernel", qm31_leaf_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_vals, n,
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
    ZS_HIP(ctx, hipMemcpyAsync(r.ptr, reqs.data(), reqs.size() * sizeof(PathReq), hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t n = (uint32_t)reqs.size();
    hipLaunchKernelGGL(gather_paths_kernel, dim3((n + 127) / 128), dim3(128), 0, ctx->stream, d_trees,
                       r.as<PathReq>(), n, d_out);
    ZS_HIP(ctx, hipGetLastError());
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

// roots of up to 28 trees (8 words each at off[k]) gathered into out, one word per thread
int32_t launch_gather_words(Ctx *ctx, const uint32_t *d_src, const std::vector<uint64_t> &pairs, uint32_t words,
                            uint32_t *d_out) {
    if (pairs.empty()) return ZS_OK;
    DevBuf r;
    ZS_HIP(ctx, hipMalloc(&r.ptr, pairs.size() * 8));
    ZS_HIP(ctx, hipMemcpyAsync(r.ptr, pairs.data(), pairs.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t n = (uint32_t)(pairs.size() / 2);
    hipLaunchKernelGGL(gather_words_kernel, dim3((n + 127) / 128), dim3(128), 0, ctx->stream, d_src,
                       r.as<uint64_t>(), n, words, d_out);
    ZS_HIP(ctx, hipGetLastError());
    ZS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ZS_OK;
}

}  // namespace zs
