/* abi_smoke.c -- a C caller of libzs_hip.so, bound the way a Zig or C++ host
 * binds it (include/zs_hip.h, no Python): LDE + row commitment of host
 * columns, a Merkle opening, and an FRI commit whose challenges come from the
 * channel.  Inputs are a fixed integer sequence; the program prints the
 * results as hex lines for tests/test_gpu_c_abi.py to compare with the oracle.
 *
 *   abi_smoke N_COLS LOG_N LOG_BLOWUP OPEN_INDEX FRI_LOG
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "zs_hip.h"

#define P 0x7fffffffu

static void hex(const char *tag, const uint8_t *b, size_t n) {
    printf("%s ", tag);
    for (size_t i = 0; i < n; i++) printf("%02x", b[i]);
    printf("\n");
}

/* element i of column c (the test regenerates the same values) */
static uint32_t cell(uint64_t c, uint64_t i) { return (uint32_t)((c * 2654435761u + i * 40503u + 7u) % P); }

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s N_COLS LOG_N LOG_BLOWUP OPEN_INDEX FRI_LOG\n", argv[0]);
        return 2;
    }
    const uint32_t n_cols = (uint32_t)atoi(argv[1]), log_n = (uint32_t)atoi(argv[2]);
    const uint32_t log_blowup = (uint32_t)atoi(argv[3]), fri_log = (uint32_t)atoi(argv[5]);
    const size_t open_index = (size_t)atol(argv[4]);
    const size_t n = (size_t)1 << log_n, N = n << log_blowup;
    printf("version %s\n", zs_hip_version());

    zs_hip_ctx *ctx = NULL;
    int32_t rc = zs_hip_init(0, 0, &ctx);
    if (rc != ZS_OK) {
        fprintf(stderr, "zs_hip_init: %d\n", rc);
        return 3;
    }
    uint32_t *cols = (uint32_t *)malloc(n_cols * n * 4);
    uint32_t *lde = (uint32_t *)malloc(n_cols * N * 4);
    uint8_t *path = (uint8_t *)malloc(32 * (log_n + log_blowup + 1));
    if (!cols || !lde || !path) return 4;
    for (uint32_t c = 0; c < n_cols; c++)
        for (size_t i = 0; i < n; i++) cols[c * n + i] = cell(c, i);

    zs_tree *tree = NULL;
    uint8_t root[32];
    rc = zs_m31_lde_commit(ctx, cols, n_cols, log_n, log_blowup, &tree, root, lde);
    if (rc != ZS_OK) {
        fprintf(stderr, "zs_m31_lde_commit: %d (%s)\n", rc, zs_hip_last_error(ctx));
        return 5;
    }
    hex("root", root, 32);
    uint64_t sum = 0;
    for (size_t i = 0; i < n_cols * N; i++) sum = (sum * 31 + lde[i]) % 0xffffffffffc5ull;
    printf("lde_checksum %llu\n", (unsigned long long)sum);
    rc = zs_merkle_open(tree, open_index, path);
    if (rc != ZS_OK) return 6;
    hex("path", path, 32 * (log_n + log_blowup));
    zs_tree_free(tree);

    /* FRI: a QM31 codeword of 2^fri_log from the same sequence, challenges from the channel */
    const size_t m = (size_t)1 << fri_log;
    uint32_t *cw = (uint32_t *)malloc(m * 16);
    uint8_t *roots = (uint8_t *)malloc(32 * fri_log);
    uint32_t rem[4 * 64];
    if (!cw || !roots) return 4;
    for (size_t i = 0; i < 4 * m; i++) cw[i] = cell(1000, i);
    zs_channel *ch = zs_channel_new((const uint8_t *)"c-abi", 5);
    const uint32_t remainder_log = 3;
    rc = zs_fri_commit(ctx, cw, fri_log, 2, remainder_log, NULL, ch, roots, rem, NULL);
    if (rc != ZS_OK) {
        fprintf(stderr, "zs_fri_commit: %d (%s)\n", rc, zs_hip_last_error(ctx));
        return 7;
    }
    hex("fri_roots", roots, 32 * (fri_log - remainder_log));
    hex("remainder", (const uint8_t *)rem, 16 * ((size_t)1 << (remainder_log - 2)));
    uint8_t after[32];
    zs_channel_sample_bytes(ch, after, 32);
    hex("channel_after", after, 32);
    zs_channel_free(ch);

    free(cw);
    free(roots);
    free(path);
    free(lde);
    free(cols);
    zs_hip_destroy(ctx);
    printf("ok\n");
    return 0;
}
