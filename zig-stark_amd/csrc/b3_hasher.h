// b3_hasher.h -- incremental BLAKE3 (chunk state + chaining-value stack) and the
// Fiat-Shamir channel operations over it (src/core/channel/channel.zig), for
// host and device.
//
// The transcript is "zig-stark:channel" || domain_separator || absorbed bytes
// (channel.zig:16-43).  sampleBytes (channel.zig:49-69) hashes a snapshot of
// the transcript extended by "zig-stark:sample" || u64le(counter) and appends
// the whole 32-byte block to the transcript; sample(QM31) reduces each u32 of
// 16 sampled bytes mod p (qm31.zig:95-101); sampleIndex (channel.zig:79-101)
// reads floor(log2 n) + 1 bits big-endian and rejects values >= n.
//
// The state is plain data (no pointers), so a host channel can be copied to
// the device, advanced there by a kernel, and copied back, with the same code
// on both sides.  A snapshot is a struct copy: the transcript is never re-hashed.
#pragma once
#include "blake3.h"

namespace zs {

struct B3Hasher {
    uint32_t cv[8];
    uint64_t chunk_counter;
    uint8_t block[64];
    uint32_t block_len;
    uint32_t blocks_compressed;
    uint32_t stack[54][8];
    uint32_t stack_len;

    ZS_HD void init() {
        b3_iv(cv);
        chunk_counter = 0;
        for (int i = 0; i < 64; i++) block[i] = 0;
        block_len = 0;
        blocks_compressed = 0;
        stack_len = 0;
    }
    ZS_HD static void words(const uint8_t *b, uint32_t m[16]) {
        for (int i = 0; i < 16; i++)
            m[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
                   ((uint32_t)b[4 * i + 3] << 24);
    }
    ZS_HD uint32_t chunk_len() const { return blocks_compressed * 64 + block_len; }
    ZS_HD uint32_t start_flag() const { return blocks_compressed == 0 ? B3_CHUNK_START : 0; }
    ZS_HD void push_chunk_cv(uint32_t new_cv[8], uint64_t total_chunks) {
        while ((total_chunks & 1) == 0) {
            uint32_t t[8];
            stack_len--;
            b3_parent(stack[stack_len], new_cv, false, t);
            for (int i = 0; i < 8; i++) new_cv[i] = t[i];
            total_chunks >>= 1;
        }
        for (int i = 0; i < 8; i++) stack[stack_len][i] = new_cv[i];
        stack_len++;
    }
    ZS_HD void update(const uint8_t *in, size_t len) {
        while (len > 0) {
            if (chunk_len() == 1024) {
                // the chunk is complete and more input follows: finish it as a non-root chunk
                uint32_t m[16], c[8];
                words(block, m);
                for (int i = 0; i < 8; i++) c[i] = cv[i];
                b3_compress_cv(c, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), 64,
                               start_flag() | B3_CHUNK_END);
                push_chunk_cv(c, chunk_counter + 1);
                chunk_counter++;
                b3_iv(cv);
                block_len = 0;
                blocks_compressed = 0;
                for (int i = 0; i < 64; i++) block[i] = 0;
            }
            if (block_len == 64) {
                uint32_t m[16];
                words(block, m);
                b3_compress_cv(cv, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), 64, start_flag());
                blocks_compressed++;
                block_len = 0;
                for (int i = 0; i < 64; i++) block[i] = 0;
            }
            uint32_t want = 1024 - chunk_len();
            uint32_t room = 64 - block_len;
            size_t take = len < room ? len : room;
            if (take > want) take = want;
            for (size_t i = 0; i < take; i++) block[block_len + i] = in[i];
            block_len += (uint32_t)take;
            in += take;
            len -= take;
        }
    }
    ZS_HD void finalize(uint8_t out[32]) const {
        uint32_t m[16], r[8];
        words(block, m);
        for (int i = 0; i < 8; i++) r[i] = cv[i];
        uint32_t flags = start_flag() | B3_CHUNK_END;
        if (stack_len == 0) {
            b3_compress_cv(r, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), block_len, flags | B3_ROOT);
        } else {
            b3_compress_cv(r, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), block_len, flags);
            for (int i = (int)stack_len - 1; i >= 0; i--) {
                uint32_t t[8];
                b3_parent(stack[i], r, i == 0, t);
                for (int w = 0; w < 8; w++) r[w] = t[w];
            }
        }
        for (int i = 0; i < 8; i++) {
            out[4 * i] = (uint8_t)r[i];
            out[4 * i + 1] = (uint8_t)(r[i] >> 8);
            out[4 * i + 2] = (uint8_t)(r[i] >> 16);
            out[4 * i + 3] = (uint8_t)(r[i] >> 24);
        }
    }
};

// copy of the parts of a hasher that are in use (the stack up to stack_len)
ZS_HD void b3_copy_live(B3Hasher &dst, const B3Hasher &src) {
    for (int i = 0; i < 8; i++) dst.cv[i] = src.cv[i];
    dst.chunk_counter = src.chunk_counter;
    for (int i = 0; i < 64; i++) dst.block[i] = src.block[i];
    dst.block_len = src.block_len;
    dst.blocks_compressed = src.blocks_compressed;
    dst.stack_len = src.stack_len;
    for (uint32_t k = 0; k < src.stack_len; k++)
        for (int i = 0; i < 8; i++) dst.stack[k][i] = src.stack[k][i];
}

// channel.zig:49-69; `snap` is working storage for the transcript snapshot
ZS_HD void channel_sample_bytes(B3Hasher &h, B3Hasher &snap, uint8_t *out, size_t len) {
    const uint8_t tag[16] = {'z', 'i', 'g', '-', 's', 't', 'a', 'r', 'k', ':', 's', 'a', 'm', 'p', 'l', 'e'};
    uint64_t counter = 0;
    while (len > 0) {
        b3_copy_live(snap, h);
        snap.update(tag, 16);
        uint8_t cbuf[8];
        for (int i = 0; i < 8; i++) cbuf[i] = (uint8_t)(counter >> (8 * i));
        snap.update(cbuf, 8);
        uint8_t blk[32];
        snap.finalize(blk);
        h.update(blk, 32);
        size_t k = len < 32 ? len : 32;
        for (size_t i = 0; i < k; i++) out[i] = blk[i];
        out += k;
        len -= k;
        counter++;
    }
}

// sample(QM31): qm31.zig:95-101 (each u32 reduced mod p)
ZS_HD void channel_sample_qm31(B3Hasher &h, B3Hasher &snap, uint32_t out[4]) {
    uint8_t b[16];
    channel_sample_bytes(h, snap, b, 16);
    for (int i = 0; i < 4; i++) {
        uint32_t v = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
                     ((uint32_t)b[4 * i + 3] << 24);
        out[i] = m31_from_u32(v);
    }
}

// absorb a 32-byte digest held as 8 little-endian words
ZS_HD void channel_absorb_digest(B3Hasher &h, const uint32_t d[8]) {
    uint8_t b[32];
    for (int i = 0; i < 8; i++) {
        b[4 * i] = (uint8_t)d[i];
        b[4 * i + 1] = (uint8_t)(d[i] >> 8);
        b[4 * i + 2] = (uint8_t)(d[i] >> 16);
        b[4 * i + 3] = (uint8_t)(d[i] >> 24);
    }
    h.update(b, 32);
}

}  // namespace zs
