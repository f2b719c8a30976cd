// channel.cpp -- host-side Fiat-Shamir channel (src/core/channel/channel.zig)
// over an incremental BLAKE3 hasher.
//
// The transcript is "zig-stark:channel" || domain_separator || absorbed bytes
// (channel.zig:16-43).  sampleBytes (channel.zig:49-69) hashes a snapshot of
// the transcript extended by "zig-stark:sample" || u64le(counter) and appends
// the whole 32-byte block to the transcript; sample(QM31) reduces each u32 of
// 16 sampled bytes mod p (qm31.zig:95-101); sampleIndex (channel.zig:79-101)
// reads floor(log2 n) + 1 bits big-endian and rejects values >= n.
//
// The hasher keeps BLAKE3's chunk state and chaining-value stack, so a
// snapshot is a struct copy and the transcript is never re-hashed.
#include <cstring>

#include "blake3.h"
#include "../../include/zs_hip.h"

namespace zs {

struct B3Hasher {
    uint32_t cv[8];
    uint64_t chunk_counter = 0;
    uint8_t block[64];
    uint32_t block_len = 0;
    uint32_t blocks_compressed = 0;
    uint32_t stack[54][8];
    uint32_t stack_len = 0;

    B3Hasher() {
        b3_iv(cv);
        std::memset(block, 0, sizeof block);
    }
    static void words(const uint8_t *b, uint32_t m[16]) {
        for (int i = 0; i < 16; i++)
            m[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
                   ((uint32_t)b[4 * i + 3] << 24);
    }
    uint32_t chunk_len() const { return blocks_compressed * 64 + block_len; }
    uint32_t start_flag() const { return blocks_compressed == 0 ? B3_CHUNK_START : 0; }
    void push_chunk_cv(uint32_t new_cv[8], uint64_t total_chunks) {
        while ((total_chunks & 1) == 0) {
            uint32_t t[8];
            stack_len--;
            b3_parent(stack[stack_len], new_cv, false, t);
            std::memcpy(new_cv, t, 32);
            total_chunks >>= 1;
        }
        std::memcpy(stack[stack_len++], new_cv, 32);
    }
    void update(const uint8_t *in, size_t len) {
        while (len > 0) {
            if (chunk_len() == 1024) {
                // the chunk is complete and more input follows: finish it as a non-root chunk
                uint32_t m[16], c[8];
                words(block, m);
                std::memcpy(c, cv, 32);
                b3_compress_cv(c, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), 64,
                               start_flag() | B3_CHUNK_END);
                push_chunk_cv(c, chunk_counter + 1);
                chunk_counter++;
                b3_iv(cv);
                block_len = 0;
                blocks_compressed = 0;
                std::memset(block, 0, sizeof block);
            }
            if (block_len == 64) {
                uint32_t m[16];
                words(block, m);
                b3_compress_cv(cv, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), 64, start_flag());
                blocks_compressed++;
                block_len = 0;
                std::memset(block, 0, sizeof block);
            }
            uint32_t want = 1024 - chunk_len();
            uint32_t room = 64 - block_len;
            size_t take = len < room ? len : room;
            if (take > want) take = want;
            std::memcpy(block + block_len, in, take);
            block_len += (uint32_t)take;
            in += take;
            len -= take;
        }
    }
    void finalize(uint8_t out[32]) const {
        uint32_t m[16], r[8];
        words(block, m);
        std::memcpy(r, cv, 32);
        uint32_t flags = start_flag() | B3_CHUNK_END;
        if (stack_len == 0) {
            b3_compress_cv(r, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), block_len, flags | B3_ROOT);
        } else {
            b3_compress_cv(r, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), block_len, flags);
            for (int i = (int)stack_len - 1; i >= 0; i--) {
                uint32_t t[8];
                b3_parent(stack[i], r, i == 0, t);
                std::memcpy(r, t, 32);
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

}  // namespace zs

struct zs_channel {
    zs::B3Hasher h;
};

extern "C" {

zs_channel *zs_channel_new(const uint8_t *sep, size_t len) {
    auto *c = new zs_channel();
    c->h.update(reinterpret_cast<const uint8_t *>("zig-stark:channel"), 17);
    if (len) c->h.update(sep, len);
    return c;
}

zs_channel *zs_channel_clone(const zs_channel *ch) { return new zs_channel(*ch); }

void zs_channel_free(zs_channel *ch) { delete ch; }

void zs_channel_absorb(zs_channel *ch, const uint8_t *data, size_t len) { ch->h.update(data, len); }

void zs_channel_sample_bytes(zs_channel *ch, uint8_t *out, size_t len) {
    uint64_t counter = 0;
    while (len > 0) {
        zs::B3Hasher snap = ch->h;
        snap.update(reinterpret_cast<const uint8_t *>("zig-stark:sample"), 16);
        uint8_t cbuf[8];
        for (int i = 0; i < 8; i++) cbuf[i] = (uint8_t)(counter >> (8 * i));
        snap.update(cbuf, 8);
        uint8_t blk[32];
        snap.finalize(blk);
        ch->h.update(blk, 32);
        size_t k = len < 32 ? len : 32;
        std::memcpy(out, blk, k);
        out += k;
        len -= k;
        counter++;
    }
}

void zs_channel_sample_qm31(zs_channel *ch, uint32_t out[4]) {
    uint8_t b[16];
    zs_channel_sample_bytes(ch, b, 16);
    for (int i = 0; i < 4; i++) {
        uint32_t v = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
                     ((uint32_t)b[4 * i + 3] << 24);
        out[i] = zs::m31_from_u32(v);
    }
}

size_t zs_channel_sample_index(zs_channel *ch, size_t n) {
    if (n == 0) return 0;
    unsigned bits = 0;
    while (bits < 63 && (size_t(1) << (bits + 1)) <= n) bits++;
    bits += 1;
    size_t bytes = (bits + 7) / 8;
    size_t mask = bits >= 64 ? ~size_t(0) : ((size_t(1) << bits) - 1);
    uint8_t buf[8];
    for (;;) {
        zs_channel_sample_bytes(ch, buf, bytes);
        size_t v = 0;
        for (size_t i = 0; i < bytes; i++) v = (v << 8) | buf[i];
        v &= mask;
        if (v < n) return v;
    }
}

}  // extern "C"
