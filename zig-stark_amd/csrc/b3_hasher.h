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

// The compression a hasher runs: cv <- compress(cv, m, counter, block_len, flags).
// B3Hasher's methods take one as a template argument; the default is the
// single-lane b3_compress_cv (the device channel passes a quad-lane one).
// lane() / lanes() / sync(): the threads that run the hasher together split its
// byte and word copies (lane k takes every lanes()-th element), then sync().
struct B3Single {
    ZS_HD void operator()(uint32_t cv[8], const uint32_t m[16], uint32_t counter_lo, uint32_t counter_hi,
                          uint32_t block_len, uint32_t flags) const {
        b3_compress_cv(cv, m, counter_lo, counter_hi, block_len, flags);
    }
    ZS_HD uint32_t lane() const { return 0; }
    ZS_HD uint32_t lanes() const { return 1; }
    ZS_HD void sync() const {}
};

// parent node of BLAKE3's chunk tree through a given compression
template <typename CF>
ZS_HD void b3_parent_with(const uint32_t l[8], const uint32_t r[8], bool root, uint32_t out[8], CF &cf) {
    uint32_t m[16];
    for (int i = 0; i < 8; i++) {
        m[i] = l[i];
        m[8 + i] = r[i];
    }
    b3_iv(out);
    cf(out, m, 0, 0, 64, B3_PARENT | (root ? B3_ROOT : 0));
}

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
    // the block as 16 little-endian words (block sits at a word-aligned offset)
    ZS_HD static void words(const uint8_t *b, uint32_t m[16]) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(b);
        for (int i = 0; i < 16; i++) m[i] = w[i];
    }
    ZS_HD uint32_t chunk_len() const { return blocks_compressed * 64 + block_len; }
    ZS_HD uint32_t start_flag() const { return blocks_compressed == 0 ? B3_CHUNK_START : 0; }
    template <typename CF>
    ZS_HD void push_chunk_cv(uint32_t new_cv[8], uint64_t total_chunks, CF &cf) {
        while ((total_chunks & 1) == 0) {
            uint32_t t[8];
            stack_len--;
            b3_parent_with(stack[stack_len], new_cv, false, t, cf);
            for (int i = 0; i < 8; i++) new_cv[i] = t[i];
            total_chunks >>= 1;
        }
        for (int i = 0; i < 8; i++) stack[stack_len][i] = new_cv[i];
        stack_len++;
    }
    ZS_HD void update(const uint8_t *in, size_t len) {
        B3Single cf;
        update(in, len, cf);
    }
    template <typename CF>
    ZS_HD void zero_block(CF &cf) {
        uint32_t *bw = reinterpret_cast<uint32_t *>(block);
        for (uint32_t i = cf.lane(); i < 16; i += cf.lanes()) bw[i] = 0;
        cf.sync();
    }
    template <typename CF>
    ZS_HD void update(const uint8_t *in, size_t len, CF &cf) {
        while (len > 0) {
            if (chunk_len() == 1024) {
                // the chunk is complete and more input follows: finish it as a non-root chunk
                uint32_t m[16], c[8];
                words(block, m);
                for (int i = 0; i < 8; i++) c[i] = cv[i];
                cf(c, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), 64, start_flag() | B3_CHUNK_END);
                push_chunk_cv(c, chunk_counter + 1, cf);
                chunk_counter++;
                b3_iv(cv);
                block_len = 0;
                blocks_compressed = 0;
                zero_block(cf);
            }
            if (block_len == 64) {
                uint32_t m[16];
                words(block, m);
                cf(cv, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), 64, start_flag());
                blocks_compressed++;
                block_len = 0;
                zero_block(cf);
            }
            uint32_t want = 1024 - chunk_len();
            uint32_t room = 64 - block_len;
            size_t take = len < room ? len : room;
            if (take > want) take = want;
            for (size_t i = cf.lane(); i < take; i += cf.lanes()) block[block_len + i] = in[i];
            cf.sync();
            block_len += (uint32_t)take;
            in += take;
            len -= take;
        }
    }
    ZS_HD void finalize(uint8_t out[32]) const {
        B3Single cf;
        finalize(out, cf);
    }
    template <typename CF>
    ZS_HD void finalize(uint8_t out[32], CF &cf) const {
        uint32_t m[16], r[8];
        words(block, m);
        for (int i = 0; i < 8; i++) r[i] = cv[i];
        uint32_t flags = start_flag() | B3_CHUNK_END;
        if (stack_len == 0) {
            cf(r, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), block_len, flags | B3_ROOT);
        } else {
            cf(r, m, (uint32_t)chunk_counter, (uint32_t)(chunk_counter >> 32), block_len, flags);
            for (int i = (int)stack_len - 1; i >= 0; i--) {
                uint32_t t[8];
                b3_parent_with(stack[i], r, i == 0, t, cf);
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
template <typename CF>
ZS_HD void b3_copy_live(B3Hasher &dst, const B3Hasher &src, CF &cf) {
    const uint32_t l = cf.lane(), n = cf.lanes();
    for (uint32_t i = l; i < 8; i += n) dst.cv[i] = src.cv[i];
    const uint32_t *sb = reinterpret_cast<const uint32_t *>(src.block);
    uint32_t *db = reinterpret_cast<uint32_t *>(dst.block);
    for (uint32_t i = l; i < 16; i += n) db[i] = sb[i];
    const uint32_t *ss = &src.stack[0][0];
    uint32_t *ds = &dst.stack[0][0];
    for (uint32_t i = l; i < 8 * src.stack_len; i += n) ds[i] = ss[i];
    dst.chunk_counter = src.chunk_counter;
    dst.block_len = src.block_len;
    dst.blocks_compressed = src.blocks_compressed;
    dst.stack_len = src.stack_len;
    cf.sync();
}

// One sampled block (channel.zig:49-69): the transcript snapshot extended by
// "zig-stark:sample" || u64le(counter) is hashed and the whole 32-byte block is
// appended to the transcript.  `work` (14 words) receives the 24 suffix bytes
// and then the block, at word 6; `snap` is working storage for the snapshot.
template <typename CF>
ZS_HD void channel_sample_block(B3Hasher &h, B3Hasher &snap, uint64_t counter, uint32_t *work, CF &cf) {
    b3_copy_live(snap, h, cf);
    work[0] = 0x2d67697au;  // "zig-"
    work[1] = 0x72617473u;  // "star"
    work[2] = 0x61733a6bu;  // "k:sa"
    work[3] = 0x656c706du;  // "mple"
    work[4] = (uint32_t)counter;
    work[5] = (uint32_t)(counter >> 32);
    const uint8_t *wb = reinterpret_cast<const uint8_t *>(work);
    snap.update(wb, 24, cf);
    snap.finalize(reinterpret_cast<uint8_t *>(work + 6), cf);
    h.update(wb + 24, 32, cf);
}

template <typename CF>
ZS_HD void channel_sample_bytes(B3Hasher &h, B3Hasher &snap, uint8_t *out, size_t len, uint32_t *work, CF &cf) {
    const uint8_t *blk = reinterpret_cast<const uint8_t *>(work + 6);
    for (uint64_t counter = 0; len > 0; counter++) {
        channel_sample_block(h, snap, counter, work, cf);
        size_t k = len < 32 ? len : 32;
        for (size_t i = 0; i < k; i++) out[i] = blk[i];
        out += k;
        len -= k;
    }
}
ZS_HD void channel_sample_bytes(B3Hasher &h, B3Hasher &snap, uint8_t *out, size_t len) {
    uint32_t work[14];
    B3Single cf;
    channel_sample_bytes(h, snap, out, len, work, cf);
}

// sample(QM31): 16 sampled bytes, each u32 reduced mod p (qm31.zig:95-101)
template <typename CF>
ZS_HD void channel_sample_qm31(B3Hasher &h, B3Hasher &snap, uint32_t out[4], uint32_t *work, CF &cf) {
    channel_sample_block(h, snap, 0, work, cf);
    for (int i = 0; i < 4; i++) out[i] = m31_from_u32(work[6 + i]);
}
ZS_HD void channel_sample_qm31(B3Hasher &h, B3Hasher &snap, uint32_t out[4]) {
    uint32_t work[14];
    B3Single cf;
    channel_sample_qm31(h, snap, out, work, cf);
}

// absorb a 32-byte digest held as 8 little-endian words (through `work`, 8 words)
template <typename CF>
ZS_HD void channel_absorb_digest(B3Hasher &h, const uint32_t d[8], uint32_t *work, CF &cf) {
    for (int i = 0; i < 8; i++) work[i] = d[i];
    h.update(reinterpret_cast<const uint8_t *>(work), 32, cf);
}
ZS_HD void channel_absorb_digest(B3Hasher &h, const uint32_t d[8]) {
    uint32_t work[8];
    B3Single cf;
    channel_absorb_digest(h, d, work, cf);
}

}  // namespace zs
