// channel.cpp -- host-side Fiat-Shamir channel (src/core/channel/channel.zig).
// The hasher and the channel operations live in b3_hasher.h, shared with the
// device kernels that advance a channel copied to device memory.
#include <cstring>

#include "b3_hasher.h"
#include "internal.h"


extern "C" {

zs_channel *zs_channel_new(const uint8_t *sep, size_t len) {
    auto *c = new zs_channel();
    c->h.init();
    c->h.update(reinterpret_cast<const uint8_t *>("zig-stark:channel"), 17);
    if (len) c->h.update(sep, len);
    return c;
}

zs_channel *zs_channel_clone(const zs_channel *ch) { return new zs_channel(*ch); }

void zs_channel_free(zs_channel *ch) { delete ch; }

void zs_channel_absorb(zs_channel *ch, const uint8_t *data, size_t len) { ch->h.update(data, len); }

void zs_channel_sample_bytes(zs_channel *ch, uint8_t *out, size_t len) {
    zs::B3Hasher snap;
    zs::channel_sample_bytes(ch->h, snap, out, len);
}

void zs_channel_sample_qm31(zs_channel *ch, uint32_t out[4]) {
    zs::B3Hasher snap;
    zs::channel_sample_qm31(ch->h, snap, out);
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
