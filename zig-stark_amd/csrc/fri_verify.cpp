// fri_verify.cpp -- Fri.verify (src/m31/fri.zig:277-336) on the host, over a
// proof in the reference's wire format (src/core/serialization.zig,
// docs/wire.md), as zs_fri_prove_dev writes it.  A verifier is cheap
// (num_queries x layers hash paths and folds), so it stays on the CPU; it
// shares the field, BLAKE3 and channel code with the device path.
#include <cstring>
#include <vector>

#include "internal.h"

namespace {

using namespace zs;

// readValue of serialization.zig:120-175 for the Fri.Proof types: unsigned ints
// little-endian (usize as 8 bytes), QM31 as four u32 reduced mod p
// (QM31.fromBytes, qm31.zig:95-101), slices as a u64 length then the elements.
struct WireReader {
    const uint8_t *p;
    size_t left;
    bool ok = true;
    bool take(size_t n, const uint8_t **out) {
        if (!ok || n > left) return ok = false;
        *out = p;
        p += n;
        left -= n;
        return true;
    }
    uint64_t uint(unsigned n) {
        const uint8_t *b;
        if (!take(n, &b)) return 0;
        uint64_t v = 0;
        for (unsigned k = 0; k < n; k++) v |= uint64_t(b[k]) << (8 * k);
        return v;
    }
    qm31 qm() {
        uint32_t w[4];
        for (int k = 0; k < 4; k++) w[k] = m31_from_u32((uint32_t)uint(4));
        return qm_from_words(w);
    }
    // a slice length, checked against the bytes left for elements of at least min_size bytes
    uint64_t len(size_t min_size) {
        uint64_t n = uint(8);
        if (ok && min_size && n > left / min_size) ok = false;
        return ok ? n : 0;
    }
};

struct Digest {
    uint8_t b[32];
};
struct Pair {
    qm31 v0, v1;
    std::vector<Digest> path0, path1;
};
struct Query {
    uint64_t index;
    std::vector<Pair> pairs;
};
struct FriProof {
    uint32_t log_size, log_blowup, remainder_log;
    uint64_t num_queries;
    std::vector<Digest> roots;
    std::vector<qm31> remainder;
    std::vector<Query> queries;
};

void read_digests(WireReader &r, std::vector<Digest> &out) {
    out.resize(r.len(32));
    for (auto &d : out) {
        const uint8_t *b;
        if (r.take(32, &b)) std::memcpy(d.b, b, 32);
    }
}

bool parse(const uint8_t *bytes, size_t n, FriProof &pf) {
    WireReader r{bytes, n};
    pf.log_size = (uint32_t)r.uint(1);  // FriParams (fri.zig:38-46)
    pf.log_blowup = (uint32_t)r.uint(1);
    pf.num_queries = r.uint(8);
    pf.remainder_log = (uint32_t)r.uint(1);
    read_digests(r, pf.roots);
    pf.remainder.resize(r.len(16));
    for (auto &c : pf.remainder) c = r.qm();
    // a query is at least its index and its pair count; a pair at least two values and two lengths
    pf.queries.resize(r.len(16));
    for (auto &q : pf.queries) {
        q.index = r.uint(8);
        q.pairs.resize(r.len(48));
        for (auto &pr : q.pairs) {
            pr.v0 = r.qm();
            pr.v1 = r.qm();
            read_digests(r, pr.path0);
            read_digests(r, pr.path1);
        }
        if (!r.ok) return false;
    }
    return r.ok && r.left == 0;  // deserialize rejects trailing bytes (serialization.zig:44-49)
}

// MerkleTree.verify (src/core/merkle/merkle.zig:72-89)
bool merkle_verify(const Digest &root, uint64_t index, const uint32_t leaf[8], const std::vector<Digest> &path) {
    uint32_t d[8];
    std::memcpy(d, leaf, 32);
    for (const Digest &s : path) {
        uint32_t sib[8], t[8];
        std::memcpy(sib, s.b, 32);
        if ((index & 1) == 0)
            b3_hash2(d, sib, t);
        else
            b3_hash2(sib, d, t);
        std::memcpy(d, t, 32);
        index >>= 1;
    }
    return std::memcmp(d, root.b, 32) == 0;
}

void hash_qm31(qm31 v, uint32_t out[8]) {  // hashQM31 (src/m31/hash.zig:33-37)
    uint32_t w[4];
    qm_to_words(v, w);
    b3_hash_qm31(w, out);
}

// domainPoint (fri.zig:96-101): FRI_OFFSET^(2^layer) * w^j, w of order 2^(dlog - layer)
qm31 domain_point(uint32_t dlog, uint32_t layer, uint64_t j) {
    return qm_mul(fri_offset_pow(layer), qm_pow(qm31_root_of_unity(dlog - layer), j));
}

// foldValue (fri.zig:125-130)
qm31 fold_value(qm31 g, qm31 y0, qm31 y1, qm31 alpha) {
    const qm31 half_sum = qm_half(qm_add(y0, y1));
    const qm31 half_diff = qm_half(qm_mul(qm_sub(y0, y1), qm_inv(g)));
    return qm_add(half_sum, qm_mul(alpha, half_diff));
}

// UnivariateQM31.eval (src/m31/poly/univariate.zig:11-19), Horner
qm31 poly_eval(const std::vector<qm31> &poly, qm31 x) {
    qm31 r = {{0, 0}, {0, 0}};
    for (size_t i = poly.size(); i-- > 0;) r = qm_add(qm_mul(r, x), poly[i]);
    return r;
}

bool verify(const FriProof &pf, zs_channel *chan) {
    const uint32_t dlog = pf.log_size + pf.log_blowup;
    const uint32_t L = dlog - pf.remainder_log;  // numFoldRounds
    const uint64_t n0 = uint64_t(1) << dlog;
    if (pf.roots.size() != L) return false;
    if (pf.queries.size() != pf.num_queries) return false;
    // replay the transcript: absorb roots, resample alphas, absorb the remainder
    std::vector<qm31> alphas(L);
    for (uint32_t i = 0; i < L; i++) {
        zs_channel_absorb(chan, pf.roots[i].b, 32);
        uint32_t a[4];
        zs_channel_sample_qm31(chan, a);
        alphas[i] = qm_from_words(a);
    }
    for (const qm31 &c : pf.remainder) {  // absorbMany: each coefficient's toBytes
        uint32_t w[4];
        qm_to_words(c, w);
        zs_channel_absorb(chan, reinterpret_cast<const uint8_t *>(w), 16);
    }
    for (const Query &q : pf.queries) {
        if (zs_channel_sample_index(chan, n0) != q.index) return false;
        if (q.pairs.size() < L) return false;  // the reference reads pairs[0..L)
        uint64_t p = q.index;
        qm31 prev{};
        for (uint32_t i = 0; i < L; i++) {
            const uint64_t nn = (n0 >> i) / 2, pn = p % nn;
            const Pair &pr = q.pairs[i];
            uint32_t leaf[8];
            hash_qm31(pr.v0, leaf);
            if (!merkle_verify(pf.roots[i], pn, leaf, pr.path0)) return false;
            hash_qm31(pr.v1, leaf);
            if (!merkle_verify(pf.roots[i], pn + nn, leaf, pr.path1)) return false;
            // the previous fold produced the value at position p of this layer
            if (i > 0 && !qm_eq(prev, p == pn ? pr.v0 : pr.v1)) return false;
            const qm31 fv = fold_value(domain_point(dlog, i, pn), pr.v0, pr.v1, alphas[i]);
            if (i == L - 1 && !qm_eq(poly_eval(pf.remainder, domain_point(dlog, i + 1, pn)), fv)) return false;
            prev = fv;
            p = pn;
        }
    }
    return true;
}

}  // namespace

extern "C" int32_t zs_fri_verify(const uint8_t *proof, size_t proof_len, zs_channel *chan, uint8_t *out_ok) {
    if (!chan || !out_ok || (!proof && proof_len)) return ZS_ERR_INPUT;
    FriProof pf;
    if (!parse(proof, proof_len, pf)) return ZS_ERR_INPUT;
    // numFoldRounds would underflow, or the domain exceed QM31's 2-adic subgroup (2^32)
    const uint32_t dlog = pf.log_size + pf.log_blowup;
    if (dlog > 32 || pf.remainder_log > dlog) {
        *out_ok = 0;
        return ZS_OK;
    }
    *out_ok = verify(pf, chan) ? 1 : 0;
    return ZS_OK;
}
