// verify.cpp -- the verifiers, on the host, over proofs in the reference's wire
// format (src/core/serialization.zig, docs/wire.md) as the device provers write
// them: Fri.verify (src/m31/fri.zig:277-336) and GenericStark(FibSeededAir)
// .verify (src/m31/stark.zig:660-780).  A verifier is cheap (num_queries x
// layers hash paths and folds), so it stays on the CPU; it shares the field,
// BLAKE3 and channel code with the device path.
#include <cstring>
#include <memory>
#include <vector>

#include "air.h"
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

// Fri.Proof fields in order (fri.zig:48-64)
bool read_fri_proof(WireReader &r, FriProof &pf) {
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
    return r.ok;
}

bool parse(const uint8_t *bytes, size_t n, FriProof &pf) {
    WireReader r{bytes, n};
    return read_fri_proof(r, pf) && r.left == 0;  // deserialize rejects trailing bytes (serialization.zig:44-49)
}

// the FriParams a verification runs under (the reference passes them beside the proof).
// Beyond the reference's verify, which only needs numFoldRounds not to underflow,
// the parameters must be ones its prover accepts (fri.zig:221-227 asserts
// remainder_log >= log_blowup) and fold at least once: with remainder_log == dlog
// or a remainder as long as the last layer, every query would pass without any
// degree bound.
struct FriParams {
    uint32_t log_size, log_blowup, remainder_log;
    uint64_t num_queries;
    bool ok() const {
        return log_size + log_blowup <= 32 && remainder_log >= log_blowup && remainder_log < log_size + log_blowup;
    }
    // coefficients the prover writes: the interpolant of 2^(remainder_log - log_blowup) points
    uint64_t remainder_len() const { return uint64_t(1) << (remainder_log - log_blowup); }
};

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

bool fri_verify(const FriParams &fp, const FriProof &pf, zs_channel *chan) {
    const uint32_t dlog = fp.log_size + fp.log_blowup;
    const uint32_t L = dlog - fp.remainder_log;  // numFoldRounds
    const uint64_t n0 = uint64_t(1) << dlog;
    if (!fp.ok() || pf.roots.size() != L) return false;
    // a longer remainder could interpolate any last layer (no degree bound); a shorter one
    // is not what the prover writes
    if (pf.remainder.size() != fp.remainder_len()) return false;
    if (pf.queries.size() != fp.num_queries) return false;
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

// ---- GenericStark.Proof (stark.zig:107-130) ----
struct Reveal {
    uint64_t index;
    std::vector<qm31> values;
    std::vector<std::vector<Digest>> paths;
};
struct StarkProof {
    FriParams params;  // StarkParams: trace_log, log_blowup, num_queries, remainder_log
    std::vector<Digest> trace_roots, pre_roots, acc_roots;
    Digest quotient_root;
    std::vector<qm31> deep_evals;
    FriProof fri;
    std::vector<Reveal> queries;
};

bool read_optional_digests(WireReader &r, std::vector<Digest> &out) {  // ?[]Digest: a presence byte
    const uint64_t present = r.uint(1);
    if (present > 1) return r.ok = false;
    if (present) read_digests(r, out);
    return r.ok;
}

bool parse_stark(const uint8_t *bytes, size_t n, StarkProof &sp) {
    WireReader r{bytes, n};
    sp.params.log_size = (uint32_t)r.uint(1);
    sp.params.log_blowup = (uint32_t)r.uint(1);
    sp.params.num_queries = r.uint(8);
    sp.params.remainder_log = (uint32_t)r.uint(1);
    read_digests(r, sp.trace_roots);
    if (!read_optional_digests(r, sp.pre_roots) || !read_optional_digests(r, sp.acc_roots)) return false;
    const uint8_t *q;
    if (r.take(32, &q)) std::memcpy(sp.quotient_root.b, q, 32);
    sp.deep_evals.resize(r.len(16));
    for (auto &v : sp.deep_evals) v = r.qm();
    if (!r.ok || !read_fri_proof(r, sp.fri)) return false;
    sp.queries.resize(r.len(24));  // index, values length, paths length
    for (auto &qv : sp.queries) {
        qv.index = r.uint(8);
        qv.values.resize(r.len(16));
        for (auto &v : qv.values) v = r.qm();
        qv.paths.resize(r.len(8));
        for (auto &pth : qv.paths) read_digests(r, pth);
        if (!r.ok) return false;
    }
    return r.ok && r.left == 0;
}

// GenericStark(Air).verify (stark.zig:660-780) for an AIR given as data
bool stark_verify_air(const StarkProof &sp, const AirDesc &d, zs_channel *chan) {
    const FriParams &pp = sp.params;
    const uint32_t tl = pp.log_size, dlog = pp.log_size + pp.log_blowup;
    uint32_t fls, flb;
    if (tl < 1 || dlog > 24 || !air_fri_params(d, tl, pp.log_blowup, &fls, &flb)) return false;
    const FriParams fp{fls, flb, pp.remainder_log, pp.num_queries};  // friParams(params)
    if (!fp.ok()) return false;
    const uint64_t n = uint64_t(1) << tl, N = uint64_t(1) << dlog, shift = uint64_t(1) << pp.log_blowup;
    const uint32_t mt = d.m_total;
    if (sp.trace_roots.size() != d.m || sp.deep_evals.size() != d.n_deep) return false;
    if (sp.pre_roots.size() != d.n_pre || sp.acc_roots.size() != d.n_rel) return false;
    const qm31 w = qm31_root_of_unity(tl), w_ev = qm31_root_of_unity(dlog), one = {{1, 0}, {0, 0}};
    auto sample = [&]() {
        uint32_t a[4];
        zs_channel_sample_qm31(chan, a);
        return qm_from_words(a);
    };
    // the transcript in the order the prover wrote it (stark.zig:676-705)
    AirWeights wt{};
    for (const Digest &dg : sp.pre_roots) zs_channel_absorb(chan, dg.b, 32);
    for (const Digest &dg : sp.trace_roots) zs_channel_absorb(chan, dg.b, 32);
    for (uint32_t j = 0; j < d.n_lc; j++) wt.alpha_cols[j] = sample();
    for (uint32_t r = 0; r < d.n_rel; r++) wt.alpha_rels[r] = sample();
    for (const Digest &dg : sp.acc_roots) zs_channel_absorb(chan, dg.b, 32);
    for (uint32_t k = 0; k < d.n_trans_total; k++) wt.alpha[k] = sample();
    for (uint32_t k = 0; k < d.total_bound; k++) wt.pk[k] = qm_pow(w, d.bstep[k]);
    for (uint32_t k = 0; k < d.total_bound; k++) wt.beta[k] = sample();
    wt.last = qm_pow(w, n - 1);
    zs_channel_absorb(chan, sp.quotient_root.b, 32);
    const qm31 z = sample();
    for (const qm31 &v : sp.deep_evals) {
        uint32_t wds[4];
        qm_to_words(v, wds);
        zs_channel_absorb(chan, reinterpret_cast<const uint8_t *>(wds), 16);
    }
    std::vector<qm31> gam(d.n_deep);
    for (auto &g : gam) g = sample();
    // FRI under friParams; the nested proof's own copy of them must be the same
    const FriProof &fr = sp.fri;
    if (fr.log_size != fls || fr.log_blowup != flb || fr.num_queries != pp.num_queries ||
        fr.remainder_log != pp.remainder_log)
        return false;
    if (!fri_verify(fp, fr, chan)) return false;
    if (sp.queries.size() != fr.queries.size()) return false;
    const uint32_t R = d.reveal_len;
    for (size_t qi = 0; qi < sp.queries.size(); qi++) {
        const Reveal &qv = sp.queries[qi];
        if (qv.index != fr.queries[qi].index) return false;
        if (qv.values.size() != R || qv.paths.size() != R) return false;
        const uint64_t p0 = qv.index, pn = (p0 + shift) % N;
        const qm31 *v = qv.values.data();
        // Merkle checks (stark.zig:716-731)
        for (uint32_t k = 0; k < R; k++) {
            const Digest *root;
            uint64_t at = p0;
            if (k < 2 * mt) {
                const uint32_t j = k % mt;
                root = j < d.m ? &sp.trace_roots[j] : &sp.acc_roots[j - d.m];
                if (k >= mt) at = pn;
            } else if (k == 2 * mt) {
                root = &sp.quotient_root;
            } else {
                root = &sp.pre_roots[k - 2 * mt - 1];
            }
            uint32_t leaf[8];
            hash_qm31(v[k], leaf);
            if (!merkle_verify(*root, at, leaf, qv.paths[k])) return false;
        }
        const qm31 x0 = qm_mul(fri_offset_pow(0), qm_pow(w_ev, p0));
        // DEEP identity: g(x0) equals the FRI layer-0 value at p0 (stark.zig:733-751)
        const qm31 inv_dz = qm_inv(qm_sub(x0, z));
        qm31 g = {{0, 0}, {0, 0}};
        for (uint32_t j = 0; j < mt; j++) {
            g = qm_add(g, qm_mul(gam[j], qm_mul(qm_sub(v[j], sp.deep_evals[j]), inv_dz)));
            g = qm_add(g, qm_mul(gam[mt + j], qm_mul(qm_sub(v[mt + j], sp.deep_evals[mt + j]), inv_dz)));
        }
        g = qm_add(g, qm_mul(gam[2 * mt], qm_mul(qm_sub(v[2 * mt], sp.deep_evals[2 * mt]), inv_dz)));
        if (fr.queries[qi].pairs.empty()) return false;
        const auto &pair0 = fr.queries[qi].pairs[0];
        if (!qm_eq(g, p0 == p0 % (N / 2) ? pair0.v0 : pair0.v1)) return false;
        // constraints: Hc(x0) == Z_H(x0) Q(x0) (stark.zig:753-779)
        const qm31 zh = qm_sub(qm_pow(x0, n), one);
        const qm31 hv = air_composition_host(d, wt, x0, v, v + mt, v + 2 * mt + 1, zh);
        if (!qm_eq(hv, qm_mul(zh, v[2 * mt]))) return false;
    }
    return true;
}

}  // namespace

extern "C" int32_t zs_stark_verify(const uint8_t *proof, size_t proof_len, const zs_air *air, zs_channel *chan,
                                   uint8_t *out_ok) {
    if (!chan || !out_ok || !air || (!proof && proof_len)) return ZS_ERR_INPUT;
    auto d = std::make_unique<AirDesc>();
    ZS_TRY(air_from_abi(air, d.get()));
    StarkProof sp;
    if (!parse_stark(proof, proof_len, sp)) return ZS_ERR_INPUT;
    *out_ok = stark_verify_air(sp, *d, chan) ? 1 : 0;
    return ZS_OK;
}

extern "C" int32_t zs_fri_verify(const uint8_t *proof, size_t proof_len, zs_channel *chan, uint8_t *out_ok) {
    if (!chan || !out_ok || (!proof && proof_len)) return ZS_ERR_INPUT;
    FriProof pf;
    if (!parse(proof, proof_len, pf)) return ZS_ERR_INPUT;
    const FriParams fp{pf.log_size, pf.log_blowup, pf.remainder_log, pf.num_queries};
    // numFoldRounds would underflow, or the domain exceed QM31's 2-adic subgroup (2^32)
    *out_ok = fp.ok() && fri_verify(fp, pf, chan) ? 1 : 0;
    return ZS_OK;
}

extern "C" int32_t zs_fib_stark_verify(const uint8_t *proof, size_t proof_len, const uint32_t a0[4],
                                       const uint32_t a1[4], const uint32_t claimed[4], zs_channel *chan,
                                       uint8_t *out_ok) {
    if (!chan || !out_ok || !a0 || !a1 || !claimed || (!proof && proof_len)) return ZS_ERR_INPUT;
    StarkProof sp;
    if (!parse_stark(proof, proof_len, sp)) return ZS_ERR_INPUT;
    // FibAir (stark.zig:877-904) as data with the seeded boundary; the trace length
    // comes from the proof's parameters
    const uint32_t tl = sp.params.log_size;
    if (!sp.params.ok() || tl < 1 || tl > 24) {
        *out_ok = 0;
        return ZS_OK;
    }
    zs_air_term terms[5] = {};
    const uint32_t con[5] = {0, 0, 1, 1, 1}, kind[5] = {ZS_AIR_NEXT, ZS_AIR_CUR, ZS_AIR_NEXT, ZS_AIR_CUR, ZS_AIR_CUR};
    const uint32_t col[5] = {0, 1, 1, 0, 1}, coef[5] = {1, P - 1, 1, P - 1, P - 1};
    for (int t = 0; t < 5; t++) {
        terms[t].constraint = con[t];
        terms[t].coeff[0] = coef[t];
        terms[t].n_factors = 1;
        terms[t].factor[0] = {kind[t], col[t]};
    }
    zs_boundary bnd[3] = {};
    const uint32_t *vals[3] = {a0, a1, claimed};
    const uint32_t bcol[3] = {0, 1, 0};
    const uint64_t bstep[3] = {0, 0, (uint64_t(1) << tl) - 1};
    for (int k = 0; k < 3; k++) {
        bnd[k].column = bcol[k];
        bnd[k].step = bstep[k];
        std::memcpy(bnd[k].value, vals[k], 16);
    }
    zs_air air{};
    air.num_columns = 2;
    air.num_transition_constraints = 2;
    air.num_terms = 5;
    air.terms = terms;
    air.degree_mul = 1;
    air.num_boundary = 3;
    air.boundary = bnd;
    auto d = std::make_unique<AirDesc>();
    ZS_TRY(air_from_abi(&air, d.get()));
    *out_ok = stark_verify_air(sp, *d, chan) ? 1 : 0;
    return ZS_OK;
}
