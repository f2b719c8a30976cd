"""The library's host verifiers (zs_fri_verify: Fri.verify of src/m31/fri.zig:277-336;
zs_fib_stark_verify: GenericStark(FibSeededAir).verify of src/m31/stark.zig:660-780)
on proofs the oracle writes in the wire format: accepted as the oracle's verify
accepts them, rejected on any tampered value, path, root, remainder or index,
ZS_ERR_INPUT for bytes that do not deserialize.  CPU only: no device is used."""
import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z
from test_oracle_commit import low_degree_codeword

P = O.P


def proof_bytes(seed, log_size, beta, rem_log, nq, label):
    rng = np.random.default_rng(seed)
    cw = low_degree_codeword(rng, log_size, beta)
    proof = O.fri_prove(cw, beta, rem_log, nq, O.Channel(label))
    return proof, O.fri_proof_bytes(proof, beta, rem_log)


@pytest.mark.parametrize("log_size,beta,rem_log,nq", [(4, 3, 3, 8), (6, 2, 3, 5), (7, 1, 2, 4)])
def test_host_verify_accepts_oracle_proofs(log_size, beta, rem_log, nq):
    label = b"zig-stark:fri:host"
    proof, b = proof_bytes(log_size * 10 + beta, log_size, beta, rem_log, nq, label)
    assert O.fri_verify(proof, beta, rem_log, O.Channel(label))
    ch, och = Z.Channel(label), O.Channel(label)
    assert Z.fri_verify(b, ch)
    O.fri_verify(proof, beta, rem_log, och)
    assert ch.sample_bytes(32) == och.sample_bytes(32)  # the transcript advanced the same way
    # another transcript draws other challenges and indices
    assert not Z.fri_verify(b, Z.Channel(b"zig-stark:fri:other"))


def test_host_verify_rejects_tampering():
    label = b"zig-stark:fri:tamper"
    log_size, beta, rem_log, nq = 5, 2, 3, 6
    _, b = proof_bytes(11, log_size, beta, rem_log, nq, label)
    dlog, L = log_size + beta, log_size + beta - rem_log
    head = 1 + 1 + 8 + 1
    roots_at = head + 8
    rem_at = roots_at + 32 * L + 8
    rc = 1 << (rem_log - beta)
    q0 = rem_at + 16 * rc + 8  # first query: index, pair count, then pair 0
    flips = {
        "root": roots_at + 5,
        "remainder": rem_at + 3,
        "index": q0,
        "value0": q0 + 16 + 2,
        "value1": q0 + 16 + 16 + 1,
        "path0": q0 + 16 + 32 + 8 + 7,
    }
    for what, at in flips.items():
        t = bytearray(b)
        t[at] ^= 0x01
        assert not Z.fri_verify(bytes(t), Z.Channel(label)), what
    assert Z.fri_verify(b, Z.Channel(label))


def test_host_verify_rejects_malformed_bytes():
    label = b"zig-stark:fri:wire"
    _, b = proof_bytes(5, 4, 2, 3, 3, label)
    for bad in (b[:-1], b + b"\0", b[:10], b""):
        with pytest.raises(Z.ZsError) as e:
            Z.fri_verify(bad, Z.Channel(label))
        assert e.value.code == Z.ZS_ERR_INPUT
    # a slice length larger than the bytes that follow
    t = bytearray(b)
    t[11:19] = (1 << 40).to_bytes(8, "little")
    with pytest.raises(Z.ZsError):
        Z.fri_verify(bytes(t), Z.Channel(label))


# ---- GenericStark(FibSeededAir).verify (stark.zig:660-780) -------------------------------
def fib_case(tl, a1=5):
    t = O.fib_trace(1, a1, 1 << tl)
    return t, np.array([1, 0, 0, 0], np.uint32), np.array([a1, 0, 0, 0], np.uint32), t[0, -1]


@pytest.mark.parametrize("tl,beta,nq,rl", [(4, 3, 12, 3), (5, 2, 8, 3), (6, 1, 5, 2)])
def test_host_stark_verify_matches_oracle(tl, beta, nq, rl):
    t, a0, a1, claimed = fib_case(tl)
    proof = O.fib_stark_prove(t, beta, nq, rl, a0, a1, claimed, O.Channel(b"fib"))
    ch, och = Z.Channel(b"fib"), O.Channel(b"fib")
    assert Z.fib_stark_verify(proof, a0, a1, claimed, ch)
    assert O.fib_stark_verify(proof, a0, a1, claimed, och)
    assert ch.sample_bytes(32) == och.sample_bytes(32)
    assert not Z.fib_stark_verify(proof, a0, a1, claimed, Z.Channel(b"other"))
    bad = claimed.copy()
    bad[0] = (int(bad[0]) + 1) % P
    assert not Z.fib_stark_verify(proof, a0, a1, bad, Z.Channel(b"fib"))
    for cut in (proof[:-1], proof + b"\0"):
        with pytest.raises(Z.ZsError) as e:
            Z.fib_stark_verify(cut, a0, a1, claimed, Z.Channel(b"fib"))
        assert e.value.code == Z.ZS_ERR_INPUT


def test_host_stark_verify_rejects_tampering_and_false_claims():
    t, a0, a1, claimed = fib_case(4)
    # stark.zig:1282-1298: a proof for a wrong claimed value does not verify
    bad = claimed.copy()
    bad[0] = (int(bad[0]) + 1) % P
    wrong = O.fib_stark_prove(t, 3, 12, 3, a0, a1, bad, O.Channel(b"fib"))
    assert not Z.fib_stark_verify(wrong, a0, a1, bad, Z.Channel(b"fib"))
    proof = O.fib_stark_prove(t, 3, 12, 3, a0, a1, claimed, O.Channel(b"fib"))
    # the quotient root, a DEEP evaluation, a revealed value, a path digest near the end
    for off in (11 + 8 + 64 + 2 + 5, 11 + 8 + 64 + 2 + 32 + 8 + 3, len(proof) - 200, len(proof) - 7):
        t2 = bytearray(proof)
        t2[off] ^= 1
        assert not Z.fib_stark_verify(bytes(t2), a0, a1, claimed, Z.Channel(b"fib")), off
        assert not O.fib_stark_verify(bytes(t2), a0, a1, claimed, O.Channel(b"fib")), off


def test_host_verifiers_survive_random_corruption():
    """Random byte flips, truncations and extensions of valid proofs: every call
    returns (accept, reject or ZS_ERR_INPUT) and the corrupted proofs never verify."""
    rng = np.random.default_rng(99)
    _, fri_b = proof_bytes(3, 5, 2, 3, 4, b"fuzz")
    t, a0, a1, claimed = fib_case(4)
    stark_b = O.fib_stark_prove(t, 2, 6, 3, a0, a1, claimed, O.Channel(b"fuzz"))
    cases = [(fri_b, lambda b: Z.fri_verify(b, Z.Channel(b"fuzz"))),
             (stark_b, lambda b: Z.fib_stark_verify(b, a0, a1, claimed, Z.Channel(b"fuzz")))]
    for good, verify in cases:
        assert verify(good)
        for _ in range(150):
            b = bytearray(good)
            kind = rng.integers(0, 3)
            if kind == 0:
                for _ in range(int(rng.integers(1, 4))):
                    b[int(rng.integers(0, len(b)))] ^= 1 << int(rng.integers(0, 8))
            elif kind == 1:
                b = b[: int(rng.integers(0, len(b)))]
            else:
                b += bytes(rng.integers(0, 256, int(rng.integers(1, 40)), dtype=np.uint8))
            if bytes(b) == good:
                continue
            try:
                assert not verify(bytes(b))
            except Z.ZsError as e:
                assert e.code == Z.ZS_ERR_INPUT


def test_host_stark_verify_rejects_altered_nested_fri_parameters():
    """The FRI proof inside a STARK proof repeats the STARK's parameters; a copy
    that differs is rejected (the honest prover writes the same 11 bytes)."""
    t, a0, a1, claimed = fib_case(4)
    proof = O.fib_stark_prove(t, 2, 6, 3, a0, a1, claimed, O.Channel(b"nested"))
    # params 11 B, trace roots 8 + 2 * 32, two absent optional root lists, the
    # quotient root 32, DEEP evaluations 8 + 5 * 16: the FRI proof starts here
    at = 11 + 8 + 64 + 2 + 32 + 8 + 5 * 16
    assert proof[at:at + 11] == proof[:11]
    assert Z.fib_stark_verify(proof, a0, a1, claimed, Z.Channel(b"nested"))
    for off in (0, 1, 2, 10):  # log_size, log_blowup, num_queries, remainder_log
        t2 = bytearray(proof)
        t2[at + off] = (t2[at + off] + 1) & 0xFF
        try:
            assert not Z.fib_stark_verify(bytes(t2), a0, a1, claimed, Z.Channel(b"nested")), off
        except Z.ZsError as e:
            assert e.code == Z.ZS_ERR_INPUT


def forge_fri_proof(cw, beta, rem_log, nq, label, rem_points):
    """A proveCodeword (fri.zig:157-273) written out in Python over oracle parts,
    with one freedom a cheating prover has: the remainder interpolates the first
    `rem_points` values of the last layer.  rem_points = 2^(rem_log - beta) is
    the honest prover; rem_points = 2^rem_log interpolates the whole last layer,
    so the final fold check holds for any codeword.  Returns the wire bytes."""
    cw = np.asarray(cw, np.uint32).reshape(-1, 4)
    n0 = cw.shape[0]
    dlog = n0.bit_length() - 1
    L = dlog - rem_log
    ch = O.Channel(label)
    layers, trees, roots = [cw], [], []
    for i in range(L):
        nodes = O.merkle_build(b"".join(O.hash_qm31(v) for v in layers[i]))
        trees.append(nodes)
        roots.append(O.tree_root(nodes))
        ch.absorb(roots[-1])
        layers.append(O.fri_fold(layers[i], dlog, i, ch.sample_qm31()))
    xs = np.array([O.fri_domain_point(dlog, L, j) for j in range(rem_points)])
    rem = O.poly_interpolate(xs, layers[L][:rem_points])
    for c in rem:
        ch.absorb(np.asarray(c, np.uint32).tobytes())
    out = bytearray()
    out += bytes([dlog - beta, beta]) + nq.to_bytes(8, "little") + bytes([rem_log])
    out += L.to_bytes(8, "little") + b"".join(roots)
    out += len(rem).to_bytes(8, "little") + np.asarray(rem, np.uint32).tobytes()
    out += nq.to_bytes(8, "little")
    for _ in range(nq):
        p = ch.sample_index(n0)
        out += p.to_bytes(8, "little") + L.to_bytes(8, "little")
        for i in range(L):
            ni = n0 >> i
            pn = p % (ni // 2)
            out += layers[i][pn].tobytes() + layers[i][pn + ni // 2].tobytes()
            for idx in (pn, pn + ni // 2):
                out += (dlog - i).to_bytes(8, "little") + O.merkle_open(trees[i], ni, idx)
            p = pn
    return bytes(out)


def test_host_verify_rejects_full_length_remainder():
    """A remainder longer than 2^(remainder_log - log_blowup) coefficients can
    interpolate any last layer, so a codeword of full degree would pass every
    fold and remainder check.  The forger is first pinned to the oracle's honest
    prover (same bytes), then given a random codeword and a remainder through
    all 2^remainder_log points of the last layer: rejected."""
    label = b"zig-stark:fri:cheat"
    log_size, beta, rem_log, nq = 4, 2, 3, 5
    rng = np.random.default_rng(7)
    honest_cw = low_degree_codeword(rng, log_size, beta)
    _, want = proof_bytes(7, log_size, beta, rem_log, nq, label)
    honest = forge_fri_proof(honest_cw, beta, rem_log, nq, label, 1 << (rem_log - beta))
    assert honest == want
    assert Z.fri_verify(honest, Z.Channel(label))
    junk = rng.integers(0, P, ((1 << (log_size + beta)), 4), dtype=np.uint32)
    cheat = forge_fri_proof(junk, beta, rem_log, nq, label, 1 << rem_log)
    assert not Z.fri_verify(cheat, Z.Channel(label))
    # one coefficient more than the honest length is rejected as well
    longer = forge_fri_proof(honest_cw, beta, rem_log, nq, label, (1 << (rem_log - beta)) + 1)
    assert not Z.fri_verify(longer, Z.Channel(label))


def test_host_verify_rejects_parameters_the_prover_refuses():
    """remainder_log = log_size + log_blowup (no fold round: the reference's
    verify checks nothing per query) and remainder_log < log_blowup (the
    reference prover asserts against it, fri.zig:221-227) never verify."""
    label = b"zig-stark:fri:params"
    rng = np.random.default_rng(8)
    log_size, beta, nq = 3, 2, 4
    dlog = log_size + beta
    for rem_log in (dlog, beta - 1):
        ch = O.Channel(label)
        k = 1 << max(rem_log - beta, 0)
        rem = rng.integers(0, P, (k, 4), dtype=np.uint32)
        L = max(dlog - rem_log, 0)
        forged = bytearray(bytes([log_size, beta]) + nq.to_bytes(8, "little") + bytes([rem_log]))
        forged += L.to_bytes(8, "little")
        for i in range(L):  # roots of whatever layers: never looked at before the parameters
            forged += bytes(32)
        forged += k.to_bytes(8, "little") + rem.tobytes()
        forged += nq.to_bytes(8, "little")
        for c in rem:
            ch.absorb(c.tobytes())
        for _ in range(nq):
            forged += ch.sample_index(1 << dlog).to_bytes(8, "little") + (0).to_bytes(8, "little")
        assert not Z.fri_verify(bytes(forged), Z.Channel(label)), rem_log
