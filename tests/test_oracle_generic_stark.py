"""The oracle's GenericStark for an AIR given as data (proveWithPreprocessed and
verify, src/m31/stark.zig:311-780), pinned by:
  * the FibSeeded AIR written as data proves to the same bytes as the oracle's
    dedicated FibSeededAir prover (itself pinned by tests/test_oracle_fib_stark.py);
  * the reference's lookup tests (stark.zig:1214-1280): the range-check, AND-table
    and multiplicity AIRs prove and verify (trace_log 4, blowup 8, 16 queries,
    remainder 2^3), and one tampered revealed value is rejected;
  * LogUp's algebra: for an honest trace the running sum telescopes to zero
    (acc[n-1] + contribution(n-1) = 0), for a trace whose lookups are not in the
    table it does not, and that trace's proof is rejected;
  * friParams (stark.zig:197-222): lookup AIRs commit FRI with log_size
    trace_log + 1 (constraint degree 3(n-1)), FibAir with trace_log."""
import numpy as np
import pytest

import airs
import oracle_lib as O
from wire import Reader

P = O.P


@pytest.mark.parametrize("tl,beta,nq,rl", [(4, 3, 12, 3), (5, 2, 8, 3), (6, 1, 5, 2)])
def test_generic_fib_equals_dedicated_fib_prover(tl, beta, nq, rl):
    n = 1 << tl
    t = O.fib_trace(1, 5, n)
    a0, a1, claimed = [1, 0, 0, 0], [5, 0, 0, 0], t[0, -1]
    air = airs.fib_seeded_air(n, a0, a1, claimed)
    want = O.fib_stark_prove(t, beta, nq, rl, a0, a1, claimed, O.Channel(b"fib"))
    ch = O.Channel(b"fib")
    got = O.stark_prove(air, None, t, beta, nq, rl, ch)
    assert got == want
    assert O.stark_verify(air, got, O.Channel(b"fib"))
    assert O.fib_stark_verify(got, a0, a1, claimed, O.Channel(b"fib"))
    # a wrong claim does not verify (stark.zig:1282-1298)
    bad = airs.fib_seeded_air(n, a0, a1, (int(claimed[0]) + 1) % P)
    assert not O.stark_verify(bad, got, O.Channel(b"fib"))


def test_fri_params_of_lookup_and_fib_airs():
    assert O.stark_fri_params(airs.range_check_air(), 4, 3) == (5, 2)  # degree 3(n-1) -> g < 2n
    assert O.stark_fri_params(airs.fib_seeded_air(16, 0, 1, 0), 4, 3) == (4, 3)
    assert O.stark_fri_params(airs.range_check_air(), 4, 0) is None  # log_size > domain


def reveal_offsets(proof):
    """Byte offsets of each query reveal's values in GenericStark.Proof bytes."""
    r = Reader(proof)
    r.uint(1), r.uint(1), r.uint(8), r.uint(1)
    r.digests()
    for _ in range(2):  # ?[]Digest
        if r.uint(1):
            r.digests()
    r.take(32)
    for _ in range(r.uint(8)):
        r.qm31()
    r.uint(1), r.uint(1), r.uint(8), r.uint(1)  # nested Fri.Proof
    r.digests()
    for _ in range(r.uint(8)):
        r.qm31()
    for _ in range(r.uint(8)):
        r.uint(8)
        for _ in range(r.uint(8)):
            r.qm31(), r.qm31()
            r.digests(), r.digests()
    offs = []
    for _ in range(r.uint(8)):
        r.uint(8)
        k = r.uint(8)
        offs.append(r.pos)
        for _ in range(k):
            r.qm31()
        for _ in range(r.uint(8)):
            r.digests()
    assert r.pos == len(proof)
    return offs


@pytest.mark.parametrize("name,tamper", [("range_check", 0), ("and_table", 0), ("multiplicity", 1)])
def test_reference_lookup_round_trip_and_tamper(name, tamper):
    """stark.zig:1214-1280: StarkParams{trace_log 4, log_blowup 3, num_queries 16}
    (remainder_log 3), Channel "zig-stark:stark:lookup"."""
    mk_air, mk_trace = airs.LOOKUP_AIRS[name]
    air = mk_air()
    trace, table = mk_trace(16)
    label = b"zig-stark:stark:lookup"
    proof = O.stark_prove(air, table, trace, 3, 16, 3, O.Channel(label))
    assert proof is not None
    assert O.stark_verify(air, proof, O.Channel(label))
    # queries[0].values[tamper] + 1 (QM31.one added to the first word)
    off = reveal_offsets(proof)[0] + 16 * tamper
    b = bytearray(proof)
    v = (int.from_bytes(b[off:off + 4], "little") + 1) % P
    b[off:off + 4] = v.to_bytes(4, "little")
    assert not O.stark_verify(air, bytes(b), O.Channel(label))
    # another transcript
    assert not O.stark_verify(air, proof, O.Channel(b"other"))


def contribution(air, r, ac, ar, row, row_pre):
    sel, mult, keys, tabs = air.lookups[r]
    key = np.zeros(4, np.uint32)
    for j, c in enumerate(keys):
        key = (key.astype(np.uint64) + O.qm31_mul(ac[j], row[c])) % P
    tkey = np.zeros(4, np.uint32)
    for j, c in enumerate(tabs):
        tkey = (tkey.astype(np.uint64) + O.qm31_mul(ac[j], row_pre[c])) % P
    sub = lambda a, b: ((a.astype(np.int64) - b) % P).astype(np.uint32)
    one = np.array([1, 0, 0, 0], np.uint32)
    lt = O.qm31_mul(row[sel], O.qm31_mul(row[mult], O.qm31_inv(sub(ar[r], key.astype(np.uint32)))))
    tt = O.qm31_mul(sub(one, row[sel]), O.qm31_inv(sub(ar[r], tkey.astype(np.uint32))))
    return sub(lt, tt)


@pytest.mark.parametrize("name", list(airs.LOOKUP_AIRS))
def test_logup_running_sum_telescopes(name):
    mk_air, mk_trace = airs.LOOKUP_AIRS[name]
    air = mk_air()
    n = 16  # the size the reference's traces are written for (the multiplicity table holds 3 x 4 rows)
    trace, table = mk_trace(n)
    rng = np.random.default_rng(len(name))
    ac = rng.integers(0, P, (air.num_lookup_columns, 4), dtype=np.uint32)
    ar = rng.integers(0, P, (1, 4), dtype=np.uint32)
    acc = O.stark_accumulators(air, table, trace, ac, ar)
    assert not acc[0, 0].any()
    last = contribution(air, 0, ac, ar, trace[:, n - 1], table[:, n - 1])
    total = (acc[0, n - 1].astype(np.uint64) + last) % P
    assert not total.any()
    # a lookup of a value the table does not hold: the sum does not close, the proof fails
    bad = trace.copy()
    bad[0, 1, 0] = 5
    acc = O.stark_accumulators(air, table, bad, ac, ar)
    last = contribution(air, 0, ac, ar, bad[:, n - 1], table[:, n - 1])
    assert ((acc[0, n - 1].astype(np.uint64) + last) % P).any()
    label = b"zig-stark:stark:lookup"
    proof = O.stark_prove(air, table, bad, 3, 16, 3, O.Channel(label))
    assert proof is not None and not O.stark_verify(air, proof, O.Channel(label))


def test_prover_refuses_parameters_fri_params_rejects():
    air = airs.range_check_air()
    trace, table = airs.range_check_trace(16)
    assert O.stark_prove(air, table, trace, 3, 4, 1, O.Channel(b"x")) is None  # remainder_log < FRI blowup 2
    assert O.stark_prove(air, table, trace, 3, 4, 7, O.Channel(b"x")) is None  # remainder_log = domain


# ---- the library's host verifier (zs_stark_verify) on the oracle's proofs: CPU only ----
import zig_stark_amd as Z  # noqa: E402


@pytest.mark.parametrize("name", list(airs.LOOKUP_AIRS))
def test_library_verifier_decides_as_oracle(name):
    mk_air, mk_trace = airs.LOOKUP_AIRS[name]
    air = mk_air()
    trace, table = mk_trace(16)
    label = b"zig-stark:stark:lookup"
    proof = O.stark_prove(air, table, trace, 3, 16, 3, O.Channel(label))
    ch, och = Z.Channel(label), O.Channel(label)
    assert Z.stark_verify(proof, air.struct(), ch)
    assert O.stark_verify(air, proof, och)
    assert ch.sample_bytes(32) == och.sample_bytes(32)  # the transcript advanced the same way
    assert not Z.stark_verify(proof, air.struct(), Z.Channel(b"other"))
    # every revealed value of the first query, the quotient root, a path byte
    offs = reveal_offsets(proof)
    for k in range(air.num_columns * 2 + 2 + 1 + air.num_preprocessed):
        b = bytearray(proof)
        b[offs[0] + 16 * k] ^= 1
        assert not Z.stark_verify(bytes(b), air.struct(), Z.Channel(label)), k
    for cut in (proof[:-1], proof + b"\0"):
        with pytest.raises(Z.ZsError) as e:
            Z.stark_verify(cut, air.struct(), Z.Channel(label))
        assert e.value.code == Z.ZS_ERR_INPUT
    # a proof of another AIR shape (no preprocessed columns declared) does not verify
    other = airs.Air(3, num_preprocessed=0)
    assert not Z.stark_verify(proof, other.struct(), Z.Channel(label))


def test_library_verifier_on_generic_fib_proofs():
    n = 16
    t = O.fib_trace(1, 5, n)
    a0, a1, claimed = [1, 0, 0, 0], [5, 0, 0, 0], t[0, -1]
    air = airs.fib_seeded_air(n, a0, a1, claimed)
    proof = O.stark_prove(air, None, t, 3, 12, 3, O.Channel(b"fib"))
    assert Z.stark_verify(proof, air.struct(), Z.Channel(b"fib"))
    assert Z.fib_stark_verify(proof, a0, a1, claimed, Z.Channel(b"fib"))
    bad = airs.fib_seeded_air(n, a0, a1, (int(claimed[0]) + 1) % P)
    assert not Z.stark_verify(proof, bad.struct(), Z.Channel(b"fib"))


def test_library_rejects_malformed_air_descriptions():
    proof = O.stark_prove(airs.range_check_air(), airs.range_check_trace(16)[1], airs.range_check_trace(16)[0],
                          3, 4, 3, O.Channel(b"x"))
    for bad in (airs.Air(0), airs.Air(3, lookups=[(2, 1, [0], [0])], num_lookup_columns=1),  # lookups, no table
                airs.Air(2, constraints=[[(1, [("cur", 5)])]]),  # a column the trace does not have
                airs.Air(3, num_preprocessed=1, lookups=[(2, 1, [0], [3])], num_lookup_columns=1)):
        with pytest.raises(Z.ZsError) as e:
            Z.stark_verify(proof, bad.struct(), Z.Channel(b"x"))
        assert e.value.code == Z.ZS_ERR_INPUT


def combined(n, a1=12345):
    import airs as A
    fib = O.fib_trace(1, a1, n)
    a0, a1v, claimed = [1, 0, 0, 0], [a1, 0, 0, 0], fib[0, -1]
    trace, table = A.combined_trace(fib, n)
    return A.combined_air(n, a0, a1v, claimed), trace, table


@pytest.mark.parametrize("tl,beta,nq,rl", [(4, 3, 16, 3), (6, 2, 12, 3)])
def test_two_relations_with_transitions_round_trip(tl, beta, nq, rl):
    """Transitions, boundary assertions and two LogUp relations in one AIR: each
    relation's running sum telescopes to zero, the proof verifies in the oracle and in
    the library's host verifier (same transcript afterwards), and a tampered value or
    a wrong claim is rejected."""
    n = 1 << tl
    air, trace, table = combined(n)
    rng = np.random.default_rng(tl)
    ac = rng.integers(0, P, (air.num_lookup_columns, 4), dtype=np.uint32)
    ar = rng.integers(0, P, (2, 4), dtype=np.uint32)
    acc = O.stark_accumulators(air, table, trace, ac, ar)
    for r in range(2):
        assert not acc[r, 0].any()
        last = contribution(air, r, ac, ar, trace[:, n - 1], table[:, n - 1])
        assert not ((acc[r, n - 1].astype(np.uint64) + last) % P).any(), r
    label = b"zig-stark:stark:combined"
    proof = O.stark_prove(air, table, trace, beta, nq, rl, O.Channel(label))
    och, ch = O.Channel(label), Z.Channel(label)
    assert O.stark_verify(air, proof, och)
    assert Z.stark_verify(proof, air.struct(), ch)
    assert ch.sample_bytes(32) == och.sample_bytes(32)
    b = bytearray(proof)
    off = reveal_offsets(proof)[0] + 16 * 9  # the first query's value of column 9 (acc of relation 1 at x0)
    b[off] ^= 1
    assert not O.stark_verify(air, bytes(b), O.Channel(label))
    assert not Z.stark_verify(bytes(b), air.struct(), Z.Channel(label))
    wrong = combined(n, a1=777)[0]  # another claimed value (and seed): the proof is of this trace only
    assert not Z.stark_verify(proof, wrong.struct(), Z.Channel(label))
    # a range-check value the table lacks breaks relation 0 only
    bad = trace.copy()
    bad[2, 1, 0] = 6
    acc = O.stark_accumulators(air, table, bad, ac, ar)
    sums = [((acc[r, n - 1].astype(np.uint64) +
              contribution(air, r, ac, ar, bad[:, n - 1], table[:, n - 1])) % P).any() for r in range(2)]
    assert sums == [True, False]
    assert not O.stark_verify(air, O.stark_prove(air, table, bad, beta, nq, rl, O.Channel(label)), O.Channel(label))
