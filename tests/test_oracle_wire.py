"""CPU pins for the wire-format restatement (src/core/serialization.zig):
the golden layout example of the reference's own test, and Fri.Proof bytes
that parse back to the proof they came from (and reject trailing or missing
bytes, as deserialize does)."""
import numpy as np
import pytest

import oracle_lib as O
from wire import WireError, parse_fri_proof

P = O.P


def test_golden_wire_layout():
    # serialization.zig:219-247 ("golden wire layout is stable")
    expected = bytes([ord("a"), ord("b"), ord("c"), 0x04, 0x03, 0x02, 0x01, 0x02, 0, 0, 0, 0, 0, 0, 0,
                      0x02, 0x01, 0x04, 0x03, 0x01, 0x01, 0xde, 0xad])
    assert O.wire_golden_example() == expected


def test_fri_proof_bytes_round_trip():
    rng = np.random.default_rng(5)
    dlog, beta, rem_log, nq = 9, 2, 4, 6
    cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
    proof = O.fri_prove(cw, beta, rem_log, nq, O.Channel(b"wire"))
    b = O.fri_proof_bytes(proof, beta, rem_log)
    p = parse_fri_proof(b)
    L = dlog - rem_log
    assert (p["log_size"], p["log_blowup"], p["num_queries"], p["remainder_log"]) == (dlog - beta, beta, nq, rem_log)
    assert b"".join(p["roots"]) == proof["roots"] and len(p["roots"]) == L
    assert np.array_equal(np.array(p["remainder"], np.uint32).reshape(-1), proof["remainder"])
    qv = proof["qvals"].reshape(nq, L, 2, 4)
    per_q = sum(2 * (dlog - i) * 32 for i in range(L))
    for q in range(nq):
        assert p["queries"][q]["index"] == proof["qidx"][q]
        pp = proof["qpaths"][q * per_q:(q + 1) * per_q]
        off = 0
        for i, pair in enumerate(p["queries"][q]["pairs"]):
            assert pair["value0"] == qv[q, i, 0].tolist() and pair["value1"] == qv[q, i, 1].tolist()
            plen = (dlog - i) * 32
            assert b"".join(pair["path0"]) == pp[off:off + plen] and b"".join(pair["path1"]) == pp[off + plen:off + 2 * plen]
            off += 2 * plen
    with pytest.raises(WireError):
        parse_fri_proof(b + b"\x00")
    with pytest.raises(WireError):
        parse_fri_proof(b[:-1])
