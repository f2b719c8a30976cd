"""CPU pins for the GenericStark(FibSeededAir) restatement (stark.zig:303-786):
the reference's own prove/verify tests (round trip; wrong claimed value
rejected, stark.zig:933-950, 1282-1298), tampered and truncated proofs
rejected, and the proof bytes parsing with the wire reader's rules."""
import numpy as np
import pytest

import oracle_lib as O

P = O.P


def qv(x):
    return np.array([x, 0, 0, 0], np.uint32)


def fib_case(tl, a1=5):
    t = O.fib_trace(1, a1, 1 << tl)
    return t, qv(1), qv(a1), t[0, -1]


@pytest.mark.parametrize("tl,beta,nq,rl", [(4, 3, 12, 3), (5, 2, 8, 3), (6, 1, 5, 2)])
def test_prove_verify_round_trip(tl, beta, nq, rl):
    t, a0, a1, claimed = fib_case(tl)
    proof = O.fib_stark_prove(t, beta, nq, rl, a0, a1, claimed, O.Channel(b"fib"))
    assert O.fib_stark_verify(proof, a0, a1, claimed, O.Channel(b"fib"))
    assert not O.fib_stark_verify(proof, a0, a1, claimed, O.Channel(b"other"))
    assert not O.fib_stark_verify(proof + b"\x00", a0, a1, claimed, O.Channel(b"fib"))
    assert not O.fib_stark_verify(proof[:-1], a0, a1, claimed, O.Channel(b"fib"))


def test_wrong_claimed_value_rejected():
    # stark.zig:1282-1298: a proof for a wrong claimed fib does not verify
    t, a0, a1, claimed = fib_case(4)
    bad = claimed.copy()
    bad[0] = (int(bad[0]) + 1) % P
    proof = O.fib_stark_prove(t, 3, 12, 3, a0, a1, bad, O.Channel(b"fib"))
    assert not O.fib_stark_verify(proof, a0, a1, bad, O.Channel(b"fib"))


def test_tampered_proof_rejected():
    t, a0, a1, claimed = fib_case(4)
    proof = bytearray(O.fib_stark_prove(t, 3, 12, 3, a0, a1, claimed, O.Channel(b"fib")))
    # flip one byte in the quotient root, one in a deep evaluation and one in a revealed value
    for off in (11 + 8 + 64 + 2 + 5, 11 + 8 + 64 + 2 + 32 + 8 + 3, len(proof) - 200):
        bad = bytearray(proof)
        bad[off] ^= 1
        assert not O.fib_stark_verify(bytes(bad), a0, a1, claimed, O.Channel(b"fib")), off
