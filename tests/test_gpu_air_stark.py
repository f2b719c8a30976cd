"""GenericStark(Air) proofs from the device (zs_stark_prove_dev) for AIRs given as
data: byte-identical to the oracle's restatement of proveWithPreprocessed
(src/m31/stark.zig:311-657) for the reference's LogUp test AIRs (range check,
AND table, multiplicity; stark.zig:957-1280) at the reference's parameters and
at larger traces, and for the seeded Fibonacci AIR (equal to the dedicated
zs_fib_stark_prove_dev); the proofs verify on the host and in the oracle, and a
trace whose lookups are missing from the table yields a proof that does not."""
import numpy as np
import pytest

import airs
import oracle_lib as O
import zig_stark_amd as Z
from conftest import to_dev

pytestmark = pytest.mark.gpu
LABEL = b"zig-stark:stark:lookup"


def gpu_prove(zctx, torch_dev, air, trace, table, lb, nq, rl, label):
    tl = trace.shape[1].bit_length() - 1
    d_trace = to_dev(trace.reshape(-1), torch_dev)
    d_pre = to_dev(table.reshape(-1), torch_dev) if table is not None else None
    ch = Z.Channel(label)
    proof = zctx.stark_prove_dev(air.struct(), d_pre.data_ptr() if d_pre is not None else None, d_trace.data_ptr(),
                                 tl, lb, nq, rl, ch)
    return proof, ch


@pytest.mark.parametrize("name", list(airs.LOOKUP_AIRS))
def test_reference_lookup_airs_match_oracle(zctx, torch_dev, name):
    mk_air, mk_trace = airs.LOOKUP_AIRS[name]
    air = mk_air()
    trace, table = mk_trace(16)
    got, ch = gpu_prove(zctx, torch_dev, air, trace, table, 3, 16, 3, LABEL)
    och = O.Channel(LABEL)
    want = O.stark_prove(air, table, trace, 3, 16, 3, och)
    assert got == want
    assert ch.sample_bytes(32) == och.sample_bytes(32)
    assert Z.stark_verify(got, air.struct(), Z.Channel(LABEL))
    assert O.stark_verify(air, got, O.Channel(LABEL))


@pytest.mark.parametrize("name,tl,lb,nq,rl", [("range_check", 10, 2, 20, 3), ("and_table", 9, 3, 12, 4),
                                              ("range_check", 6, 1, 8, 2)])
def test_larger_lookup_traces_match_oracle(zctx, torch_dev, name, tl, lb, nq, rl):
    mk_air, mk_trace = airs.LOOKUP_AIRS[name]
    air = mk_air()
    trace, table = mk_trace(1 << tl)
    got, _ = gpu_prove(zctx, torch_dev, air, trace, table, lb, nq, rl, LABEL)
    assert got == O.stark_prove(air, table, trace, lb, nq, rl, O.Channel(LABEL))
    assert Z.stark_verify(got, air.struct(), Z.Channel(LABEL))


def test_missing_lookup_value_does_not_verify(zctx, torch_dev):
    air = airs.range_check_air()
    trace, table = airs.range_check_trace(64)
    trace[0, 3, 0] = 7  # looked up, absent from the table
    got, _ = gpu_prove(zctx, torch_dev, air, trace, table, 3, 16, 3, LABEL)
    assert got == O.stark_prove(air, table, trace, 3, 16, 3, O.Channel(LABEL))
    assert not Z.stark_verify(got, air.struct(), Z.Channel(LABEL))


@pytest.mark.parametrize("tl,lb,nq,rl", [(6, 2, 10, 3), (12, 2, 20, 3)])
def test_generic_fib_equals_dedicated_prover(zctx, torch_dev, tl, lb, nq, rl):
    n = 1 << tl
    a1 = 987654321
    t = O.fib_trace(1, a1, n)
    a0v, a1v, claimed = np.array([1, 0, 0, 0], np.uint32), np.array([a1, 0, 0, 0], np.uint32), t[0, -1]
    air = airs.fib_seeded_air(n, a0v, a1v, claimed)
    got, _ = gpu_prove(zctx, torch_dev, air, t, None, lb, nq, rl, b"zig-stark:fib")
    d_t = to_dev(t.reshape(-1), torch_dev)
    want = zctx.fib_stark_prove_dev(d_t.data_ptr(), tl, lb, nq, rl, a0v, a1v, claimed, Z.Channel(b"zig-stark:fib"))
    assert got == want
    if tl <= 6:
        assert got == O.stark_prove(air, None, t, lb, nq, rl, O.Channel(b"zig-stark:fib"))
    assert Z.stark_verify(got, air.struct(), Z.Channel(b"zig-stark:fib"))


def test_prover_rejects_parameters_and_airs(zctx, torch_dev):
    air = airs.range_check_air()
    trace, table = airs.range_check_trace(16)
    for rl in (1, 7):  # below the FRI blowup (2); the whole domain
        with pytest.raises(Z.ZsError) as e:
            gpu_prove(zctx, torch_dev, air, trace, table, 3, 4, rl, LABEL)
        assert e.value.code == Z.ZS_ERR_INPUT
    with pytest.raises(Z.ZsError) as e:  # lookups without the table columns
        gpu_prove(zctx, torch_dev, airs.Air(3, lookups=[(2, 1, [0], [0])], num_lookup_columns=1), trace, None, 3, 4,
                  3, LABEL)
    assert e.value.code == Z.ZS_ERR_INPUT


@pytest.mark.parametrize("name,tl", [("range_check", 13), ("and_table", 14), ("range_check", 16)])
def test_multi_tile_traces_verify(zctx, torch_dev, name, tl):
    """Above 2^12 rows the accumulator scan and the evaluations at z run over
    several tiles (2048 and 4096 rows each); the oracle's O(n^2) interpolation is
    too slow to compare bytes here, so the host verifier (itself pinned on the
    oracle's proofs) decides: an honest trace's proof verifies, and one looked-up
    value missing from the table in the last tile gives a proof that does not."""
    mk_air, mk_trace = airs.LOOKUP_AIRS[name]
    air = mk_air()
    trace, table = mk_trace(1 << tl)
    got, _ = gpu_prove(zctx, torch_dev, air, trace, table, 2, 20, 3, LABEL)
    assert Z.stark_verify(got, air.struct(), Z.Channel(LABEL))
    bad = trace.copy()
    bad[0, (1 << (tl - 1)) - 3, 0] = 9  # a selected row of the last tile of the lookup half
    got, _ = gpu_prove(zctx, torch_dev, air, bad, table, 2, 20, 3, LABEL)
    assert not Z.stark_verify(got, air.struct(), Z.Channel(LABEL))


def combined(n, a1=12345):
    fib = O.fib_trace(1, a1, n)
    trace, table = airs.combined_trace(fib, n)
    return airs.combined_air(n, [1, 0, 0, 0], [a1, 0, 0, 0], fib[0, -1]), trace, table


@pytest.mark.parametrize("tl,lb,nq,rl", [(4, 3, 16, 3), (9, 2, 20, 3), (7, 1, 10, 2)])
def test_two_relations_with_transitions_match_oracle(zctx, torch_dev, tl, lb, nq, rl):
    """Fibonacci transitions and boundary assertions beside two LogUp relations (range
    check, AND table) with shared column challenges: the device proof is the oracle's,
    byte for byte, and verifies on the host and in the oracle."""
    air, trace, table = combined(1 << tl)
    label = b"zig-stark:stark:combined"
    got, ch = gpu_prove(zctx, torch_dev, air, trace, table, lb, nq, rl, label)
    och = O.Channel(label)
    assert got == O.stark_prove(air, table, trace, lb, nq, rl, och)
    assert ch.sample_bytes(32) == och.sample_bytes(32)
    assert Z.stark_verify(got, air.struct(), Z.Channel(label))
    assert O.stark_verify(air, got, O.Channel(label))


def test_two_relations_multi_tile_verify(zctx, torch_dev):
    """2^14 rows (several accumulator-scan and evaluation tiles, two scan rows): honest
    proof verifies; one AND-table key missing from the table does not."""
    air, trace, table = combined(1 << 14)
    label = b"zig-stark:stark:combined"
    got, _ = gpu_prove(zctx, torch_dev, air, trace, table, 2, 20, 3, label)
    assert Z.stark_verify(got, air.struct(), Z.Channel(label))
    bad = trace.copy()
    bad[5, (1 << 13) - 5, 0] = 3  # x = 3 is in no (ta, tb) row
    got, _ = gpu_prove(zctx, torch_dev, air, bad, table, 2, 20, 3, label)
    assert not Z.stark_verify(got, air.struct(), Z.Channel(label))
