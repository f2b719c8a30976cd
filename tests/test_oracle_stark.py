"""CPU pins for the GenericStark trace commitment restatement (stark.zig:343-414)
and the seeded Fibonacci trace (stark.zig:877-931), plus the library's host-side
trace generator against the oracle."""
import numpy as np
import pytest

import oracle_lib as O

P = O.P


def test_fib_trace_known_values():
    # seeds (0, 1) give the Fibonacci numbers: FibAir.generateTrace (stark.zig:918-935)
    t = O.fib_trace(0, 1, 64)
    a = t[0, :, 0].astype(int)
    assert list(a[:11]) == [0, 1, 1, 2, 3, 5, 8, 13, 21, 34, 55]
    assert a[20] == 6765 and a[30] == 832040
    assert a[47] == 2971215073 % P  # first Fibonacci number above p
    assert np.array_equal(t[1, :-1], t[0, 1:])  # column 1 is column 0 shifted by one step
    assert not t[:, :, 1:].any()  # M31 values embedded in the first QM31 coordinate


def test_fib_transition_constraints_hold():
    # FibAir.evalTransition: next[0] = cur[1], next[1] = cur[0] + cur[1]
    t = O.fib_trace(1, 123456789, 256)[:, :, 0].astype(np.int64)
    assert np.array_equal(t[0, 1:], t[1, :-1])
    assert np.array_equal(t[1, 1:], (t[0, :-1] + t[1, :-1]) % P)


def test_library_fib_trace_matches_oracle():
    import zig_stark_amd as Z
    for lg, a1 in [(0, 5), (4, 7), (12, 2**31 - 2)]:
        assert np.array_equal(Z.fib_trace(1, a1, lg), O.fib_trace(1, a1, 1 << lg))


@pytest.mark.parametrize("trace_log,beta", [(2, 1), (3, 2), (5, 2)])
def test_trace_commit_codeword_is_generating_polynomial_on_D(trace_log, beta):
    """A trace that is a degree < n polynomial p sampled on H = <w> commits to a
    codeword equal to p on D = FRI_OFFSET * <w_ev> (uniqueness of the interpolant)."""
    rng = np.random.default_rng(trace_log * 10 + beta)
    n, N = 1 << trace_log, 1 << (trace_log + beta)
    coeffs = rng.integers(0, P, (n, 4), dtype=np.uint32)
    w = O.qm31_root(trace_log)
    trace = np.array([O.poly_eval(coeffs, O.qm31_pow(w, i)) for i in range(n)], np.uint32)
    roots, cw = O.trace_commit(trace[None], beta)
    w_ev = O.qm31_root(trace_log + beta)
    off = np.array([1, 0, 1, 0], np.uint32)
    d = [O.qm31_mul(off, O.qm31_pow(w_ev, i)) for i in range(N)]
    want = np.array([O.poly_eval(coeffs, x) for x in d], np.uint32)
    assert np.array_equal(cw[0], want)
    # D is a coset of the order-N group: d[i + N/2] = -d[i]
    for i in range(N // 2):
        assert np.array_equal((np.asarray(d[i], np.int64) + d[i + N // 2]) % P, np.zeros(4))
    # root = MerkleTree over hashQM31 leaves of the codeword
    leaves = b"".join(O.hash_qm31(v) for v in cw[0])
    nodes = O.merkle_build(leaves)
    assert O.tree_root(nodes) == roots[0]


def test_trace_commit_absorbs_roots_in_column_order():
    t = O.fib_trace(1, 99, 8)
    ch = O.Channel(b"stark")
    roots, _ = O.trace_commit(t, 2, channel=ch, want_codewords=False)
    ref = O.Channel(b"stark")
    for r in roots:
        ref.absorb(r)
    assert ch.sample_bytes(32) == ref.sample_bytes(32)
