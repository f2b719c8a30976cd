"""GPU parity for GenericStark's trace commitment (stark.zig:343-414): QM31 DFT
interpolation over H and evaluation on D against the oracle's O(n^2) Lagrange
+ Horner restatement, column trees of hashQM31 leaves, roots absorbed into the
channel in column order; and the seeded Fibonacci AIR trace of the STARK
configuration (2 columns x 2^12, blowup 4)."""
import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu
P = O.P


def run_gpu(zctx, torch_dev, cols, beta, channel=None):
    import torch
    n_cols, n = cols.shape[0], cols.shape[1]
    tl = n.bit_length() - 1
    N = n << beta
    d_cols = to_dev(cols.reshape(-1), torch_dev)
    d_cw = torch.empty(n_cols * N * 4, dtype=torch.int32, device=torch_dev)
    d_nodes = torch.empty(n_cols * (2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
    roots = zctx.trace_commit_dev(d_cols.data_ptr(), n_cols, tl, beta, d_cw.data_ptr(), d_nodes.data_ptr(), channel)
    return roots, from_dev(d_cw).reshape(n_cols, N, 4), from_dev(d_nodes).reshape(n_cols, -1)


@pytest.mark.parametrize("trace_log,beta,n_cols", [(1, 1, 1), (2, 2, 1), (3, 1, 2), (4, 3, 2), (6, 2, 3),
                                                   (9, 2, 2), (10, 1, 1), (11, 2, 1)])
def test_trace_commit_matches_oracle(zctx, torch_dev, trace_log, beta, n_cols):
    rng = np.random.default_rng(trace_log * 100 + beta * 10 + n_cols)
    cols = rng.integers(0, P, (n_cols, 1 << trace_log, 4), dtype=np.uint32)
    ch = Z.Channel(b"zig-stark:stark")
    roots, cw, nodes = run_gpu(zctx, torch_dev, cols, beta, ch)
    och = O.Channel(b"zig-stark:stark")
    oroots, ocw = O.trace_commit(cols, beta, channel=och)
    assert np.array_equal(cw, ocw)
    assert roots == oroots
    assert ch.sample_bytes(32) == och.sample_bytes(32)
    # every level of each column tree
    for c in range(n_cols):
        leaves = b"".join(O.hash_qm31(v) for v in ocw[c])
        assert nodes[c].view(np.uint32).tobytes() == O.merkle_build(leaves)


def test_fibonacci_stark_trace_commitment(zctx, torch_dev):
    """The STARK configuration's trace: seeded Fibonacci AIR, 2 columns x 2^12,
    a0 = 1, a1 = the seed-1 uniform M31 value, blowup 4 (codewords of 2^14)."""
    a1 = int(O.fill_m31(1, 1)[0])
    trace = Z.fib_trace(1, a1, 12)
    assert np.array_equal(trace, O.fib_trace(1, a1, 1 << 12))
    ch = Z.Channel(b"zig-stark:fib")
    roots, cw, _ = run_gpu(zctx, torch_dev, trace, 2, ch)
    och = O.Channel(b"zig-stark:fib")
    oroots, ocw = O.trace_commit(trace, 2, channel=och)
    assert roots == oroots
    assert np.array_equal(cw, ocw)
    assert ch.sample_qm31().tolist() == och.sample_qm31().tolist()


def test_trace_commit_large_size_property(zctx, torch_dev):
    """2^16 rows, blowup 4: the trace is a degree-8 polynomial p sampled on H, so the
    codeword must equal p at every point of D (checked at sampled points) and the
    root must be the tree over the codeword's hashQM31 leaves."""
    rng = np.random.default_rng(16)
    tl, beta = 16, 2
    n = 1 << tl
    coeffs = rng.integers(0, P, (8, 4), dtype=np.uint32)
    w = O.qm31_root(tl)
    trace = np.zeros((1, n, 4), np.uint32)
    x = np.array([1, 0, 0, 0], np.uint32)
    for i in range(n):
        trace[0, i] = O.poly_eval(coeffs, x)
        x = O.qm31_mul(x, w)
    roots, cw, _ = run_gpu(zctx, torch_dev, trace, beta)
    w_ev = O.qm31_root(tl + beta)
    off = np.array([1, 0, 1, 0], np.uint32)
    for i in rng.integers(0, n << beta, 32).tolist():
        d = O.qm31_mul(off, O.qm31_pow(w_ev, i))
        assert np.array_equal(cw[0, i], O.poly_eval(coeffs, d)), i
    leaves = b"".join(O.hash_qm31(v) for v in cw[0])
    assert O.tree_root(O.merkle_build(leaves)) == roots[0]


@pytest.mark.parametrize("tl,beta,nq,rl", [(4, 3, 12, 3), (6, 2, 8, 3), (9, 1, 5, 2)])
def test_fib_stark_proof_matches_oracle(zctx, torch_dev, tl, beta, nq, rl):
    """The whole GenericStark(FibAir) proof, byte for byte: trace roots, quotient
    root, DEEP evaluations, every FRI layer root, remainder, query indices,
    revealed values and paths; the oracle's verifier accepts it."""
    a1 = 5 + tl
    trace = Z.fib_trace(1, a1, tl)
    a0v, a1v = np.array([1, 0, 0, 0], np.uint32), np.array([a1, 0, 0, 0], np.uint32)
    claimed = trace[0, -1]
    d_tr = to_dev(trace.reshape(-1), torch_dev)
    ch = Z.Channel(b"zig-stark:fib")
    got = zctx.fib_stark_prove_dev(d_tr.data_ptr(), tl, beta, nq, rl, a0v, a1v, claimed, ch)
    och = O.Channel(b"zig-stark:fib")
    want = O.fib_stark_prove(trace, beta, nq, rl, a0v, a1v, claimed, och)
    assert got == want
    assert ch.sample_bytes(32) == och.sample_bytes(32)
    assert O.fib_stark_verify(got, a0v, a1v, claimed, O.Channel(b"zig-stark:fib"))
    assert Z.fib_stark_verify(got, a0v, a1v, claimed, Z.Channel(b"zig-stark:fib"))


def test_fib_stark_configuration_proof(zctx, torch_dev):
    """The STARK configuration: seeded Fibonacci AIR, 2 columns x 2^12, blowup 4,
    20 queries, remainder 2^3: the GPU proof equals the oracle's and verifies."""
    a1 = int(O.fill_m31(1, 1)[0])
    trace = Z.fib_trace(1, a1, 12)
    a0v, a1v = np.array([1, 0, 0, 0], np.uint32), np.array([a1, 0, 0, 0], np.uint32)
    claimed = trace[0, -1]
    d_tr = to_dev(trace.reshape(-1), torch_dev)
    got = zctx.fib_stark_prove_dev(d_tr.data_ptr(), 12, 2, 20, 3, a0v, a1v, claimed, Z.Channel(b"zig-stark:fib"))
    want = O.fib_stark_prove(trace, 2, 20, 3, a0v, a1v, claimed, O.Channel(b"zig-stark:fib"))
    assert got == want
    assert O.fib_stark_verify(got, a0v, a1v, claimed, O.Channel(b"zig-stark:fib"))
    bad = np.array(claimed)
    bad[0] = (int(bad[0]) + 1) % P
    assert not O.fib_stark_verify(got, a0v, a1v, bad, O.Channel(b"zig-stark:fib"))
    # the library's host verifier decides the same
    assert Z.fib_stark_verify(got, a0v, a1v, claimed, Z.Channel(b"zig-stark:fib"))
    assert not Z.fib_stark_verify(got, a0v, a1v, bad, Z.Channel(b"zig-stark:fib"))
