"""GPU parity: FRI commit phase (src/m31/fri.zig:157-256) against the oracle:
every layer, every root, the remainder, the channel-drawn challenges, query
indices, revealed values and Merkle paths; and a full-size 2^22 commit whose
proof the oracle's verifier accepts."""
import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z

pytestmark = pytest.mark.gpu
P = O.P


@pytest.mark.parametrize("dlog,beta,rem_log", [(3, 0, 1), (6, 2, 3), (10, 2, 5), (14, 2, 5), (16, 3, 4)])
def test_fri_fixed_alphas_match_oracle(zctx, dlog, beta, rem_log):
    rng = np.random.default_rng(dlog)
    cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
    L = dlog - rem_log
    alphas = rng.integers(0, P, (L, 4), dtype=np.uint32)
    roots, rem, res = zctx.fri_commit(cw, beta, rem_log, alphas=alphas)
    oroots, olayers, orem, _ = O.fri_commit(cw, beta, rem_log, alphas=alphas)
    assert roots == oroots
    assert np.array_equal(rem, orem)
    for i in range(L + 1):
        assert np.array_equal(res.layer(i), olayers[i]), i


@pytest.mark.parametrize("log_size,beta,rem_log,queries", [(4, 3, 3, 8), (8, 2, 5, 20), (12, 2, 5, 20)])
def test_fri_channel_prove_matches_oracle(zctx, log_size, beta, rem_log, queries):
    rng = np.random.default_rng(log_size)
    dlog = log_size + beta
    cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
    ch = Z.Channel(b"zig-stark:fri:test")
    roots, rem, res = zctx.fri_commit(cw, beta, rem_log, channel=ch)
    idx, vals, paths = res.query(ch, queries)
    proof = O.fri_prove(cw, beta, rem_log, queries, O.Channel(b"zig-stark:fri:test"))
    assert b"".join(roots) == proof["roots"]
    assert np.array_equal(rem.reshape(-1), proof["remainder"])
    assert idx == proof["qidx"]
    assert np.array_equal(vals, proof["qvals"])
    assert paths == proof["qpaths"]


def low_degree_codeword_on_gpu(zctx, rng, log_size, beta):
    """A QM31 codeword of degree < 2^log_size on D_0: evaluate random
    coefficients by Horner over the FRI domain (oracle domain points)."""
    coeffs = rng.integers(0, P, (1 << log_size, 4), dtype=np.uint32)
    dlog = log_size + beta
    return coeffs, np.array([O.poly_eval(coeffs, O.fri_domain_point(dlog, 0, j)) for j in range(1 << dlog)])


def test_fri_low_degree_proof_verifies(zctx):
    rng = np.random.default_rng(3)
    _, cw = low_degree_codeword_on_gpu(zctx, rng, 8, 2)
    ch = Z.Channel(b"zig-stark:fri:test")
    roots, rem, res = zctx.fri_commit(cw, 2, 3, channel=ch)
    idx, vals, paths = res.query(ch, 16)
    proof = dict(dlog=10, L=7, roots=b"".join(roots), remainder=rem.reshape(-1), qidx=idx, qvals=vals,
                 qpaths=paths, num_queries=16)
    assert O.fri_verify(proof, 2, 3, O.Channel(b"zig-stark:fri:test"))
    bad = dict(proof, qvals=vals.copy())
    bad["qvals"][5] = (int(bad["qvals"][5]) + 1) % P
    assert not O.fri_verify(bad, 2, 3, O.Channel(b"zig-stark:fri:test"))
    # the same proof in wire bytes, through the library's host verifier
    assert Z.fri_verify(O.fri_proof_bytes(proof, 2, 3), Z.Channel(b"zig-stark:fri:test"))
    assert not Z.fri_verify(O.fri_proof_bytes(bad, 2, 3), Z.Channel(b"zig-stark:fri:test"))


def test_fri_prove_dev_low_degree_verifies_on_host(zctx, torch_dev):
    """zs_fri_prove_dev's wire bytes for a low-degree codeword pass zs_fri_verify (and the
    oracle's verify); one flipped value bit fails both."""
    from conftest import to_dev
    rng = np.random.default_rng(77)
    log_size, beta, rem_log, nq = 9, 2, 4, 12
    _, cw = low_degree_codeword_on_gpu(zctx, rng, log_size, beta)
    d_cw = to_dev(cw.reshape(-1), torch_dev)
    b = zctx.fri_prove_dev(d_cw.data_ptr(), log_size, beta, rem_log, nq, Z.Channel(b"zig-stark:fri:host"))
    assert Z.fri_verify(b, Z.Channel(b"zig-stark:fri:host"))
    L = log_size + beta - rem_log
    q0 = 11 + 8 + 32 * L + 8 + 16 * (1 << (rem_log - beta)) + 8  # first query
    t = bytearray(b)
    t[q0 + 16 + 3] ^= 0x10  # pair 0, value0
    assert not Z.fri_verify(bytes(t), Z.Channel(b"zig-stark:fri:host"))


def test_fri_full_size_2p22(zctx):
    """One QM31 column of 2^22, folded to 2^5 (17 committed layers), seeded alphas."""
    rng = np.random.default_rng(3)
    dlog, beta, rem_log = 22, 2, 5
    cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
    L = dlog - rem_log
    alphas = rng.integers(0, P, (L, 4), dtype=np.uint32)
    roots, rem, res = zctx.fri_commit(cw, beta, rem_log, alphas=alphas)
    assert len(roots) == 17
    # each layer is the fold of the previous one (oracle fold on sampled slices)
    prev = res.layer(0)
    assert np.array_equal(prev, cw)
    for i in range(L):
        nxt = res.layer(i + 1)
        n = prev.shape[0]
        for j in rng.integers(0, n // 2, 4).tolist():
            # fold a 2-element layer made of the pair (j, j + n/2): fold index 0 uses D_i[j]
            y0, y1 = prev[j], prev[j + n // 2]
            x = O.fri_domain_point(22, i, j)
            xi = O.qm31_inv(x)
            half = (P + 1) // 2
            hs = (O.qm31_mul((y0.astype(np.uint64) + y1) % P, [half, 0, 0, 0]))
            hd = O.qm31_mul(O.qm31_mul(((y0.astype(np.int64) - y1) % P).astype(np.uint32), xi), [half, 0, 0, 0])
            v = (hs.astype(np.uint64) + O.qm31_mul(alphas[i], hd)) % P
            assert np.array_equal(nxt[j], v.astype(np.uint32)), (i, j)
        prev = nxt
    # the last layer's remainder interpolant reproduces its first points
    last = res.layer(L)
    for j in range(1 << (rem_log - beta)):
        assert np.array_equal(O.poly_eval(rem, O.fri_domain_point(22, L, j)), last[j])
    # the last committed layers' roots equal the oracle's tree over the same values
    for i in range(L - 4, L):
        leaves = b"".join(O.hash_qm31(v) for v in res.layer(i))
        assert roots[i] == O.tree_root(O.merkle_build(leaves))


@pytest.mark.parametrize("dlog,beta,rem_log", [(10, 2, 5), (15, 2, 5), (18, 1, 3)])
def test_fri_graph_replays_match_eager_and_oracle(zctx, torch_dev, dlog, beta, rem_log):
    """Challenges given up front: the captured graph is replayed on later calls.
    Every replay (new challenges, new codeword contents in the same buffer, eager
    work and synchronisations in between) equals the eager launches and the oracle."""
    import torch
    from conftest import to_dev
    rng = np.random.default_rng(dlog * 7 + beta)
    L = dlog - rem_log
    d_cw = torch.empty((1 << dlog) * 4, dtype=torch.int32, device=torch_dev)
    for rep in range(4):
        cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
        d_cw.copy_(to_dev(cw.reshape(-1), torch_dev))
        alphas = rng.integers(0, P, (L, 4), dtype=np.uint32)
        g_roots, g_rem, _ = zctx.fri_commit_dev(d_cw.data_ptr(), dlog, beta, rem_log, alphas=alphas,
                                                flags=Z.ZS_FLAG_GRAPH)
        e_roots, e_rem, _ = zctx.fri_commit_dev(d_cw.data_ptr(), dlog, beta, rem_log, alphas=alphas)
        assert g_roots == e_roots and np.array_equal(g_rem, e_rem), rep
        if dlog <= 15:
            oroots, _, orem, _ = O.fri_commit(cw, beta, rem_log, alphas=alphas, want_layers=False)
            assert g_roots == oroots and np.array_equal(g_rem, orem), rep
        torch.cuda.synchronize()
        _ = (d_cw.view(-1)[:1024].long().sum()).item()  # eager work between replays


@pytest.mark.parametrize("dlog,beta,rem_log", [(9, 2, 5), (14, 2, 4)])
def test_fri_device_channel_graph_replays(zctx, torch_dev, dlog, beta, rem_log):
    """Challenges drawn on the device from the channel, no result kept: the
    commit phase is captured once and replayed; roots, remainder and the
    channel state afterwards equal the oracle's on every call."""
    from conftest import to_dev
    rng = np.random.default_rng(dlog + 100)
    d_cw = to_dev(np.zeros((1 << dlog) * 4, np.uint32), torch_dev)
    for rep in range(3):
        cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
        d_cw.copy_(to_dev(cw.reshape(-1), torch_dev))
        for flags in (Z.ZS_FLAG_GRAPH, 0):
            ch = Z.Channel(b"zig-stark:fri:dev")
            roots, rem, _ = zctx.fri_commit_dev(d_cw.data_ptr(), dlog, beta, rem_log, channel=ch, flags=flags)
            och = O.Channel(b"zig-stark:fri:dev")
            oroots, _, orem, _ = O.fri_commit(cw, beta, rem_log, channel=och, want_layers=False)
            assert roots == oroots and np.array_equal(rem, orem), (rep, flags)
            assert ch.sample_bytes(32) == och.sample_bytes(32), (rep, flags)


def test_fri_device_channel_2p20_matches_oracle(zctx, torch_dev):
    """Channel challenges at 2^20: layer trees wide enough for the per-lane 3-level
    subtrees, the workgroup subtrees (one node per lane, then quads) and the draw
    in the launch that computes each root, then the one-workgroup tail."""
    from conftest import to_dev
    dlog, beta, rem_log = 20, 2, 5
    cw = np.random.default_rng(2020).integers(0, P, (1 << dlog, 4), dtype=np.uint32)
    d_cw = to_dev(cw.reshape(-1), torch_dev)
    och = O.Channel(b"zig-stark:fri:2p20")
    oroots, _, orem, _ = O.fri_commit(cw, beta, rem_log, channel=och, want_layers=False)
    for flags in (Z.ZS_FLAG_GRAPH, 0):
        ch = Z.Channel(b"zig-stark:fri:2p20")
        roots, rem, _ = zctx.fri_commit_dev(d_cw.data_ptr(), dlog, beta, rem_log, channel=ch, flags=flags)
        assert roots == oroots and np.array_equal(rem, orem), flags
    assert ch.sample_bytes(32) == och.sample_bytes(32)


@pytest.mark.parametrize("log_size,beta,rem_log,queries", [(8, 2, 5, 20), (12, 2, 3, 20), (6, 3, 4, 7)])
def test_fri_prove_wire_bytes_match_oracle(zctx, torch_dev, log_size, beta, rem_log, queries):
    """proveCodeword's Fri.Proof in the reference's wire format, byte for byte."""
    from conftest import to_dev
    from wire import parse_fri_proof
    rng = np.random.default_rng(log_size * 3 + rem_log)
    dlog = log_size + beta
    cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
    d_cw = to_dev(cw.reshape(-1), torch_dev)
    got = zctx.fri_prove_dev(d_cw.data_ptr(), log_size, beta, rem_log, queries, Z.Channel(b"zig-stark:wire"))
    proof = O.fri_prove(cw, beta, rem_log, queries, O.Channel(b"zig-stark:wire"))
    assert got == O.fri_proof_bytes(proof, beta, rem_log)
    # the host verifier decides as the oracle's verify (a random codeword is not low-degree)
    assert Z.fri_verify(got, Z.Channel(b"zig-stark:wire")) == O.fri_verify(proof, beta, rem_log,
                                                                           O.Channel(b"zig-stark:wire"))
    p = parse_fri_proof(got)
    assert len(p["queries"]) == queries and p["log_size"] == log_size


def test_fri_2p22_layer_trees_match_per_level_trees(zctx, torch_dev):
    """Every layer tree of a 2^22 seeded commit (built together, several trees per
    launch) equals the same layer's tree built alone by one launch per level.
    hashQM31(v) is BLAKE3 of v's 16 bytes, the row leaf of a 4-column row, so
    the per-level row commitment over the layer's coordinates is an independent
    builder of the same tree (src/m31/fri.zig:195-233, src/core/merkle/merkle.zig:17-40)."""
    import torch
    rng = np.random.default_rng(31)
    dlog, beta, rem_log = 22, 2, 5
    L = dlog - rem_log
    cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
    alphas = rng.integers(0, P, (L, 4), dtype=np.uint32)
    roots, _, res = zctx.fri_commit(cw, beta, rem_log, alphas=alphas)
    d_nodes = torch.empty(((2 << dlog) - 1) * 8, dtype=torch.int32, device=torch_dev)
    for i in range(L):
        vals = res.layer(i)
        lg = dlog - i
        d_cols = torch.from_numpy(np.ascontiguousarray(vals.T).view(np.int32)).to(torch_dev)
        zctx.row_commit_dev(d_cols.data_ptr(), 4, lg, 1 << lg, d_nodes.data_ptr(), Z.ZS_FLAG_PLAIN)
        zctx.synchronize()
        top = ((2 << lg) - 1) * 8
        root = d_nodes[top - 8:top].cpu().numpy().view(np.uint32).tobytes()
        assert root == roots[i], i


def test_fri_commit_dev_ignores_flags_it_does_not_take(zctx, torch_dev):
    """Only ZS_FLAG_GRAPH and ZS_FLAG_NO_GRAPH apply to zs_fri_commit_dev; other bits a caller passes
    (here the bit the commit job uses internally to say layer 0's leaves are already
    hashed) leave the commitment unchanged."""
    from conftest import to_dev
    dlog, beta, rem_log = 12, 2, 5
    rng = np.random.default_rng(77)
    cw = rng.integers(0, P, (1 << dlog, 4), dtype=np.uint32)
    alphas = rng.integers(0, P, (dlog - rem_log, 4), dtype=np.uint32)
    d_cw = to_dev(cw.reshape(-1), torch_dev)
    oroots, _, orem, _ = O.fri_commit(cw, beta, rem_log, alphas=alphas, want_layers=False)
    for flags in (0, 1 << 30, (1 << 30) | Z.ZS_FLAG_GRAPH, (1 << 30) | Z.ZS_FLAG_NO_GRAPH, 0xFFFFFFF0):
        roots, rem, _ = zctx.fri_commit_dev(d_cw.data_ptr(), dlog, beta, rem_log, alphas=alphas, flags=flags)
        assert roots == oroots and np.array_equal(rem, orem), hex(flags)
