"""GPU parity: blowup LDE + row-leaf BLAKE3 Merkle commitment against the
oracle (circleIFFT/circleFFT composition, hashM31s leaves, MerkleTree),
bit-exact; plus the full 64 x 2^20 -> 2^22 workload through
size-independent properties."""
import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu
P = O.P


def oracle_commit(cols, beta):
    lde = np.stack([O.lde(c, beta) for c in cols])
    return lde, O.row_commit(lde)


@pytest.mark.parametrize("n_cols,lg,beta", [
    (1, 1, 1), (2, 1, 2), (3, 2, 2), (5, 3, 1), (64, 6, 2), (16, 10, 2), (7, 11, 2),
    (4, 12, 3), (17, 8, 1), (300, 4, 2), (16, 9, 4), (4, 10, 0), (2, 16, 0),
])
def test_lde_commit_matches_oracle(zctx, n_cols, lg, beta):
    rng = np.random.default_rng(n_cols * 1000 + lg * 10 + beta)
    cols = rng.integers(0, P, (n_cols, 1 << lg), dtype=np.uint32)
    tree, root, lde = zctx.lde_commit(cols, beta)
    ref_lde, ref_nodes = oracle_commit(cols, beta)
    assert np.array_equal(lde, ref_lde)
    assert root == O.tree_root(ref_nodes)
    N = 1 << (lg + beta)
    # every level equals the oracle's
    off = 0
    for lvl in range(lg + beta + 1):
        cnt = N >> lvl
        assert tree.level(lvl) == ref_nodes[32 * off: 32 * (off + cnt)], lvl
        off += cnt
    # openings equal MerkleTree.open and verify
    for i in rng.integers(0, N, 6).tolist():
        path = tree.open(i)
        assert path == O.merkle_open(ref_nodes, N, i)
        assert O.merkle_verify(root, i, ref_nodes[32 * i: 32 * i + 32], path)


@pytest.mark.parametrize("n_cols,lg,beta", [(8, 3, 2), (64, 8, 2), (9, 12, 2)])
def test_plain_path_matches_fused(zctx, torch_dev, n_cols, lg, beta):
    import torch
    rng = np.random.default_rng(lg)
    cols = rng.integers(0, P, (n_cols, 1 << lg), dtype=np.uint32)
    N = 1 << (lg + beta)
    d_cols = to_dev(cols, torch_dev)
    outs = []
    for flags in (0, Z.ZS_FLAG_PLAIN):
        d_lde = torch.zeros((n_cols, N), dtype=torch.int32, device=torch_dev)
        d_nodes = torch.zeros((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
        zctx.lde_commit_dev(d_cols.data_ptr(), n_cols, lg, beta, d_lde.data_ptr(), d_nodes.data_ptr(), flags)
        zctx.synchronize()
        outs.append((from_dev(d_lde), from_dev(d_nodes)))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    ref_lde, ref_nodes = oracle_commit(cols, beta)
    assert np.array_equal(outs[0][0], ref_lde)
    assert outs[0][1].tobytes() == ref_nodes


def test_full_size_64x2p20_properties(zctx, torch_dev):
    """64 columns x 2^20 -> blowup 4 -> 2^22 rows, as the benchmark runs it."""
    import torch
    n_cols, lg, beta = 64, 20, 2
    n, N = 1 << lg, 1 << (lg + beta)
    d_cols = torch.empty((n_cols, n), dtype=torch.int32, device=torch_dev)
    zctx.fill_m31_dev(2, d_cols.data_ptr(), n_cols * n)
    d_lde = torch.empty((n_cols, N), dtype=torch.int32, device=torch_dev)
    d_nodes = torch.empty((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
    zctx.lde_commit_dev(d_cols.data_ptr(), n_cols, lg, beta, d_lde.data_ptr(), d_nodes.data_ptr())
    zctx.synchronize()
    cols_h = from_dev(d_cols)
    # the device generator is the oracle's seeded stream
    assert np.array_equal(cols_h[0, :4096], O.fill_m31(2, 4096))
    # (1) round trip: the IFFT of each LDE column is the zero-padded interpolant
    d_back = torch.empty_like(d_lde)
    zctx.circle_ifft_batch_dev(d_lde.data_ptr(), d_back.data_ptr(), lg + beta, n_cols, N)
    d_coef = torch.empty_like(d_cols)
    zctx.circle_ifft_batch_dev(d_cols.data_ptr(), d_coef.data_ptr(), lg, n_cols, n)
    zctx.synchronize()
    back, coef = from_dev(d_back), from_dev(d_coef)
    h, H = n // 2, N // 2
    assert np.array_equal(back[:, :h], coef[:, :h])
    assert np.array_equal(back[:, H:H + h], coef[:, h:])
    assert not back[:, h:H].any() and not back[:, H + h:].any()
    # (2) two whole columns against the oracle LDE
    lde_h = from_dev(d_lde)
    for c in (0, 63):
        assert np.array_equal(lde_h[c], O.lde(cols_h[c], beta))
    # (3) sampled leaves and nodes against the oracle hashes, and the tree is consistent
    nodes = from_dev(d_nodes).tobytes()
    rng = np.random.default_rng(5)
    for i in rng.integers(0, N, 64).tolist():
        assert nodes[32 * i: 32 * i + 32] == O.hash_m31s(lde_h[:, i])
    off = 0
    for lvl in range(lg + beta):
        cnt = N >> lvl
        for j in rng.integers(0, cnt // 2, 8).tolist():
            l = nodes[32 * (off + 2 * j): 32 * (off + 2 * j + 1)]
            r = nodes[32 * (off + 2 * j + 1): 32 * (off + 2 * j + 2)]
            assert nodes[32 * (off + cnt + j): 32 * (off + cnt + j + 1)] == O.hash2(l, r)
        off += cnt
    assert off == 2 * N - 2


@pytest.mark.parametrize("n_cols,lg,beta", [(3, 16, 2), (2, 17, 3), (2, 18, 1), (1, 21, 1), (5, 10, 3), (3, 11, 3)])
def test_lde_commit_register_group_path(zctx, n_cols, lg, beta):
    """Shapes that take the register-group passes and the fused interpolation/evaluation kernel
    (single-workgroup form for lg 8..11, two passes for lg 16..22)."""
    rng = np.random.default_rng(n_cols * 7 + lg)
    cols = rng.integers(0, P, (n_cols, 1 << lg), dtype=np.uint32)
    tree, root, lde = zctx.lde_commit(cols, beta)
    ref_lde = np.stack([O.lde(c, beta) for c in cols])
    assert np.array_equal(lde, ref_lde)
    N = 1 << (lg + beta)
    leaves = b"".join(O.hash_m31s(ref_lde[:, i]) for i in range(0, N, max(1, N // 512)))
    got = tree.level(0)
    step = max(1, N // 512)
    assert b"".join(got[32 * i: 32 * i + 32] for i in range(0, N, step)) == leaves
    if N <= 1 << 17:
        assert root == O.tree_root(O.row_commit(ref_lde))


@pytest.mark.parametrize("n_cols,lg,beta", [(8, 10, 2), (24, 13, 2), (3, 19, 2), (2, 20, 2)])
def test_autotuned_lde_is_bit_exact(zctx, torch_dev, n_cols, lg, beta):
    """The autotuner keeps a configuration per shape; LDE and root stay equal to the oracle's.
    Every candidate launch shape must give the same root (the autotuner checks), so the
    two-pass shapes (lg 19, 20) cover each tile size of the pipelined last pass."""
    import torch
    cols = O.fill_m31(lg * 31 + n_cols, n_cols << lg).reshape(n_cols, -1)
    tuned = zctx.autotune_lde(n_cols, lg, beta)
    for role in ("interp", "mid", "eval"):
        el, cpw = tuned[role]
        assert 11 <= el <= 14 and cpw in (1, 2, 4, 8, 16, 32)
    assert tuned["us"] > 0
    d_cols = to_dev(cols.reshape(-1), torch_dev)
    N = 1 << (lg + beta)
    d_lde = torch.empty(n_cols * N, dtype=torch.int32, device=torch_dev)
    d_nodes = torch.empty((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
    zctx.lde_commit_dev(d_cols.data_ptr(), n_cols, lg, beta, d_lde.data_ptr(), d_nodes.data_ptr())
    lde = from_dev(d_lde).reshape(n_cols, N)
    for c in (0, n_cols - 1):
        assert np.array_equal(lde[c], O.lde(cols[c], beta))
    assert from_dev(d_nodes)[-8:].view(np.uint32).tobytes() == O.tree_root(O.row_commit(lde))


_TUNE_SCRIPT = r"""
import json, sys, time
sys.path.insert(0, sys.argv[1])
import __graft_entry__
Z = __graft_entry__.load_package()
ctx = Z.Context(0)
t0 = time.perf_counter()
r = ctx.autotune_lde(8, 12, 2)
print(json.dumps({"s": time.perf_counter() - t0, "picks": [list(r["interp"]), list(r["mid"]), list(r["eval"])]}))
ctx.close()
"""


def test_autotune_picks_reused_by_a_second_process(tmp_path):
    """The first process tunes a shape and appends its picks to the autotune file;
    a second process on the same device reads them instead of tuning again."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cache = tmp_path / "autotune.txt"
    env = dict(os.environ, ZS_AUTOTUNE_CACHE=str(cache))
    env.pop("ZS_AUTOTUNE_REFRESH", None)

    def run():
        r = subprocess.run([sys.executable, "-c", _TUNE_SCRIPT, root], env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        return json.loads(r.stdout.strip().splitlines()[-1])

    first = run()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1
    second = run()
    assert cache.read_text().splitlines() == lines  # nothing tuned, nothing appended
    assert second["picks"] == first["picks"]
    assert second["s"] < first["s"]


@pytest.mark.parametrize("n_cols,lg", [(24, 21), (12, 22)])
def test_lde_commit_11_stage_passes(zctx, torch_dev, n_cols, lg):
    """2^21 and 2^22 rows at blowup 4 take 11-stage passes (first interpolation pass,
    fused middle pass, last evaluation pass) with several columns per workgroup:
    every LDE word and tree node equals the reference-shaped plain path's, and the
    first and last columns equal the oracle's LDE."""
    import torch
    beta = 2
    n, N = 1 << lg, 1 << (lg + beta)
    d_cols = torch.empty((n_cols, n), dtype=torch.int32, device=torch_dev)
    zctx.fill_m31_dev(lg, d_cols.data_ptr(), n_cols * n)
    outs = []
    for flags in (0, Z.ZS_FLAG_PLAIN):
        d_lde = torch.zeros((n_cols, N), dtype=torch.int32, device=torch_dev)
        d_nodes = torch.zeros((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
        zctx.lde_commit_dev(d_cols.data_ptr(), n_cols, lg, beta, d_lde.data_ptr(), d_nodes.data_ptr(), flags)
        zctx.synchronize()
        outs.append((d_lde, d_nodes))
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])
    lde_h = from_dev(outs[0][0])
    cols_h = from_dev(d_cols)
    for c in (0, n_cols - 1):
        assert np.array_equal(lde_h[c], O.lde(cols_h[c], beta)), c
