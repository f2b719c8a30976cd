"""The full-size workloads compared whole with the oracle, through the HIP path:
the committed fixture tests/golden/full_size.json holds the oracle's results
(written by tests/golden/make_full_size.py; the inputs are the seeded streams
that zs_fill_m31_dev reproduces on the device).

* 64 columns x 2^20 -> blowup 4 -> 2^22-row tree: the root (a commitment to every
  LDE value), five authentication paths, and a BLAKE3 digest of each LDE column;
* the whole commitment of that trace (gamma, QM31 combination, FRI to 2^5 with
  channel challenges): every FRI root and the remainder;
* FRI on a 2^22 QM31 codeword to 2^5, challenges given and drawn from the channel
  (fused draws on the device, HIP graph and launches): all 17 roots, the
  remainder and the channel state afterwards (src/m31/fri.zig:157-233)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z
from conftest import from_dev

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "full_size.json")))


@pytest.fixture(scope="module")
def trace_64x2p20(zctx, torch_dev):
    import torch
    f = FIX["lde_commit_64x2p20"]
    n_cols, lg, beta = f["n_cols"], f["log_n"], f["log_blowup"]
    n, N = 1 << lg, 1 << (lg + beta)
    d_cols = torch.empty((n_cols, n), dtype=torch.int32, device=torch_dev)
    zctx.fill_m31_dev(f["seed"], d_cols.data_ptr(), n_cols * n)
    d_lde = torch.empty((n_cols, N), dtype=torch.int32, device=torch_dev)
    d_nodes = torch.empty((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
    yield d_cols, d_lde, d_nodes
    del d_cols, d_lde, d_nodes
    torch.cuda.empty_cache()


def test_lde_commit_64x2p20_equals_oracle(zctx, trace_64x2p20):
    f = FIX["lde_commit_64x2p20"]
    d_cols, d_lde, d_nodes = trace_64x2p20
    n_cols, lg, beta = f["n_cols"], f["log_n"], f["log_blowup"]
    N = 1 << (lg + beta)
    for flags in (0, Z.ZS_FLAG_PLAIN):
        d_nodes.fill_(0)
        zctx.lde_commit_dev(d_cols.data_ptr(), n_cols, lg, beta, d_lde.data_ptr(), d_nodes.data_ptr(), flags)
        zctx.synchronize()
        nodes = from_dev(d_nodes).tobytes()
        assert nodes[-32:].hex() == f["root"], flags
        for i, path in f["paths"].items():
            assert O.merkle_open(nodes, N, int(i)).hex() == path, (flags, i)
    lde = from_dev(d_lde)
    for c in range(n_cols):
        assert O.blake3(np.ascontiguousarray(lde[c]).tobytes()).hex() == f["column_blake3"][c], c


def test_commit_job_64x2p20_equals_oracle(zctx, trace_64x2p20):
    f, g = FIX["lde_commit_64x2p20"], FIX["commit_job_64x2p20"]
    d_cols, d_lde, d_nodes = trace_64x2p20
    root, roots, rem = zctx.commit_job_dev(d_cols.data_ptr(), f["n_cols"], f["log_n"], f["log_blowup"],
                                           g["remainder_log"], d_lde.data_ptr(), d_nodes.data_ptr())
    assert root.hex() == g["trace_root"]
    assert [r.hex() for r in roots] == g["fri_roots"]
    assert np.array_equal(rem, np.array(g["remainder"], np.uint32))


@pytest.fixture(scope="module")
def codeword_2p22(zctx, torch_dev):
    import torch
    f = FIX["fri_2p22_seeded"]
    d_cw = torch.empty((1 << f["dlog"]) * 4, dtype=torch.int32, device=torch_dev)
    zctx.fill_m31_dev(f["codeword_seed"], d_cw.data_ptr(), (1 << f["dlog"]) * 4)
    yield d_cw
    del d_cw
    torch.cuda.empty_cache()


@pytest.mark.parametrize("flags", [Z.ZS_FLAG_GRAPH, 0])
def test_fri_2p22_seeded_equals_oracle(zctx, codeword_2p22, flags):
    f = FIX["fri_2p22_seeded"]
    L = f["dlog"] - f["remainder_log"]
    alphas = O.fill_m31(f["alpha_seed"], L * 4).reshape(L, 4)
    for _ in range(2):  # with ZS_FLAG_GRAPH the second call replays the captured graph
        roots, rem, _ = zctx.fri_commit_dev(codeword_2p22.data_ptr(), f["dlog"], f["log_blowup"],
                                            f["remainder_log"], alphas=alphas, flags=flags)
        assert [r.hex() for r in roots] == f["roots"]
        assert np.array_equal(rem, np.array(f["remainder"], np.uint32))


@pytest.mark.parametrize("flags", [Z.ZS_FLAG_GRAPH, 0])
def test_fri_2p22_channel_equals_oracle(zctx, codeword_2p22, flags):
    f = FIX["fri_2p22_channel"]
    for _ in range(2):
        ch = Z.Channel(f["channel"].encode())
        roots, rem, _ = zctx.fri_commit_dev(codeword_2p22.data_ptr(), f["dlog"], f["log_blowup"],
                                            f["remainder_log"], channel=ch, flags=flags)
        assert [r.hex() for r in roots] == f["roots"]
        assert np.array_equal(rem, np.array(f["remainder"], np.uint32))
        assert ch.sample_bytes(32).hex() == f["next_sample_32"]
