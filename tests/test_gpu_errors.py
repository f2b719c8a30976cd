"""GPU entry points reject shapes they cannot run with ZS_ERR_INPUT, before any
launch (zig-capi.h's error convention: a status code, outputs untouched)."""
import numpy as np
import pytest

import zig_stark_amd as Z

pytestmark = pytest.mark.gpu


def _code(fn):
    with pytest.raises(Z.ZsError) as e:
        fn()
    return e.value.code


def test_lde_commit_rejects_bad_shapes(zctx, torch_dev):
    import torch
    buf = torch.zeros(1 << 12, dtype=torch.int32, device=torch_dev)
    out = torch.full((1 << 12,), 7, dtype=torch.int32, device=torch_dev)
    p, q = buf.data_ptr(), out.data_ptr()
    # no columns; a column of 2^0 rows; 2^(log_n + log_blowup) beyond 2^28 rows
    for n_cols, lg, beta in [(0, 4, 2), (1, 0, 2), (1, 27, 2)]:
        assert _code(lambda: zctx.lde_commit_dev(p, n_cols, lg, beta, q, q)) == Z.ZS_ERR_INPUT
        assert _code(lambda: zctx.lde_dev(p, n_cols, lg, beta, q)) == Z.ZS_ERR_INPUT
    zctx.synchronize()
    assert bool((out == 7).all())  # nothing was written


def test_fri_commit_rejects_bad_parameters(zctx, torch_dev):
    import torch
    d_cw = torch.zeros(4 << 10, dtype=torch.int32, device=torch_dev)
    alphas = np.zeros((10, 4), np.uint32)
    # remainder as large as the codeword: no committed layer
    assert _code(lambda: zctx.fri_commit_dev(d_cw.data_ptr(), 10, 2, 10, alphas=alphas)) == Z.ZS_ERR_INPUT
    # neither challenges nor a channel
    assert _code(lambda: zctx.fri_commit_dev(d_cw.data_ptr(), 10, 2, 5)) == Z.ZS_ERR_INPUT


def test_merkle_open_rejects_out_of_range_index(zctx):
    rng = np.random.default_rng(3)
    cols = rng.integers(0, Z.P, (2, 1 << 4), dtype=np.uint32)
    tree, root, _ = zctx.lde_commit(cols, 1)
    assert len(tree.open((1 << 5) - 1)) == 32 * 5
    assert _code(lambda: tree.open(1 << 5)) == Z.ZS_ERR_INPUT


def test_log_sizes_that_wrap_are_rejected(zctx, torch_dev):
    """log_n + log_blowup is checked without the 32-bit wrap: a huge log_blowup
    whose sum with log_n wraps below 28 is still ZS_ERR_INPUT."""
    import torch
    out = torch.full((1 << 12,), 7, dtype=torch.int32, device=torch_dev)
    q = out.data_ptr()
    for lg, beta in [(20, (1 << 32) - 16), ((1 << 32) - 1, 2)]:
        assert _code(lambda: zctx.lde_commit_dev(q, 1, lg, beta, q, q)) == Z.ZS_ERR_INPUT
        assert _code(lambda: zctx.lde_dev(q, 1, lg, beta, q)) == Z.ZS_ERR_INPUT
    zctx.synchronize()
    assert bool((out == 7).all())


def test_batch_rejects_too_many_workers(zctx):
    assert _code(lambda: zctx.commit_batch_dev([], 2, 3, n_workers=65)) == Z.ZS_ERR_INPUT
    assert zctx.commit_batch_dev([], 2, 3, n_workers=4) == []
