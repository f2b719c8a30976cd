"""zs_m31_lde_commit_dev_noalloc (no reference counterpart: the LDE + row-leaf
commitment for callers that cannot allocate device memory) against
zs_m31_lde_commit_dev on the same inputs: every LDE value and every tree level
equal, on consecutive calls whose second input is smaller than the first (the
fixed array keeps the first call's contents), with no change in free device
memory across the calls, and shapes beyond the fixed array rejected."""
import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu


def _buffers(torch_dev, cols, beta):
    import torch
    n_cols, n = cols.shape
    N = n << beta
    return (to_dev(cols, torch_dev), torch.empty((n_cols, N), dtype=torch.int32, device=torch_dev),
            torch.empty((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev))


# the first shape needs the two-pass forward transform's intermediate buffer (log_n > 11),
# the second is smaller, the third runs in one pass
SHAPES = [(8, 20, 2), (16, 12, 2), (5, 9, 3)]


def test_noalloc_equals_lde_commit_dev(torch_dev):
    import torch
    ctx = Z.Context(0)
    try:
        cases = []
        for k, (c, lg, beta) in enumerate(SHAPES):
            cols = O.fill_m31(900 + k, c << lg).reshape(c, 1 << lg)
            d_cols, d_lde, d_nodes = _buffers(torch_dev, cols, beta)
            ctx.lde_commit_dev(d_cols.data_ptr(), c, lg, beta, d_lde.data_ptr(), d_nodes.data_ptr())
            ctx.synchronize()
            want = (from_dev(d_lde).copy(), from_dev(d_nodes).copy())
            d_lde.fill_(-1)
            d_nodes.fill_(-1)
            cases.append((c, lg, beta, cols, d_cols, d_lde, d_nodes, want))
        torch.cuda.synchronize()
        free_before, _ = torch.cuda.mem_get_info()
        for c, lg, beta, _, d_cols, d_lde, d_nodes, _ in cases:
            print(f"\nnoalloc: {c} x 2^{lg}, blowup 2^{beta}", flush=True)
            ctx.lde_commit_dev_noalloc(d_cols.data_ptr(), c, lg, beta, d_lde.data_ptr(), d_nodes.data_ptr())
        ctx.synchronize()
        free_after, _ = torch.cuda.mem_get_info()
        assert free_after == free_before
        for c, lg, beta, cols, _, d_lde, d_nodes, want in cases:
            assert np.array_equal(from_dev(d_lde), want[0]), (c, lg, beta)
            assert np.array_equal(from_dev(d_nodes), want[1]), (c, lg, beta)
        # the smallest case also against the oracle
        c, lg, beta, cols, _, d_lde, d_nodes, _ = cases[2]
        lde = np.stack([O.lde(x, beta) for x in cols])
        assert np.array_equal(from_dev(d_lde).reshape(c, -1), lde)
        assert from_dev(d_nodes)[-8:].tobytes() == O.tree_root(O.row_commit(lde))
    finally:
        ctx.close()


def test_noalloc_rejects_shapes_beyond_the_array(torch_dev):
    import torch
    ctx = Z.Context(0)
    try:
        d = torch.empty(8, dtype=torch.int32, device=torch_dev)
        p = d.data_ptr()
        for c, lg, beta in [(16, 20, 2), (1, 21, 2), (1, 20, 3)]:  # 320 MiB of buffers; 2^23-point domains
            with pytest.raises(Z.ZsError) as e:
                ctx.lde_commit_dev_noalloc(p, c, lg, beta, p, p)
            assert e.value.code == Z.ZS_ERR_INPUT
            assert "allocation-free" in str(e.value)  # the clear reason, from zs_hip_last_error
    finally:
        ctx.close()
