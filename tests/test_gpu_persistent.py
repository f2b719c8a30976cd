"""The persistent batch launch (zs_commit_batch_dev with ZS_FLAG_PERSISTENT): the
LDE, row leaves and node levels of every trace of an uneven batch in one launch,
then each trace's gamma draw, combination and FRI.  Every LDE word, every tree
node and every commitment equals what the per-trace worker path writes (whose
small traces equal the oracle's composition, checked here too), for pipelined
shapes (log rows 16..20), generic-pass shapes (3..15 and 21), rows of more than
256 columns, several blowups, and traces the launch does not take (log rows 1, 2)."""
import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z
from conftest import to_dev

pytestmark = pytest.mark.gpu
P = O.P


def make_batch(torch_dev, shapes, beta, seed):
    import torch
    rng = np.random.default_rng(seed)
    inputs, bufs = [], []
    for c, lg in shapes:
        x = rng.integers(0, P, (c, 1 << lg), dtype=np.uint32)
        N = 1 << (lg + beta)
        d_cols = to_dev(x, torch_dev)
        outs = []
        for _ in range(2):  # one output set per path
            outs.append((torch.empty((c, N), dtype=torch.int32, device=torch_dev),
                         torch.empty((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)))
        inputs.append(x)
        bufs.append((d_cols, outs))
    return inputs, bufs


def specs_of(inputs, bufs, which):
    return [(b[0].data_ptr(), x.shape[0], x.shape[1].bit_length() - 1, b[1][which][0].data_ptr(),
             b[1][which][1].data_ptr()) for x, b in zip(inputs, bufs)]


CASES = {
    # (n_cols, log rows) per trace
    "mixed": ([(8, 16), (3, 17), (33, 18), (5, 19), (2, 20), (16, 12), (7, 3), (9, 5), (12, 8), (4, 11),
               (6, 13), (10, 15), (1, 14), (300, 6), (3, 1), (2, 2)], 2),
    "blowup1": ([(4, 16), (9, 18), (3, 10), (5, 17)], 1),
    "blowup3": ([(4, 16), (6, 19), (2, 9), (3, 13)], 3),
    "large": ([(64, 20), (3, 21), (128, 16), (8, 17)], 2),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("workers", [4, 1])
def test_persistent_batch_equals_worker_path(zctx, torch_dev, case, workers):
    import torch
    shapes, beta = CASES[case]
    if workers == 1 and case == "large":
        pytest.skip("the sequential tails add nothing at this size")
    print(f"persistent batch {case}: beta {beta}, workers {workers}, shapes {shapes}", flush=True)
    rem_log = beta if case == "mixed" else beta + 2  # the 2- and 4-row traces leave room for no more
    inputs, bufs = make_batch(torch_dev, shapes, beta, seed=len(case) * 31 + workers)
    ref = zctx.commit_batch_dev(specs_of(inputs, bufs, 0), beta, rem_log, n_workers=4)
    got = zctx.commit_batch_dev(specs_of(inputs, bufs, 1), beta, rem_log, n_workers=workers,
                                flags=Z.ZS_FLAG_PERSISTENT)
    torch.cuda.synchronize()
    for k, (x, b) in enumerate(zip(inputs, bufs)):
        assert got[k][0] == ref[k][0], (k, shapes[k])
        assert got[k][1] == ref[k][1], (k, shapes[k])
        assert np.array_equal(got[k][2], ref[k][2]), (k, shapes[k])
        assert torch.equal(b[1][0][0], b[1][1][0]), ("LDE", k, shapes[k])
        assert torch.equal(b[1][0][1], b[1][1][1]), ("tree", k, shapes[k])
        if x.shape[1] <= 1 << 8 and x.shape[0] <= 16 and case == "mixed":
            oroot, oroots, orem, _, _ = O.commit_job(x, beta, rem_log)
            assert got[k][0] == oroot and got[k][1] == oroots and np.array_equal(got[k][2], orem)


def test_persistent_batch_repeats_and_reuses_its_pool(zctx, torch_dev):
    """Two calls in a row, the second smaller: the pool and plan slots are reused."""
    import torch
    for shapes in ([(16, 18), (4, 12), (8, 16)], [(3, 16), (2, 7)]):
        print(f"persistent batch repeat: {shapes}", flush=True)
        inputs, bufs = make_batch(torch_dev, shapes, 2, seed=5)
        ref = zctx.commit_batch_dev(specs_of(inputs, bufs, 0), 2, 4, n_workers=2)
        got = zctx.commit_batch_dev(specs_of(inputs, bufs, 1), 2, 4, n_workers=2, flags=Z.ZS_FLAG_PERSISTENT)
        torch.cuda.synchronize()
        for k in range(len(shapes)):
            assert got[k][0] == ref[k][0] and got[k][1] == ref[k][1]
            assert torch.equal(bufs[k][1][0][1], bufs[k][1][1][1])


def test_persistent_barrier_bound_fails_cleanly(zctx, torch_dev, monkeypatch):
    """A grid-barrier wait past its bound (forced here from the host: a bound of 0 ms)
    makes every workgroup leave, and the call returns a device error instead of
    waiting; the next call, with the default bound, gives the worker path's results."""
    import torch
    shapes = [(8, 16), (4, 12), (6, 17)]
    inputs, bufs = make_batch(torch_dev, shapes, 2, seed=11)
    print("persistent batch, barrier bound 0 ms", flush=True)
    monkeypatch.setenv("ZS_PERSIST_WAIT_MS", "0")
    with pytest.raises(Z.ZsError) as e:
        zctx.commit_batch_dev(specs_of(inputs, bufs, 1), 2, 4, n_workers=2, flags=Z.ZS_FLAG_PERSISTENT)
    assert e.value.code == Z.ZS_ERR_DEVICE
    monkeypatch.delenv("ZS_PERSIST_WAIT_MS")
    print("persistent batch, default bound", flush=True)
    ref = zctx.commit_batch_dev(specs_of(inputs, bufs, 0), 2, 4, n_workers=2)
    got = zctx.commit_batch_dev(specs_of(inputs, bufs, 1), 2, 4, n_workers=2, flags=Z.ZS_FLAG_PERSISTENT)
    torch.cuda.synchronize()
    for k in range(len(shapes)):
        assert got[k][0] == ref[k][0] and got[k][1] == ref[k][1] and np.array_equal(got[k][2], ref[k][2])
        assert torch.equal(bufs[k][1][0][1], bufs[k][1][1][1])
