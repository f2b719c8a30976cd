"""GPU parity for the full trace commitment (LDE -> row tree -> channel gamma ->
QM31 combination -> FRI commit), batches of uneven traces, and the serving loop:
every result equals the oracle composition, and a job's result is the same
alone, in a batch, and in a mix of concurrent requests."""
import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu
P = O.P


def dev_job(torch_dev, cols, beta):
    import torch
    n_cols, n = cols.shape
    N = n << beta
    d_cols = to_dev(cols, torch_dev)
    d_lde = torch.empty((n_cols, N), dtype=torch.int32, device=torch_dev)
    d_nodes = torch.empty((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
    return d_cols, d_lde, d_nodes


@pytest.mark.parametrize("n_cols,lg", [(3, 4), (8, 6), (16, 9), (20, 11), (5, 12), (8, 16)])
def test_commit_job_matches_oracle(zctx, torch_dev, n_cols, lg):
    rng = np.random.default_rng(n_cols * 100 + lg)
    cols = rng.integers(0, P, (n_cols, 1 << lg), dtype=np.uint32)
    d_cols, d_lde, d_nodes = dev_job(torch_dev, cols, 2)
    root, roots, rem = zctx.commit_job_dev(d_cols.data_ptr(), n_cols, lg, 2, 5, d_lde.data_ptr(), d_nodes.data_ptr())
    oroot, oroots, orem, olde, onodes = O.commit_job(cols, 2, 5)
    assert root == oroot
    assert roots == oroots
    assert np.array_equal(rem, orem)
    assert np.array_equal(from_dev(d_lde), olde)
    # the reference-shaped plain path gives the same commitment
    root2, roots2, rem2 = zctx.commit_job_dev(d_cols.data_ptr(), n_cols, lg, 2, 5, d_lde.data_ptr(),
                                              d_nodes.data_ptr(), Z.ZS_FLAG_PLAIN)
    assert (root2, roots2) == (root, roots) and np.array_equal(rem2, rem)


@pytest.mark.parametrize("n_cols,lg,beta", [(96, 12, 2), (130, 13, 2), (97, 12, 3), (128, 14, 1), (256, 12, 2)])
def test_commit_job_combination_before_blowup(zctx, torch_dev, n_cols, lg, beta):
    """From 96 columns on the QM31 combination is formed over the input rows and blown up
    as 4 M31 columns (the LDE is linear); the commitment equals the oracle's, which
    combines the LDE columns, and the plain path's, which combines after the blowup too."""
    rng = np.random.default_rng(n_cols * 1000 + lg * 10 + beta)
    cols = rng.integers(0, P, (n_cols, 1 << lg), dtype=np.uint32)
    d_cols, d_lde, d_nodes = dev_job(torch_dev, cols, beta)
    rem_log = beta + 3
    got = zctx.commit_job_dev(d_cols.data_ptr(), n_cols, lg, beta, rem_log, d_lde.data_ptr(), d_nodes.data_ptr())
    plain = zctx.commit_job_dev(d_cols.data_ptr(), n_cols, lg, beta, rem_log, d_lde.data_ptr(), d_nodes.data_ptr(),
                                Z.ZS_FLAG_PLAIN)
    oroot, oroots, orem, _, _ = O.commit_job(cols, beta, rem_log)
    for root, roots, rem in (got, plain):
        assert root == oroot
        assert roots == oroots
        assert np.array_equal(rem, orem)


def test_uneven_batch_equals_single_jobs(zctx, torch_dev):
    rng = np.random.default_rng(7)
    shapes = [(int(rng.integers(8, 40)), int(rng.integers(6, 13))) for _ in range(14)]
    inputs = [rng.integers(0, P, (c, 1 << lg), dtype=np.uint32) for c, lg in shapes]
    bufs = [dev_job(torch_dev, x, 2) for x in inputs]
    specs = [(b[0].data_ptr(), x.shape[0], x.shape[1].bit_length() - 1, b[1].data_ptr(), b[2].data_ptr())
             for x, b in zip(inputs, bufs)]
    batched = zctx.commit_batch_dev(specs, 2, 5, n_workers=4)
    plain = zctx.commit_batch_dev(specs, 2, 5, n_workers=1, flags=Z.ZS_FLAG_PLAIN)
    for k, x in enumerate(inputs):
        alone = zctx.commit_job_dev(*specs[k][:3], 2, 5, specs[k][3], specs[k][4])
        for got in (batched[k], plain[k]):
            assert got[0] == alone[0] and got[1] == alone[1] and np.array_equal(got[2], alone[2])
        if x.shape[1] <= 1 << 9:
            oroot, oroots, orem, _, _ = O.commit_job(x, 2, 5)
            assert alone[0] == oroot and alone[1] == oroots and np.array_equal(alone[2], orem)


def test_batch_with_a_rejected_job_keeps_the_others(zctx, torch_dev):
    """One job of a worker-stream batch has a shape commit jobs reject (no columns):
    the call reports ZS_ERR_INPUT and that job's status says so, while every other job
    (their FRI tails deferred to the launch after the workers) still gets its full
    result, equal to the same job alone."""
    import ctypes
    rng = np.random.default_rng(11)
    shapes = [(12, 11), (9, 8), (20, 12), (7, 13), (16, 10)]
    inputs = [rng.integers(0, P, (c, 1 << lg), dtype=np.uint32) for c, lg in shapes]
    bufs = [dev_job(torch_dev, x, 2) for x in inputs]
    bad = 2
    jobs = (Z.Job * len(inputs))()
    keep = []
    for k, (x, b) in enumerate(zip(inputs, bufs)):
        lg = x.shape[1].bit_length() - 1
        L = lg + 2 - 5
        fr = ctypes.create_string_buffer(32 * L)
        rem = np.zeros(4 << (5 - 2), np.uint32)
        keep.append((fr, rem, L))
        jobs[k].d_cols, jobs[k].n_cols, jobs[k].log_n = b[0].data_ptr(), 0 if k == bad else x.shape[0], lg
        jobs[k].d_lde, jobs[k].d_nodes = b[1].data_ptr(), b[2].data_ptr()
        jobs[k].fri_roots = ctypes.cast(fr, ctypes.c_void_p)
        jobs[k].remainder = rem.ctypes.data
    rc = Z.lib().zs_commit_batch_dev(zctx.h, jobs, len(inputs), 2, 5, 4, 0)
    assert rc == Z.ZS_ERR_INPUT
    assert jobs[bad].status == Z.ZS_ERR_INPUT
    for k, (x, b) in enumerate(zip(inputs, bufs)):
        if k == bad:
            continue
        assert jobs[k].status == 0, k
        alone = zctx.commit_job_dev(b[0].data_ptr(), x.shape[0], x.shape[1].bit_length() - 1, 2, 5,
                                    b[1].data_ptr(), b[2].data_ptr())
        fr, rem, L = keep[k]
        assert bytes(jobs[k].trace_root) == alone[0], k
        assert [fr.raw[32 * i:32 * (i + 1)] for i in range(L)] == alone[1], k
        assert np.array_equal(rem.reshape(-1, 4), alone[2]), k


def test_server_results_same_alone_and_mixed(zctx, torch_dev):
    """Requests submitted together (overlapping on the device under a memory budget)
    give the results each one gives alone."""
    rng = np.random.default_rng(11)
    reqs = [(int(rng.integers(16, 64)), int(rng.integers(8, 15)), 1000 + k) for k in range(16)]
    budget = 2 * 10**9
    srv = Z.Server(0, n_workers=4, hbm_budget=budget)
    tickets = [srv.submit(c, lg, 2, 5, seed=s) for c, lg, s in reqs]
    mixed = [srv.wait(t) for t in tickets]
    _, peak = srv.memory()
    assert peak <= budget
    for (c, lg, s), resp in zip(reqs, mixed):
        assert resp.status == 0
        assert resp.t_submit <= resp.t_admit <= resp.t_done
        alone_t = srv.submit(c, lg, 2, 5, seed=s)
        alone = srv.wait(alone_t)
        assert alone.result(2, 5)[0] == resp.result(2, 5)[0]
        assert alone.result(2, 5)[1] == resp.result(2, 5)[1]
        assert np.array_equal(alone.result(2, 5)[2], resp.result(2, 5)[2])
    # one request against the oracle, with the seeded input regenerated on the host
    c, lg, s = reqs[0]
    cols = O.fill_m31(s, c << lg).reshape(c, 1 << lg)
    oroot, oroots, orem, _, _ = O.commit_job(cols, 2, 5)
    got = mixed[0].result(2, 5)
    assert got[0] == oroot and got[1] == oroots and np.array_equal(got[2], orem)
    # host-buffer request path, answered through poll
    t = srv.submit(c, lg, 2, 5, host_cols=cols)
    import time
    for _ in range(6000):
        resp = srv.poll(t)
        if resp is not None:
            break
        time.sleep(0.005)
    assert resp is not None and resp.result(2, 5)[0] == oroot
    in_use, _ = srv.memory()
    assert in_use == 0
    srv.close()


def test_server_rejects_request_over_budget():
    srv = Z.Server(0, n_workers=1, hbm_budget=10**6)
    with pytest.raises(Z.ZsError) as e:
        srv.submit(64, 16, 2, 5, seed=1)
    assert e.value.code == Z.ZS_ERR_OOM
    srv.close()
