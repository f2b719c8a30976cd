"""The batch and serving workloads at their full sizes, through the HIP path:

* the uneven batch: 256 traces, log rows U{12..20}, columns U{8..128} (seed 4,
  trace k filled from seed (4 << 32) + k), blowup 4, FRI to 2^5.  Every trace's
  commitment is the same batched on worker streams, committed alone, and on the
  reference-shaped plain loop; every trace of at most 2^13 rows equals the
  oracle composition (src/m31/stark.zig:396-414 trace trees, src/m31/fri.zig:
  157-233 FRI).
* the serving loop: Poisson arrivals (mean 200/s) of requests with log rows
  U{14..22} and columns U{16..256} under a 200 GB device budget.  Each response
  equals the same request served alone, the two smallest equal the oracle, and
  the admitted memory never exceeds the budget.
* admission: a large request queued behind a stream of small ones is passed at
  most a bounded number of times; a ticket is answered to one caller only."""
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z

pytestmark = pytest.mark.gpu
BETA, REM = 2, 5
MAX_BYPASS = 8  # job.cpp kMaxBypass


def _pool():
    return ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1)))


def _same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2])


def test_uneven_batch_full_workload(torch_dev):
    import torch
    rng = np.random.default_rng(4)
    n_traces = 256
    lgs = rng.integers(12, 21, n_traces)
    cols = rng.integers(8, 129, n_traces)
    ctx = Z.Context(0)  # its own workers and scratch, released at the end
    try:
        bufs, specs = [], []
        for k in range(n_traces):
            c, lg = int(cols[k]), int(lgs[k])
            n, N = 1 << lg, 1 << (lg + BETA)
            d_c = torch.empty((c, n), dtype=torch.int32, device=torch_dev)
            d_l = torch.empty((c, N), dtype=torch.int32, device=torch_dev)
            d_t = torch.empty((2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
            ctx.fill_m31_dev((4 << 32) + k, d_c.data_ptr(), c * n)
            bufs.append((d_c, d_l, d_t))
            specs.append((d_c.data_ptr(), c, lg, d_l.data_ptr(), d_t.data_ptr()))
        torch.cuda.synchronize()
        print(f"\nbatch: {n_traces} traces, {sum(int(c) << int(l) for c, l in zip(cols, lgs)) / 1e9:.2f} Gelem",
              flush=True)
        batched = ctx.commit_batch_dev(specs, BETA, REM, n_workers=8)
        alone = ctx.commit_batch_dev(specs, BETA, REM, n_workers=1)
        plain = ctx.commit_batch_dev(specs, BETA, REM, n_workers=1, flags=Z.ZS_FLAG_PLAIN)
        for k in range(n_traces):
            assert _same(batched[k], alone[k]), k
            assert _same(plain[k], alone[k]), k
        small = [k for k in range(n_traces) if lgs[k] <= 13]
        assert len(small) >= 40

        def oracle(k):
            c, lg = int(cols[k]), int(lgs[k])
            x = O.fill_m31((4 << 32) + k, c << lg).reshape(c, 1 << lg)
            root, roots, rem, _, _ = O.commit_job(x, BETA, REM)
            return k, (root, roots, rem)

        with _pool() as ex:
            for k, want in ex.map(oracle, small):
                assert _same(alone[k], want), k
        print(f"batch: {len(small)} traces of <= 2^13 rows equal the oracle", flush=True)
    finally:
        ctx.close()
        del bufs, specs
        torch.cuda.empty_cache()


def test_serving_full_workload(torch_dev):
    import torch
    torch.cuda.empty_cache()
    budget = 200 * 10**9
    n = 64
    rng = np.random.default_rng(5)
    gaps = rng.exponential(1 / 200.0, n)
    lgs = rng.integers(14, 23, n)
    cols = rng.integers(16, 257, n)
    seeds = [(5 << 32) + k for k in range(n)]
    srv = Z.Server(0, n_workers=8, hbm_budget=budget)
    try:
        t0 = time.perf_counter()
        at = np.cumsum(gaps)
        tickets = []
        for k in range(n):
            d = t0 + at[k] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit(int(cols[k]), int(lgs[k]), BETA, REM, seed=seeds[k]))
        mixed = [srv.wait(t) for t in tickets]
        _, peak = srv.memory()
        assert peak <= budget
        assert all(r.status == 0 for r in mixed)
        overlapped = sum(1 for a, b in zip(mixed, mixed[1:]) if b.t_admit < a.t_done)
        print(f"\nserve: {n} requests, peak {peak / 1e9:.1f} GB admitted, {overlapped} admitted while the "
              f"previous one ran", flush=True)
        assert overlapped > 0  # requests did share the device
        for k in range(n):
            alone = srv.wait(srv.submit(int(cols[k]), int(lgs[k]), BETA, REM, seed=seeds[k]))
            assert alone.status == 0
            assert _same(alone.result(BETA, REM), mixed[k].result(BETA, REM)), k
        smallest = sorted(range(n), key=lambda k: int(cols[k]) << int(lgs[k]))[:2]

        def oracle(k):
            c, lg = int(cols[k]), int(lgs[k])
            x = O.fill_m31(seeds[k], c << lg).reshape(c, 1 << lg)
            root, roots, rem, _, _ = O.commit_job(x, BETA, REM)
            return k, (root, roots, rem)

        with _pool() as ex:
            for k, want in ex.map(oracle, smallest):
                assert _same(mixed[k].result(BETA, REM), want), k
    finally:
        srv.close()


def _footprint(c, lg):
    probe = Z.Server(0, n_workers=1, hbm_budget=64 * 10**9)
    try:
        return probe.wait(probe.submit(c, lg, BETA, REM, seed=1)).device_bytes
    finally:
        probe.close()


def test_large_request_not_starved_by_small_ones():
    """The large request fits only while no small one runs; two workers keep small
    requests running.  Small requests queued behind it pass it at most
    MAX_BYPASS times, after which it is admitted before the rest."""
    small_c, small_lg, large_c, large_lg = 16, 14, 64, 18
    s_bytes, l_bytes = _footprint(small_c, small_lg), _footprint(large_c, large_lg)
    budget = l_bytes + s_bytes - 4096  # the large one needs every small one to be done
    assert 2 * s_bytes <= budget
    srv = Z.Server(0, n_workers=2, hbm_budget=budget)
    try:
        before = [srv.submit(small_c, small_lg, BETA, REM, seed=10 + k) for k in range(2)]
        big = srv.submit(large_c, large_lg, BETA, REM, seed=99)
        after = [srv.submit(small_c, small_lg, BETA, REM, seed=100 + k) for k in range(40)]
        for t in before:
            assert srv.wait(t).status == 0
        rb = srv.wait(big)
        ra = [srv.wait(t) for t in after]
        assert rb.status == 0 and all(r.status == 0 for r in ra)
        passed = sum(1 for r in ra if r.t_admit < rb.t_admit)
        print(f"\n{passed} of {len(ra)} later small requests passed the large one", flush=True)
        assert passed <= MAX_BYPASS
        assert any(r.t_admit > rb.t_admit for r in ra)
        _, peak = srv.memory()
        assert peak <= budget
    finally:
        srv.close()


def test_ticket_answered_once():
    srv = Z.Server(0, n_workers=1, hbm_budget=8 * 10**9)
    try:
        t = srv.submit(64, 16, BETA, REM, seed=3)
        out = {}

        def waiter(name):
            try:
                out[name] = srv.wait(t).status
            except Z.ZsError as e:
                out[name] = e.code

        th = [threading.Thread(target=waiter, args=(k,)) for k in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert sorted(out.values()) == [Z.ZS_ERR_INPUT, 0]
        with pytest.raises(Z.ZsError) as e:  # answered and released
            srv.wait(t)
        assert e.value.code == Z.ZS_ERR_INPUT
    finally:
        srv.close()
