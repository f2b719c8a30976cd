"""Regression vectors (tests/golden/regression_vectors.json, written by
tests/golden/make_fixtures.py from the pinned oracle): the oracle must keep
reproducing them on the CPU, and the HIP library through the C ABI on the GPU."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

V = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "regression_vectors.json")))


def fri_inputs():
    f = V["fri_2p10"]
    vals = O.fill_m31(f["seed"], (1 << f["dlog"]) * 4 + (f["dlog"] - f["remainder_log"]) * 4)
    cw = vals[: (1 << f["dlog"]) * 4].reshape(-1, 4)
    return cw, vals[(1 << f["dlog"]) * 4:].reshape(-1, 4)


def test_oracle_reproduces_vectors():
    v = V["lde_2p4"]
    col = O.fill_m31(v["seed"], 1 << v["log_n"])
    assert col.tolist() == v["input"]
    assert O.lde(col, v["log_blowup"]).tolist() == v["lde"]
    r = V["row_tree_64x2p6"]
    cols = O.fill_m31(r["seed"], r["n_cols"] << r["log_n"]).reshape(r["n_cols"], -1)
    nodes = O.row_commit(np.stack([O.lde(c, r["log_blowup"]) for c in cols]))
    assert O.tree_root(nodes).hex() == r["root"]
    n_rows = 1 << (r["log_n"] + r["log_blowup"])
    for i, p in r["paths"].items():
        assert O.merkle_open(nodes, n_rows, int(i)).hex() == p
    f = V["fri_2p10"]
    cw, alphas = fri_inputs()
    roots, _, rem, _ = O.fri_commit(cw, f["log_blowup"], f["remainder_log"], alphas=alphas, want_layers=False)
    assert [x.hex() for x in roots] == f["roots"] and rem.tolist() == f["remainder"]
    t = V["fib_trace_log4"]
    troots, _ = O.trace_commit(O.fib_trace(t["a0"], t["a1"], 1 << t["trace_log"]), t["log_blowup"],
                               want_codewords=False)
    assert [x.hex() for x in troots] == t["roots"]


@pytest.mark.gpu
def test_library_reproduces_vectors(zctx, torch_dev):
    import torch
    import zig_stark_amd as Z
    from conftest import from_dev, to_dev
    v = V["lde_2p4"]
    d_col = to_dev(np.array(v["input"], np.uint32), torch_dev)
    d_lde = torch.empty(16 << 2, dtype=torch.int32, device=torch_dev)
    zctx.lde_dev(d_col.data_ptr(), 1, v["log_n"], v["log_blowup"], d_lde.data_ptr())
    assert from_dev(d_lde).tolist() == v["lde"]
    r = V["row_tree_64x2p6"]
    cols = O.fill_m31(r["seed"], r["n_cols"] << r["log_n"]).reshape(r["n_cols"], -1)
    tree, root, _ = zctx.lde_commit(cols, r["log_blowup"])
    assert root.hex() == r["root"]
    for i, p in r["paths"].items():
        assert tree.open(int(i)).hex() == p
    f = V["fri_2p10"]
    cw, alphas = fri_inputs()
    roots, rem, _ = zctx.fri_commit(cw, f["log_blowup"], f["remainder_log"], alphas=alphas)
    assert [x.hex() for x in roots] == f["roots"] and rem.tolist() == f["remainder"]
    t = V["fib_trace_log4"]
    trace = Z.fib_trace(t["a0"], t["a1"], t["trace_log"])
    N = 1 << (t["trace_log"] + t["log_blowup"])
    d_tr = to_dev(trace.reshape(-1), torch_dev)
    d_cw = torch.empty(2 * N * 4, dtype=torch.int32, device=torch_dev)
    d_nodes = torch.empty(2 * (2 * N - 1) * 8, dtype=torch.int32, device=torch_dev)
    troots = zctx.trace_commit_dev(d_tr.data_ptr(), 2, t["trace_log"], t["log_blowup"], d_cw.data_ptr(),
                                   d_nodes.data_ptr())
    assert [x.hex() for x in troots] == t["roots"]
