"""Writes tests/golden/regression_vectors.json from the oracle (oracle/liboracle.so).

Regression vectors, not independent oracles: the oracle is pinned by the tests
in tests/test_oracle_*.py first, and these freeze its outputs for small cases
so later changes to the oracle or the library are caught.  Run from the repo
root:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import oracle_lib as O  # noqa: E402


def main():
    out = {}
    # LDE of one 2^4 column, blowup 4 (seed 11)
    col = O.fill_m31(11, 16)
    out["lde_2p4"] = {"seed": 11, "log_n": 4, "log_blowup": 2, "input": col.tolist(),
                      "lde": O.lde(col, 2).tolist()}
    # row tree of the LDE of 64 columns x 2^6 (blowup 4 -> 2^8 rows), seed 12: root and three paths
    cols = O.fill_m31(12, 64 * 64).reshape(64, 64)
    lde = np.stack([O.lde(c, 2) for c in cols])
    nodes = O.row_commit(lde)
    paths = {str(i): O.merkle_open(nodes, 256, i).hex() for i in (0, 77, 255)}
    out["row_tree_64x2p6"] = {"seed": 12, "n_cols": 64, "log_n": 6, "log_blowup": 2,
                              "root": O.tree_root(nodes).hex(), "paths": paths}
    # FRI commit of a 2^10 QM31 codeword with fixed challenges (seed 13: codeword, then alphas)
    vals = O.fill_m31(13, (1 << 10) * 4 + 5 * 4)
    cw = vals[: (1 << 10) * 4].reshape(-1, 4)
    alphas = vals[(1 << 10) * 4:].reshape(5, 4)
    roots, _, rem, _ = O.fri_commit(cw, 2, 5, alphas=alphas, want_layers=False)
    out["fri_2p10"] = {"seed": 13, "dlog": 10, "log_blowup": 2, "remainder_log": 5,
                       "alphas_follow_codeword": True, "roots": [r.hex() for r in roots],
                       "remainder": rem.tolist()}
    # seeded Fibonacci AIR trace, trace_log 4, blowup 4: column roots (a0 = 1, a1 = 5)
    trace = O.fib_trace(1, 5, 16)
    troots, _ = O.trace_commit(trace, 2, want_codewords=False)
    out["fib_trace_log4"] = {"a0": 1, "a1": 5, "trace_log": 4, "log_blowup": 2,
                             "roots": [r.hex() for r in troots]}
    with open(os.path.join(HERE, "regression_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
