"""Writes tests/golden/full_size.json: the oracle's results for the full-size
workloads, which the GPU tests compare whole instead of sampling.

  * 64 columns x 2^20 (the seed-2 stream, column c = elements [c 2^20, (c+1) 2^20)),
    LDE with blowup 4 to 2^22 rows, row-leaf Merkle tree (src/m31/stark.zig:396-414,
    src/m31/hash.zig:15-25, src/core/merkle/merkle.zig:17-45): the root, a few
    authentication paths, and per column a BLAKE3 digest of its LDE values;
  * the whole commitment of the same trace (channel "zig-stark:hip:commit",
    gamma from the trace root, QM31 combination, FRI to 2^5 with channel
    challenges): the FRI roots and the remainder;
  * one QM31 codeword of 2^22 (the seed-3 stream, 4 words per value), FRI commit to
    2^5 (src/m31/fri.zig:157-233) with challenges given (the seed-33 stream, 17
    values) and with challenges drawn from Channel("zig-stark:fri:2p22"): every
    layer root, the remainder, and the channel's next 32 sampled bytes.

The oracle is the C restatement under oracle/ (pinned by tests/test_oracle_*.py);
the LDE columns are computed on parallel threads.  Run from the repo root after
building the oracle:  python tests/golden/make_full_size.py   (about a minute on
8 cores, 2-3 GiB of memory).
"""
import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import oracle_lib as O  # noqa: E402

N_COLS, LOG_N, BETA, REM = 64, 20, 2, 5
PATH_INDICES = (0, 1, 12345, 1 << 21, (1 << 22) - 1)


def lde_all(cols):
    with ThreadPoolExecutor(os.cpu_count() or 1) as ex:
        return np.stack(list(ex.map(lambda c: O.lde(c, BETA), cols)))


def main():
    out = {}
    n, N = 1 << LOG_N, 1 << (LOG_N + BETA)
    cols = O.fill_m31(2, N_COLS * n).reshape(N_COLS, n)
    lde = lde_all(cols)
    nodes = O.row_commit(lde)
    root = O.tree_root(nodes)
    out["lde_commit_64x2p20"] = {
        "seed": 2, "n_cols": N_COLS, "log_n": LOG_N, "log_blowup": BETA, "root": root.hex(),
        "paths": {str(i): O.merkle_open(nodes, N, i).hex() for i in PATH_INDICES},
        "column_blake3": [O.blake3(np.ascontiguousarray(lde[c]).tobytes()).hex() for c in range(N_COLS)]}
    # the whole commitment of the same trace, composed as oracle_lib.commit_job does
    ch = O.Channel(b"zig-stark:hip:commit")
    ch.absorb(root)
    gamma = ch.sample_qm31()
    del nodes
    g = O.qm31_combine(lde, gamma)
    del lde
    roots, _, rem, _ = O.fri_commit(g, BETA, REM, channel=ch, want_layers=False)
    out["commit_job_64x2p20"] = {"seed": 2, "remainder_log": REM, "trace_root": root.hex(),
                                 "fri_roots": [r.hex() for r in roots], "remainder": rem.tolist()}
    del g
    # FRI on a 2^22 codeword, challenges given and drawn
    dlog = 22
    L = dlog - REM
    cw = O.fill_m31(3, (1 << dlog) * 4).reshape(-1, 4)
    alphas = O.fill_m31(33, L * 4).reshape(L, 4)
    roots, _, rem, _ = O.fri_commit(cw, BETA, REM, alphas=alphas, want_layers=False)
    out["fri_2p22_seeded"] = {"codeword_seed": 3, "alpha_seed": 33, "dlog": dlog, "log_blowup": BETA,
                              "remainder_log": REM, "roots": [r.hex() for r in roots], "remainder": rem.tolist()}
    label = b"zig-stark:fri:2p22"
    ch = O.Channel(label)
    roots, _, rem, _ = O.fri_commit(cw, BETA, REM, channel=ch, want_layers=False)
    out["fri_2p22_channel"] = {"codeword_seed": 3, "channel": label.decode(), "dlog": dlog, "log_blowup": BETA,
                               "remainder_log": REM, "roots": [r.hex() for r in roots], "remainder": rem.tolist(),
                               "next_sample_32": ch.sample_bytes(32).hex()}
    with open(os.path.join(HERE, "full_size.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
