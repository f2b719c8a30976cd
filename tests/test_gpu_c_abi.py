"""A C program (tests/c/abi_smoke.c, built by build() with gcc against
include/zs_hip.h and libzs_hip.so) drives the library the way a Zig or C++ host
would: LDE + row commitment of host columns, a Merkle opening, and an FRI
commit with channel-drawn challenges (src/m31/ntt/circle.zig, src/m31/hash.zig,
src/core/merkle/merkle.zig:57-68, src/m31/fri.zig:157-233).  Every result it
prints must equal the oracle's on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "c", "abi_smoke")
P = 0x7FFFFFFF


def cells(c, n):
    """cell(c, i) of abi_smoke.c for i < n."""
    i = np.arange(n, dtype=np.uint64)
    return ((np.uint64(c) * np.uint64(2654435761) + i * np.uint64(40503) + np.uint64(7)) % np.uint64(P)).astype(
        np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("n_cols,lg,beta,index,fri_log", [(5, 9, 2, 777, 9), (3, 12, 1, 4095, 11)])
def test_c_caller_matches_oracle(n_cols, lg, beta, index, fri_log):
    assert os.path.exists(EXE), "C caller not built: run __graft_entry__.build()"
    print("case", n_cols, lg, beta, index, fri_log, flush=True)
    r = subprocess.run([EXE, str(n_cols), str(lg), str(beta), str(index), str(fri_log)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = dict(line.split(" ", 1) for line in r.stdout.splitlines() if " " in line)
    assert r.stdout.rstrip().endswith("ok")

    cols = np.stack([cells(c, 1 << lg) for c in range(n_cols)])
    lde = np.stack([O.lde(col, beta) for col in cols])
    nodes = O.row_commit(lde)
    assert out["root"] == O.tree_root(nodes).hex()
    assert out["path"] == O.merkle_open(nodes, lde.shape[1], index).hex()
    s = 0
    for v in lde.reshape(-1).tolist():
        s = (s * 31 + v) % 0xFFFFFFFFFFC5
    assert int(out["lde_checksum"]) == s

    cw = cells(1000, 4 << fri_log).reshape(-1, 4)
    ch = O.Channel(b"c-abi")
    roots, _, rem, _ = O.fri_commit(cw, 2, 3, channel=ch, want_layers=False)
    assert out["fri_roots"] == b"".join(roots).hex()
    assert out["remainder"] == rem.astype("<u4").tobytes().hex()
    assert out["channel_after"] == ch.sample_bytes(32).hex()
