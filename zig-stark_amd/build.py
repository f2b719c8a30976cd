"""Builds libzs_hip.so (the C-ABI library of include/zs_hip.h) for gfx950.

Every source in csrc/ is compiled with hipcc for --offload-arch=gfx950 and
linked into zig-stark_amd/libzs_hip.so.  Run directly or through
__graft_entry__.build().
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "libzs_hip.so")
OBJ = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["fft.hip", "fft_tile.hip", "merkle.hip", "fri.hip", "qntt.hip", "stark.hip", "capi.cpp", "channel.cpp", "job.cpp", "verify.cpp", "noalloc.hip", "air.hip", "persist.hip"]
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build (device code is
# never instrumented: each -fsanitize= applies to the host pass only).  It goes to
# _build_san/libzs_hip_san.so and is loaded through ZS_HIP_LIBRARY with the clang
# ASan runtime preloaded (tests/run_sanitized.sh).
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
             "-shared-libasan", "-g"]
SAN_OBJ = os.path.join(PKG, "_build_san")
SAN_OUT = os.path.join(SAN_OBJ, "libzs_hip_san.so")


def _compile(src, sanitize=False):
    obj = os.path.join(SAN_OBJ if sanitize else OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *(SAN_FLAGS if sanitize else []), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    warn = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    return obj, warn


def build(verbose=False, sanitize=False):
    os.makedirs(SAN_OBJ if sanitize else OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        results = list(ex.map(lambda s: _compile(s, sanitize), SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, w in results:
            for line in w:
                print(line, file=sys.stderr)
    out = SAN_OUT if sanitize else OUT
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    if sanitize:
        cmd[1:1] = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libasan"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True, sanitize="--sanitize" in sys.argv[1:]))
