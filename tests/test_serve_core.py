"""The serving loop's host core (zig-stark_amd/csrc/serve_core.h: queue, admission
with a bounded bypass, arena, tickets, shutdown), which zs_server_* runs on the
GPU, stress-tested on the CPU by tests/c/serve_core_stress.cpp under
ThreadSanitizer and under AddressSanitizer + UndefinedBehaviorSanitizer."""
import os
import subprocess

import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(TESTS, "c", "serve_core_stress.cpp")

SANITIZERS = {
    "thread": ["-fsanitize=thread"],
    "address,undefined": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
}


@pytest.mark.parametrize("kind", list(SANITIZERS))
def test_serve_core_under_sanitizer(tmp_path, kind):
    exe = str(tmp_path / "serve_core_stress")
    build = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread", *SANITIZERS[kind], SRC,
             "-o", exe]
    r = subprocess.run(build, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    # the harness links its own sanitizer runtime, so ASan is told not to insist on
    # coming first among preloaded libraries; under tests/run_sanitized.sh (whose ASan
    # runtime is preloaded into this process) it gets the preload list that script found
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    if "ZS_BASE_PRELOAD" in env:
        env["LD_PRELOAD"] = env["ZS_BASE_PRELOAD"]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-5000:]
    assert r.stdout.strip() == "ok"
    assert "Sanitizer" not in r.stderr, r.stderr[-5000:]
