"""CPU-side checks of the C-ABI library: it loads, exports every symbol
include/zs_hip.h declares, and its host logic (canonic cosets, the
Fiat-Shamir channel) agrees with the oracle.  No device compute."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z


def test_library_exports_every_header_symbol():
    L = Z.lib()
    names = Z.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_string():
    assert "gfx950" in Z.version()


@pytest.mark.parametrize("lg", [1, 2, 3, 8, 16, 22, 28])
def test_canonic_half_matches_oracle(lg):
    assert Z.canonic_half(lg).as_list() == [int(v) for v in O.canonic_half(lg)]


def test_canonic_half_rejects_out_of_range():
    c = Z.Coset()
    assert Z.lib().zs_circle_coset_canonic_half(0, ctypes.byref(c)) == Z.ZS_ERR_INPUT
    assert Z.lib().zs_circle_coset_canonic_half(29, ctypes.byref(c)) == Z.ZS_ERR_INPUT


def test_channel_matches_oracle():
    rng = np.random.default_rng(12)
    a, b = Z.Channel(b"zig-stark:hip"), O.Channel(b"zig-stark:hip")
    for step in range(120):
        d = rng.integers(0, 256, int(rng.integers(0, 300)), dtype=np.uint8).tobytes()
        a.absorb(d)
        b.absorb(d)
        k = step % 4
        if k == 0:
            assert a.sample_bytes(16) == b.sample_bytes(16)
        elif k == 1:
            n = int(rng.integers(1, 2**30))
            assert a.sample_index(n) == b.sample_index(n)
        elif k == 2:
            assert np.array_equal(a.sample_qm31(), b.sample_qm31())
        else:
            assert a.sample_bytes(100) == b.sample_bytes(100)
    assert b.transcript_len() > 20000  # many BLAKE3 chunks deep
    c = a.clone()
    assert c.sample_bytes(32) == a.sample_bytes(32)


def test_init_without_device_fails_loudly():
    """With no HIP device the library reports ZS_ERR_DEVICE (no CPU fallback)."""
    h = ctypes.c_void_p()
    rc = Z.lib().zs_hip_init(0, 0, ctypes.byref(h))
    if rc == Z.ZS_OK:
        Z.lib().zs_hip_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc == Z.ZS_ERR_DEVICE
    assert Z.lib().zs_hip_init(-1, 0, ctypes.byref(h)) == Z.ZS_ERR_DEVICE
