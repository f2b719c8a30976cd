"""GPU parity: circle FFT / IFFT through the C ABI against the oracle
(src/m31/ntt/circle.zig), bit-exact.  Mirrors the reference's GPU test
(src/cuda/circlefft_test.zig:23-79: canonicHalf(3..16), inputs reduced from
random u32) and extends it to every size from 2 to 2^22, arbitrary cosets
and batched device columns."""
import numpy as np
import pytest

import oracle_lib as O
import zig_stark_amd as Z
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu
P = O.P


@pytest.mark.parametrize("lg", list(range(1, 17)) + [18, 20])
def test_fft_ifft_match_oracle(zctx, lg):
    rng = np.random.default_rng(0xC1E22026 + lg)
    n = 1 << lg
    # circlefft_test.zig:32 draws M31.fromInt(random u32)
    coeffs = (rng.integers(0, 2**32, n, dtype=np.uint64) % P).astype(np.uint32)
    evals = (rng.integers(0, 2**32, n, dtype=np.uint64) % P).astype(np.uint32)
    assert np.array_equal(zctx.circle_fft(coeffs), O.circle_fft(coeffs))
    assert np.array_equal(zctx.circle_ifft(evals), O.circle_ifft(evals))


@pytest.mark.parametrize("lg", [22])
def test_fft_ifft_round_trip_large(zctx, lg):
    rng = np.random.default_rng(lg)
    n = 1 << lg
    coeffs = rng.integers(0, P, n, dtype=np.uint32)
    ev = zctx.circle_fft(coeffs)
    assert np.array_equal(zctx.circle_ifft(ev), coeffs)
    c5 = O.canonic_half(lg)
    for i in rng.integers(0, n, 8).tolist():
        assert ev[i] == O.eval_at_point(coeffs, O.domain_point(c5, i))


@pytest.mark.parametrize("lg", [1, 2, 3, 5, 9, 12])
def test_fft_arbitrary_coset(zctx, lg):
    """Any half-coset, not only canonicHalf: CircleCoset.standard-like offsets."""
    rng = np.random.default_rng(100 + lg)
    n = 1 << lg
    gen = (1268011823, 2)
    off = O.point_mul(gen, 12345)
    step = O.point_mul(gen, 1 << (31 - (lg - 1))) if lg > 1 else (1, 0)
    c5 = np.array([off[0], off[1], step[0], step[1], lg - 1], np.uint32)
    hc = Z.Coset(*[int(v) for v in c5])
    coeffs = rng.integers(0, P, n, dtype=np.uint32)
    assert np.array_equal(zctx.circle_fft(coeffs, hc), O.circle_fft(coeffs, c5))
    assert np.array_equal(zctx.circle_ifft(coeffs, hc), O.circle_ifft(coeffs, c5))


@pytest.mark.parametrize("lg,cols", [(4, 3), (11, 5), (13, 2)])
def test_batched_device_columns(zctx, torch_dev, lg, cols):
    rng = np.random.default_rng(lg * 7 + cols)
    n = 1 << lg
    stride = n + 37
    host = rng.integers(0, P, (cols, stride), dtype=np.uint32)
    d_in = to_dev(host, torch_dev)
    d_out = to_dev(np.zeros_like(host), torch_dev)
    zctx.circle_fft_batch_dev(d_in.data_ptr(), d_out.data_ptr(), lg, cols, stride)
    zctx.synchronize()
    out = from_dev(d_out)
    for c in range(cols):
        assert np.array_equal(out[c, :n], O.circle_fft(host[c, :n]))
    zctx.circle_ifft_batch_dev(d_in.data_ptr(), d_out.data_ptr(), lg, cols, stride)
    zctx.synchronize()
    out = from_dev(d_out)
    for c in range(cols):
        assert np.array_equal(out[c, :n], O.circle_ifft(host[c, :n]))


def test_invalid_sizes_are_rejected(zctx):
    with pytest.raises(Z.ZsError) as e:
        zctx.circle_fft(np.zeros(6, np.uint32))
    assert e.value.code == Z.ZS_ERR_INPUT
