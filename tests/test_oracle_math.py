"""Pins for the oracle's field, circle-group and circle-FFT restatements.

Each check compares the oracle with something it does not share code with:
Python big-integer arithmetic, a schoolbook restatement of the tower
definition, double-and-add scalar multiplication, brute-force evaluation in
the FFT basis from its product definition, and the reference's own
known-answer tests (src/m31/ntt/circle.zig:358-453, src/m31/field/*.zig tests,
src/m31/circle/*.zig tests).
"""
import numpy as np
import pytest

import oracle_lib as O

P = O.P


# ---- M31: src/m31/field/m31.zig tests (:147-189) ----------------------------------------
def test_m31_against_bigint():
    rng = np.random.default_rng(42)
    a = rng.integers(0, P, 2000, dtype=np.uint64)
    b = rng.integers(0, P, 2000, dtype=np.uint64)
    edge = [0, 1, P - 1, P - 2, 2, 1 << 30, (1 << 30) + 1]
    pairs = list(zip(a.tolist(), b.tolist())) + [(x, y) for x in edge for y in edge]
    for x, y in pairs:
        assert O.m31_add(x, y) == (x + y) % P
        assert O.m31_sub(x, y) == (x - y) % P
        assert O.m31_mul(x, y) == (x * y) % P


def test_m31_from_int_and_inverse():
    L = O.lib()
    for x in [0, 5, P, P + 5, 2**32 - 1, 2**31, P - 1]:
        assert L.zo_m31_from_u32(x) == x % P
    rng = np.random.default_rng(7)
    for x in rng.integers(1, P, 200).tolist():
        assert O.m31_mul(x, O.m31_inv(x)) == 1
        assert L.zo_m31_pow(x, P - 1) == 1
    assert O.m31_inv(1) == 1


# ---- CM31/QM31: cm31.zig:87-117, qm31.zig:104-209 ----------------------------------------
def qm_schoolbook(x, y):
    """(a + b u)(c + d u) with a..d in CM31 = F[i]/(i^2+1) and u^2 = -i, by expanding
    over the basis {1, i, u, iu} with Python integers."""
    a0, a1, b0, b1 = [int(v) for v in x]
    c0, c1, d0, d1 = [int(v) for v in y]

    def cmul(p0, p1, q0, q1):
        return ((p0 * q0 - p1 * q1) % P, (p0 * q1 + p1 * q0) % P)

    ac = cmul(a0, a1, c0, c1)
    bd = cmul(b0, b1, d0, d1)
    ad = cmul(a0, a1, d0, d1)
    bc = cmul(b0, b1, c0, c1)
    # u^2 = -i: bd * u^2 = -i * bd = (bd1, -bd0)
    re = ((ac[0] + bd[1]) % P, (ac[1] - bd[0]) % P)
    im = ((ad[0] + bc[0]) % P, (ad[1] + bc[1]) % P)
    return np.array([re[0], re[1], im[0], im[1]], np.uint32)


def test_qm31_mul_against_schoolbook():
    rng = np.random.default_rng(21)
    for _ in range(500):
        x = rng.integers(0, P, 4, dtype=np.uint32)
        y = rng.integers(0, P, 4, dtype=np.uint32)
        assert np.array_equal(O.qm31_mul(x, y), qm_schoolbook(x, y))


def test_qm31_tower_identities_and_inverse():
    j = np.array([0, 0, 1, 0], np.uint32)
    minus_i = np.array([0, P - 1, 0, 0], np.uint32)
    assert np.array_equal(O.qm31_mul(j, j), minus_i)
    rng = np.random.default_rng(22)
    one = np.array([1, 0, 0, 0], np.uint32)
    for _ in range(200):
        x = rng.integers(0, P, 4, dtype=np.uint32)
        if not x.any():
            continue
        assert np.array_equal(O.qm31_mul(x, O.qm31_inv(x)), one)


def test_qm31_two_adic_root():
    root = O.qm31_root(32)
    # bytes of TWO_ADIC_ROOT (qm31.zig:80-83), little-endian words
    raw = bytes([0x60, 0x1f, 0x3d, 0x34, 0x32, 0x39, 0xa0, 0x6a,
                 0xa0, 0xd6, 0x51, 0x61, 0x91, 0xe7, 0x86, 0x3e])
    assert np.array_equal(root, np.frombuffer(raw, "<u4").astype(np.uint32))
    one = np.array([1, 0, 0, 0], np.uint32)
    assert np.array_equal(O.qm31_pow(root, 1 << 32), one)
    assert np.array_equal(O.qm31_pow(root, 1 << 31), np.array([P - 1, 0, 0, 0], np.uint32))
    for k in range(0, 25):
        w = O.qm31_root(k)
        assert np.array_equal(O.qm31_pow(w, 1 << k), one)
        if k:
            assert not np.array_equal(O.qm31_pow(w, 1 << (k - 1)), one)


# ---- circle group: point.zig:126-169, coset.zig:68-143 ----------------------------------
def padd(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def pmul_naive(p, s):
    r, b = (1, 0), p
    while s:
        if s & 1:
            r = padd(r, b)
        b = padd(b, b)
        s >>= 1
    return r


GEN = (1268011823, 2)


def test_generator_order_and_mulscalar():
    assert (GEN[0] ** 2 + GEN[1] ** 2) % P == 1
    assert O.point_mul(GEN, 1 << 31) == (1, 0)
    assert O.point_mul(GEN, 1 << 30) != (1, 0)
    rng = np.random.default_rng(77)
    for s in rng.integers(0, 1 << 62, 60).tolist() + [0, 1, 2, 15, 16, 17, (1 << 64) - 1]:
        assert O.point_mul(GEN, s) == pmul_naive(GEN, s)
    for k in range(1, 21):
        g = O.point_mul(GEN, 1 << (31 - k))
        assert O.point_mul(g, 1 << k) == (1, 0)
        if k > 1:
            assert O.point_mul(g, 1 << (k - 1)) != (1, 0)


@pytest.mark.parametrize("lg", range(1, 9))
def test_canonic_half_domain(lg):
    c5 = O.canonic_half(lg)
    gneg = (GEN[0], (P - GEN[1]) % P)
    assert (int(c5[0]), int(c5[1])) == pmul_naive(gneg, 1 << (30 - lg))
    assert (int(c5[2]), int(c5[3])) == pmul_naive(gneg, (1 << (32 - lg)) % (1 << 31))
    n = 1 << lg
    pts = [O.domain_point(c5, i) for i in range(n)]
    assert len(set(pts)) == n
    for x, y in pts:
        assert (x * x + y * y) % P == 1
    h = n // 2
    for i in range(h):
        assert pts[h + i] == (pts[i][0], (P - pts[i][1]) % P)


# ---- circle FFT: circle.zig:358-453 ------------------------------------------------------
def pi_map(x):
    return (2 * x * x - 1) % P


def eval_basis_product(coeffs, pt):
    """f(P) = sum_k c_k * prod_b basis_b(P)^{bit b of k}, interleaved index k, with
    basis_0 = y, basis_1 = x, basis_b = pi^(b-1)(x) (circle.zig:8-26, 112-127)."""
    n = len(coeffs)
    lg = n.bit_length() - 1
    half = n // 2
    x, y = pt
    basis = [y]
    xb = x
    for _ in range(1, lg):
        basis.append(xb)
        xb = pi_map(xb)
    tot = 0
    for k in range(n):
        c = int(coeffs[k // 2]) if k % 2 == 0 else int(coeffs[half + k // 2])
        term = c
        for b in range(lg):
            if (k >> b) & 1:
                term = term * basis[b] % P
        tot += term
    return tot % P


@pytest.mark.parametrize("lg", range(1, 8))
def test_eval_at_point_matches_basis_product(lg):
    rng = np.random.default_rng(100 + lg)
    coeffs = rng.integers(0, P, 1 << lg, dtype=np.uint32)
    c5 = O.canonic_half(lg + 1)
    for i in range(0, 1 << (lg + 1), max(1, (1 << (lg + 1)) // 16)):
        pt = O.domain_point(c5, i)
        assert O.eval_at_point(coeffs, pt) == eval_basis_product(coeffs, pt)


@pytest.mark.parametrize("lg", range(1, 12))
def test_fft_matches_brute_force_and_roundtrips(lg):
    rng = np.random.default_rng(31 + lg)
    n = 1 << lg
    c5 = O.canonic_half(lg)
    coeffs = rng.integers(0, P, n, dtype=np.uint32)
    ev = O.circle_fft(coeffs, c5)
    idx = range(n) if lg <= 9 else rng.integers(0, n, 128).tolist()
    for i in idx:
        assert ev[i] == O.eval_at_point(coeffs, O.domain_point(c5, i))
    assert np.array_equal(O.circle_ifft(ev, c5), coeffs)


def test_fft_known_answers():
    c5 = O.canonic_half(3)
    n = 8
    pts = [O.domain_point(c5, i) for i in range(n)]
    c = np.zeros(n, np.uint32)
    c[0] = 7
    assert np.all(O.circle_fft(c, c5) == 7)
    c = np.zeros(n, np.uint32)
    c[1] = 1
    assert list(O.circle_fft(c, c5)) == [p[0] for p in pts]
    c = np.zeros(n, np.uint32)
    c[n // 2] = 3
    assert list(O.circle_fft(c, c5)) == [3 * p[1] % P for p in pts]


def test_twiddle_tree_layout():
    """slowPrecomputeTwiddles (circle.zig:69-82): level k = bit-reversed x of the
    first 2^(k-1) points of the halved coset, last slot 1."""
    for lg in range(2, 9):
        c5 = O.canonic_half(lg)
        tw = O.twiddles(c5)
        off = (int(c5[0]), int(c5[1]))
        step = (int(c5[2]), int(c5[3]))
        pos = 0
        for k in range(lg - 1, 0, -1):
            m = 1 << (k - 1)
            xs = [padd(off, pmul_naive(step, i))[0] for i in range(m)]
            w = k - 1
            rev = [xs[int(format(i, f"0{w}b")[::-1] or "0", 2)] for i in range(m)] if w else xs
            assert list(tw[pos:pos + m]) == rev
            pos += m
            off, step = padd(off, off), padd(step, step)
        assert tw[pos] == 1 and pos + 1 == tw.size


@pytest.mark.parametrize("lg,beta", [(1, 1), (1, 2), (2, 2), (3, 2), (4, 2), (5, 1), (5, 3), (6, 2), (8, 2)])
def test_lde_is_evaluation_of_padded_interpolant(lg, beta):
    rng = np.random.default_rng(lg * 10 + beta)
    col = rng.integers(0, P, 1 << lg, dtype=np.uint32)
    l = O.lde(col, beta)
    pad = O.lde_padded(col, beta)
    c5 = O.canonic_half(lg + beta)
    N = 1 << (lg + beta)
    for i in range(0, N, max(1, N // 64)):
        assert l[i] == O.eval_at_point(pad, O.domain_point(c5, i))
    assert np.array_equal(O.circle_ifft(l, c5), pad)
    h, H = 1 << (lg - 1), N // 2
    assert not pad[h:H].any() and not pad[H + h:].any()
    # pad holds exactly the interpolant of the column on canonicHalf(lg)
    assert np.array_equal(pad[:h], O.circle_ifft(col)[:h])
    assert np.array_equal(pad[H:H + h], O.circle_ifft(col)[h:])


def test_fill_m31_stream():
    a = O.fill_m31(5, 1000)
    assert a.max() < P
    assert np.array_equal(O.fill_m31(5, 10), a[:10])
    assert not np.array_equal(O.fill_m31(6, 10), a[:10])
