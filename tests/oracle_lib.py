"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker for the HIP path, never
the thing measured or shipped.  Imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("ZS_ORACLE_LIBRARY") or os.path.join(ROOT, "oracle", "liboracle.so")
P = 2147483647

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


u32p = ctypes.POINTER(ctypes.c_uint32)
u8p = ctypes.c_char_p
sz = ctypes.c_size_t


def _declare(L):
    L.zo_m31_mul.restype = ctypes.c_uint32
    L.zo_m31_mul.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.zo_m31_add.restype = ctypes.c_uint32
    L.zo_m31_add.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.zo_m31_sub.restype = ctypes.c_uint32
    L.zo_m31_sub.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.zo_m31_inv.restype = ctypes.c_uint32
    L.zo_m31_inv.argtypes = [ctypes.c_uint32]
    L.zo_m31_pow.restype = ctypes.c_uint32
    L.zo_m31_pow.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.zo_m31_from_u32.restype = ctypes.c_uint32
    L.zo_m31_from_u32.argtypes = [ctypes.c_uint32]
    L.zo_eval_at_point.restype = ctypes.c_uint32
    L.zo_eval_at_point.argtypes = [u32p, sz, ctypes.c_uint32, ctypes.c_uint32]
    L.zo_circle_fft.argtypes = [u32p, sz, u32p, u32p]
    L.zo_circle_ifft.argtypes = [u32p, sz, u32p, u32p]
    L.zo_lde.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p]
    L.zo_lde_padded_coeffs.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p]
    L.zo_precompute_twiddles.argtypes = [u32p, u32p]
    L.zo_coset_canonic_half_p.argtypes = [ctypes.c_uint32, u32p]
    L.zo_coset_at_p.argtypes = [u32p, sz, u32p]
    L.zo_domain_point_p.argtypes = [u32p, sz, u32p]
    L.zo_point_mul_scalar_p.argtypes = [u32p, ctypes.c_uint64, u32p]
    L.zo_blake3.argtypes = [u8p, sz, ctypes.c_char_p]
    L.zo_blake2s_from_g.argtypes = [u8p, sz, ctypes.c_char_p]
    L.zo_blake3_compress.argtypes = [u32p, u32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, u32p]
    L.zo_hash2.argtypes = [u8p, u8p, ctypes.c_char_p]
    L.zo_hash_m31s.argtypes = [u32p, sz, ctypes.c_char_p]
    L.zo_hash_qm31.argtypes = [u32p, ctypes.c_char_p]
    L.zo_merkle_build.argtypes = [u8p, sz, ctypes.c_char_p]
    L.zo_merkle_open.argtypes = [u8p, sz, sz, ctypes.c_char_p]
    L.zo_merkle_verify.argtypes = [u8p, sz, u8p, u8p, sz]
    L.zo_row_commit.argtypes = [u32p, sz, sz, ctypes.c_char_p]
    L.zo_channel_new.restype = ctypes.c_void_p
    L.zo_channel_new.argtypes = [u8p, sz]
    L.zo_channel_clone.restype = ctypes.c_void_p
    L.zo_channel_clone.argtypes = [ctypes.c_void_p]
    L.zo_channel_free.argtypes = [ctypes.c_void_p]
    L.zo_channel_absorb.argtypes = [ctypes.c_void_p, u8p, sz]
    L.zo_channel_sample_bytes.argtypes = [ctypes.c_void_p, ctypes.c_char_p, sz]
    L.zo_channel_sample_qm31.argtypes = [ctypes.c_void_p, u32p]
    L.zo_channel_sample_index.restype = sz
    L.zo_channel_sample_index.argtypes = [ctypes.c_void_p, sz]
    L.zo_channel_transcript_len.restype = sz
    L.zo_channel_transcript_len.argtypes = [ctypes.c_void_p]
    L.zo_qm31_mul_p.argtypes = [u32p, u32p, u32p]
    L.zo_qm31_inv_p.argtypes = [u32p, u32p]
    L.zo_qm31_pow_p.argtypes = [u32p, ctypes.c_uint64, u32p]
    L.zo_qm31_root_of_unity_p.argtypes = [ctypes.c_uint32, u32p]
    L.zo_poly_eval.argtypes = [u32p, sz, u32p, u32p]
    L.zo_poly_interpolate.argtypes = [u32p, u32p, sz, u32p]
    L.zo_fri_domain_point.argtypes = [ctypes.c_uint32, ctypes.c_uint32, sz, u32p]
    L.zo_fri_fold.argtypes = [u32p, sz, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p]
    L.zo_fri_commit.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                ctypes.c_void_p, ctypes.c_char_p, u32p, u32p, u32p]
    L.zo_fri_prove.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, sz,
                               ctypes.c_void_p, ctypes.c_char_p, u32p, ctypes.POINTER(sz), u32p,
                               ctypes.c_char_p]
    L.zo_fri_query_path_bytes.restype = sz
    L.zo_fri_query_path_bytes.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.zo_fri_verify.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, sz, ctypes.c_void_p,
                                u8p, u32p, ctypes.POINTER(sz), u32p, u8p]
    L.zo_fill_m31.argtypes = [ctypes.c_uint64, u32p, sz]
    L.zo_trace_commit.argtypes = [u32p, sz, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, u32p, ctypes.c_char_p]
    L.zo_fib_trace.argtypes = [ctypes.c_uint32, ctypes.c_uint32, sz, u32p]
    L.zo_fib_stark_prove.restype = ctypes.c_void_p
    L.zo_fib_stark_prove.argtypes = [ctypes.c_uint32, ctypes.c_uint32, sz, ctypes.c_uint32, u32p, u32p, u32p, u32p,
                                     ctypes.c_void_p, ctypes.POINTER(sz)]
    L.zo_fib_stark_verify.argtypes = [u8p, sz, u32p, u32p, u32p, ctypes.c_void_p]
    L.zo_free.argtypes = [ctypes.c_void_p]
    L.zo_wire_golden_example.restype = sz
    L.zo_wire_golden_example.argtypes = [ctypes.c_char_p]
    L.zo_stark_prove.restype = ctypes.c_void_p
    L.zo_stark_prove.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, sz, ctypes.c_uint32, u32p, u32p,
                                 ctypes.c_void_p, ctypes.POINTER(sz)]
    L.zo_stark_verify.argtypes = [ctypes.c_void_p, u8p, sz, ctypes.c_void_p]
    L.zo_stark_fri_params.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p]
    L.zo_stark_accumulators.argtypes = [ctypes.c_void_p, ctypes.c_uint32, u32p, u32p, u32p, u32p, u32p]
    L.zo_fri_proof_serialize.restype = sz
    L.zo_fri_proof_serialize.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, sz, u8p, u32p,
                                         ctypes.POINTER(sz), u32p, u8p, ctypes.c_char_p]


def _p(a):
    return a.ctypes.data_as(u32p)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---- field ---------------------------------------------------------------------
def m31_mul(a, b):
    return lib().zo_m31_mul(a, b)


def m31_add(a, b):
    return lib().zo_m31_add(a, b)


def m31_sub(a, b):
    return lib().zo_m31_sub(a, b)


def m31_inv(a):
    return lib().zo_m31_inv(a)


def qm31_mul(x, y):
    x, y = u32(x), u32(y)
    o = np.zeros(4, np.uint32)
    lib().zo_qm31_mul_p(_p(x), _p(y), _p(o))
    return o


def qm31_inv(x):
    x = u32(x)
    o = np.zeros(4, np.uint32)
    lib().zo_qm31_inv_p(_p(x), _p(o))
    return o


def qm31_pow(x, e):
    x = u32(x)
    o = np.zeros(4, np.uint32)
    lib().zo_qm31_pow_p(_p(x), e, _p(o))
    return o


def qm31_root(log_size):
    o = np.zeros(4, np.uint32)
    lib().zo_qm31_root_of_unity_p(log_size, _p(o))
    return o


# ---- circle --------------------------------------------------------------------
def canonic_half(log_size):
    o = np.zeros(5, np.uint32)
    lib().zo_coset_canonic_half_p(log_size, _p(o))
    return o


def coset_at(c5, i):
    c5 = u32(c5)
    o = np.zeros(2, np.uint32)
    lib().zo_coset_at_p(_p(c5), i, _p(o))
    return int(o[0]), int(o[1])


def domain_point(c5, i):
    c5 = u32(c5)
    o = np.zeros(2, np.uint32)
    lib().zo_domain_point_p(_p(c5), i, _p(o))
    return int(o[0]), int(o[1])


def point_mul(p, s):
    p = u32(p)
    o = np.zeros(2, np.uint32)
    lib().zo_point_mul_scalar_p(_p(p), s, _p(o))
    return int(o[0]), int(o[1])


def twiddles(c5):
    c5 = u32(c5)
    o = np.zeros(1 << int(c5[4]), np.uint32)
    lib().zo_precompute_twiddles(_p(c5), _p(o))
    return o


def circle_fft(coeffs, coset5=None):
    coeffs = u32(coeffs)
    n = coeffs.size
    if coset5 is None:
        coset5 = canonic_half(n.bit_length() - 1)
    coset5 = u32(coset5)
    out = np.zeros(n, np.uint32)
    rc = lib().zo_circle_fft(_p(coeffs), n, _p(coset5), _p(out))
    assert rc == 0, rc
    return out


def circle_ifft(evals, coset5=None):
    evals = u32(evals)
    n = evals.size
    if coset5 is None:
        coset5 = canonic_half(n.bit_length() - 1)
    coset5 = u32(coset5)
    out = np.zeros(n, np.uint32)
    rc = lib().zo_circle_ifft(_p(evals), n, _p(coset5), _p(out))
    assert rc == 0, rc
    return out


def eval_at_point(coeffs, pt):
    coeffs = u32(coeffs)
    return lib().zo_eval_at_point(_p(coeffs), coeffs.size, pt[0], pt[1])


def lde(col, log_blowup):
    col = u32(col)
    lg = col.size.bit_length() - 1
    out = np.zeros(col.size << log_blowup, np.uint32)
    rc = lib().zo_lde(_p(col), lg, log_blowup, _p(out))
    assert rc == 0, rc
    return out


def lde_padded(col, log_blowup):
    col = u32(col)
    lg = col.size.bit_length() - 1
    out = np.zeros(col.size << log_blowup, np.uint32)
    rc = lib().zo_lde_padded_coeffs(_p(col), lg, log_blowup, _p(out))
    assert rc == 0, rc
    return out


# ---- hashing / merkle ----------------------------------------------------------
def blake3(b):
    b = bytes(b)
    out = ctypes.create_string_buffer(32)
    lib().zo_blake3(b, len(b), out)
    return out.raw


def blake2s_from_g(b):
    b = bytes(b)
    out = ctypes.create_string_buffer(32)
    lib().zo_blake2s_from_g(b, len(b), out)
    return out.raw


def hash2(a, b):
    out = ctypes.create_string_buffer(32)
    lib().zo_hash2(bytes(a), bytes(b), out)
    return out.raw


def hash_m31s(vals):
    vals = u32(vals)
    out = ctypes.create_string_buffer(32)
    lib().zo_hash_m31s(_p(vals), vals.size, out)
    return out.raw


def hash_qm31(v):
    v = u32(v)
    out = ctypes.create_string_buffer(32)
    lib().zo_hash_qm31(_p(v), out)
    return out.raw


def merkle_build(leaves):
    """leaves: bytes of n*32. Returns all levels, (2n-1)*32 bytes."""
    n = len(leaves) // 32
    out = ctypes.create_string_buffer((2 * n - 1) * 32)
    rc = lib().zo_merkle_build(bytes(leaves), n, out)
    assert rc == 0
    return out.raw


def merkle_open(nodes, n, index):
    depth = n.bit_length() - 1
    out = ctypes.create_string_buffer(max(depth, 1) * 32)
    rc = lib().zo_merkle_open(bytes(nodes), n, index, out)
    assert rc == 0
    return out.raw[: depth * 32]


def merkle_verify(root, index, leaf, path):
    return bool(lib().zo_merkle_verify(bytes(root), index, bytes(leaf), bytes(path), len(path) // 32))


def row_commit(cols):
    """cols: (n_cols, n_rows) uint32 array. Returns all tree levels as bytes."""
    cols = u32(cols)
    c, n = cols.shape
    out = ctypes.create_string_buffer((2 * n - 1) * 32)
    rc = lib().zo_row_commit(_p(cols), c, n, out)
    assert rc == 0
    return out.raw


def tree_root(nodes):
    return bytes(nodes[-32:])


# ---- channel -------------------------------------------------------------------
class Channel:
    def __init__(self, sep=b"", _handle=None):
        self.h = _handle if _handle is not None else lib().zo_channel_new(bytes(sep), len(sep))

    def __del__(self):
        if getattr(self, "h", None):
            lib().zo_channel_free(self.h)
            self.h = None

    def clone(self):
        return Channel(_handle=lib().zo_channel_clone(self.h))

    def absorb(self, data):
        data = bytes(data)
        lib().zo_channel_absorb(self.h, data, len(data))

    def sample_bytes(self, n):
        out = ctypes.create_string_buffer(n)
        lib().zo_channel_sample_bytes(self.h, out, n)
        return out.raw

    def sample_qm31(self):
        o = np.zeros(4, np.uint32)
        lib().zo_channel_sample_qm31(self.h, _p(o))
        return o

    def sample_index(self, n):
        return lib().zo_channel_sample_index(self.h, n)

    def transcript_len(self):
        return lib().zo_channel_transcript_len(self.h)


# ---- polynomials / FRI ---------------------------------------------------------
def poly_eval(coeffs, x):
    coeffs, x = u32(coeffs).reshape(-1), u32(x)
    o = np.zeros(4, np.uint32)
    lib().zo_poly_eval(_p(coeffs), coeffs.size // 4, _p(x), _p(o))
    return o


def poly_interpolate(xs, ys):
    xs, ys = u32(xs).reshape(-1), u32(ys).reshape(-1)
    n = xs.size // 4
    o = np.zeros(4 * n, np.uint32)
    rc = lib().zo_poly_interpolate(_p(xs), _p(ys), n, _p(o))
    assert rc == 0
    return o.reshape(n, 4)


def fri_domain_point(dlog, layer, j):
    o = np.zeros(4, np.uint32)
    lib().zo_fri_domain_point(dlog, layer, j, _p(o))
    return o


def fri_fold(vals, dlog, layer, alpha):
    vals, alpha = u32(vals).reshape(-1), u32(alpha)
    n = vals.size // 4
    o = np.zeros(2 * n, np.uint32)
    rc = lib().zo_fri_fold(_p(vals), n, dlog, layer, _p(alpha), _p(o))
    assert rc == 0
    return o.reshape(n // 2, 4)


def fri_commit(codeword, log_blowup, remainder_log, alphas=None, channel=None, want_layers=True):
    """Returns (roots[L] bytes list, layers list of (n_i,4) arrays or None, remainder (k,4), alphas (L,4))."""
    cw = u32(codeword).reshape(-1)
    n0 = cw.size // 4
    dlog = n0.bit_length() - 1
    L = dlog - remainder_log
    roots = ctypes.create_string_buffer(32 * L)
    total = sum(n0 >> i for i in range(L + 1))
    layers = np.zeros(4 * total, np.uint32) if want_layers else None
    rem = np.zeros(4 << (remainder_log - log_blowup), np.uint32)
    al_out = np.zeros(4 * L, np.uint32)
    al_in = u32(alphas).reshape(-1) if alphas is not None else None
    rc = lib().zo_fri_commit(_p(cw), dlog, log_blowup, remainder_log,
                             _p(al_in) if al_in is not None else None,
                             channel.h if channel is not None else None,
                             roots, _p(layers) if layers is not None else None, _p(rem), _p(al_out))
    assert rc == 0, rc
    lay = None
    if want_layers:
        lay, off = [], 0
        for i in range(L + 1):
            ni = n0 >> i
            lay.append(layers[4 * off: 4 * (off + ni)].reshape(ni, 4))
            off += ni
    return ([roots.raw[32 * i: 32 * (i + 1)] for i in range(L)], lay, rem.reshape(-1, 4),
            al_out.reshape(L, 4))


def fri_prove(codeword, log_blowup, remainder_log, num_queries, channel):
    cw = u32(codeword).reshape(-1)
    n0 = cw.size // 4
    dlog = n0.bit_length() - 1
    L = dlog - remainder_log
    roots = ctypes.create_string_buffer(32 * L)
    rem = np.zeros(4 << (remainder_log - log_blowup), np.uint32)
    qidx = (sz * num_queries)()
    qvals = np.zeros(num_queries * L * 2 * 4, np.uint32)
    per_q = lib().zo_fri_query_path_bytes(dlog, remainder_log)
    qpaths = ctypes.create_string_buffer(max(1, per_q * num_queries))
    rc = lib().zo_fri_prove(_p(cw), dlog, log_blowup, remainder_log, num_queries, channel.h, roots,
                            _p(rem), qidx, _p(qvals), qpaths)
    assert rc == 0, rc
    return dict(dlog=dlog, L=L, roots=roots.raw, remainder=rem, qidx=list(qidx), qvals=qvals,
                qpaths=qpaths.raw[: per_q * num_queries], num_queries=num_queries)


def fri_verify(proof, log_blowup, remainder_log, channel):
    q = proof["num_queries"]
    qidx = (sz * q)(*proof["qidx"])
    return bool(lib().zo_fri_verify(proof["dlog"], log_blowup, remainder_log, q, channel.h,
                                    bytes(proof["roots"]), _p(u32(proof["remainder"])), qidx,
                                    _p(u32(proof["qvals"])), bytes(proof["qpaths"])))


def fill_m31(seed, n):
    o = np.zeros(n, np.uint32)
    lib().zo_fill_m31(seed, _p(o), n)
    return o


def qm31_combine(lde, gamma):
    """g[i] = sum_c gamma^c * lde[c][i] as (N, 4) QM31 values (lde: (n_cols, N) M31)."""
    n_cols, N = lde.shape
    g = np.zeros((N, 4), np.uint64)
    pw = np.array([1, 0, 0, 0], np.uint32)
    col = lde.astype(np.uint64)
    for c in range(n_cols):
        for k in range(4):
            g[:, k] = (g[:, k] + (int(pw[k]) * col[c]) % P) % P
        pw = qm31_mul(pw, gamma)
    return g.astype(np.uint32)


def commit_job(cols, log_blowup, remainder_log):
    """The trace commitment the library's zs_commit_job_dev computes, from oracle parts:
    LDE -> row tree -> Channel("zig-stark:hip:commit") absorbs the root -> gamma ->
    g = sum_c gamma^c lde_c -> FRI commit phase with the same channel.
    Returns (trace_root, fri_roots, remainder, lde, nodes)."""
    lde_ = np.stack([lde(c, log_blowup) for c in cols])
    nodes = row_commit(lde_)
    root = tree_root(nodes)
    ch = Channel(b"zig-stark:hip:commit")
    ch.absorb(root)
    gamma = ch.sample_qm31()
    g = qm31_combine(lde_, gamma)
    roots, _, rem, _ = fri_commit(g, log_blowup, remainder_log, channel=ch, want_layers=False)
    return root, roots, rem, lde_, nodes


def trace_commit(cols, log_blowup, channel=None, want_codewords=True):
    """GenericStark trace commitment (stark.zig:343-414): cols (n_cols, n, 4) QM31.
    Returns (roots list, codewords (n_cols, N, 4) or None)."""
    cols = u32(cols)
    n_cols, n = cols.shape[0], cols.shape[1]
    tl = n.bit_length() - 1
    N = n << log_blowup
    cw = np.zeros((n_cols, N, 4), np.uint32) if want_codewords else None
    roots = ctypes.create_string_buffer(32 * n_cols)
    rc = lib().zo_trace_commit(_p(cols.reshape(-1)), n_cols, tl, log_blowup,
                               channel.h if channel is not None else None,
                               _p(cw) if cw is not None else None, roots)
    assert rc == 0, rc
    return [roots.raw[32 * c:32 * (c + 1)] for c in range(n_cols)], cw


def fib_trace(a0, a1, n):
    out = np.zeros((2, n, 4), np.uint32)
    lib().zo_fib_trace(a0, a1, n, _p(out))
    return out


def wire_golden_example():
    out = ctypes.create_string_buffer(64)
    n = lib().zo_wire_golden_example(out)
    return out.raw[:n]


def fri_proof_bytes(proof, log_blowup, remainder_log):
    """Fri.Proof wire bytes (serialization.zig) of a fri_prove() result."""
    dlog = proof["dlog"]
    q = proof["num_queries"]
    qidx = (sz * q)(*proof["qidx"])
    args = (dlog - log_blowup, log_blowup, remainder_log, q, bytes(proof["roots"]), _p(u32(proof["remainder"])),
            qidx, _p(u32(proof["qvals"])), bytes(proof["qpaths"]))
    n = lib().zo_fri_proof_serialize(*args, None)
    out = ctypes.create_string_buffer(n)
    lib().zo_fri_proof_serialize(*args, out)
    return out.raw


def fib_stark_prove(trace, log_blowup, num_queries, remainder_log, a0, a1, claimed, channel):
    """GenericStark(FibSeededAir).prove -> GenericStark.Proof wire bytes."""
    trace = u32(trace)
    tl = trace.shape[1].bit_length() - 1
    q = lambda v: _p(u32(v).reshape(4))
    n = sz()
    ptr = lib().zo_fib_stark_prove(tl, log_blowup, num_queries, remainder_log, _p(trace.reshape(-1)), q(a0), q(a1),
                                   q(claimed), channel.h, ctypes.byref(n))
    assert ptr, "prove failed"
    try:
        return ctypes.string_at(ptr, n.value)
    finally:
        lib().zo_free(ptr)


def fib_stark_verify(proof, a0, a1, claimed, channel):
    q = lambda v: _p(u32(v).reshape(4))
    return bool(lib().zo_fib_stark_verify(bytes(proof), len(proof), q(a0), q(a1), q(claimed), channel.h))


# ---- GenericStark(Air) for an AIR given as data (tests/airs.py) -------------------
def stark_prove(air, pre, trace, log_blowup, num_queries, remainder_log, channel):
    """proveWithPreprocessed (stark.zig:311-657) -> GenericStark.Proof wire bytes, or None
    when friParams rejects the parameters.  pre / trace: (cols, n, 4) QM31 arrays."""
    trace = u32(trace)
    tl = trace.shape[1].bit_length() - 1
    pre = u32(pre) if pre is not None and len(pre) else np.zeros(4, np.uint32)
    st = air.struct()
    n = sz()
    ptr = lib().zo_stark_prove(ctypes.byref(st), tl, log_blowup, num_queries, remainder_log, _p(pre.reshape(-1)),
                               _p(trace.reshape(-1)), channel.h, ctypes.byref(n))
    if not ptr:
        return None
    try:
        return ctypes.string_at(ptr, n.value)
    finally:
        lib().zo_free(ptr)


def stark_verify(air, proof, channel):
    st = air.struct()
    return bool(lib().zo_stark_verify(ctypes.byref(st), bytes(proof), len(proof), channel.h))


def stark_fri_params(air, trace_log, log_blowup):
    a, b = ctypes.c_uint32(), ctypes.c_uint32()
    st = air.struct()
    rc = lib().zo_stark_fri_params(ctypes.byref(st), trace_log, log_blowup, ctypes.byref(a), ctypes.byref(b))
    return (a.value, b.value) if rc == 0 else None


def stark_accumulators(air, pre, trace, alpha_cols, alpha_rels):
    trace = u32(trace)
    n = trace.shape[1]
    tl = n.bit_length() - 1
    pre = u32(pre) if pre is not None and len(pre) else np.zeros(4, np.uint32)
    st = air.struct()
    out = np.zeros((max(1, len(air.lookups)), n, 4), np.uint32)
    lib().zo_stark_accumulators(ctypes.byref(st), tl, _p(pre.reshape(-1)), _p(trace.reshape(-1)),
                                _p(u32(alpha_cols).reshape(-1)), _p(u32(alpha_rels).reshape(-1)), _p(out))
    return out[:len(air.lookups)]
