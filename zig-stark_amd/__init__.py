"""Python binding of libzs_hip.so (include/zs_hip.h), the MI355X M31
trace-commitment path of zig-stark.

The names mirror the reference interface: circle_fft / circle_ifft follow
src/cuda/circlefft_gpu.zig:46,111 (and src/m31/ntt/circle.zig:195,272),
lde_commit follows the per-column commitment of src/m31/stark.zig:396-414
with row leaves (src/m31/hash.zig:15-25), fri_commit follows
src/m31/fri.zig:157-233 and fri_query fri.zig:236-256, Channel follows
src/core/channel/channel.zig.

There is no CPU fallback: if the library or a HIP device is missing, the
calls raise ZsError.  Device buffers are torch tensors on a ROCm device
(torch is only used for device memory and streams).
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# ZS_HIP_LIBRARY selects another build of the same library (the sanitizer build of
# tests/run_sanitized.sh); by default the in-tree libzs_hip.so
LIB_PATH = os.environ.get("ZS_HIP_LIBRARY") or os.path.join(PKG, "libzs_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "zs_hip.h")
P = 2147483647

ZS_OK, ZS_ERR_GENERIC, ZS_ERR_INPUT, ZS_ERR_OOM, ZS_ERR_PROTOCOL, ZS_ERR_DEVICE = 0, -1, -2, -3, -4, -5
ZS_FLAG_PLAIN = 1
ZS_FLAG_NO_GRAPH = 2
ZS_FLAG_PERSISTENT = 4
ZS_FLAG_GRAPH = 8

_lib = None


class ZsError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"zs_hip error {code}: {msg}")
        self.code = code


class Coset(ctypes.Structure):
    """CircleCoset (src/m31/circle/coset.zig:10-14)."""
    _fields_ = [("offset_x", ctypes.c_uint32), ("offset_y", ctypes.c_uint32), ("step_x", ctypes.c_uint32),
                ("step_y", ctypes.c_uint32), ("log_size", ctypes.c_uint32)]

    def as_list(self):
        return [self.offset_x, self.offset_y, self.step_x, self.step_y, self.log_size]


u32p = ctypes.POINTER(ctypes.c_uint32)
vp = ctypes.c_void_p
sz = ctypes.c_size_t
i32 = ctypes.c_int32
u32 = ctypes.c_uint32


def lib():
    """Load libzs_hip.so; raise if it is not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ZsError(ZS_ERR_GENERIC, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # torch ships its own libamdhip64.so.7 under the same soname as the system
        # ROCm runtime; whichever is loaded first serves the whole process.  Load
        # torch first (when installed) so torch and this library share its runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    L.zs_hip_version.restype = ctypes.c_char_p
    L.zs_hip_init.argtypes = [i32, u32, ctypes.POINTER(vp)]
    L.zs_hip_init.restype = i32
    L.zs_hip_destroy.argtypes = [vp]
    L.zs_hip_last_error.argtypes = [vp]
    L.zs_hip_last_error.restype = ctypes.c_char_p
    L.zs_hip_set_stream.argtypes = [vp, vp]
    L.zs_hip_synchronize.argtypes = [vp]
    L.zs_circle_coset_canonic_half.argtypes = [u32, ctypes.POINTER(Coset)]
    L.zs_m31_circle_fft.argtypes = [vp, u32p, sz, ctypes.POINTER(Coset), u32p]
    L.zs_m31_circle_ifft.argtypes = [vp, u32p, sz, ctypes.POINTER(Coset), u32p]
    L.zs_m31_circle_fft_batch_dev.argtypes = [vp, vp, vp, u32, u32, sz]
    L.zs_m31_circle_ifft_batch_dev.argtypes = [vp, vp, vp, u32, u32, sz]
    L.zs_m31_lde_commit_dev.argtypes = [vp, vp, u32, u32, u32, vp, vp, u32]
    L.zs_m31_lde_commit_dev_noalloc.argtypes = [vp, vp, u32, u32, u32, vp, vp]
    L.zs_stark_prove_dev.argtypes = [vp, vp, vp, vp, u32, u32, sz, u32, vp, HostAllocator, ctypes.POINTER(vp),
                                     ctypes.POINTER(sz)]
    L.zs_stark_verify.argtypes = [ctypes.c_char_p, sz, vp, vp, ctypes.POINTER(ctypes.c_uint8)]
    L.zs_m31_lde_commit.argtypes = [vp, u32p, u32, u32, u32, ctypes.POINTER(vp), ctypes.c_char_p, u32p]
    L.zs_m31_row_commit_dev.argtypes = [vp, vp, u32, u32, sz, vp, u32]
    L.zs_m31_lde_dev.argtypes = [vp, vp, u32, u32, u32, vp]
    L.zs_m31_row_leaves_dev.argtypes = [vp, vp, u32, u32, sz, vp]
    L.zs_tree_log_leaves.argtypes = [vp]
    L.zs_tree_log_leaves.restype = u32
    L.zs_tree_root.argtypes = [vp, ctypes.c_char_p]
    L.zs_merkle_open.argtypes = [vp, sz, ctypes.c_char_p]
    L.zs_tree_read_level.argtypes = [vp, u32, ctypes.c_char_p]
    L.zs_tree_free.argtypes = [vp]
    L.zs_channel_new.argtypes = [ctypes.c_char_p, sz]
    L.zs_channel_new.restype = vp
    L.zs_channel_clone.argtypes = [vp]
    L.zs_channel_clone.restype = vp
    L.zs_channel_free.argtypes = [vp]
    L.zs_channel_absorb.argtypes = [vp, ctypes.c_char_p, sz]
    L.zs_channel_sample_bytes.argtypes = [vp, ctypes.c_char_p, sz]
    L.zs_channel_sample_qm31.argtypes = [vp, u32p]
    L.zs_channel_sample_index.argtypes = [vp, sz]
    L.zs_channel_sample_index.restype = sz
    L.zs_fri_commit_dev.argtypes = [vp, vp, u32, u32, u32, u32p, vp, u32, ctypes.c_char_p, u32p, ctypes.POINTER(vp)]
    L.zs_fri_commit.argtypes = [vp, u32p, u32, u32, u32, u32p, vp, ctypes.c_char_p, u32p, ctypes.POINTER(vp)]
    L.zs_fri_query.argtypes = [vp, vp, sz, ctypes.POINTER(sz), u32p, ctypes.c_char_p]
    L.zs_fri_read_layer.argtypes = [vp, u32, u32p]
    L.zs_fri_result_free.argtypes = [vp]
    L.zs_fill_m31_dev.argtypes = [vp, ctypes.c_uint64, vp, sz]
    L.zs_commit_job_dev.argtypes = [vp, vp, u32, u32, u32, u32, vp, vp, ctypes.c_char_p, ctypes.c_char_p, u32p, u32]
    L.zs_commit_batch_dev.argtypes = [vp, ctypes.c_void_p, sz, u32, u32, u32, u32]
    L.zs_fri_prove_dev.argtypes = [vp, vp, u32, u32, u32, sz, vp, HostAllocator, ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(sz)]
    L.zs_fib_stark_prove_dev.argtypes = [vp, vp, u32, u32, sz, u32, u32p, u32p, u32p, vp, HostAllocator,
                                         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(sz)]
    L.zs_fri_verify.argtypes = [ctypes.c_char_p, sz, vp, ctypes.POINTER(ctypes.c_uint8)]
    L.zs_fib_stark_verify.argtypes = [ctypes.c_char_p, sz, u32p, u32p, u32p, vp, ctypes.POINTER(ctypes.c_uint8)]
    L.zs_free.argtypes = [HostAllocator, ctypes.c_void_p, sz]
    L.zs_free.restype = None
    L.zs_autotune_lde.argtypes = [vp, u32, u32, u32, u32p]
    L.zs_qm31_trace_commit_dev.argtypes = [vp, vp, u32, u32, u32, vp, vp, ctypes.c_char_p, vp]
    L.zs_fib_trace.argtypes = [u32, u32, u32, u32p]
    L.zs_server_new.argtypes = [i32, u32, ctypes.c_uint64]
    L.zs_server_new.restype = vp
    L.zs_server_submit.argtypes = [vp, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    L.zs_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_void_p]
    L.zs_server_poll.argtypes = [vp, ctypes.c_uint64, ctypes.c_void_p, ctypes.POINTER(i32)]
    L.zs_server_memory.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    L.zs_server_free.argtypes = [vp]
    L.zs_hip_profile.argtypes = [vp, i32]
    L.zs_hip_profile_read.argtypes = [vp, ctypes.c_char_p, sz, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(u32), u32, ctypes.POINTER(u32)]
    for name in ["zs_hip_set_stream", "zs_hip_synchronize", "zs_circle_coset_canonic_half", "zs_m31_circle_fft",
                 "zs_m31_circle_ifft", "zs_m31_circle_fft_batch_dev", "zs_m31_circle_ifft_batch_dev",
                 "zs_m31_lde_commit_dev", "zs_m31_lde_commit_dev_noalloc", "zs_m31_lde_commit",
                 "zs_m31_row_commit_dev", "zs_m31_lde_dev",
                 "zs_m31_row_leaves_dev", "zs_tree_root",
                 "zs_merkle_open", "zs_tree_read_level", "zs_fri_commit_dev", "zs_fri_commit", "zs_fri_query",
                 "zs_fri_read_layer", "zs_fill_m31_dev", "zs_hip_profile", "zs_hip_profile_read",
                 "zs_commit_job_dev", "zs_commit_batch_dev", "zs_server_submit", "zs_server_wait",
                 "zs_server_memory", "zs_server_poll", "zs_qm31_trace_commit_dev", "zs_fib_trace", "zs_autotune_lde", "zs_fri_prove_dev", "zs_fib_stark_prove_dev",
                 "zs_fri_verify", "zs_fib_stark_verify", "zs_stark_prove_dev", "zs_stark_verify"]:
        getattr(L, name).restype = i32


def header_functions():
    """Names of the functions include/zs_hip.h declares."""
    import re
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zs_[a-z0-9_]+)\s*\(", text)))


def _p(a):
    return a.ctypes.data_as(u32p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def canonic_half(log_size):
    """CircleCoset.canonicHalf (coset.zig:55-65)."""
    c = Coset()
    rc = lib().zs_circle_coset_canonic_half(log_size, ctypes.byref(c))
    if rc:
        raise ZsError(rc, "canonic_half")
    return c


def version():
    return lib().zs_hip_version().decode()


class Channel:
    """Fiat-Shamir transcript (src/core/channel/channel.zig), host side."""

    def __init__(self, sep=b"", _handle=None):
        self.h = _handle if _handle is not None else lib().zs_channel_new(bytes(sep), len(sep))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.zs_channel_free(self.h)
            self.h = None

    def clone(self):
        return Channel(_handle=lib().zs_channel_clone(self.h))

    def absorb(self, data):
        data = bytes(data)
        lib().zs_channel_absorb(self.h, data, len(data))

    def sample_bytes(self, n):
        out = ctypes.create_string_buffer(n)
        lib().zs_channel_sample_bytes(self.h, out, n)
        return out.raw

    def sample_qm31(self):
        o = np.zeros(4, np.uint32)
        lib().zs_channel_sample_qm31(self.h, _p(o))
        return o

    def sample_index(self, n):
        return lib().zs_channel_sample_index(self.h, n)


class Tree:
    """A device-resident MerkleTree (src/core/merkle/merkle.zig)."""

    def __init__(self, handle):
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.zs_tree_free(self.h)
            self.h = None

    @property
    def log_leaves(self):
        return lib().zs_tree_log_leaves(self.h)

    def root(self):
        out = ctypes.create_string_buffer(32)
        _check(lib().zs_tree_root(self.h, out))
        return out.raw

    def open(self, index):
        n = self.log_leaves
        out = ctypes.create_string_buffer(max(1, 32 * n))
        _check(lib().zs_merkle_open(self.h, index, out))
        return out.raw[: 32 * n]

    def level(self, level):
        cnt = (1 << self.log_leaves) >> level
        out = ctypes.create_string_buffer(32 * cnt)
        _check(lib().zs_tree_read_level(self.h, level, out))
        return out.raw


class FriResult:
    def __init__(self, handle, dlog, L, ctx):
        self.h, self.dlog, self.L, self.ctx = handle, dlog, L, ctx

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.zs_fri_result_free(self.h)
            self.h = None

    def layer(self, i):
        o = np.zeros(4 << (self.dlog - i), np.uint32)
        _check(lib().zs_fri_read_layer(self.h, i, _p(o)), self.ctx)
        return o.reshape(-1, 4)

    def query(self, channel, num_queries):
        """fri.zig:236-256: indices, revealed pair values and both paths per layer."""
        idx = (sz * max(1, num_queries))()
        vals = np.zeros(max(1, num_queries * self.L * 8), np.uint32)
        per_q = sum(2 * (self.dlog - i) for i in range(self.L)) * 32
        paths = ctypes.create_string_buffer(max(1, per_q * num_queries))
        _check(lib().zs_fri_query(self.h, channel.h, num_queries, idx, _p(vals), paths), self.ctx)
        return list(idx)[:num_queries], vals[: num_queries * self.L * 8], paths.raw[: per_q * num_queries]


def _check(rc, ctx=None):
    if rc != ZS_OK:
        msg = ""
        if ctx is not None:
            msg = lib().zs_hip_last_error(ctx.h).decode(errors="replace")
        raise ZsError(rc, msg)


class Job(ctypes.Structure):
    """zs_job: one trace of a batch (device input and outputs, host results)."""
    _fields_ = [("d_cols", vp), ("n_cols", u32), ("log_n", u32), ("d_lde", vp), ("d_nodes", vp),
                ("trace_root", ctypes.c_uint8 * 32), ("fri_roots", vp), ("remainder", vp), ("status", i32)]


class Request(ctypes.Structure):
    _fields_ = [("n_cols", u32), ("log_n", u32), ("log_blowup", u32), ("remainder_log", u32),
                ("seed", ctypes.c_uint64), ("host_cols", u32p)]


class Response(ctypes.Structure):
    _fields_ = [("status", i32), ("fri_layers", u32), ("trace_root", ctypes.c_uint8 * 32),
                ("fri_roots", ctypes.c_uint8 * (28 * 32)), ("remainder", u32 * (4 * 64)),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("device_bytes", ctypes.c_uint64)]

    def result(self, log_blowup, remainder_log):
        L = self.fri_layers
        rc = 1 << (remainder_log - log_blowup)
        return (bytes(self.trace_root), [bytes(self.fri_roots[32 * i:32 * (i + 1)]) for i in range(L)],
                np.array(self.remainder[:4 * rc], np.uint32).reshape(rc, 4))


def stark_verify(proof, air, channel):
    """GenericStark(Air).verify (src/m31/stark.zig:660-780) of GenericStark.Proof wire
    bytes on the host; `air` is a zs_air ctypes structure.  True or False;
    ZsError(ZS_ERR_INPUT) for bytes that do not deserialize."""
    ok = ctypes.c_uint8()
    _check(lib().zs_stark_verify(bytes(proof), len(proof), ctypes.byref(air), channel.h, ctypes.byref(ok)))
    return bool(ok.value)


def fri_verify(proof, channel):
    """Fri.verify (src/m31/fri.zig:277-336) of Fri.Proof wire bytes, on the host;
    `channel` is advanced as the reference's verify advances it.  True or False;
    ZsError(ZS_ERR_INPUT) when the bytes do not deserialize."""
    ok = ctypes.c_uint8()
    _check(lib().zs_fri_verify(bytes(proof), len(proof), channel.h, ctypes.byref(ok)))
    return bool(ok.value)


def fib_stark_verify(proof, a0, a1, claimed, channel):
    """GenericStark(FibSeededAir).verify (src/m31/stark.zig:660-780) of the proof's
    wire bytes, on the host.  True or False; ZsError(ZS_ERR_INPUT) for bytes that
    do not deserialize."""
    ok = ctypes.c_uint8()
    q = lambda v: _p(_u32(v).reshape(4))
    _check(lib().zs_fib_stark_verify(bytes(proof), len(proof), q(a0), q(a1), q(claimed), channel.h,
                                     ctypes.byref(ok)))
    return bool(ok.value)


def fib_trace(a0, a1, log_n):
    """Seeded Fibonacci AIR trace: (2, 2^log_n, 4) QM31 columns (stark.zig:918-935)."""
    out = np.zeros((2, 1 << log_n, 4), np.uint32)
    rc = lib().zs_fib_trace(a0, a1, log_n, _p(out))
    if rc:
        raise ZsError(rc, "zs_fib_trace")
    return out


class Server:
    """Continuous-batching request loop with HBM-budget admission (zs_server_*)."""

    def __init__(self, device=0, n_workers=4, hbm_budget=200 * 10**9):
        self.h = lib().zs_server_new(device, n_workers, hbm_budget)
        if not self.h:
            raise ZsError(ZS_ERR_DEVICE, "zs_server_new")

    def submit(self, n_cols, log_n, log_blowup=2, remainder_log=5, seed=0, host_cols=None):
        r = Request(n_cols, log_n, log_blowup, remainder_log, seed, None)
        keep = None
        if host_cols is not None:
            keep = _u32(host_cols).reshape(-1)
            r.host_cols = _p(keep)
        t = ctypes.c_uint64()
        rc = lib().zs_server_submit(self.h, ctypes.byref(r), ctypes.byref(t))
        if rc:
            raise ZsError(rc, "zs_server_submit")
        return t.value

    def wait(self, ticket):
        out = Response()
        rc = lib().zs_server_wait(self.h, ticket, ctypes.byref(out))
        if rc:
            raise ZsError(rc, "zs_server_wait")
        return out

    def poll(self, ticket):
        """The response if the request is answered, else None."""
        out, ready = Response(), i32()
        rc = lib().zs_server_poll(self.h, ticket, ctypes.byref(out), ctypes.byref(ready))
        if rc:
            raise ZsError(rc, "zs_server_poll")
        return out if ready.value else None

    def memory(self):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        lib().zs_server_memory(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib().zs_server_free(self.h)
            self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()


_ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)
_FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)


class HostAllocator(ctypes.Structure):
    """zs_host_allocator (zig-capi.h:35-39) backed by the C library's malloc/free."""
    _fields_ = [("ctx", ctypes.c_void_p), ("alloc", _ALLOC_FN), ("free", _FREE_FN)]


_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.free.argtypes = [ctypes.c_void_p]
_alloc_cb = _ALLOC_FN(lambda ctx, n: _libc.malloc(max(int(n), 1)))
_free_cb = _FREE_FN(lambda ctx, ptr, n: _libc.free(ptr))
HOST = HostAllocator(None, _alloc_cb, _free_cb)


class Context:
    """A zs_hip_ctx on one device (circlefft_gpu.zig:20-39's lazily-initialised context)."""

    def __init__(self, device=0):
        h = vp()
        rc = lib().zs_hip_init(device, 0, ctypes.byref(h))
        if rc != ZS_OK:
            raise ZsError(rc, "zs_hip_init: no usable HIP device")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().zs_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def check(self, rc):
        _check(rc, self)

    def set_stream(self, stream_ptr):
        self.check(lib().zs_hip_set_stream(self.h, stream_ptr))

    def synchronize(self):
        self.check(lib().zs_hip_synchronize(self.h))

    def profile(self, enable=True):
        """Bracket every hot kernel launch with HIP events (clears earlier records)."""
        self.check(lib().zs_hip_profile(self.h, 1 if enable else 0))

    def profile_read(self):
        """{kernel name: (total ms, launches)} of the launches recorded since profile(True)."""
        cap, mx = 1 << 16, 256
        names = ctypes.create_string_buffer(cap)
        ms = (ctypes.c_double * mx)()
        cnt = (u32 * mx)()
        n = u32()
        self.check(lib().zs_hip_profile_read(self.h, names, cap, ms, cnt, mx, ctypes.byref(n)))
        keys = names.raw.split(b"\0")[: n.value]
        return {k.decode(): (ms[i], cnt[i]) for i, k in enumerate(keys)}

    # ---- circle FFT (circlefft_gpu.zig:46, :111) ----
    def circle_fft(self, coeffs, half_coset=None):
        coeffs = _u32(coeffs)
        n = coeffs.size
        hc = half_coset if half_coset is not None else canonic_half(n.bit_length() - 1)
        out = np.zeros(n, np.uint32)
        self.check(lib().zs_m31_circle_fft(self.h, _p(coeffs), n, ctypes.byref(hc), _p(out)))
        return out

    def circle_ifft(self, evals, half_coset=None):
        evals = _u32(evals)
        n = evals.size
        hc = half_coset if half_coset is not None else canonic_half(n.bit_length() - 1)
        out = np.zeros(n, np.uint32)
        self.check(lib().zs_m31_circle_ifft(self.h, _p(evals), n, ctypes.byref(hc), _p(out)))
        return out

    def circle_fft_batch_dev(self, d_in, d_out, log_n, n_cols, col_stride):
        self.check(lib().zs_m31_circle_fft_batch_dev(self.h, d_in, d_out, log_n, n_cols, col_stride))

    def circle_ifft_batch_dev(self, d_in, d_out, log_n, n_cols, col_stride):
        self.check(lib().zs_m31_circle_ifft_batch_dev(self.h, d_in, d_out, log_n, n_cols, col_stride))

    # ---- LDE + row commitment ----
    def lde_commit(self, cols, log_blowup, want_lde=True):
        """cols: (n_cols, 2^log_n) array. Returns (Tree, root, lde or None)."""
        cols = _u32(cols)
        n_cols, n = cols.shape
        lg = n.bit_length() - 1
        lde = np.zeros((n_cols, n << log_blowup), np.uint32) if want_lde else None
        t = vp()
        root = ctypes.create_string_buffer(32)
        self.check(lib().zs_m31_lde_commit(self.h, _p(cols), n_cols, lg, log_blowup, ctypes.byref(t), root,
                                           _p(lde) if lde is not None else None))
        return Tree(t), root.raw, lde

    def lde_commit_dev(self, d_cols, n_cols, log_n, log_blowup, d_lde, d_nodes, flags=0):
        self.check(lib().zs_m31_lde_commit_dev(self.h, d_cols, n_cols, log_n, log_blowup, d_lde, d_nodes, flags))

    def lde_commit_dev_noalloc(self, d_cols, n_cols, log_n, log_blowup, d_lde, d_nodes):
        """zs_m31_lde_commit_dev_noalloc: the same commitment with no device allocation
        (fixed per-GPU array; shapes beyond it raise ZsError(ZS_ERR_INPUT))."""
        self.check(lib().zs_m31_lde_commit_dev_noalloc(self.h, d_cols, n_cols, log_n, log_blowup, d_lde, d_nodes))

    def lde_dev(self, d_cols, n_cols, log_n, log_blowup, d_lde):
        self.check(lib().zs_m31_lde_dev(self.h, d_cols, n_cols, log_n, log_blowup, d_lde))

    def row_leaves_dev(self, d_cols, n_cols, log_rows, col_stride, d_leaves):
        self.check(lib().zs_m31_row_leaves_dev(self.h, d_cols, n_cols, log_rows, col_stride, d_leaves))

    def row_commit_dev(self, d_cols, n_cols, log_rows, col_stride, d_nodes, flags=0):
        self.check(lib().zs_m31_row_commit_dev(self.h, d_cols, n_cols, log_rows, col_stride, d_nodes, flags))

    def fill_m31_dev(self, seed, d_out, n):
        self.check(lib().zs_fill_m31_dev(self.h, seed, d_out, n))

    def autotune_lde(self, n_cols, log_n, log_blowup):
        """Pick the fastest launch configuration of each kind of LDE pass for this
        shape (kept on the context); returns {"interp": (tile elements log2,
        columns per workgroup), "mid": (...), "eval": (...), "us": one LDE}."""
        out = np.zeros(7, np.uint32)
        self.check(lib().zs_autotune_lde(self.h, n_cols, log_n, log_blowup, _p(out)))
        o = [int(x) for x in out]
        return {"interp": (o[0], o[1]), "mid": (o[2], o[3]), "eval": (o[4], o[5]), "us": o[6]}

    # ---- GenericStark trace commitment (stark.zig:343-414) ----
    def trace_commit_dev(self, d_cols, n_cols, trace_log, log_blowup, d_codewords, d_nodes, channel=None):
        """Returns the n_cols column roots (absorbed into `channel` in order if given)."""
        roots = ctypes.create_string_buffer(32 * n_cols)
        self.check(lib().zs_qm31_trace_commit_dev(self.h, d_cols, n_cols, trace_log, log_blowup, d_codewords,
                                                  d_nodes, roots, channel.h if channel is not None else None))
        return [roots.raw[32 * c:32 * (c + 1)] for c in range(n_cols)]

    # ---- one trace's commitment, and batches of uneven traces ----
    def commit_job_dev(self, d_cols, n_cols, log_n, log_blowup, remainder_log, d_lde, d_nodes, flags=0):
        """LDE + row tree + gamma from the trace root + QM31 combination + FRI commit.
        Returns (trace_root, fri_roots, remainder)."""
        L = log_n + log_blowup - remainder_log
        root = ctypes.create_string_buffer(32)
        roots = ctypes.create_string_buffer(32 * L)
        rem = np.zeros(4 << (remainder_log - log_blowup), np.uint32)
        self.check(lib().zs_commit_job_dev(self.h, d_cols, n_cols, log_n, log_blowup, remainder_log, d_lde, d_nodes,
                                           root, roots, _p(rem), flags))
        return root.raw, [roots.raw[32 * i:32 * (i + 1)] for i in range(L)], rem.reshape(-1, 4)

    def commit_batch_dev(self, specs, log_blowup, remainder_log, n_workers=4, flags=0):
        """specs: list of (d_cols, n_cols, log_n, d_lde, d_nodes).  Returns per job
        (trace_root, fri_roots, remainder)."""
        jobs = (Job * len(specs))()
        keep = []
        for k, (d_cols, n_cols, log_n, d_lde, d_nodes) in enumerate(specs):
            L = log_n + log_blowup - remainder_log
            fr = ctypes.create_string_buffer(32 * L)
            rem = np.zeros(4 << (remainder_log - log_blowup), np.uint32)
            keep.append((fr, rem, L))
            jobs[k].d_cols, jobs[k].n_cols, jobs[k].log_n = d_cols, n_cols, log_n
            jobs[k].d_lde, jobs[k].d_nodes = d_lde, d_nodes
            jobs[k].fri_roots = ctypes.cast(fr, vp)
            jobs[k].remainder = rem.ctypes.data
        self.check(lib().zs_commit_batch_dev(self.h, jobs, len(specs), log_blowup, remainder_log, n_workers, flags))
        return [(bytes(jobs[k].trace_root), [keep[k][0].raw[32 * i:32 * (i + 1)] for i in range(keep[k][2])],
                 keep[k][1].reshape(-1, 4)) for k in range(len(specs))]

    def fri_prove_dev(self, d_codeword, log_size, log_blowup, remainder_log, num_queries, channel):
        """proveCodeword on a device codeword; returns the Fri.Proof wire bytes."""
        ptr, n = ctypes.c_void_p(), sz()
        self.check(lib().zs_fri_prove_dev(self.h, d_codeword, log_size, log_blowup, remainder_log, num_queries,
                                          channel.h, HOST, ctypes.byref(ptr), ctypes.byref(n)))
        try:
            return ctypes.string_at(ptr.value, n.value)
        finally:
            lib().zs_free(HOST, ptr, n)

    def stark_prove_dev(self, air, d_pre, d_trace, trace_log, log_blowup, num_queries, remainder_log, channel):
        """GenericStark(Air).proveWithPreprocessed on device columns (`air`: a zs_air ctypes
        structure; d_pre: device pointer or None); returns the wire bytes."""
        ptr, n = ctypes.c_void_p(), sz()
        self.check(lib().zs_stark_prove_dev(self.h, ctypes.byref(air), d_pre, d_trace, trace_log, log_blowup,
                                            num_queries, remainder_log, channel.h, HOST, ctypes.byref(ptr),
                                            ctypes.byref(n)))
        try:
            return ctypes.string_at(ptr.value, n.value)
        finally:
            lib().zs_free(HOST, ptr, n)

    def fib_stark_prove_dev(self, d_trace, trace_log, log_blowup, num_queries, remainder_log, a0, a1, claimed,
                            channel):
        """GenericStark(FibAir) proof of a device trace (2 QM31 columns); returns the wire bytes."""
        q = lambda v: _p(_u32(v).reshape(4))
        av, bv, cv = _u32(a0).reshape(4), _u32(a1).reshape(4), _u32(claimed).reshape(4)
        ptr, n = ctypes.c_void_p(), sz()
        self.check(lib().zs_fib_stark_prove_dev(self.h, d_trace, trace_log, log_blowup, num_queries, remainder_log,
                                                _p(av), _p(bv), _p(cv), channel.h, HOST, ctypes.byref(ptr),
                                                ctypes.byref(n)))
        try:
            return ctypes.string_at(ptr.value, n.value)
        finally:
            lib().zs_free(HOST, ptr, n)

    # ---- FRI (fri.zig:157-233) ----
    def fri_commit(self, codeword, log_blowup, remainder_log, alphas=None, channel=None, keep=True):
        cw = _u32(codeword).reshape(-1)
        n0 = cw.size // 4
        dlog = n0.bit_length() - 1
        L = dlog - remainder_log
        roots = ctypes.create_string_buffer(32 * L)
        rem = np.zeros(4 << (remainder_log - log_blowup), np.uint32)
        al = _u32(alphas).reshape(-1) if alphas is not None else None
        res = vp()
        self.check(lib().zs_fri_commit(self.h, _p(cw), dlog, log_blowup, remainder_log,
                                       _p(al) if al is not None else None,
                                       channel.h if channel is not None else None, roots, _p(rem),
                                       ctypes.byref(res) if keep else None))
        fr = FriResult(res, dlog, L, self) if keep else None
        return [roots.raw[32 * i:32 * (i + 1)] for i in range(L)], rem.reshape(-1, 4), fr

    def fri_commit_dev(self, d_codeword, dlog, log_blowup, remainder_log, alphas=None, channel=None, keep=False,
                       flags=0):
        L = dlog - remainder_log
        roots = ctypes.create_string_buffer(32 * L)
        rem = np.zeros(4 << (remainder_log - log_blowup), np.uint32)
        al = _u32(alphas).reshape(-1) if alphas is not None else None
        res = vp()
        self.check(lib().zs_fri_commit_dev(self.h, d_codeword, dlog, log_blowup, remainder_log,
                                           _p(al) if al is not None else None,
                                           channel.h if channel is not None else None, flags, roots, _p(rem),
                                           ctypes.byref(res) if keep else None))
        fr = FriResult(res, dlog, L, self) if keep else None
        return [roots.raw[32 * i:32 * (i + 1)] for i in range(L)], rem.reshape(-1, 4), fr
