"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. The product path never imports it.
The C sources beside this file restate the reference algorithm; each cites the
reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

Q = 1152921504606830593
Q36 = 68718428161  # MOD_SWITCH_TARGET_36BIT, src/pir/mod_switch.rs:47


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    return ctypes.CDLL(_LIB)


_lib = _load()
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class _Pack(ctypes.Structure):
    _fields_ = [("opaque", ctypes.c_uint8 * 512)]


class _Ntt(ctypes.Structure):
    _fields_ = [("opaque", ctypes.c_uint8 * 256)]


def _p64(a):
    return a.ctypes.data_as(_u64p)


_lib.orc_pack_init.argtypes = [ctypes.POINTER(_Pack), ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t]
_lib.orc_pack_free.argtypes = [ctypes.POINTER(_Pack)]
_lib.orc_respond.argtypes = [ctypes.POINTER(_Pack), _u8p, _u64p, _u64p, _u64p, _u64p, _u64p, _u64p, _u64p]
_lib.orc_store_build.argtypes = [ctypes.POINTER(_Pack), _u64p, _u64p]
_lib.orc_fold_b.argtypes = [ctypes.POINTER(_Pack), _u64p, _u64p, _u64p]
_lib.orc_automorph_table.argtypes = [ctypes.POINTER(_Pack), ctypes.c_size_t, _u32p]
_lib.orc_expand_w_mask.argtypes = [ctypes.POINTER(_Pack), _u8p, _u64p]
_lib.orc_mod_switch.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64, _u64p]
_lib.orc_encode_response.argtypes = [ctypes.c_size_t, ctypes.c_uint64, _u64p, _u64p, ctypes.c_size_t, _u8p]
_lib.orc_encode_response.restype = ctypes.c_size_t
_lib.orc_expand_seed.argtypes = [_u8p, ctypes.c_size_t, ctypes.c_uint64, _u64p]
_lib.orc_gauss_stream_hash.argtypes = [ctypes.c_double, ctypes.c_uint64, ctypes.c_size_t]
_lib.orc_gauss_stream_hash.restype = ctypes.c_uint64
_lib.orc_gauss_centered_hash_seed7.argtypes = [ctypes.c_size_t, ctypes.c_uint64]
_lib.orc_gauss_centered_hash_seed7.restype = ctypes.c_uint64
_lib.orc_ntt_init.argtypes = [ctypes.POINTER(_Ntt), ctypes.c_size_t, ctypes.c_uint64]
_lib.orc_ntt_free.argtypes = [ctypes.POINTER(_Ntt)]
_lib.orc_ntt_forward.argtypes = [ctypes.POINTER(_Ntt), _u64p]
_lib.orc_ntt_inverse.argtypes = [ctypes.POINTER(_Ntt), _u64p]
_lib.orc_mont_mul.argtypes = [ctypes.POINTER(_Ntt), ctypes.c_uint64, ctypes.c_uint64]
_lib.orc_mont_mul.restype = ctypes.c_uint64
_lib.orc_negacyclic_mul_naive.argtypes = [ctypes.c_size_t, ctypes.c_uint64, _u64p, _u64p, _u64p]
_lib.orc_chacha_block.argtypes = [_u32p, ctypes.c_uint64, ctypes.c_uint64, _u32p]


def _seed(seed):
    buf = (ctypes.c_uint8 * 32).from_buffer_copy(bytes(seed))
    return ctypes.cast(buf, _u8p), buf


class Ntt:
    """NttContext over DEFAULT_Q (src/math/ntt.rs)."""

    def __init__(self, n, q=Q):
        self._c = _Ntt()
        if _lib.orc_ntt_init(ctypes.byref(self._c), n, q) != 0:
            raise ValueError("no primitive root")
        self.n, self.q = n, q

    def __del__(self):
        _lib.orc_ntt_free(ctypes.byref(self._c))

    def forward(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        _lib.orc_ntt_forward(ctypes.byref(self._c), _p64(a))
        return a

    def inverse(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        _lib.orc_ntt_inverse(ctypes.byref(self._c), _p64(a))
        return a

    def mont_mul(self, a, b):
        return _lib.orc_mont_mul(ctypes.byref(self._c), a, b)


def negacyclic_mul_naive(a, b, q=Q):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros_like(a)
    _lib.orc_negacyclic_mul_naive(len(a), q, _p64(a), _p64(b), _p64(out))
    return out


class Pack:
    """PackParams + offline keys for one (ring_dim, gamma, gadget) (inspiring2.rs:209-286)."""

    def __init__(self, n, gamma, gadget_len=3, dropped=0):
        self._p = _Pack()
        rc = _lib.orc_pack_init(ctypes.byref(self._p), n, gamma, gadget_len, dropped)
        if rc != 0:
            raise ValueError(f"pack init failed {rc}")
        self.n, self.gamma, self.rows = n, gamma, gadget_len - dropped

    def __del__(self):
        _lib.orc_pack_free(ctypes.byref(self._p))

    def automorph_table(self, t):
        out = np.zeros(self.n, dtype=np.uint32)
        _lib.orc_automorph_table(ctypes.byref(self._p), t, out.ctypes.data_as(_u32p))
        return out

    def w_mask(self, w_seed):
        out = np.zeros((self.rows, self.n), dtype=np.uint64)
        s, _keep = _seed(w_seed)
        _lib.orc_expand_w_mask(ctypes.byref(self._p), s, _p64(out))
        return out

    def store_build(self, columns):
        columns = np.ascontiguousarray(columns, dtype=np.uint64)
        out = np.zeros((self.gamma, self.n), dtype=np.uint64)
        _lib.orc_store_build(ctypes.byref(self._p), _p64(columns), _p64(out))
        return out

    def fold_b(self, columns, b):
        columns = np.ascontiguousarray(columns, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        out = np.zeros(self.n, dtype=np.uint64)
        _lib.orc_fold_b(ctypes.byref(self._p), _p64(columns), _p64(b), _p64(out))
        return out

    def respond(self, w_seed, a, b, columns, y_body, sums_ntt=None):
        """(a_hat, b) of pack_shard (respond.rs:770-863); sums_ntt selects the store path."""
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        columns = np.ascontiguousarray(columns, dtype=np.uint64)
        y_body = np.ascontiguousarray(y_body, dtype=np.uint64)
        a_hat = np.zeros(self.n, dtype=np.uint64)
        b_out = np.zeros(self.n, dtype=np.uint64)
        s, _keep = _seed(w_seed)
        sp = None
        if sums_ntt is not None:
            sums_ntt = np.ascontiguousarray(sums_ntt, dtype=np.uint64)
            sp = _p64(sums_ntt)
        _lib.orc_respond(ctypes.byref(self._p), s, _p64(a), _p64(b), _p64(columns), _p64(y_body), sp,
                         _p64(a_hat), _p64(b_out))
        return a_hat, b_out


def mod_switch(values, q_old=Q, q_new=Q36):
    v = np.ascontiguousarray(values, dtype=np.uint64)
    out = np.zeros_like(v)
    _lib.orc_mod_switch(_p64(v), len(v), q_old, q_new, _p64(out))
    return out


def encode_response(a_hat, b, retained, modulus=Q):
    a_hat = np.ascontiguousarray(a_hat, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    n = len(a_hat)
    size = _lib.orc_encode_response(n, modulus, _p64(a_hat), _p64(b), retained, None)
    buf = (ctypes.c_uint8 * size)()
    _lib.orc_encode_response(n, modulus, _p64(a_hat), _p64(b), retained, ctypes.cast(buf, _u8p))
    return bytes(buf)


def expand_seed(seed, n, q=Q):
    out = np.zeros(n, dtype=np.uint64)
    s, _keep = _seed(seed)
    _lib.orc_expand_seed(s, n, q, _p64(out))
    return out


def gauss_stream_hash(sigma, seed, count):
    return _lib.orc_gauss_stream_hash(sigma, seed, count)


def gauss_centered_hash_seed7(count, q=Q):
    return _lib.orc_gauss_centered_hash_seed7(count, q)


def chacha_block(key_words, counter, stream):
    k = np.ascontiguousarray(key_words, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    _lib.orc_chacha_block(k.ctypes.data_as(_u32p), counter, stream, out.ctypes.data_as(_u32p))
    return out


def fnv1a(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for x in data:
        h ^= x
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h
