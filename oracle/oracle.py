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
_lib.orc_decompose_digits.argtypes = [ctypes.POINTER(_Pack), _u64p, _u64p]
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
_lib.orc_pack_set_simd.argtypes = [ctypes.POINTER(_Pack), ctypes.c_int]
_lib.orc_ifma_available.restype = ctypes.c_int
_lib.orc_ifma_mont_mul_acc.argtypes = [_u64p, _u64p, _u64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64]
_lib.orc_ifma_mont_mul_acc_gather.argtypes = [_u64p, _u64p, _u64p, _u32p, ctypes.c_size_t, ctypes.c_uint64,
                                              ctypes.c_uint64]


def ifma_available():
    """Whether this host has the AVX-512 F/DQ/IFMA features of the reference's SIMD path."""
    return bool(_lib.orc_ifma_available())


def ifma_mont_mul_acc(acc, a, b, q_inv_neg, q=Q, tab=None):
    """acc + a * b[tab] * 2^-64 mod q through the IFMA kernel (tab None: b itself)."""
    acc = np.array(acc, dtype=np.uint64)
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    if tab is None:
        _lib.orc_ifma_mont_mul_acc(_p64(acc), _p64(a), _p64(b), acc.size, q, q_inv_neg)
    else:
        tab = np.ascontiguousarray(tab, dtype=np.uint32)
        _lib.orc_ifma_mont_mul_acc_gather(_p64(acc), _p64(a), _p64(b), tab.ctypes.data_as(_u32p), acc.size, q,
                                          q_inv_neg)
    return acc


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

    def set_simd(self, on=True):
        """Use the reference's AVX-512 IFMA multiply-accumulates in the recursion (the
        simd-packing-offline feature); returns whether the host runs them."""
        return bool(_lib.orc_pack_set_simd(ctypes.byref(self._p), 1 if on else 0))

    def automorph_table(self, t):
        out = np.zeros(self.n, dtype=np.uint32)
        _lib.orc_automorph_table(ctypes.byref(self._p), t, out.ctypes.data_as(_u32p))
        return out

    def w_mask(self, w_seed):
        out = np.zeros((self.rows, self.n), dtype=np.uint64)
        s, _keep = _seed(w_seed)
        _lib.orc_expand_w_mask(ctypes.byref(self._p), s, _p64(out))
        return out

    def decompose(self, coeffs):
        """The recursion's digit split of canonical coefficients: key_rows x len(coeffs)."""
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        assert c.size == self.n
        out = np.zeros((self.rows, self.n), dtype=np.uint64)
        _lib.orc_decompose_digits(ctypes.byref(self._p), _p64(c), _p64(out))
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


_lib.orc_expand_seed_zone.argtypes = [_u8p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64, _u64p]


def expand_seed_zone(seed, n, zone, q=Q):
    """expand_seed with an explicit acceptance bound on the low product word (test hook: a
    narrow bound forces the rejections the reference's bound makes rare)."""
    out = np.zeros(n, dtype=np.uint64)
    s, _keep = _seed(seed)
    _lib.orc_expand_seed_zone(s, n, q, zone, _p64(out))
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


# ---------------- special modulus (special.c) ----------------

SM_Q = 536_813_569        # src/special_modulus/mod.rs:49
SM_P = 2_147_389_441      # src/special_modulus/mod.rs:52
SM_D, SM_GAMMA, SM_CELL, SM_RECORD, SM_SIGMA = 2048, 512, 256, 246, 6.4  # mod.rs:84-96


class _SmRing(ctypes.Structure):
    _fields_ = [("opaque", ctypes.c_uint8 * 256)]


_i8p = ctypes.POINTER(ctypes.c_int8)
_lib.sm_ring_init.argtypes = [ctypes.POINTER(_SmRing), ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint64,
                              ctypes.c_uint64, ctypes.c_int]
_lib.sm_ring_free.argtypes = [ctypes.POINTER(_SmRing)]
_lib.sm_key_mask.argtypes = [ctypes.POINTER(_SmRing), _u8p, _u32p, _u32p]
_lib.sm_expand_query_mask.argtypes = [ctypes.POINTER(_SmRing), _u8p, _u32p]
_lib.sm_expand_uniform_zone.argtypes = [_u8p, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t, _u32p, _u64p, _u32p]


def sm_expand_uniform_zone(seed, stream, moduli, zones, count):
    """expand_uniform over consecutive moduli of one ChaCha20 stream, each draw at or above its
    zone rejected (test hook: narrow zones force the rejections). Returns limbs x count."""
    m = np.ascontiguousarray(moduli, dtype=np.uint32)
    z = np.ascontiguousarray(zones, dtype=np.uint64)
    out = np.zeros((len(m), count), dtype=np.uint32)
    s, _keep = _seed(seed)
    _lib.sm_expand_uniform_zone(s, stream, len(m), count, m.ctypes.data_as(_u32p), _p64(z), out.ctypes.data_as(_u32p))
    return out


_lib.sm_p_to_q_centred.argtypes = [ctypes.POINTER(_SmRing), ctypes.c_uint32]
_lib.sm_p_to_q_centred.restype = ctypes.c_uint32
_lib.sm_q_to_p_centred.argtypes = [ctypes.POINTER(_SmRing), ctypes.c_uint32]
_lib.sm_q_to_p_centred.restype = ctypes.c_uint32
_lib.sm_mod_down.argtypes = [ctypes.POINTER(_SmRing), ctypes.c_uint32, ctypes.c_uint32]
_lib.sm_mod_down.restype = ctypes.c_uint32
_lib.sm_table.argtypes = [ctypes.POINTER(_SmRing), ctypes.c_size_t, ctypes.c_size_t, _u8p, ctypes.c_size_t,
                          _u32p, _i8p]
_lib.sm_seat.argtypes = [ctypes.POINTER(_SmRing), _u32p, _u32p, _u8p, _u32p, _u32p, _u32p, _u32p]
_lib.sm_respond.argtypes = [ctypes.POINTER(_SmRing), _u32p, _i8p, ctypes.c_size_t, _u32p, _u32p, _u32p, _u32p,
                            _u8p, _u32p, _u32p, _u32p, ctypes.c_size_t]
_lib.sm_respond_avx512.argtypes = _lib.sm_respond.argtypes
_lib.sm_avx512_available.restype = ctypes.c_int
_lib.sm_client_from_seed.argtypes = [ctypes.POINTER(_SmRing), ctypes.c_double, _u8p]
_lib.sm_client_from_seed.restype = ctypes.c_void_p
_lib.sm_client_free.argtypes = [ctypes.c_void_p]
_lib.sm_client_key.argtypes = [ctypes.c_void_p, _u8p, _u32p, _u32p]
_lib.sm_client_query.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u8p, _u32p]
_lib.sm_client_decode.argtypes = [ctypes.c_void_p, _u32p, _u32p, ctypes.c_size_t, _u8p]
_lib.sm_pack.argtypes = [_u32p, ctypes.c_size_t, ctypes.c_uint64, _u8p]
_lib.sm_pack.restype = ctypes.c_size_t
_lib.sm_packed_bytes.argtypes = [ctypes.c_size_t, ctypes.c_uint64]
_lib.sm_packed_bytes.restype = ctypes.c_size_t
_lib.sm_kat_cells.argtypes = [ctypes.c_uint64, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, _u8p]
_lib.sm_ntt_forward.argtypes = [ctypes.c_void_p, _u32p]
_lib.sm_ntt_inverse.argtypes = [ctypes.c_void_p, _u32p]


def _p32(a):
    return a.ctypes.data_as(_u32p)


def _u32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    assert n is None or a.size == n
    return a


def sm_pack(values, m):
    """The tight codec at the bit width of m (src/special_modulus/wire.rs:10-21)."""
    v = _u32(values)
    out = (ctypes.c_uint8 * _lib.sm_packed_bytes(v.size, m))()
    _lib.sm_pack(_p32(v), v.size, m, ctypes.cast(out, _u8p))
    return bytes(out)


def sm_kat_cells(seed, rows, worst=False, cell_bytes=SM_CELL, record_bytes=SM_RECORD):
    out = np.zeros(rows * cell_bytes, dtype=np.uint8)
    _lib.sm_kat_cells(seed, rows, cell_bytes, record_bytes, int(worst), out.ctypes.data_as(_u8p))
    return out


def sm_avx512_available():
    """Kernel::Avx512's feature test (server.rs:50-58) on this CPU."""
    return bool(_lib.sm_avx512_available())


class SmRing:
    """Ring of one special-modulus parameter set (src/special_modulus/mod.rs:256-363), with the
    server's table, seat and respond (server.rs)."""

    def __init__(self, d=SM_D, gamma=SM_GAMMA, q=SM_Q, p=SM_P, tables=True, cell_bytes=None, record_bytes=None):
        self._r = _SmRing()
        if _lib.sm_ring_init(ctypes.byref(self._r), d, gamma, q, p, int(tables)) != 0:
            raise ValueError("special-modulus ring init failed")
        self.d, self.gamma, self.q, self.p = d, gamma, q, p
        self.cell_bytes = cell_bytes if cell_bytes is not None else gamma // 2
        self.record_bytes = record_bytes if record_bytes is not None else self.cell_bytes
        self.nb = 2 * self.record_bytes

    def __del__(self):
        _lib.sm_ring_free(ctypes.byref(self._r))

    def _lq(self):
        # sm_ring = {d, gamma, g, lq, lp, ...}: lq sits after three size_t words
        return ctypes.addressof(self._r) + 3 * 8

    def _lp(self):
        return self._lq() + 48

    def ntt_q(self, a, inverse=False):
        a = _u32(a).copy()
        (_lib.sm_ntt_inverse if inverse else _lib.sm_ntt_forward)(self._lq(), _p32(a))
        return a

    def ntt_p(self, a, inverse=False):
        a = _u32(a).copy()
        (_lib.sm_ntt_inverse if inverse else _lib.sm_ntt_forward)(self._lp(), _p32(a))
        return a

    def key_mask(self, seed):
        wq = np.zeros(self.d, np.uint32)
        wp = np.zeros(self.d, np.uint32)
        s, _k = _seed(seed)
        _lib.sm_key_mask(ctypes.byref(self._r), s, _p32(wq), _p32(wp))
        return wq, wp

    def query_mask(self, seed):
        a = np.zeros(self.d, np.uint32)
        s, _k = _seed(seed)
        _lib.sm_expand_query_mask(ctypes.byref(self._r), s, _p32(a))
        return a

    def p_to_q_centred(self, x):
        return _lib.sm_p_to_q_centred(ctypes.byref(self._r), x)

    def q_to_p_centred(self, c):
        return _lib.sm_q_to_p_centred(ctypes.byref(self._r), c)

    def mod_down(self, xq, xp):
        return _lib.sm_mod_down(ctypes.byref(self._r), xq, xp)

    def table(self, cells):
        """(step1 gamma x d, values d x gamma, rows) of Server::table (server.rs:211-250)."""
        cells = np.ascontiguousarray(np.frombuffer(bytes(cells), np.uint8))
        step1 = np.zeros((self.gamma, self.d), np.uint32)
        values = np.zeros((self.d, self.gamma), np.int8)
        rc = _lib.sm_table(ctypes.byref(self._r), self.cell_bytes, self.record_bytes, cells.ctypes.data_as(_u8p),
                           cells.size, _p32(step1), values.ctypes.data_as(_i8p))
        if rc == -1:
            raise ValueError(f"{cells.size} bytes are not at most {self.d} cells of {self.cell_bytes} bytes")
        if rc < -1:
            raise ValueError(f"cell {-2 - rc} has bytes past its {self.record_bytes}-byte record")
        return step1, values, cells.size // self.cell_bytes

    def seat(self, key_q, key_p, mask_seed):
        outs = [np.zeros(self.d, np.uint32) for _ in range(4)]
        s, _k = _seed(mask_seed)
        kq, kp = _u32(key_q, self.d), _u32(key_p, self.d)
        if _lib.sm_seat(ctypes.byref(self._r), _p32(kq), _p32(kp), s, *[_p32(o) for o in outs]) != 0:
            raise ValueError("packing key coefficient is not a residue")
        return tuple(outs)  # (key_q, key_p, w_q, w_p), transform form, Montgomery operands

    def respond(self, table, seat, seed, body, avx512=False):
        """respond_impl; avx512=True runs the same code compiled for AVX-512F/BW/DQ/VL, the
        reference's Kernel::Avx512 (refused where the CPU lacks those features)."""
        step1, values, rows = table
        mask = np.zeros(self.d, np.uint32)
        out = np.zeros(self.nb, np.uint32)
        s, _k = _seed(seed)
        b = _u32(body, self.d)
        if avx512 and not sm_avx512_available():
            raise RuntimeError("this CPU lacks AVX-512F/BW/DQ/VL")
        fn = _lib.sm_respond_avx512 if avx512 else _lib.sm_respond
        if fn(ctypes.byref(self._r), _p32(step1), values.ctypes.data_as(_i8p), rows,
              *[_p32(x) for x in seat], s, _p32(b), _p32(mask), _p32(out), self.nb) != 0:
            raise ValueError("query does not fit the set")
        return mask, out

    def response_bytes(self, mask, body):
        return sm_pack(mask, self.q) + sm_pack(body, self.q)


class SmClient:
    """Client::from_seed (src/special_modulus/client.rs:137-244) and its queries."""

    def __init__(self, ring, seed, sigma=SM_SIGMA):
        self.ring = ring
        s, _k = _seed(seed)
        self._c = _lib.sm_client_from_seed(ctypes.byref(ring._r), sigma, s)

    def __del__(self):
        _lib.sm_client_free(self._c)

    def packing_key(self):
        d = self.ring.d
        seed = (ctypes.c_uint8 * 32)()
        kq, kp = np.zeros(d, np.uint32), np.zeros(d, np.uint32)
        _lib.sm_client_key(self._c, ctypes.cast(seed, _u8p), _p32(kq), _p32(kp))
        return bytes(seed), kq, kp

    def key_bytes(self, version=2):
        seed, kq, kp = self.packing_key()
        return version.to_bytes(2, "big") + seed + sm_pack(kq, self.ring.q) + sm_pack(kp, self.ring.p)

    def query(self, row):
        seed = (ctypes.c_uint8 * 32)()
        body = np.zeros(self.ring.d, np.uint32)
        if _lib.sm_client_query(self._c, row, ctypes.cast(seed, _u8p), _p32(body)) != 0:
            raise ValueError(f"row {row} is past the {self.ring.d} rows of a shard")
        return bytes(seed), body

    def decode(self, mask, body):
        body = _u32(body)
        out = (ctypes.c_uint8 * (body.size // 2))()
        _lib.sm_client_decode(self._c, _p32(_u32(mask)), _p32(body), body.size, ctypes.cast(out, _u8p))
        return bytes(out)


def _sm_tables(ring):
    """The ring's automorphism tables (gamma x d u16), as the reference searched them."""
    addr = ctypes.addressof(ring._r) + 3 * 8 + 2 * 48 + 5 * 4 + 4
    ptr = ctypes.c_void_p.from_address(addr).value
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint16)), shape=(ring.gamma, ring.d)).copy()


SmRing.tables = _sm_tables


# ---------------- client side of the packing-key KAT cells (client.c) ----------------

_lib.orc_client_key_kat.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint64, _u64p,
                                    ctypes.POINTER(ctypes.c_int64)]
_lib.orc_packing_keys_bytes.argtypes = [ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_size_t, _u8p]
_lib.orc_packing_keys_bytes.restype = ctypes.c_size_t
_lib.orc_cell_new.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint64]
_lib.orc_cell_new.restype = ctypes.c_void_p
_lib.orc_cell_new_gadget.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_size_t]
_lib.orc_cell_new_gadget.restype = ctypes.c_void_p
_lib.orc_client_packing_keys_gadget.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _u64p, _u8p,
                                                ctypes.c_void_p, _u64p]
_lib.orc_seeded_row.argtypes = [ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, _u8p,
                                _u64p]
_lib.orc_gauss_new.argtypes = [ctypes.c_double, ctypes.c_uint64]
_lib.orc_gauss_new.restype = ctypes.c_void_p
_lib.orc_gauss_delete.argtypes = [ctypes.c_void_p]
_lib.orc_sample_poly.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, _u64p]
_lib.orc_chacha_new_u64.argtypes = [ctypes.c_uint64]
_lib.orc_chacha_new_u64.restype = ctypes.c_void_p
_lib.orc_chacha_delete.argtypes = [ctypes.c_void_p]
_lib.orc_chacha_fill_bytes.argtypes = [ctypes.c_void_p, _u8p, ctypes.c_size_t]
_lib.orc_phase.argtypes = [ctypes.c_size_t, ctypes.c_uint64, _u64p, _u64p, _u64p, ctypes.c_size_t, _u64p]
_lib.orc_round_plaintext.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
_lib.orc_round_plaintext.restype = ctypes.c_uint64
_lib.orc_decrypt.argtypes = [ctypes.c_size_t, ctypes.c_uint64, _u64p, _u64p, _u64p, ctypes.c_size_t,
                             ctypes.c_uint64, ctypes.c_uint64, _u64p]
_lib.orc_switch_secret.argtypes = [ctypes.c_size_t, _u64p, ctypes.c_uint64, ctypes.c_uint64, _u64p]
_lib.orc_reconstruct_entry.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_size_t, _u8p]
_lib.orc_cell_free.argtypes = [ctypes.c_void_p]
_lib.orc_cell_info.argtypes = [ctypes.c_void_p, _u8p, _u64p, _u8p, _u64p]
_lib.orc_cell_columns.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _u64p]
_lib.orc_cell_query.argtypes = [ctypes.c_void_p, ctypes.c_uint64, _u32p, _u8p, _u64p]
_lib.orc_cell_gamma.argtypes = [ctypes.c_void_p]
_lib.orc_cell_gamma.restype = ctypes.c_size_t
_lib.orc_apply_automorphism.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint64, _u64p]


def packing_keys_bytes(y_body, gamma):
    """bincode of ClientPackingKeys (exact gadget, coefficient-form rows)."""
    y = np.ascontiguousarray(y_body, dtype=np.uint64)
    rows, n = y.shape
    size = _lib.orc_packing_keys_bytes(n, _p64(y), rows, gamma, None)
    out = (ctypes.c_uint8 * size)()
    _lib.orc_packing_keys_bytes(n, _p64(y), rows, gamma, ctypes.cast(out, _u8p))
    return bytes(out)


def client_key_kat(n, gamma, seed):
    """One client_packing_key_kat cell: (key bytes, the sampler's next sample)."""
    y = np.zeros((3, n), np.uint64)
    nxt = ctypes.c_int64()
    if _lib.orc_client_key_kat(n, gamma, seed, _p64(y), ctypes.byref(nxt)) != 0:
        raise ValueError(f"width {gamma} is not legal at ring {n}")
    return packing_keys_bytes(y, gamma), nxt.value


def apply_automorphism(poly, t, q=Q):
    a = np.ascontiguousarray(poly, dtype=np.uint64)
    out = np.zeros_like(a)
    _lib.orc_apply_automorphism(_p64(a), a.size, t, q, _p64(out))
    return out


class Cell:
    """A packing-key KAT cell (three shards, one client session, a reproducible query
    stream), built in the reference fixture's draw order (tests/packing_key_fixture/mod.rs:89-113).
    approximate=True builds the cell under d2048_approximate_gadget (two key rows)."""

    def __init__(self, n, entry, seed, approximate=False):
        self._c = _lib.orc_cell_new_gadget(n, entry, seed, 1 if approximate else 0)
        if not self._c:
            raise ValueError(f"no cell at ring {n}, entry {entry}")
        self.n, self.entry, self.seed = n, entry, seed
        self.approximate = approximate
        self.rows = 2 if approximate else 3
        self.gamma = _lib.orc_cell_gamma(self._c)
        w = (ctypes.c_uint8 * 32)()
        self.y_body = np.zeros((self.rows, n), np.uint64)
        self.db = np.zeros(3 * n * entry, np.uint8)
        self.secret = np.zeros(n, np.uint64)
        _lib.orc_cell_info(self._c, ctypes.cast(w, _u8p), _p64(self.y_body),
                           self.db.ctypes.data_as(_u8p), _p64(self.secret))
        self.w_seed = bytes(w)

    def __del__(self):
        if getattr(self, "_c", None):
            _lib.orc_cell_free(self._c)
            self._c = None

    def key_bytes(self):
        return packing_keys_bytes(self.y_body, self.gamma)

    def columns(self, shard):
        cols = np.zeros((self.gamma, self.n), np.uint64)
        _lib.orc_cell_columns(self._c, shard, _p64(cols))
        return cols

    def query(self, index):
        """next_query(index): (shard id, row seed, b)."""
        shard = ctypes.c_uint32()
        seed = (ctypes.c_uint8 * 32)()
        b = np.zeros(self.n, np.uint64)
        _lib.orc_cell_query(self._c, index, ctypes.byref(shard), ctypes.cast(seed, _u8p), _p64(b))
        return shard.value, bytes(seed), b


# ---------------- the client's decode (PIR.Extract) ----------------

P = 65537                               # InspireParams::d2048().p, src/params.rs:363
SIGMA = 6.4
MOD_SWITCH_TARGET_45BIT = 35_184_372_060_161   # src/pir/mod_switch.rs:33
IMPLEMENTED_TARGETS = (MOD_SWITCH_TARGET_45BIT, Q36)  # src/pir/mod_switch.rs:53


class Gauss:
    """GaussianSampler::with_seed(sigma, seed) (src/math/gaussian.rs)."""

    def __init__(self, seed, sigma=SIGMA):
        self._g = _lib.orc_gauss_new(sigma, seed)

    def __del__(self):
        if getattr(self, "_g", None):
            _lib.orc_gauss_delete(self._g)
            self._g = None

    def poly(self, n, q=Q):
        """Poly::sample_gaussian_moduli: n draws, each lifted into [0, q)."""
        out = np.zeros(n, np.uint64)
        _lib.orc_sample_poly(self._g, n, q, _p64(out))
        return out


class ChaCha:
    """ChaCha20Rng::seed_from_u64(seed)."""

    def __init__(self, seed):
        self._r = _lib.orc_chacha_new_u64(seed)

    def __del__(self):
        if getattr(self, "_r", None):
            _lib.orc_chacha_delete(self._r)
            self._r = None

    def fill(self, count):
        buf = (ctypes.c_uint8 * count)()
        _lib.orc_chacha_fill_bytes(self._r, ctypes.cast(buf, _u8p), count)
        return bytes(buf)


def client_packing_keys(n, gamma, secret, w_seed, gauss, approximate=False):
    """ClientPackingKeys::generate_once (src/inspiring/inspiring2.rs:794-831): the key rows
    y_body, drawing one error polynomial per kept row from gauss."""
    rows = 2 if approximate else 3
    y = np.zeros((rows, n), np.uint64)
    s_ = np.ascontiguousarray(secret, dtype=np.uint64)
    wp, _wbuf = _seed(w_seed)
    if _lib.orc_client_packing_keys_gadget(n, gamma, 1 if approximate else 0, _p64(s_), wp, gauss._g,
                                           _p64(y)) != 0:
        raise ValueError(f"width {gamma} is not legal at ring {n}")
    return y


def seeded_row(secret, local, rng, gauss):
    """SeededRgswCiphertext::encrypt_with_rng of delta * X^(-local) at a one-row gadget
    (src/rgsw/types.rs:345-386): (row seed, b)."""
    s_ = np.ascontiguousarray(secret, dtype=np.uint64)
    n = s_.size
    seed = (ctypes.c_uint8 * 32)()
    b = np.zeros(n, np.uint64)
    _lib.orc_seeded_row(n, _p64(s_), local, rng._r, gauss._g, ctypes.cast(seed, _u8p), _p64(b))
    return bytes(seed), b


def phase(secret, a, b, count, q=Q):
    """Coefficients 0..count-1 of a*s + b mod (X^n + 1, q); b may hold only count coefficients."""
    s_ = np.ascontiguousarray(secret, dtype=np.uint64)
    a_ = np.ascontiguousarray(a, dtype=np.uint64)
    n = a_.size
    bb = np.zeros(n, np.uint64)
    bb[:min(count, len(b))] = np.asarray(b, dtype=np.uint64)[:count]
    out = np.zeros(n, np.uint64)
    _lib.orc_phase(n, q, _p64(s_), _p64(a_), _p64(bb), count, _p64(out))
    return out[:count]


def round_plaintext(value, delta, p=P):
    """PlaintextRounding::round (src/rlwe/enc.rs:35-43)."""
    return _lib.orc_round_plaintext(value, delta, p)


def switch_secret(secret, q_new, q_old=Q):
    """mod_switch_secret_key (src/pir/mod_switch.rs:221-252)."""
    s_ = np.ascontiguousarray(secret, dtype=np.uint64)
    out = np.zeros_like(s_)
    _lib.orc_switch_secret(s_.size, _p64(s_), q_old, q_new, _p64(out))
    return out


def reconstruct_entry(values, entry):
    """reconstruct_entry (src/pir/encode_db.rs:161-176)."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    out = (ctypes.c_uint8 * entry)()
    _lib.orc_reconstruct_entry(_p64(v), v.size, entry, ctypes.cast(out, _u8p))
    return bytes(out)


def check_mod_switch_noise_budget(n, target, q=Q, p=P, sigma=SIGMA):
    """check_mod_switch_noise_budget (src/pir/mod_switch.rs:111-144), with the target checks of
    ModSwitchParams::new (:67-98). Raises ValueError on a refusal."""
    if target == 0:
        raise ValueError("mod-switch target_modulus must be positive")
    if target > q:
        raise ValueError(f"mod-switch target_modulus {target} exceeds source modulus {q}; "
                         "widening is not a mod-switch")
    if target < p:
        raise ValueError(f"mod-switch target_modulus {target} < plaintext modulus p={p}; "
                         "decryption impossible")
    if target % (2 * n) != 1:
        raise ValueError(f"mod-switch target_modulus {target} is not NTT-friendly: must satisfy "
                         f"q' == 1 (mod 2*d) where d={n}")
    s_max = 7.0 * sigma
    round_noise = 0.5 + n * 0.5 * s_max
    scaled_old = (target / q) * (q / (4.0 * p))
    total = round_noise + scaled_old + float(target % p)
    budget = target / (2.0 * p)
    if total >= budget:
        raise ValueError(f"mod-switch noise-budget violation: total_noise={total:.0f} >= budget={budget:.0f}")
    return budget / total


def extract_inspiring(secret, a, b_prefix, entry, retained=None, modulus=Q, p=P):
    """extract_inspiring (src/pir/extract.rs:229-263) and, for a response on another modulus,
    extract_inspiring_mod_switched (src/pir/mod_switch.rs:573-639): the entry bytes the
    response decodes to under the client's secret (in [0, Q))."""
    a_ = np.ascontiguousarray(a, dtype=np.uint64)
    n = a_.size
    cols = (entry + 1) // 2
    if retained is not None and retained != cols:
        raise ValueError(f"extract_inspiring: packed response retains {retained} coefficients, "
                         f"{cols} required")
    if modulus == Q:
        s_, delta = np.asarray(secret, np.uint64), Q // p
    else:
        if modulus not in IMPLEMENTED_TARGETS:
            raise ValueError(f"response modulus {modulus} is not an implemented mod-switch target")
        check_mod_switch_noise_budget(n, modulus, p=p)
        s_, delta = switch_secret(secret, modulus), modulus // p
    values = [round_plaintext(int(v), delta, p) for v in phase(s_, a_, b_prefix, cols, modulus)]
    return reconstruct_entry(np.array(values, np.uint64), entry)


def served_post_switch_error(secret, a, b_prefix, modulus, record, p=P):
    """served_post_switch_error (src/pir/mod_switch.rs:1027-1059): the largest distance, over the
    response's retained coefficients, between the phase under the switched secret and
    m * floor(q'/p) for the record's column value m."""
    rec = np.frombuffer(bytes(record), np.uint8)
    cols = len(b_prefix)
    msg = rec[0:2 * cols:2].astype(np.uint64) | (rec[1:2 * cols:2].astype(np.uint64) << 8)
    s_ = secret if modulus == Q else switch_secret(secret, modulus)
    noisy = phase(s_, a, b_prefix, cols, modulus)
    dp = modulus // p
    worst = 0
    for i in range(cols):
        d = abs(int(noisy[i]) - int(msg[i]) * dp % modulus)
        worst = max(worst, min(d, modulus - d))
    return worst


def post_switch_error_kat_inputs():
    """The client side of production_post_switch_error_kat_at_45_36_and_33_bits
    (src/pir/mod_switch.rs:1070-1100), in its draw order: setup_with_rng at sampler seed 0x5002
    and setup rng 0x5002_0001 (secret, 11 tree-packing key matrices of 3 error rows, w_seed,
    v_seed), query_seeded's own row error and its inline packing keys, then the row request
    re-encrypted under sampler 0x5002_0002 and rng 0x5002_0003 for record 17 of a 2048 x 512-byte
    database whose byte k is (37 k + 11) mod 256."""
    n, entry, gamma, target = 2048, 512, 256, 17
    g = Gauss(0x5002)
    secret = g.poly(n)
    for _ in range(11 * 3):
        g.poly(n)
    rng = ChaCha(0x5002_0001)
    w_seed = rng.fill(32)
    rng.fill(32)
    g.poly(n)
    y_body = client_packing_keys(n, gamma, secret, w_seed, g)
    seed, b = seeded_row(secret, target, ChaCha(0x5002_0003), Gauss(0x5002_0002))
    db = ((np.arange(n * entry, dtype=np.int64) * 37 + 11) % 256).astype(np.uint8)
    rec = db.reshape(n, entry)
    columns = (rec[:, 0::2].astype(np.uint64) | (rec[:, 1::2].astype(np.uint64) << 8)).T.copy()
    return {"n": n, "entry": entry, "gamma": gamma, "target": target, "secret": secret, "w_seed": w_seed,
            "y_body": y_body, "seed": seed, "b": b, "db": db, "columns": columns,
            "record": rec[target].tobytes()}


# (target modulus, recorded post-switch error) of src/pir/mod_switch.rs:1102-1107
POST_SWITCH_ERROR_KAT = ((MOD_SWITCH_TARGET_45BIT, 317_711), (Q36, 652), (68_719_464_449, 53_503),
                         (8_589_905_921, 36_836))


def parse_response(data):
    """The packed InspiRING ServerResponse bincode (respond.rs:111-176, poly.rs:166-225,
    286-310): {"a", "b_prefix", "modulus", "retained"}. Raises ValueError on another layout."""
    import struct
    at = 0

    def take(fmt):
        nonlocal at
        if at + struct.calcsize(fmt) > len(data):
            raise ValueError("truncated response")
        v = struct.unpack_from("<" + fmt, data, at)[0]
        at += struct.calcsize(fmt)
        return v

    def tight(nbytes, count, w):
        nonlocal at
        raw = np.frombuffer(bytes(data[at:at + nbytes]), np.uint8)
        at += nbytes
        bits = np.unpackbits(raw, bitorder="little")[:count * w].reshape(count, w).astype(np.uint64)
        return (bits << np.arange(w, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)

    if take("I") != 1:
        raise ValueError("not a packed response")
    a_len = take("Q")
    a_at = at
    at += a_len
    if take("Q") != 1:
        raise ValueError("a: one limb expected")
    modulus = take("Q")
    if take("Q") != modulus:
        raise ValueError("a: q differs from its limb")
    n = take("Q")
    if take("Q") != 0 or take("B") != 0:
        raise ValueError("a: crt inverse or NTT flag set")
    w = int(modulus - 1).bit_length()
    b_len = take("Q")
    b_at = at
    at += b_len
    retained = take("I")
    if take("Q") != 0 or take("B") != 1 or take("I") != 0:
        raise ValueError("trailer is not Some(Inspiring) with no column ciphertexts")
    if at != len(data) or a_len != (n * w + 7) // 8 or b_len != (retained * w + 7) // 8:
        raise ValueError("length fields do not match the layout")
    at = a_at
    a = tight(a_len, n, w)
    at = b_at
    b = tight(b_len, retained, w)
    return {"a": a, "b_prefix": b, "modulus": modulus, "retained": retained}


def load_packing_seat_kat(path):
    """Parse packing_seat_response_kat.txt: {(tag, entry, seed): {"key": (len, fnv),
    "resp": {index: (len, fnv)}}}."""
    cells = {}
    for line in open(path):
        f = line.split()
        if not f:
            continue
        kv = dict(x.split("=", 1) for x in f[2:])
        key = (f[1], int(kv["entry"]), int(kv["seed"], 16))
        c = cells.setdefault(key, {"key": None, "resp": {}})
        rec = (int(kv["len"]), int(kv["fnv"], 16))
        if f[0] == "SESSIONKEY":
            c["key"] = rec
        else:
            c["resp"][int(kv["index"])] = rec
    return cells
