"""ctypes binding of libinspire_gpu.so (include/inspire_gpu.h).

Thin host mirror of the reference's server respond API (raven-inspire):

    Context(params, w_seed, num_columns)   ServerInspiringCache::new        src/pir/respond.rs:926
    Context.upload_columns / upload_records ShardData                        src/pir/setup.rs:280
    Shard.precompute_packing()             ShardData::precompute_packing    src/pir/setup.rs:335
    Context.seat(y_body, num_to_pack)      PackingKeySeat::from_wire        src/pir/session.rs:592
    Context.respond_seeded_batch(...)      respond_seeded_batch             src/pir/respond.rs:1189
    Context.respond_seeded_inspiring(...)  respond_seeded_inspiring         src/pir/respond.rs:1218

Errors raise PirError carrying the library's message, as the reference returns
Err(PirError(String)). There is no CPU fallback: without the built library (or
without a GPU) every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libinspire_gpu.so")

Q = 1152921504606830593
Q36 = 68718428161

OK, ERR_INVALID, ERR_DEVICE, ERR_STORE = 0, 1, 2, 3


class PirError(Exception):
    """The reference's PirError(String) (src/pir/error.rs:6)."""

    def __init__(self, message, code=ERR_INVALID):
        super().__init__(message)
        self.code = code


class InspireParams(ctypes.Structure):
    _fields_ = [
        ("ring_dim", ctypes.c_uint32),
        ("q", ctypes.c_uint64),
        ("gadget_base", ctypes.c_uint64),
        ("packing_gadget_len", ctypes.c_uint32),
        ("approximate_packing_gadget", ctypes.c_uint32),
    ]


def params_d2048(approximate=False, ring_dim=2048):
    """InspireParams::d2048 / d2048_approximate_gadget (src/params.rs:350-400), respond fields."""
    return InspireParams(ring_dim, Q, 1 << 20, 3, 1 if approximate else 0)


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, u32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint64
        cp = ctypes.c_char_p
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.inspire_ctx_create.argtypes = [ctypes.POINTER(InspireParams), vp, u32, pp, cp, sz]
        L.inspire_ctx_destroy.argtypes = [vp]
        L.inspire_ctx_key_rows.argtypes = [vp]
        L.inspire_ctx_key_rows.restype = u32
        L.inspire_ctx_set_queries_per_cu.argtypes = [vp, u32, cp, sz]
        L.inspire_shard_upload_columns.argtypes = [vp, u32, vp, u32, pp, cp, sz]
        L.inspire_shard_upload_records.argtypes = [vp, u32, vp, sz, sz, pp, cp, sz]
        L.inspire_shard_precompute_packing.argtypes = [vp, vp, cp, sz]
        L.inspire_shard_has_store.argtypes = [vp]
        L.inspire_shard_write_value.argtypes = [vp, u32, u32, ctypes.c_uint16, cp, sz]
        L.inspire_shard_resident_bytes.argtypes = [vp]
        L.inspire_shard_resident_bytes.restype = sz
        L.inspire_shard_destroy.argtypes = [vp]
        L.inspire_seat_create.argtypes = [vp, vp, u32, u32, pp, cp, sz]
        L.inspire_seat_resident_bytes.argtypes = [vp]
        L.inspire_seat_resident_bytes.restype = sz
        L.inspire_seat_destroy.argtypes = [vp]
        L.inspire_respond_batch.argtypes = [vp, vp, vp, vp, vp, sz, u64, vp, vp, cp, sz]
        L.inspire_respond_batch_expanded.argtypes = [vp, vp, vp, vp, vp, sz, u64, vp, vp, cp, sz]
        L.inspire_workspace_create.argtypes = [vp, sz, pp, cp, sz]
        L.inspire_workspace_destroy.argtypes = [vp]
        L.inspire_respond_batch_device.argtypes = [vp, vp, vp, vp, vp, vp, sz, u64, vp, vp, vp, cp, sz]
        L.inspire_respond_batch_device_timed.argtypes = [vp, vp, vp, vp, vp, vp, sz, u64, vp, vp, vp, vp, cp, sz]
        L.inspire_server_create.argtypes = [vp, sz, sz, pp, cp, sz]
        L.inspire_server_destroy.argtypes = [vp]
        L.inspire_server_register_seat.argtypes = [vp, vp, u32, u32, ctypes.POINTER(u64), cp, sz]
        L.inspire_server_unregister_seat.argtypes = [vp, u64, cp, sz]
        L.inspire_server_submit.argtypes = [vp, u64, vp, vp, vp, sz, u64, ctypes.POINTER(u64), cp, sz]
        L.inspire_server_wait.argtypes = [vp, u64, vp, vp, ctypes.POINTER(ctypes.c_double), cp, sz]
        L.inspire_server_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.inspire_workspace_status.argtypes = [vp, sz, vp, cp, sz]
        _lib = L
    return _lib


def _call(fn, *args):
    err = ctypes.create_string_buffer(1024)
    rc = fn(*args, err, ctypes.sizeof(err))
    if rc != OK:
        raise PirError(err.value.decode(errors="replace"), rc)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Shard:
    def __init__(self, ctx, handle, shard_id):
        self._ctx, self._h, self.id = ctx, handle, shard_id

    def precompute_packing(self):
        _call(lib().inspire_shard_precompute_packing, self._ctx._h, self._h)

    def has_store(self):
        return bool(lib().inspire_shard_has_store(self._h))

    def write_value(self, record, column, value):
        _call(lib().inspire_shard_write_value, self._h, record, column, value)

    def resident_bytes(self):
        return lib().inspire_shard_resident_bytes(self._h)

    def close(self):
        if self._h:
            lib().inspire_shard_destroy(self._h)
            self._h = None


class Seat:
    def __init__(self, ctx, handle, num_to_pack):
        self._ctx, self._h, self.num_to_pack = ctx, handle, num_to_pack

    def resident_bytes(self):
        return lib().inspire_seat_resident_bytes(self._h)

    def close(self):
        if self._h:
            lib().inspire_seat_destroy(self._h)
            self._h = None


class Workspace:
    def __init__(self, ctx, max_count):
        h = ctypes.c_void_p()
        _call(lib().inspire_workspace_create, ctx._h, max_count, ctypes.byref(h))
        self._ctx, self._h, self.max_count = ctx, h, max_count

    def status(self, count):
        out = np.zeros(count, dtype=np.int32)
        _call(lib().inspire_workspace_status, self._h, count, _ptr(out))
        return out

    def close(self):
        if self._h:
            lib().inspire_workspace_destroy(self._h)
            self._h = None


class Context:
    """Per-CRS packing state resident on the GPU (ServerInspiringCache analogue)."""

    def __init__(self, params, w_seed, num_columns):
        h = ctypes.c_void_p()
        seed = (ctypes.c_uint8 * 32).from_buffer_copy(bytes(w_seed))
        _call(lib().inspire_ctx_create, ctypes.byref(params), seed, num_columns, ctypes.byref(h))
        self._h = h
        self.n = params.ring_dim
        self.gamma = num_columns
        self.key_rows = lib().inspire_ctx_key_rows(h)
        self.shards = {}

    def set_queries_per_cu(self, k):
        _call(lib().inspire_ctx_set_queries_per_cu, self._h, k)

    def close(self):
        for s in list(self.shards.values()):
            s.close()
        self.shards.clear()
        if self._h:
            lib().inspire_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload_columns(self, shard_id, columns):
        cols = np.ascontiguousarray(columns, dtype=np.uint64)
        h = ctypes.c_void_p()
        _call(lib().inspire_shard_upload_columns, self._h, shard_id, _ptr(cols), cols.shape[0], ctypes.byref(h))
        s = Shard(self, h, shard_id)
        self.shards[shard_id] = s
        return s

    def upload_records(self, shard_id, records, entry_size):
        rec = np.ascontiguousarray(np.frombuffer(bytes(records), dtype=np.uint8))
        h = ctypes.c_void_p()
        _call(lib().inspire_shard_upload_records, self._h, shard_id, _ptr(rec), len(rec) // entry_size, entry_size,
              ctypes.byref(h))
        s = Shard(self, h, shard_id)
        self.shards[shard_id] = s
        return s

    def seat(self, y_body, num_to_pack=None):
        y = np.ascontiguousarray(y_body, dtype=np.uint64)
        h = ctypes.c_void_p()
        ntp = self.gamma if num_to_pack is None else num_to_pack
        _call(lib().inspire_seat_create, self._h, _ptr(y), y.shape[0], ntp, ctypes.byref(h))
        return Seat(self, h, ntp)

    def respond_seeded_batch(self, seat, shard_ids, seeds, bs, switch_modulus=0):
        """Returns (a [count x n], b_prefix [count x gamma]) as uint64 arrays."""
        ids = np.ascontiguousarray(shard_ids, dtype=np.uint32)
        count = len(ids)
        sd = np.ascontiguousarray(np.frombuffer(b"".join(bytes(s) for s in seeds), dtype=np.uint8))
        b = np.ascontiguousarray(bs, dtype=np.uint64).reshape(count, self.n)
        a_out = np.zeros((count, self.n), dtype=np.uint64)
        b_out = np.zeros((count, self.gamma), dtype=np.uint64)
        _call(lib().inspire_respond_batch, self._h, seat._h, _ptr(ids), _ptr(sd), _ptr(b), count, switch_modulus,
              _ptr(a_out), _ptr(b_out))
        return a_out, b_out

    def respond_batch_expanded(self, seat, shard_ids, a_polys, bs, switch_modulus=0):
        """Unseeded queries (ClientQuery): each query's expanded a instead of its seed."""
        ids = np.ascontiguousarray(shard_ids, dtype=np.uint32)
        count = len(ids)
        a = np.ascontiguousarray(a_polys, dtype=np.uint64).reshape(count, self.n)
        b = np.ascontiguousarray(bs, dtype=np.uint64).reshape(count, self.n)
        a_out = np.zeros((count, self.n), dtype=np.uint64)
        b_out = np.zeros((count, self.gamma), dtype=np.uint64)
        _call(lib().inspire_respond_batch_expanded, self._h, seat._h, _ptr(ids), _ptr(a), _ptr(b), count,
              switch_modulus, _ptr(a_out), _ptr(b_out))
        return a_out, b_out

    def respond_inspiring(self, y_body, shard_id, a, b, switch_modulus=0):
        """respond_inspiring (src/pir/respond.rs:736): unseeded query with its key inline."""
        seat = self.seat(y_body)
        try:
            ao, bo = self.respond_batch_expanded(seat, [shard_id], np.asarray(a)[None, :], np.asarray(b)[None, :],
                                                 switch_modulus)
        finally:
            seat.close()
        return ao[0], bo[0]

    def respond_seeded_inspiring(self, y_body, shard_id, seed, b, switch_modulus=0):
        """One query carrying its key inline: the key is seated for this request only."""
        seat = self.seat(y_body)
        try:
            a, bp = self.respond_seeded_batch(seat, [shard_id], [seed], np.asarray(b)[None, :], switch_modulus)
        finally:
            seat.close()
        return a[0], bp[0]

    def respond_batch_device(self, seat, ws, d_ids, d_seeds, d_b, count, d_a_out, d_b_out, stream=0,
                             switch_modulus=0):
        """Device-pointer batch (integers are device addresses); enqueued, not waited for."""
        _call(lib().inspire_respond_batch_device, self._h, seat._h, ws._h, ctypes.c_void_p(d_ids),
              ctypes.c_void_p(d_seeds), ctypes.c_void_p(d_b), count, switch_modulus, ctypes.c_void_p(d_a_out),
              ctypes.c_void_p(d_b_out), ctypes.c_void_p(stream))


class Server:
    """Continuous-batching respond server over a Context (inspire_server_*)."""

    def __init__(self, ctx, max_launch=512, hbm_budget=0):
        h = ctypes.c_void_p()
        _call(lib().inspire_server_create, ctx._h, max_launch, hbm_budget, ctypes.byref(h))
        self._h, self.ctx = h, ctx

    def register_seat(self, y_body, num_to_pack=None):
        """ServerSessionStore::register: returns a handle."""
        y = np.ascontiguousarray(y_body, dtype=np.uint64)
        out = ctypes.c_uint64()
        ntp = self.ctx.gamma if num_to_pack is None else num_to_pack
        _call(lib().inspire_server_register_seat, self._h, _ptr(y), y.shape[0], ntp, ctypes.byref(out))
        return out.value

    def unregister_seat(self, handle):
        _call(lib().inspire_server_unregister_seat, self._h, handle)

    def submit(self, handle, shard_ids, seeds, bs, switch_modulus=0):
        ids = np.ascontiguousarray(shard_ids, dtype=np.uint32)
        sd = np.ascontiguousarray(np.frombuffer(b"".join(bytes(s) for s in seeds), dtype=np.uint8))
        b = np.ascontiguousarray(bs, dtype=np.uint64).reshape(len(ids), self.ctx.n)
        t = ctypes.c_uint64()
        _call(lib().inspire_server_submit, self._h, handle, _ptr(ids), _ptr(sd), _ptr(b), len(ids), switch_modulus,
              ctypes.byref(t))
        return t.value, len(ids)

    def wait(self, ticket):
        """(a [count x n], b_prefix [count x gamma], latency_ms) of a submitted batch."""
        t, count = ticket
        a = np.zeros((count, self.ctx.n), dtype=np.uint64)
        b = np.zeros((count, self.ctx.gamma), dtype=np.uint64)
        lat = ctypes.c_double()
        _call(lib().inspire_server_wait, self._h, t, _ptr(a), _ptr(b), ctypes.byref(lat))
        return a, b, lat.value

    def stats(self):
        l, q, r = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        lib().inspire_server_stats(self._h, ctypes.byref(l), ctypes.byref(q), ctypes.byref(r))
        return {"launches": l.value, "queries": q.value, "resident_bytes": r.value}

    def close(self):
        if self._h:
            lib().inspire_server_destroy(self._h)
            self._h = None


def respond_batch_device_timed(ctx, seat, ws, d_ids, d_seeds, d_b, count, d_a_out, d_b_out, stream=0,
                               switch_modulus=0):
    """As Context.respond_batch_device, waits; returns (scan_ms, chain_ms) from HIP events."""
    ms = (ctypes.c_float * 2)()
    _call(lib().inspire_respond_batch_device_timed, ctx._h, seat._h, ws._h, ctypes.c_void_p(d_ids),
          ctypes.c_void_p(d_seeds), ctypes.c_void_p(d_b), count, switch_modulus, ctypes.c_void_p(d_a_out),
          ctypes.c_void_p(d_b_out), ctypes.c_void_p(stream), ms)
    return ms[0], ms[1]


def encode_response(a_hat, b_prefix, modulus=Q):
    """ServerResponse::to_binary for a packed InspiRING response (src/pir/respond.rs:111-176,
    src/math/poly.rs:166-225, 286-310): bincode fixint LE, tight coefficient packing."""
    a_hat = [int(x) for x in a_hat]
    b_prefix = [int(x) for x in b_prefix]
    bits = (modulus - 1).bit_length()

    def tight(vals):
        acc = 0
        for i, v in enumerate(vals):
            if v >= modulus:
                raise PirError(f"tight coefficient {v} at index {i} is not canonical for modulus {modulus}")
            acc |= v << (i * bits)
        nbytes = (len(vals) * bits + 7) // 8
        return nbytes.to_bytes(8, "little") + acc.to_bytes(nbytes, "little")

    n = len(a_hat)
    out = bytearray()
    out += (1).to_bytes(4, "little")  # Packed
    out += tight(a_hat)
    out += (1).to_bytes(8, "little") + modulus.to_bytes(8, "little")  # moduli
    out += modulus.to_bytes(8, "little") + n.to_bytes(8, "little") + (0).to_bytes(8, "little") + b"\x00"
    out += tight(b_prefix)
    out += len(b_prefix).to_bytes(4, "little")  # retained
    out += (0).to_bytes(8, "little")  # column_ciphertexts
    out += b"\x01" + (0).to_bytes(4, "little")  # Some(PackingMode::Inspiring)
    return bytes(out)
