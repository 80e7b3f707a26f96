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
        L.inspire_ctx_attach_crs.argtypes = [vp, vp, cp, sz]
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
        L.inspire_respond_batch_device_noalloc.argtypes = [vp, vp, vp, vp, vp, sz, u64, vp, vp, vp, vp, cp, sz]
        L.inspire_respond_batch_device_timed.argtypes = [vp, vp, vp, vp, vp, vp, sz, u64, vp, vp, vp, vp, cp, sz]
        L.inspire_workspace_trace.argtypes = [vp, sz, cp, sz]
        L.inspire_workspace_kernel_times.argtypes = [vp, vp, vp, vp, sz, vp, cp, sz]
        L.inspire_server_create.argtypes = [vp, sz, sz, pp, cp, sz]
        L.inspire_server_destroy.argtypes = [vp]
        L.inspire_server_register_seat.argtypes = [vp, vp, u32, u32, ctypes.POINTER(u64), cp, sz]
        L.inspire_server_unregister_seat.argtypes = [vp, u64, cp, sz]
        L.inspire_server_submit.argtypes = [vp, u64, vp, vp, vp, sz, u64, ctypes.POINTER(u64), cp, sz]
        L.inspire_server_wait.argtypes = [vp, u64, vp, vp, ctypes.POINTER(ctypes.c_double), cp, sz]
        L.inspire_server_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64),
                                           ctypes.POINTER(u64)]
        L.inspire_workspace_status.argtypes = [vp, sz, vp, cp, sz]
        L.inspire_response_wire_bytes.argtypes = [vp, u64]
        L.inspire_response_wire_bytes.restype = sz
        L.inspire_respond_batch_device_wire.argtypes = [vp, vp, vp, vp, vp, vp, sz, u64, vp, sz, vp, cp, sz]
        L.inspire_respond_batch_wire.argtypes = [vp, vp, vp, vp, vp, sz, u64, vp, sz, cp, sz]
        L.inspire_respond_graph_create.argtypes = [vp, vp, vp, vp, vp, vp, sz, u64, vp, vp, pp, cp, sz]
        L.inspire_graph_launch.argtypes = [vp, vp, cp, sz]
        L.inspire_graph_destroy.argtypes = [vp]
        L.inspire_ctx_autotune.argtypes = [vp, vp, vp, sz, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(u32), cp, sz]
        L.inspire_ctx_set_scan_splits.argtypes = [vp, u32, cp, sz]
        L.inspire_ctx_autotune_scan.argtypes = [vp, vp, vp, sz, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(u32),
                                                cp, sz]
        smp = ctypes.POINTER(SmParams)
        L.inspire_sm_params_validate.argtypes = [smp, cp, sz]
        L.inspire_sm_server_create.argtypes = [smp, pp, cp, sz]
        L.inspire_sm_server_destroy.argtypes = [vp]
        L.inspire_sm_table_create.argtypes = [vp, vp, sz, pp, cp, sz]
        L.inspire_sm_table_rows.argtypes = [vp]
        L.inspire_sm_table_rows.restype = u32
        L.inspire_sm_table_resident_bytes.argtypes = [vp]
        L.inspire_sm_table_resident_bytes.restype = sz
        L.inspire_sm_table_destroy.argtypes = [vp]
        L.inspire_sm_seat_create.argtypes = [vp, vp, vp, vp, pp, cp, sz]
        L.inspire_sm_seat_resident_bytes.argtypes = [vp]
        L.inspire_sm_seat_resident_bytes.restype = sz
        L.inspire_sm_seat_destroy.argtypes = [vp]
        L.inspire_sm_respond.argtypes = [vp, vp, vp, vp, vp, vp, vp, cp, sz]
        L.inspire_sm_respond_batch.argtypes = [vp, sz, vp, vp, vp, vp, vp, vp, cp, sz]
        L.inspire_sm_respond_batch_device.argtypes = [vp, sz, vp, vp, vp, vp, vp, vp, vp, vp, cp, sz]
        L.inspire_sm_response_wire_bytes.argtypes = [vp]
        L.inspire_sm_response_wire_bytes.restype = sz
        L.inspire_sm_respond_batch_wire.argtypes = [vp, sz, vp, vp, vp, vp, vp, sz, cp, sz]
        L.inspire_sm_respond_batch_device_wire.argtypes = [vp, sz, vp, vp, vp, vp, vp, vp, sz, vp, cp, sz]
        L.inspire_sm_server_trace.argtypes = [vp, sz, cp, sz]
        L.inspire_sm_server_kernel_times.argtypes = [vp, vp, sz, vp, cp, sz]
        L.inspire_sm_batcher_create.argtypes = [vp, sz, sz, pp, cp, sz]
        L.inspire_sm_batcher_destroy.argtypes = [vp]
        L.inspire_sm_batcher_register_key.argtypes = [vp, vp, vp, vp, ctypes.POINTER(u64), cp, sz]
        L.inspire_sm_batcher_unregister_key.argtypes = [vp, u64, cp, sz]
        L.inspire_sm_batcher_submit.argtypes = [vp, u64, vp, vp, vp, sz, ctypes.POINTER(u64), cp, sz]
        L.inspire_sm_batcher_wait.argtypes = [vp, u64, vp, vp, ctypes.POINTER(ctypes.c_double), cp, sz]
        L.inspire_sm_batcher_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)]
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

    def trace(self, launches):
        """Record scan and chain kernel events on the next `launches` device launches."""
        _call(lib().inspire_workspace_trace, self._h, launches)
        self._traced = launches

    def kernel_times(self):
        """(group_ms, scan_ms, chain_ms) arrays of the traced launches; waits for them."""
        cap = getattr(self, "_traced", 0)
        group, scan, chain = np.zeros(cap, np.float32), np.zeros(cap, np.float32), np.zeros(cap, np.float32)
        n = ctypes.c_size_t()
        _call(lib().inspire_workspace_kernel_times, self._h, _ptr(group), _ptr(scan), _ptr(chain), cap,
              ctypes.byref(n))
        self._traced = 0
        return group[:n.value], scan[:n.value], chain[:n.value]

    def close(self):
        if self._h:
            lib().inspire_workspace_destroy(self._h)
            self._h = None


class Context:
    """Per-CRS packing state resident on the GPU (ServerInspiringCache analogue)."""

    def __init__(self, params, w_seed, num_columns):
        """w_seed None: a context without a CRS, for shards and their packing stores only
        (ShardData::precompute_packing); attach_crs binds one before any respond."""
        h = ctypes.c_void_p()
        seed = None if w_seed is None else (ctypes.c_uint8 * 32).from_buffer_copy(bytes(w_seed))
        _call(lib().inspire_ctx_create, ctypes.byref(params), seed, num_columns, ctypes.byref(h))
        self._h = h
        self.n = params.ring_dim
        self.gamma = num_columns
        self.key_rows = lib().inspire_ctx_key_rows(h)
        self.shards = {}

    def attach_crs(self, w_seed):
        """ServerInspiringCache::new(&crs, &encoded) over a context created without a CRS."""
        seed = (ctypes.c_uint8 * 32).from_buffer_copy(bytes(w_seed))
        _call(lib().inspire_ctx_attach_crs, self._h, seed)
        return self

    def set_queries_per_cu(self, k):
        _call(lib().inspire_ctx_set_queries_per_cu, self._h, k)

    def set_scan_splits(self, k):
        """Row blocks per shard of the DB scan: 8, 16 or 32 at ring 2048 and up."""
        _call(lib().inspire_ctx_set_scan_splits, self._h, k)

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

    def response_bytes(self, switch_modulus=0):
        """Length of one ServerResponse in bincode at this context's shape."""
        return lib().inspire_response_wire_bytes(self._h, switch_modulus)

    def respond_seeded_batch_wire(self, seat, shard_ids, seeds, bs, switch_modulus=0):
        """respond_seeded_batch with each response's to_binary bytes packed on the device."""
        ids = np.ascontiguousarray(shard_ids, dtype=np.uint32)
        count = len(ids)
        sd = np.ascontiguousarray(np.frombuffer(b"".join(bytes(s) for s in seeds), dtype=np.uint8))
        b = np.ascontiguousarray(bs, dtype=np.uint64).reshape(count, self.n)
        size = self.response_bytes(switch_modulus)
        out = np.zeros((count, size), dtype=np.uint8)
        _call(lib().inspire_respond_batch_wire, self._h, seat._h, _ptr(ids), _ptr(sd), _ptr(b), count, switch_modulus,
              _ptr(out), size)
        return [out[k].tobytes() for k in range(count)]

    def respond_batch_device_wire(self, seat, ws, d_ids, d_seeds, d_b, count, d_wire, wire_stride, stream=0,
                                  switch_modulus=0):
        """Device-pointer batch writing response bytes at d_wire + k * wire_stride."""
        _call(lib().inspire_respond_batch_device_wire, self._h, seat._h, ws._h, ctypes.c_void_p(d_ids),
              ctypes.c_void_p(d_seeds), ctypes.c_void_p(d_b), count, switch_modulus, ctypes.c_void_p(d_wire),
              wire_stride, ctypes.c_void_p(stream))

    def respond_batch_device(self, seat, ws, d_ids, d_seeds, d_b, count, d_a_out, d_b_out, stream=0,
                             switch_modulus=0):
        """Device-pointer batch (integers are device addresses); enqueued, not waited for."""
        _call(lib().inspire_respond_batch_device, self._h, seat._h, ws._h, ctypes.c_void_p(d_ids),
              ctypes.c_void_p(d_seeds), ctypes.c_void_p(d_b), count, switch_modulus, ctypes.c_void_p(d_a_out),
              ctypes.c_void_p(d_b_out), ctypes.c_void_p(stream))


def respond_batch_device_noalloc(ctx, seat, d_ids, d_seeds, d_b, count, d_a_out, d_b_out, d_status, stream=0,
                                 switch_modulus=0):
    """inspire_respond_batch_device with no workspace: the library's per-GPU scratch array
    holds the scan partials and d_status receives the per-query status (count int32)."""
    _call(lib().inspire_respond_batch_device_noalloc, ctx._h, seat._h, ctypes.c_void_p(d_ids),
          ctypes.c_void_p(d_seeds), ctypes.c_void_p(d_b), count, switch_modulus, ctypes.c_void_p(d_a_out),
          ctypes.c_void_p(d_b_out), ctypes.c_void_p(d_status), ctypes.c_void_p(stream))


NOALLOC_SCRATCH_BYTES = 32 << 20


class RespondGraph:
    """The respond pipeline captured once into a hipGraph over fixed device buffers; launch()
    replays it on a stream."""

    def __init__(self, ctx, seat, ws, d_ids, d_seeds, d_b, count, d_a_out, d_b_out, switch_modulus=0):
        h = ctypes.c_void_p()
        _call(lib().inspire_respond_graph_create, ctx._h, seat._h, ws._h, ctypes.c_void_p(d_ids),
              ctypes.c_void_p(d_seeds), ctypes.c_void_p(d_b), count, switch_modulus, ctypes.c_void_p(d_a_out),
              ctypes.c_void_p(d_b_out), ctypes.byref(h))
        self._h = h

    def launch(self, stream=0):
        _call(lib().inspire_graph_launch, self._h, ctypes.c_void_p(stream))

    def close(self):
        if self._h:
            lib().inspire_graph_destroy(self._h)
            self._h = None


def autotune(ctx, seat, ws, count):
    """inspire_ctx_autotune: times 1, 2 and 3 queries per CU at `count` queries, keeps the fastest."""
    ms = (ctypes.c_float * 3)()
    chosen = ctypes.c_uint32()
    _call(lib().inspire_ctx_autotune, ctx._h, seat._h, ws._h, count, ms, ctypes.byref(chosen))
    return {"queries_per_cu": chosen.value, "ms_qpc1": round(ms[0], 4), "ms_qpc2": round(ms[1], 4),
            "ms_qpc3": round(ms[2], 4)}


def autotune_scan(ctx, seat, ws, count):
    """inspire_ctx_autotune_scan: times the DB scan split into 8, 16 and 32 row blocks per shard at
    `count` queries (uniform shard ids), keeps the fastest."""
    ms = (ctypes.c_float * 3)()
    chosen = ctypes.c_uint32()
    _call(lib().inspire_ctx_autotune_scan, ctx._h, seat._h, ws._h, count, ms, ctypes.byref(chosen))
    return {"scan_splits": chosen.value, "scan_ms_ks8": round(ms[0], 4), "scan_ms_ks16": round(ms[1], 4),
            "scan_ms_ks32": round(ms[2], 4)}


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
        """Launches and queries so far, and the seated keys: resident_bytes on the device (each
        key word with its Shoup companion) and payload_bytes in the reference's accounting
        (PackingKeySeat::payload_bytes, src/pir/session.rs:630-636: rows x ring_dim x 8 per key)."""
        l, q, r, k = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        lib().inspire_server_stats(self._h, ctypes.byref(l), ctypes.byref(q), ctypes.byref(r), ctypes.byref(k))
        return {"launches": l.value, "queries": q.value, "resident_bytes": r.value, "seats": k.value,
                "payload_bytes": k.value * self.ctx.key_rows * self.ctx.n * 8}

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


# ---------------------------------------------------------------------------------------
# Special-modulus set (src/special_modulus): Server / ShardTable / KeySeat on the GPU, and
# the host-side wire codecs of wire.rs.
#
#     SmServer(params)                       Server::new / with_kernel   server.rs:145-166
#     SmServer.table(cells)                  Server::table               server.rs:205-250
#     SmServer.seat(key)                     Server::seat                server.rs:177-203
#     SmServer.respond(table, seat, query)   Server::respond             server.rs:252-279
#     SmServer.respond_batch(...)            one launch for many (table, seat, query)
#     PackingKey / Query / Response          wire.rs to_bytes / from_bytes
# ---------------------------------------------------------------------------------------

SM_Q = 536_813_569
SM_P = 2_147_389_441
SM_KEY_MASK_VERSION = 2
SM_SEED_BYTES = 32


class SmParams(ctypes.Structure):
    """special_modulus::Params (src/special_modulus/mod.rs:62-97)."""

    _fields_ = [
        ("ring_dim", ctypes.c_uint32),
        ("gamma", ctypes.c_uint32),
        ("q", ctypes.c_uint64),
        ("special_prime", ctypes.c_uint64),
        ("cell_bytes", ctypes.c_uint32),
        ("record_bytes", ctypes.c_uint32),
        ("sigma", ctypes.c_double),
        ("key_mask_version", ctypes.c_uint16),
    ]

    def body_coefficients(self):
        return 2 * self.record_bytes

    def key_bytes(self):
        return 2 + SM_SEED_BYTES + _packed_bytes(self.ring_dim, self.q) + _packed_bytes(self.ring_dim, self.special_prime)

    def query_bytes(self):
        return SM_SEED_BYTES + _packed_bytes(self.ring_dim, self.q)

    def response_bytes(self):
        return _packed_bytes(self.ring_dim, self.q) + _packed_bytes(self.body_coefficients(), self.q)

    def validate(self):
        _call(lib().inspire_sm_params_validate, ctypes.byref(self))

    def __eq__(self, other):
        return isinstance(other, SmParams) and all(getattr(self, f) == getattr(other, f) for f, _ in self._fields_)


def ring2048_cell256():
    """Params::ring2048_cell256 (mod.rs:84-97), the served set."""
    return SmParams(2048, 512, SM_Q, SM_P, 256, 246, 6.4, SM_KEY_MASK_VERSION)


def _bits(m):
    return int(m - 1).bit_length()


def _packed_bytes(count, m):
    return (count * _bits(m) + 7) // 8


def pack_tight(values, m):
    """pack_coefficients_tight at one modulus (src/math/poly.rs:166-224): bit b of value i at
    bit i*bits + b, least significant first."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    bad = np.nonzero(v >= m)[0]
    if bad.size:
        i = int(bad[0])
        raise PirError(f"tight coefficient {int(v[i])} at index {i} is not canonical for modulus {m}")
    bits = _bits(m)
    bitmat = ((v[:, None] >> np.arange(bits, dtype=np.uint64)) & 1).astype(np.uint8).reshape(-1)
    out = np.packbits(bitmat, bitorder="little")
    return out.tobytes()[: _packed_bytes(v.size, m)]


def unpack_tight(data, m, count):
    """unpack_coefficients_tight (src/math/poly.rs:227-281): refuses a wrong length, set
    padding bits and a value at or above the modulus."""
    bits = _bits(m)
    want = (count * bits + 7) // 8
    if len(data) != want:
        raise PirError(f"tight coefficient payload has {len(data)} bytes, expected {want}")
    raw = np.frombuffer(bytes(data), dtype=np.uint8)
    used = (count * bits) % 8
    if used and int(raw[-1]) & (0xFF ^ ((1 << used) - 1)):
        raise PirError("tight coefficient payload has nonzero unused high bits")
    bitv = np.unpackbits(raw, bitorder="little")[: count * bits].reshape(count, bits).astype(np.uint64)
    vals = (bitv << np.arange(bits, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
    bad = np.nonzero(vals >= m)[0]
    if bad.size:
        i = int(bad[0])
        raise PirError(f"tight coefficient {int(vals[i])} at index {i} is not canonical for modulus {m}")
    return vals.astype(np.uint32)


def _split(data, want, what):
    if len(data) != want:
        raise PirError(f"{what} has {len(data)} bytes, expected {want}")


class PackingKey:
    """wire::PackingKey (src/special_modulus/wire.rs:33-86)."""

    def __init__(self, version, mask_seed, q, p):
        self.version, self.mask_seed = int(version), bytes(mask_seed)
        self.q = np.ascontiguousarray(q, dtype=np.uint32)
        self.p = np.ascontiguousarray(p, dtype=np.uint32)

    def to_bytes(self, params):
        return (self.version.to_bytes(2, "big") + self.mask_seed + pack_tight(self.q, params.q)
                + pack_tight(self.p, params.special_prime))

    @classmethod
    def from_bytes(cls, params, data):
        params.validate()
        d = params.ring_dim
        _split(data, params.key_bytes(), "packing key")
        version = int.from_bytes(data[:2], "big")
        if version != params.key_mask_version:
            raise PirError(f"packing key is under key-mask version {version}; this set serves version "
                           f"{params.key_mask_version}")
        seed = bytes(data[2:2 + SM_SEED_BYTES])
        rest = data[2 + SM_SEED_BYTES:]
        at = _packed_bytes(d, params.q)
        return cls(version, seed, unpack_tight(rest[:at], params.q, d), unpack_tight(rest[at:], params.special_prime, d))


class Query:
    """wire::Query (wire.rs:88-117): the mask seed, then the body at q."""

    def __init__(self, seed, body):
        self.seed = bytes(seed)
        self.body = np.ascontiguousarray(body, dtype=np.uint32)

    def to_bytes(self, params):
        return self.seed + pack_tight(self.body, params.q)

    @classmethod
    def from_bytes(cls, params, data):
        _split(data, params.query_bytes(), "query")
        return cls(data[:SM_SEED_BYTES], unpack_tight(data[SM_SEED_BYTES:], params.q, params.ring_dim))


class Response:
    """wire::Response (wire.rs:119-142): the whole mask, then the body coefficients."""

    def __init__(self, mask, body):
        self.mask = np.ascontiguousarray(mask, dtype=np.uint32)
        self.body = np.ascontiguousarray(body, dtype=np.uint32)

    def to_bytes(self, params):
        return pack_tight(self.mask, params.q) + pack_tight(self.body, params.q)

    @classmethod
    def from_bytes(cls, params, data):
        at = _packed_bytes(params.ring_dim, params.q)
        _split(data, params.response_bytes(), "response")
        return cls(unpack_tight(data[:at], params.q, params.ring_dim),
                   unpack_tight(data[at:], params.q, params.body_coefficients()))

    def __eq__(self, other):
        return np.array_equal(self.mask, other.mask) and np.array_equal(self.body, other.body)


class SmTable:
    """ShardTable (server.rs:97-127): the step-1 table and the cells, both in HBM."""

    def __init__(self, server, handle):
        self._srv, self._h = server, handle

    def rows(self):
        return lib().inspire_sm_table_rows(self._h)

    def resident_bytes(self):
        return lib().inspire_sm_table_resident_bytes(self._h)

    def close(self):
        if self._h:
            lib().inspire_sm_table_destroy(self._h)
            self._h = None


class SmSeat:
    """KeySeat (server.rs:66-92): the key and its mask, both limbs, transform form, in HBM."""

    def __init__(self, server, handle):
        self._srv, self._h = server, handle

    def bytes(self):
        return lib().inspire_sm_seat_resident_bytes(self._h)

    def close(self):
        if self._h:
            lib().inspire_sm_seat_destroy(self._h)
            self._h = None


class SmServer:
    """special_modulus::Server on the GPU (server.rs:129-279)."""

    def __init__(self, params):
        h = ctypes.c_void_p()
        self.params = params
        _call(lib().inspire_sm_server_create, ctypes.byref(params), ctypes.byref(h))
        self._h = h

    def close(self):
        if self._h:
            lib().inspire_sm_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def table(self, cells):
        data = np.frombuffer(bytes(cells), dtype=np.uint8)
        h = ctypes.c_void_p()
        _call(lib().inspire_sm_table_create, self._h, _ptr(data), data.size, ctypes.byref(h))
        return SmTable(self, h)

    def seat(self, key):
        d = self.params.ring_dim
        if key.q.size != d or key.p.size != d:
            raise PirError(f"packing key needs {d} coefficients per limb")
        seed = np.frombuffer(key.mask_seed, dtype=np.uint8).copy()
        h = ctypes.c_void_p()
        _call(lib().inspire_sm_seat_create, self._h, _ptr(seed), _ptr(key.q), _ptr(key.p), ctypes.byref(h))
        return SmSeat(self, h)

    def respond(self, table, seat, query):
        masks, bodies = self.respond_batch([table], [seat], [query])
        return Response(masks[0], bodies[0])

    def respond_batch(self, tables, seats, queries):
        """count queries, each under its own table and seat, in one launch."""
        count = len(queries)
        d, nb = self.params.ring_dim, self.params.body_coefficients()
        if count == 0:
            return np.zeros((0, d), np.uint32), np.zeros((0, nb), np.uint32)
        for qy in queries:
            if qy.body.size != d:
                raise PirError("query does not fit the set")
        th = (ctypes.c_void_p * count)(*[t._h.value for t in tables])
        sh = (ctypes.c_void_p * count)(*[s._h.value for s in seats])
        seeds = np.frombuffer(b"".join(qy.seed for qy in queries), dtype=np.uint8).copy()
        bodies = np.ascontiguousarray(np.stack([qy.body for qy in queries]), dtype=np.uint32)
        masks = np.zeros((count, d), np.uint32)
        outs = np.zeros((count, nb), np.uint32)
        _call(lib().inspire_sm_respond_batch, self._h, count, th, sh, _ptr(seeds), _ptr(bodies), _ptr(masks),
              _ptr(outs))
        return masks, outs

    def response_bytes(self):
        """Length of one Response::to_bytes."""
        return lib().inspire_sm_response_wire_bytes(self._h)

    def respond_batch_wire(self, tables, seats, queries):
        """respond_batch returning each query's Response::to_bytes, packed on the device."""
        count = len(queries)
        if count == 0:
            return []
        for qy in queries:
            if qy.body.size != self.params.ring_dim:
                raise PirError("query does not fit the set")
        th = (ctypes.c_void_p * count)(*[t._h.value for t in tables])
        sh = (ctypes.c_void_p * count)(*[s._h.value for s in seats])
        seeds = np.frombuffer(b"".join(qy.seed for qy in queries), dtype=np.uint8).copy()
        bodies = np.ascontiguousarray(np.stack([qy.body for qy in queries]), dtype=np.uint32)
        size = self.response_bytes()
        out = np.zeros((count, size), np.uint8)
        _call(lib().inspire_sm_respond_batch_wire, self._h, count, th, sh, _ptr(seeds), _ptr(bodies), _ptr(out),
              size)
        return [out[k].tobytes() for k in range(count)]

    def respond_batch_device_wire(self, tables, seats, d_seeds, d_bodies, d_scratch, d_wire, wire_stride, stream=0):
        """Device-resident inputs, scratch (count x ring_dim words) and response bytes."""
        count = len(tables)
        th = (ctypes.c_void_p * count)(*[t._h.value for t in tables])
        sh = (ctypes.c_void_p * count)(*[s._h.value for s in seats])
        _call(lib().inspire_sm_respond_batch_device_wire, self._h, count, th, sh, ctypes.c_void_p(d_seeds),
              ctypes.c_void_p(d_bodies), ctypes.c_void_p(d_scratch), ctypes.c_void_p(d_wire), wire_stride,
              ctypes.c_void_p(stream))

    def trace(self, launches):
        """Record respond-kernel events on the next `launches` device launches."""
        _call(lib().inspire_sm_server_trace, self._h, launches)
        self._traced = launches

    def kernel_times(self):
        """Respond-kernel durations (ms) of the traced launches; waits for the last one."""
        cap = getattr(self, "_traced", 0)
        out = np.zeros(cap, np.float32)
        n = ctypes.c_size_t()
        _call(lib().inspire_sm_server_kernel_times, self._h, _ptr(out), cap, ctypes.byref(n))
        self._traced = 0
        return out[:n.value]

    def respond_batch_device(self, tables, seats, d_seeds, d_bodies, d_mask, d_body, stream=0, timed=False):
        """Device-resident inputs and outputs (addresses); returns the kernel ms when timed."""
        count = len(tables)
        th = (ctypes.c_void_p * count)(*[t._h.value for t in tables])
        sh = (ctypes.c_void_p * count)(*[s._h.value for s in seats])
        ms = ctypes.c_float(0.0)
        _call(lib().inspire_sm_respond_batch_device, self._h, count, th, sh, ctypes.c_void_p(d_seeds),
              ctypes.c_void_p(d_bodies), ctypes.c_void_p(d_mask), ctypes.c_void_p(d_body), ctypes.c_void_p(stream),
              ctypes.byref(ms) if timed else None)
        return ms.value if timed else None


class SmBatcher:
    """Continuous-batching queue over one SmServer: submit() from any thread, wait() for the
    (masks, bodies, latency_ms) of a ticket."""

    def __init__(self, server, max_launch=2048, hbm_budget=0):
        h = ctypes.c_void_p()
        _call(lib().inspire_sm_batcher_create, server._h, max_launch, hbm_budget, ctypes.byref(h))
        self._h, self.server = h, server
        self.n = server.params.ring_dim
        self.nb = server.params.body_coefficients()
        self._counts = {}
        self._lock = __import__("threading").Lock()

    def register_key(self, key):
        seed = np.frombuffer(key.mask_seed, dtype=np.uint8).copy()
        h = ctypes.c_uint64()
        _call(lib().inspire_sm_batcher_register_key, self._h, _ptr(seed), _ptr(key.q), _ptr(key.p), ctypes.byref(h))
        return h.value

    def unregister_key(self, handle):
        _call(lib().inspire_sm_batcher_unregister_key, self._h, handle)

    def submit(self, handle, tables, queries):
        count = len(queries)
        th = (ctypes.c_void_p * count)(*[t._h.value for t in tables])
        seeds = np.frombuffer(b"".join(qy.seed for qy in queries), dtype=np.uint8).copy()
        bodies = np.ascontiguousarray(np.stack([qy.body for qy in queries]), dtype=np.uint32)
        if bodies.shape[1] != self.n:
            raise PirError("query does not fit the set")
        t = ctypes.c_uint64()
        _call(lib().inspire_sm_batcher_submit, self._h, handle, th, _ptr(seeds), _ptr(bodies), count, ctypes.byref(t))
        with self._lock:
            self._counts[t.value] = count
        return t.value

    def wait(self, ticket):
        with self._lock:
            count = self._counts.pop(ticket)
        masks = np.zeros((count, self.n), np.uint32)
        bodies = np.zeros((count, self.nb), np.uint32)
        lat = ctypes.c_double()
        _call(lib().inspire_sm_batcher_wait, self._h, ticket, _ptr(masks), _ptr(bodies), ctypes.byref(lat))
        return masks, bodies, lat.value

    def stats(self):
        a, q, r = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        lib().inspire_sm_batcher_stats(self._h, ctypes.byref(a), ctypes.byref(q), ctypes.byref(r))
        return {"launches": a.value, "queries": q.value, "resident_bytes": r.value}

    def close(self):
        if self._h:
            lib().inspire_sm_batcher_destroy(self._h)
            self._h = None
