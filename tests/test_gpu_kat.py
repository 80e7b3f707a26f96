"""The HIP respond path against the reference's recorded response digests.

Every single-modulus cell of tests/golden/packing_seat_response_kat.txt (d2048 at 32, 256,
512 and 1024-byte records, gamma 16..512; d256 at 2..256-byte records, gamma 1..128): the
cell's database, key and query stream come from the oracle's client restatement (itself
pinned by the same file's SESSIONKEY lines), the responses from the GPU, and their
bincode bytes must hash to the recorded FNV-1a digests
(tests/packing_seat_byte_identity_kat.rs:48-146).
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = oracle.load_packing_seat_kat(os.path.join(HERE, "golden", "packing_seat_response_kat.txt"))
TAGS = {"d2048": 2048, "d256-single-l3": 256}
CELLS = sorted(k for k in KAT if k[0] in TAGS)


def _cell_inputs(cell):
    tag, entry, seed = cell
    c = oracle.Cell(TAGS[tag], entry, seed)
    recs = sorted(KAT[cell]["resp"].items())
    queries = [c.query(index) for index, _ in recs]
    return c, recs, queries


def _context(gpu, c, precompute):
    ctx = gpu.Context(gpu.params_d2048(ring_dim=c.n), c.w_seed, c.gamma)
    for shard in range(3):
        sh = ctx.upload_records(shard, c.db[shard * c.n * c.entry:(shard + 1) * c.n * c.entry].tobytes(), c.entry)
        if precompute:
            sh.precompute_packing()
    return ctx


@pytest.mark.parametrize("cell", CELLS, ids=lambda c: f"{c[0]}-e{c[1]}-{c[2]:#x}")
def test_batch_through_a_seat_serves_the_recorded_bytes(gpu, cell):
    c, recs, queries = _cell_inputs(cell)
    print(f"kat cell {cell}: gamma {c.gamma}, {len(recs)} queries", flush=True)
    with _context(gpu, c, precompute=True) as ctx:
        seat = ctx.seat(c.y_body)
        ids = [q[0] for q in queries]
        A, B = ctx.respond_seeded_batch(seat, ids, [q[1] for q in queries], np.stack([q[2] for q in queries]))
        seat.close()
    for k, (index, (length, fnv)) in enumerate(recs):
        wire = gpu.encode_response(A[k], B[k])
        assert (len(wire), oracle.fnv1a(wire)) == (length, fnv), f"{cell} index {index}"


@pytest.mark.parametrize("cell", [c for c in CELLS if c[1] in (2, 32, 256)],
                         ids=lambda c: f"{c[0]}-e{c[1]}-{c[2]:#x}")
def test_inline_key_without_stores_serves_the_recorded_bytes(gpu, cell):
    """respond_seeded_inspiring with the key inline, shards decoded without their stores."""
    c, recs, queries = _cell_inputs(cell)
    print(f"kat cell {cell}: inline, no stores", flush=True)
    with _context(gpu, c, precompute=False) as ctx:
        for (index, (length, fnv)), (shard, qseed, b) in zip(recs, queries):
            a_hat, b_prefix = ctx.respond_seeded_inspiring(c.y_body, shard, qseed, b)
            wire = gpu.encode_response(a_hat, b_prefix)
            assert (len(wire), oracle.fnv1a(wire)) == (length, fnv), f"{cell} index {index}"


def test_served_by_handle_serves_the_recorded_bytes(gpu):
    """The continuous-batching server with the key registered by handle, requests of
    different sizes from two cells' clients mixed."""
    cell = ("d2048", 32, 0x52455300)
    c, recs, queries = _cell_inputs(cell)
    print(f"kat cell {cell}: served by handle", flush=True)
    with _context(gpu, c, precompute=True) as ctx:
        server = gpu.Server(ctx, max_launch=64)
        try:
            h = server.register_seat(c.y_body)
            tickets = [server.submit(h, [q[0] for q in part], [q[1] for q in part], np.stack([q[2] for q in part]))
                       for part in (queries[:1], queries[1:4], queries[4:])]
            got = []
            for t in tickets:
                a, b, _ = server.wait(t)
                got.extend(zip(a, b))
        finally:
            server.close()
    for (index, (length, fnv)), (a, b) in zip(recs, got):
        wire = gpu.encode_response(a, b)
        assert (len(wire), oracle.fnv1a(wire)) == (length, fnv), f"index {index}"


@pytest.mark.parametrize("cell", CELLS, ids=lambda c: f"{c[0]}-e{c[1]}-{c[2]:#x}")
def test_device_packed_response_bytes_are_the_recorded_bytes(gpu, cell):
    """The packing kernel's wire epilogue writes ServerResponse::to_binary itself."""
    c, recs, queries = _cell_inputs(cell)
    print(f"kat cell {cell}: device wire bytes", flush=True)
    with _context(gpu, c, precompute=True) as ctx:
        seat = ctx.seat(c.y_body)
        wires = ctx.respond_seeded_batch_wire(seat, [q[0] for q in queries], [q[1] for q in queries],
                                              np.stack([q[2] for q in queries]))
        seat.close()
    for (index, (length, fnv)), wire in zip(recs, wires):
        assert (len(wire), oracle.fnv1a(wire)) == (length, fnv), f"{cell} index {index}"
