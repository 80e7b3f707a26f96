"""The approximate gadget (d2048_approximate_gadget) on the GPU, by the reference's own two
checks of it besides decoding (tests/approximate_packing_gadget.rs:413-466):

* every_responder_serves_one_answer_under_the_approximate_gadget: an approximate cell's
  responses have the exact cell's size and differ from its bytes (a device path that ignored
  the gadget, or drew the approximate mask on the exact stream, would answer the same bytes);
* production_cells_decode_under_the_approximate_gadget: the response from a shard's packing
  store equals the per-query response from its columns. The device always packs through its
  store, so its bytes are compared with the oracle's per-query path (packing_offline from the
  columns, no store), and decoded with the cell's secret.

The rounding and mask-stream rules that make these bytes what they are are pinned on the CPU
(tests/test_approximate_gadget_pins.py).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

D = 2048


def _context(gpu, c):
    ctx = gpu.Context(gpu.params_d2048(approximate=c.approximate), c.w_seed, c.gamma)
    for shard in range(3):
        ctx.upload_records(shard, c.db[shard * c.n * c.entry:(shard + 1) * c.n * c.entry].tobytes(),
                           c.entry).precompute_packing()
    return ctx


def _answers(gpu, c, indices, switch_modulus=0):
    queries = [c.query(i) for i in indices]
    with _context(gpu, c) as ctx:
        seat = ctx.seat(c.y_body)
        wires = ctx.respond_seeded_batch_wire(seat, [q[0] for q in queries], [q[1] for q in queries],
                                              np.stack([q[2] for q in queries]), switch_modulus=switch_modulus)
        a, b = ctx.respond_seeded_batch(seat, [q[0] for q in queries], [q[1] for q in queries],
                                        np.stack([q[2] for q in queries]))
        seat.close()
    return queries, wires, a, b


@pytest.mark.parametrize("entry", [32, 256, 512])
def test_approximate_answers_differ_from_exact_at_equal_size(gpu, entry):
    seed = 0x4150_5800 + entry
    indices = [0, 255, 256, 700, D + 3]
    print(f"approximate vs exact: entry {entry}", flush=True)
    approx = oracle.Cell(D, entry, seed, approximate=True)
    exact = oracle.Cell(D, entry, seed, approximate=False)
    for modulus in (0, oracle.Q36):
        _, wa, _, _ = _answers(gpu, approx, indices, modulus)
        _, we, _, _ = _answers(gpu, exact, indices, modulus)
        for x, y in zip(wa, we):
            assert len(x) == len(y)
            assert x != y


@pytest.mark.parametrize("entry", [32, 256, 512, 1024])
def test_store_answer_equals_the_per_query_answer(gpu, entry):
    c = oracle.Cell(D, entry, 0x4150_5244, approximate=True)
    indices = [0, D - 1, D, 2 * D + 17, 3 * D - 1] if entry <= 512 else [0, 2 * D + 17]
    print(f"approximate store vs per-query: entry {entry}, gamma {c.gamma}", flush=True)
    queries, wires, a, b = _answers(gpu, c, indices)
    p = oracle.Pack(D, c.gamma, 3, 1)
    for k, (shard, seed, bq) in enumerate(queries):
        wa, wb = p.respond(c.w_seed, oracle.expand_seed(seed, D), bq, c.columns(shard), c.y_body)
        assert np.array_equal(a[k], wa), indices[k]
        assert np.array_equal(b[k], wb[:c.gamma]), indices[k]
        assert wires[k] == oracle.encode_response(wa, wb[:c.gamma], c.gamma), indices[k]
        rec = c.db[indices[k] * entry:(indices[k] + 1) * entry].tobytes()
        assert oracle.extract_inspiring(c.secret, a[k], b[k], entry) == rec, indices[k]
