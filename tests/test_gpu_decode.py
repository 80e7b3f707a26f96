"""GPU responses decoded by the client, and the reference's post-switch error known answers
on the device path.

* production_post_switch_error_kat_at_45_36_and_33_bits (src/pir/mod_switch.rs:1070-1114): the
  KAT's key and row request (replayed by the oracle's client restatement, itself pinned by
  these same numbers on the CPU) go through the device respond at d2048 with 512-byte records
  (gamma 256), with and without the shard's packing store; the device's own modulus switch
  must leave exactly the recorded errors 317,711 / 652 / 53,503 / 36,836 at the four targets.
* production_cells_decode_under_the_approximate_gadget and
  production_cells_decode_after_the_served_switch (tests/approximate_packing_gadget.rs:435-494):
  device responses of exact and approximate-gadget cells decode (extract_inspiring /
  extract_inspiring_mod_switched, restated in the oracle) to the planted records, unswitched,
  as 36- and 45-bit wire bytes, and through the serving engine.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

D = 2048
INDICES = [0, D - 1, D, 2 * D + 17, 3 * D - 1]


@pytest.fixture(scope="module")
def kat():
    return oracle.post_switch_error_kat_inputs()


@pytest.mark.parametrize("store", [True, False], ids=["store", "no-store"])
def test_post_switch_error_kat_on_the_device(gpu, kat, store):
    gamma = kat["gamma"]
    print(f"post-switch error kat: store {store}", flush=True)
    with gpu.Context(gpu.params_d2048(), kat["w_seed"], gamma) as ctx:
        sh = ctx.upload_columns(0, kat["columns"])
        if store:
            sh.precompute_packing()
        seat = ctx.seat(kat["y_body"])
        args = (seat, [0], [kat["seed"]], kat["b"][None, :])
        a, b = ctx.respond_seeded_batch(*args)
        assert oracle.extract_inspiring(kat["secret"], a[0], b[0], kat["entry"]) == kat["record"]
        for modulus, expected in oracle.POST_SWITCH_ERROR_KAT:
            sa, sb = ctx.respond_seeded_batch(*args, switch_modulus=modulus)
            assert np.array_equal(sa[0], oracle.mod_switch(a[0], oracle.Q, modulus)), f"target {modulus}"
            assert np.array_equal(sb[0], oracle.mod_switch(b[0], oracle.Q, modulus)), f"target {modulus}"
            err = oracle.served_post_switch_error(kat["secret"], sa[0], sb[0], modulus, kat["record"])
            assert err == expected, f"target {modulus}"
        for modulus, length in ((oracle.Q36, 10_446), (oracle.MOD_SWITCH_TARGET_45BIT, 13_038), (0, 17_358)):
            wire = ctx.respond_seeded_batch_wire(*args, switch_modulus=modulus)[0]
            assert len(wire) == length
            r = oracle.parse_response(wire)
            got = oracle.extract_inspiring(kat["secret"], r["a"], r["b_prefix"], kat["entry"], r["retained"],
                                           r["modulus"])
            assert got == kat["record"], f"wire at {modulus}"
        seat.close()


def _cell_context(gpu, c):
    ctx = gpu.Context(gpu.params_d2048(approximate=c.approximate), c.w_seed, c.gamma)
    for shard in range(3):
        ctx.upload_records(shard, c.db[shard * c.n * c.entry:(shard + 1) * c.n * c.entry].tobytes(),
                           c.entry).precompute_packing()
    return ctx


@pytest.mark.parametrize("approximate", [True, False], ids=["approximate", "exact"])
@pytest.mark.parametrize("entry", [32, 256, 512, 1024])
def test_device_responses_decode_to_the_planted_records(gpu, approximate, entry):
    c = oracle.Cell(D, entry, 0x4150_5244, approximate=approximate)
    queries = [c.query(i) for i in INDICES]
    print(f"decode cell: approximate {approximate}, entry {entry}, gamma {c.gamma}", flush=True)
    with _cell_context(gpu, c) as ctx:
        seat = ctx.seat(c.y_body)
        args = (seat, [q[0] for q in queries], [q[1] for q in queries], np.stack([q[2] for q in queries]))
        a, b = ctx.respond_seeded_batch(*args)
        wires = {m: ctx.respond_seeded_batch_wire(*args, switch_modulus=m) for m in oracle.IMPLEMENTED_TARGETS}
        seat.close()
    for k, index in enumerate(INDICES):
        rec = c.db[index * entry:(index + 1) * entry].tobytes()
        assert oracle.extract_inspiring(c.secret, a[k], b[k], entry) == rec, f"index {index}"
        for m, ws in wires.items():
            r = oracle.parse_response(ws[k])
            got = oracle.extract_inspiring(c.secret, r["a"], r["b_prefix"], entry, r["retained"], r["modulus"])
            assert got == rec, f"index {index} switched to {m}"


def test_served_responses_decode(gpu):
    """The serving engine (one registered seat, batches submitted together) answers a mixed
    stream of approximate-gadget row requests with responses that decode."""
    c = oracle.Cell(D, 512, 0x4150_5357, approximate=True)
    order = [3, 2047, 4096, 6143, 0, 5000, 17, 2048]
    queries = [c.query(i) for i in order]
    with _cell_context(gpu, c) as ctx:
        srv = gpu.Server(ctx, max_launch=64)
        try:
            h = srv.register_seat(c.y_body)
            tickets = [srv.submit(h, [q[0] for q in queries[s:s + 3]], [q[1] for q in queries[s:s + 3]],
                                  np.stack([q[2] for q in queries[s:s + 3]]), switch_modulus=oracle.Q36)
                       for s in range(0, len(order), 3)]
            outs = [srv.wait(t) for t in tickets]
        finally:
            srv.close()
    k = 0
    for a, b, _ in outs:
        for j in range(a.shape[0]):
            index = order[k]
            got = oracle.extract_inspiring(c.secret, a[j], b[j], 512, modulus=oracle.Q36)
            assert got == c.db[index * 512:(index + 1) * 512].tobytes(), f"index {index}"
            k += 1
    assert k == len(order)
