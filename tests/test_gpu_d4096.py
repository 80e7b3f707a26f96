"""InspireParams::d4096 (src/params.rs:392-436) on the GPU: ring 4096 at the d2048 prime and
gadget (three base-2^20 digits), one workgroup of 1024 threads per query.

The reference commits no response vector at this ring, so the device is checked against the
oracle's restatement, which is the code the d2048 and d256 vectors pin with the ring size as a
parameter (its store and per-query paths agree at 4096, tests/test_oracle_golden.py), and through
the client: responses of a packing-key cell built at ring 4096 in the fixture's draw order
decode to the planted records, unswitched and as 36-bit wire bytes (the d4096 set clears the
36-bit gate at 1.48x, params.rs:394-396), under both gadgets.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

N = 4096
Q = oracle.Q


@pytest.mark.parametrize("gamma,approx", [(16, False), (256, False), (256, True), (1024, False)])
def test_d4096_seeded_matches_oracle(gpu, gamma, approx):
    rng = np.random.default_rng(4096 + gamma + approx)
    rows = 2 if approx else 3
    cols = rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64)
    y = rng.integers(0, Q, (rows, N), dtype=np.uint64)
    bs = rng.integers(0, Q, (2, N), dtype=np.uint64)
    seeds = [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in range(2)]
    w_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    p = oracle.Pack(N, gamma, 3, 1 if approx else 0)
    store = p.store_build(cols)
    print(f"d4096: gamma {gamma}, approximate {approx}", flush=True)
    with gpu.Context(gpu.params_d2048(approx, ring_dim=N), w_seed, gamma) as ctx:
        ctx.upload_columns(0, cols).precompute_packing()
        seat = ctx.seat(y)
        a, b = ctx.respond_seeded_batch(seat, [0, 0], seeds, bs)
        sa, sb = ctx.respond_seeded_batch(seat, [0, 0], seeds, bs, switch_modulus=oracle.Q36)
        seat.close()
    for k in range(2):
        wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k], N), bs[k], cols, y, store)
        assert np.array_equal(a[k], wa) and np.array_equal(b[k], wb[:gamma]), k
        assert np.array_equal(sa[k], oracle.mod_switch(wa)) and np.array_equal(sb[k], oracle.mod_switch(wb[:gamma]))


@pytest.mark.parametrize("approximate", [False, True], ids=["exact", "approximate"])
def test_d4096_responses_decode_to_the_planted_records(gpu, approximate):
    entry = 512
    c = oracle.Cell(N, entry, 0x4150_5244, approximate=approximate)
    indices = [0, 3, N - 1, N, N + 2047, 3 * N - 1]
    queries = [c.query(i) for i in indices]
    print(f"d4096 decode cell: approximate {approximate}, gamma {c.gamma}", flush=True)
    with gpu.Context(gpu.params_d2048(approximate=approximate, ring_dim=N), c.w_seed, c.gamma) as ctx:
        for shard in range(3):
            ctx.upload_records(shard, c.db[shard * N * entry:(shard + 1) * N * entry].tobytes(),
                               entry).precompute_packing()
        seat = ctx.seat(c.y_body)
        args = (seat, [q[0] for q in queries], [q[1] for q in queries], np.stack([q[2] for q in queries]))
        a, b = ctx.respond_seeded_batch(*args)
        wires = ctx.respond_seeded_batch_wire(*args, switch_modulus=oracle.Q36)
        seat.close()
    for k, index in enumerate(indices):
        rec = c.db[index * entry:(index + 1) * entry].tobytes()
        assert oracle.extract_inspiring(c.secret, a[k], b[k], entry) == rec, index
        r = oracle.parse_response(wires[k])
        got = oracle.extract_inspiring(c.secret, r["a"], r["b_prefix"], entry, r["retained"], r["modulus"])
        assert got == rec, f"36-bit wire, index {index}"
