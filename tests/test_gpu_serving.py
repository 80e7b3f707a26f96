"""The continuous-batching server: every batch answered in the mix equals the same batch
answered alone (and the oracle), requests from many seats merge into shared launches,
and seats are admitted against the device budget."""
import threading

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.Q
N = 2048


def _setup(gpu, approx, gamma, shards, seed):
    rng = np.random.default_rng(seed)
    w_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    ctx = gpu.Context(gpu.params_d2048(approx), w_seed, gamma)
    cols = []
    for s in range(shards):
        c = rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64)
        cols.append(c)
        ctx.upload_columns(s, c).precompute_packing()
    return rng, w_seed, ctx, cols


@pytest.mark.parametrize("approx", [True, False])
def test_mixed_requests_equal_alone(gpu, approx):
    gamma, shards = 64, 6
    rng, w_seed, ctx, cols = _setup(gpu, approx, gamma, shards, seed=7 + approx)
    rows = 2 if approx else 3
    server = gpu.Server(ctx, max_launch=48)
    keys = [rng.integers(0, Q, (rows, N), dtype=np.uint64) for _ in range(5)]
    handles = [server.register_seat(k) for k in keys]
    reqs = []
    for r in range(24):
        count = int(rng.integers(1, 17))
        kidx = int(rng.integers(0, len(keys)))
        ids = rng.integers(0, shards, count).astype(np.uint32)
        seeds = [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in range(count)]
        bs = rng.integers(0, Q, (count, N), dtype=np.uint64)
        reqs.append((kidx, ids, seeds, bs))
    mod = oracle.Q36 if approx else 0
    results = [None] * len(reqs)

    def client(part):
        for i in part:
            kidx, ids, seeds, bs = reqs[i]
            t = server.submit(handles[kidx], ids, seeds, bs, switch_modulus=mod)
            results[i] = server.wait(t)

    threads = [threading.Thread(target=client, args=(list(range(t, len(reqs), 4)),)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    stats = server.stats()
    assert stats["queries"] == sum(len(r[1]) for r in reqs)
    assert stats["launches"] < len(reqs), stats  # requests were merged
    # alone: each batch through the one-batch entry point under its own seat
    seats = [ctx.seat(k) for k in keys]
    for i, (kidx, ids, seeds, bs) in enumerate(reqs):
        a_alone, b_alone = ctx.respond_seeded_batch(seats[kidx], ids, seeds, bs, switch_modulus=mod)
        a_mix, b_mix, lat = results[i]
        assert np.array_equal(a_mix, a_alone) and np.array_equal(b_mix, b_alone), i
        assert lat > 0
    # and one query against the oracle
    kidx, ids, seeds, bs = reqs[0]
    p = oracle.Pack(N, gamma, 3, 1 if approx else 0)
    wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[0], N), bs[0], cols[ids[0]], keys[kidx])
    if mod:
        wa, wb = oracle.mod_switch(wa), oracle.mod_switch(wb)
    assert np.array_equal(results[0][0][0], wa) and np.array_equal(results[0][1][0], wb[:gamma])
    for s in seats:
        s.close()
    server.close()
    ctx.close()


def test_admission_and_handles(gpu):
    gamma = 16
    rng, w_seed, ctx, cols = _setup(gpu, True, gamma, 1, seed=3)
    seat_bytes = 2 * N * 16
    server = gpu.Server(ctx, max_launch=8, hbm_budget=3 * seat_bytes)
    hs = [server.register_seat(rng.integers(0, Q, (2, N), dtype=np.uint64)) for _ in range(3)]
    with pytest.raises(gpu.PirError, match="admission refused"):
        server.register_seat(rng.integers(0, Q, (2, N), dtype=np.uint64))
    server.unregister_seat(hs[0])
    h4 = server.register_seat(rng.integers(0, Q, (2, N), dtype=np.uint64))  # fits again
    st = server.stats()
    # device bytes carry each word's Shoup companion; the reference counts 32 KiB per
    # approximate-gadget seat (2 rows x 2048 x 8, tests/approximate_packing_gadget.rs:519)
    assert st["resident_bytes"] == 3 * seat_bytes and st["seats"] == 3
    assert st["payload_bytes"] == 3 * 32_768
    with pytest.raises(gpu.PirError, match="no entry for session_handle"):
        server.submit(hs[0], [0], [bytes(32)], np.zeros((1, N), dtype=np.uint64))
    a, b, _ = server.wait(server.submit(h4, [0], [bytes(32)], np.zeros((1, N), dtype=np.uint64)))
    assert a.shape == (1, N) and b.shape == (1, gamma)
    # two waits on one ticket: one collects the result, the other is refused (not left waiting)
    import threading
    ticket = server.submit(h4, [0, 0], [bytes(32), bytes([1]) * 32], np.zeros((2, N), dtype=np.uint64))
    outcomes = []

    def collect():
        try:
            server.wait(ticket)
            outcomes.append("ok")
        except gpu.PirError as e:
            outcomes.append(str(e))

    ths = [threading.Thread(target=collect) for _ in range(2)]
    for th in ths:
        th.start()
    for th in ths:
        th.join(timeout=60)
    assert sorted(outcomes, key=lambda o: o != "ok")[0] == "ok" and len(outcomes) == 2
    refused = [o for o in outcomes if o != "ok"][0]
    assert "collected by another wait" in refused or "unknown ticket" in refused  # which depends on timing
    with pytest.raises(gpu.PirError, match="unknown ticket"):
        server.wait(ticket)
    server.close()
    ctx.close()


def test_sessions_opened_and_closed_while_requests_run(gpu):
    """Seats registered and unregistered between queued requests: an unregistered seat stays
    alive for its queued requests, its device block returns to the seat pool only after them,
    and a seat registered afterwards (which may take that block) answers under its own key.
    Every batch equals the same batch answered alone."""
    gamma, shards = 32, 4
    rng, w_seed, ctx, cols = _setup(gpu, True, gamma, shards, seed=11)
    server = gpu.Server(ctx, max_launch=64)
    keys = [rng.integers(0, Q, (2, N), dtype=np.uint64) for _ in range(6)]
    live = {0: server.register_seat(keys[0]), 1: server.register_seat(keys[1])}
    reqs, tickets = [], []

    def submit(kidx, count):
        ids = rng.integers(0, shards, count).astype(np.uint32)
        seeds = [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in range(count)]
        bs = rng.integers(0, Q, (count, N), dtype=np.uint64)
        reqs.append((kidx, ids, seeds, bs))
        tickets.append(server.submit(live[kidx], ids, seeds, bs, switch_modulus=oracle.Q36))

    print("serving: sessions opened and closed between queued requests", flush=True)
    for k in range(2, 6):
        submit(k - 2, 24)
        submit(k - 1, 7)
        server.unregister_seat(live.pop(k - 2))  # its queued request keeps it alive
        live[k] = server.register_seat(keys[k])
        submit(k, 16)
    results = [server.wait(t) for t in tickets]
    st = server.stats()
    assert st["seats"] == len(live) == 2 and st["resident_bytes"] == 2 * 2 * N * 16
    seats = [ctx.seat(k) for k in keys]
    for i, (kidx, ids, seeds, bs) in enumerate(reqs):
        a_alone, b_alone = ctx.respond_seeded_batch(seats[kidx], ids, seeds, bs, switch_modulus=oracle.Q36)
        assert np.array_equal(results[i][0], a_alone) and np.array_equal(results[i][1], b_alone), i
    for s in seats:
        s.close()
    server.close()
    ctx.close()
