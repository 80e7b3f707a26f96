"""The special-modulus continuous-batching queue: every request answered in the mix equals
the same batch answered alone (and the oracle), requests from many keys and tables merge
into shared launches, and keys are admitted against the device budget."""
import threading

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SM_Q, SM_P = oracle.SM_Q, oracle.SM_P


def _key(gpu, rng):
    return gpu.PackingKey(2, bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                          rng.integers(0, SM_Q, 2048, dtype=np.uint32), rng.integers(0, SM_P, 2048, dtype=np.uint32))


def test_mixed_requests_equal_alone(gpu):
    params = gpu.ring2048_cell256()
    rng = np.random.default_rng(21)
    cells = [oracle.sm_kat_cells(300 + i, r).tobytes() for i, r in enumerate((2048, 900, 2048))]
    with gpu.SmServer(params) as server:
        tables = [server.table(c) for c in cells]
        batcher = gpu.SmBatcher(server, max_launch=96)
        keys = [_key(gpu, rng) for _ in range(5)]
        handles = [batcher.register_key(k) for k in keys]
        reqs = []
        for _ in range(24):
            count = int(rng.integers(1, 17))
            kidx = int(rng.integers(0, len(keys)))
            tids = [int(x) for x in rng.integers(0, len(tables), count)]
            queries = [gpu.Query(bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                                 rng.integers(0, SM_Q, 2048, dtype=np.uint32)) for _ in range(count)]
            reqs.append((kidx, tids, queries))
        results = [None] * len(reqs)

        def client(part):
            for i in part:
                kidx, tids, queries = reqs[i]
                t = batcher.submit(handles[kidx], [tables[j] for j in tids], queries)
                results[i] = batcher.wait(t)

        threads = [threading.Thread(target=client, args=(list(range(t, len(reqs), 4)),)) for t in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        stats = batcher.stats()
        assert stats["queries"] == sum(len(r[2]) for r in reqs)
        assert stats["launches"] < len(reqs), stats  # requests were merged
        assert stats["resident_bytes"] == 5 * 32_768
        seats = [server.seat(k) for k in keys]
        for i, (kidx, tids, queries) in enumerate(reqs):
            masks, bodies = server.respond_batch([tables[j] for j in tids], [seats[kidx]] * len(queries), queries)
            m_mix, b_mix, lat = results[i]
            assert np.array_equal(m_mix, masks) and np.array_equal(b_mix, bodies), i
            assert lat > 0
        # and one query against the oracle
        ring = oracle.SmRing(cell_bytes=256, record_bytes=246)
        kidx, tids, queries = reqs[0]
        k = keys[kidx]
        want = ring.respond(ring.table(cells[tids[0]]), ring.seat(k.q, k.p, k.mask_seed), queries[0].seed,
                            queries[0].body)
        assert np.array_equal(results[0][0][0], want[0]) and np.array_equal(results[0][1][0], want[1])
        for s in seats:
            s.close()
        batcher.close()
        for t in tables:
            t.close()


def test_admission_and_handles(gpu):
    params = gpu.ring2048_cell256()
    rng = np.random.default_rng(3)
    with gpu.SmServer(params) as server:
        table = server.table(b"")
        batcher = gpu.SmBatcher(server, max_launch=8, hbm_budget=3 * 32_768)
        hs = [batcher.register_key(_key(gpu, rng)) for _ in range(3)]
        with pytest.raises(gpu.PirError, match="admission refused"):
            batcher.register_key(_key(gpu, rng))
        batcher.unregister_key(hs[0])
        h4 = batcher.register_key(_key(gpu, rng))  # fits again
        assert batcher.stats()["resident_bytes"] == 3 * 32_768
        q = gpu.Query(bytes(32), np.zeros(2048, np.uint32))
        with pytest.raises(gpu.PirError, match="no key seated"):
            batcher.submit(hs[0], [table], [q])
        with pytest.raises(gpu.PirError, match="1..=8 queries"):
            batcher.submit(h4, [table] * 9, [q] * 9)
        masks, bodies, _ = batcher.wait(batcher.submit(h4, [table], [q]))
        assert masks.shape == (1, 2048) and bodies.shape == (1, 492)
        batcher.close()
        table.close()


def test_keys_opened_and_closed_while_requests_run(gpu):
    """Keys registered (their seat kernel not waited for by the registration) and unregistered
    between queued requests: each request equals the same batch answered alone under a seat of
    its key, including requests submitted right after their key's registration and requests
    under a key whose seat memory came back from an unregistered one."""
    params = gpu.ring2048_cell256()
    rng = np.random.default_rng(23)
    cells = [oracle.sm_kat_cells(310 + i, 2048).tobytes() for i in range(2)]
    with gpu.SmServer(params) as server:
        tables = [server.table(c) for c in cells]
        batcher = gpu.SmBatcher(server, max_launch=64)
        keys = [_key(gpu, rng) for _ in range(6)]
        live = {0: batcher.register_key(keys[0]), 1: batcher.register_key(keys[1])}
        reqs, tickets = [], []

        def submit(kidx, count):
            tids = [int(x) for x in rng.integers(0, len(tables), count)]
            queries = [gpu.Query(bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                                 rng.integers(0, SM_Q, 2048, dtype=np.uint32)) for _ in range(count)]
            reqs.append((kidx, tids, queries))
            tickets.append(batcher.submit(live[kidx], [tables[j] for j in tids], queries))

        print("sm serving: keys opened and closed between queued requests", flush=True)
        for k in range(2, 6):
            submit(k - 2, 20)
            submit(k - 1, 5)
            batcher.unregister_key(live.pop(k - 2))  # its queued request keeps it alive
            live[k] = batcher.register_key(keys[k])
            submit(k, 12)
        results = [batcher.wait(t) for t in tickets]
        assert batcher.stats()["resident_bytes"] == 2 * 32_768
        seats = [server.seat(k) for k in keys]
        for i, (kidx, tids, queries) in enumerate(reqs):
            masks, bodies = server.respond_batch([tables[j] for j in tids], [seats[kidx]] * len(queries), queries)
            assert np.array_equal(results[i][0], masks) and np.array_equal(results[i][1], bodies), i
        for s in seats:
            s.close()
        batcher.close()
        for t in tables:
            t.close()
