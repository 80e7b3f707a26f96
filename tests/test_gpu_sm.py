"""The special-modulus path on the GPU, through the C ABI, against the oracle
(oracle/special.c, itself pinned to the reference's served-set wire digests)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SM_Q, SM_P = oracle.SM_Q, oracle.SM_P
PINNED = [0x84FE10759FE96D5F, 0xD889F6A134B7E313, 0xEF10B409C5474E4B]


def _small(gpu, d, gamma):
    # tests/special_modulus_kat.rs:26-34: cells of gamma/2 bytes, all of them record
    return gpu.SmParams(d, gamma, SM_Q, SM_P, gamma // 2, gamma // 2, 6.4, 2)


def _oracle_ring(params):
    return oracle.SmRing(d=params.ring_dim, gamma=params.gamma, cell_bytes=params.cell_bytes,
                         record_bytes=params.record_bytes)


def _random_key(gpu, rng, d):
    return gpu.PackingKey(2, bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                          rng.integers(0, SM_Q, d, dtype=np.uint32), rng.integers(0, SM_P, d, dtype=np.uint32))


def _random_query(gpu, rng, d):
    return gpu.Query(bytes(rng.integers(0, 256, 32, dtype=np.uint8)), rng.integers(0, SM_Q, d, dtype=np.uint32))


def test_served_wire_bytes_are_pinned(gpu):
    params = gpu.ring2048_cell256()
    ring = oracle.SmRing(cell_bytes=256, record_bytes=246)
    client = oracle.SmClient(ring, bytes([42] * 32))
    cells = oracle.sm_kat_cells(42, 2048)
    with gpu.SmServer(params) as server:
        table = server.table(cells.tobytes())
        assert table.rows() == 2048
        key_bytes = client.key_bytes()
        seat = server.seat(gpu.PackingKey.from_bytes(params, key_bytes))
        qseed, qbody = client.query(1234)
        query_bytes = gpu.Query(qseed, qbody).to_bytes(params)
        response = server.respond(table, seat, gpu.Query.from_bytes(params, query_bytes))
        response_bytes = response.to_bytes(params)
        assert (len(key_bytes), len(query_bytes), len(response_bytes)) == (15_394, 7_456, 9_208)
        assert [oracle.fnv1a(key_bytes), oracle.fnv1a(query_bytes), oracle.fnv1a(response_bytes)] == PINNED
        assert client.decode(response.mask, response.body) == bytes(cells[1234 * 256:1234 * 256 + 246])
        # the same response packed by the kernel (Response::to_bytes on the device)
        assert server.response_bytes() == 9_208
        wire = server.respond_batch_wire([table], [seat], [gpu.Query.from_bytes(params, query_bytes)])[0]
        assert oracle.fnv1a(wire) == PINNED[2]
        seat.close()
        table.close()


@pytest.mark.parametrize("d,gamma", [(256, 32), (512, 64), (256, 128), (64, 8), (1024, 64), (128, 16)])
def test_small_rings_match_the_oracle(gpu, d, gamma):
    params = _small(gpu, d, gamma)
    ring = _oracle_ring(params)
    rng = np.random.default_rng(1000 + d + gamma)
    worst = np.full(d * params.cell_bytes, 0x88, dtype=np.uint8)  # every nibble -8: the largest fold
    rand_rows = int(rng.integers(1, d + 1))
    cells = [worst.tobytes(), rng.integers(0, 256, rand_rows * params.cell_bytes, dtype=np.uint8).tobytes()]
    with gpu.SmServer(params) as server:
        tables = [server.table(c) for c in cells]
        keys = [_random_key(gpu, rng, d) for _ in range(2)]
        seats = [server.seat(k) for k in keys]
        queries = [_random_query(gpu, rng, d) for _ in range(4)]
        pick_t, pick_s = [0, 1, 1, 0], [0, 0, 1, 1]
        masks, bodies = server.respond_batch([tables[i] for i in pick_t], [seats[i] for i in pick_s], queries)
        otables = [ring.table(c) for c in cells]
        oseats = [ring.seat(k.q, k.p, k.mask_seed) for k in keys]
        for j, qy in enumerate(queries):
            wm, wb = ring.respond(otables[pick_t[j]], oseats[pick_s[j]], qy.seed, qy.body)
            assert np.array_equal(masks[j], wm), (d, gamma, j)
            assert np.array_equal(bodies[j], wb), (d, gamma, j)
        for x in tables + seats:
            x.close()


def test_served_batch_mixed_tables_and_seats_equal_alone_and_oracle(gpu):
    params = gpu.ring2048_cell256()
    ring = oracle.SmRing(cell_bytes=256, record_bytes=246)
    rng = np.random.default_rng(77)
    cells = [oracle.sm_kat_cells(77, 2048).tobytes(), oracle.sm_kat_cells(78, 1500).tobytes(),
             oracle.sm_kat_cells(79, 2048, worst=True).tobytes(), b""]
    with gpu.SmServer(params) as server:
        tables = [server.table(c) for c in cells]
        assert [t.rows() for t in tables] == [2048, 1500, 2048, 0]
        keys = [_random_key(gpu, rng, 2048) for _ in range(2)]
        seats = [server.seat(k) for k in keys]
        assert seats[0].bytes() == 32_768
        pick_t, pick_s = [0, 1, 2, 3, 2, 0], [0, 1, 0, 1, 1, 1]
        queries = [_random_query(gpu, rng, 2048) for _ in pick_t]
        masks, bodies = server.respond_batch([tables[i] for i in pick_t], [seats[i] for i in pick_s], queries)
        otables = [ring.table(c) for c in cells[:3]]
        oseats = [ring.seat(k.q, k.p, k.mask_seed) for k in keys]
        for j, qy in enumerate(queries):
            alone = server.respond(tables[pick_t[j]], seats[pick_s[j]], qy)
            assert np.array_equal(alone.mask, masks[j]) and np.array_equal(alone.body, bodies[j]), j
            if pick_t[j] < 3:
                wm, wb = ring.respond(otables[pick_t[j]], oseats[pick_s[j]], qy.seed, qy.body)
                assert np.array_equal(masks[j], wm) and np.array_equal(bodies[j], wb), j
        for x in tables + seats:
            x.close()


def test_served_records_decode_at_batch_scale(gpu):
    # tests/special_modulus_kat.rs:296-336 at a full batch: every response of a 96-query
    # launch over four tables and four client keys decodes to its record
    params = gpu.ring2048_cell256()
    ring = oracle.SmRing(cell_bytes=256, record_bytes=246)
    rng = np.random.default_rng(5)
    rows = [2048, 2048, 1500, 700]
    cells = [oracle.sm_kat_cells(200 + i, r) for i, r in enumerate(rows)]
    clients = [oracle.SmClient(ring, bytes([i + 1] * 32)) for i in range(4)]
    with gpu.SmServer(params) as server:
        tables = [server.table(c.tobytes()) for c in cells]
        seats = [server.seat(gpu.PackingKey.from_bytes(params, c.key_bytes())) for c in clients]
        picks, queries = [], []
        for j in range(96):
            ti, ci = int(rng.integers(0, 4)), int(rng.integers(0, 4))
            row = int(rng.integers(0, 2048))
            seed, body = clients[ci].query(row)
            picks.append((ti, ci, row))
            queries.append(gpu.Query(seed, body))
        masks, bodies = server.respond_batch([tables[t] for t, _, _ in picks], [seats[c] for _, c, _ in picks],
                                             queries)
        for j, (ti, ci, row) in enumerate(picks):
            want = bytes(cells[ti][row * 256:row * 256 + 246]) if row < rows[ti] else bytes(246)
            assert clients[ci].decode(masks[j], bodies[j]) == want, (j, ti, ci, row)
        for x in tables + seats:
            x.close()


def test_refusals(gpu):
    params = gpu.ring2048_cell256()
    rng = np.random.default_rng(1)
    with gpu.SmServer(params) as server, gpu.SmServer(params) as other:
        cells = bytearray(oracle.sm_kat_cells(3, 4).tobytes())
        cells[3 * 256 + 250] = 1
        with pytest.raises(gpu.PirError, match="cell 3 has bytes past its 246-byte record"):
            server.table(bytes(cells))
        with pytest.raises(gpu.PirError, match="not at most 2048 cells of 256 bytes"):
            server.table(bytes(300))
        with pytest.raises(gpu.PirError, match="not at most 2048 cells"):
            server.table(bytes(2049 * 256))
        key = _random_key(gpu, rng, 2048)
        bad = gpu.PackingKey(2, key.mask_seed, key.q.copy(), key.p.copy())
        bad.p[7] = SM_P
        with pytest.raises(gpu.PirError, match="not a residue"):
            server.seat(bad)
        table, seat = server.table(b""), server.seat(key)
        q = _random_query(gpu, rng, 2048)
        q.body[5] = SM_Q
        with pytest.raises(gpu.PirError, match="query does not fit the set"):
            server.respond(table, seat, q)
        foreign = other.table(b"")
        with pytest.raises(gpu.PirError, match="another parameter set"):
            server.respond(foreign, seat, _random_query(gpu, rng, 2048))
        with pytest.raises(gpu.PirError, match="1 to 4194304 queries"):
            gpu.SmBatcher(server, max_launch=(1 << 22) + 1)
        for x in (table, seat, foreign):
            x.close()
    bad_params = gpu.ring2048_cell256()
    bad_params.key_mask_version = 1
    with pytest.raises(gpu.PirError, match="per-key-seed version"):
        gpu.SmServer(bad_params)


def test_empty_and_oversize_batches(gpu):
    """Through the C ABI: an empty batch answers nothing on all four batch entry points (the
    loop over no queries), and a count past INSPIRE_MAX_BATCH is refused before any of the
    batch's arrays is read (the placeholder address 1 is never dereferenced)."""
    import ctypes

    L = gpu.lib()
    with gpu.SmServer(gpu.ring2048_cell256()) as server:
        h, z = server._h, None
        gpu._call(L.inspire_sm_respond_batch, h, 0, z, z, z, z, z, z)
        gpu._call(L.inspire_sm_respond_batch_wire, h, 0, z, z, z, z, z, 0)
        gpu._call(L.inspire_sm_respond_batch_device, h, 0, z, z, z, z, z, z, z, z)
        gpu._call(L.inspire_sm_respond_batch_device_wire, h, 0, z, z, z, z, z, z, 0, z)
        x, big = ctypes.c_void_p(1), (1 << 22) + 1
        calls = [(L.inspire_sm_respond_batch, (x,) * 6),
                 (L.inspire_sm_respond_batch_wire, (x,) * 5 + (9216,)),
                 (L.inspire_sm_respond_batch_device, (x,) * 6 + (z, z)),
                 (L.inspire_sm_respond_batch_device_wire, (x,) * 6 + (9216, z))]
        for fn, args in calls:
            with pytest.raises(gpu.PirError, match="a batch holds at most 4194304 queries, got 4194305"):
                gpu._call(fn, h, big, *args)
        with pytest.raises(gpu.PirError, match="null argument"):
            gpu._call(L.inspire_sm_respond_batch, h, 1, z, z, z, z, z, z)


def test_kernel_trace_keeps_results(gpu):
    """Launches traced by the server (events around the respond kernel, read after the loop)
    answer exactly as untraced ones; only the armed number of launches is recorded."""
    import torch

    rng = np.random.default_rng(6)
    dev = torch.device("cuda:0")
    L, n = 4, 2048
    with gpu.SmServer(gpu.ring2048_cell256()) as server:
        table = server.table(oracle.sm_kat_cells(5, 64).tobytes())
        seat = server.seat(_random_key(gpu, rng, n))
        seeds = torch.from_numpy(rng.integers(0, 256, (L, 32), dtype=np.uint8)).to(dev)
        bodies = torch.from_numpy(rng.integers(0, SM_Q, (L, n), dtype=np.int64).astype(np.int32)).to(dev)
        outs = []
        for traced in (False, True):
            mask = torch.empty((L, n), dtype=torch.int32, device=dev)
            body = torch.empty((L, 2 * 246), dtype=torch.int32, device=dev)
            if traced:
                server.trace(1)
            for _ in range(2):
                server.respond_batch_device([table] * L, [seat] * L, seeds.data_ptr(), bodies.data_ptr(),
                                            mask.data_ptr(), body.data_ptr())
            if traced:
                t = server.kernel_times()
                assert len(t) == 1 and t[0] > 0
            torch.cuda.synchronize()
            outs.append((mask.cpu(), body.cpu()))
        assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
        with pytest.raises(gpu.PirError, match="at most 1048576 launches"):
            server.trace((1 << 20) + 1)
        table.close()
        seat.close()


@pytest.mark.parametrize("d,gamma", [(2048, 512), (256, 32), (64, 8)])
def test_device_wire_bytes_equal_the_host_codec(gpu, d, gamma):
    """Response::to_bytes packed by the kernel equals the host codec over the word arrays of
    the same batch, through host and device buffers (uneven strides, mixed tables and keys)."""
    import torch

    params = gpu.ring2048_cell256() if d == 2048 else _small(gpu, d, gamma)
    rng = np.random.default_rng(d + gamma)
    with gpu.SmServer(params) as server:
        def cells(rows):
            c = rng.integers(0, 256, (rows, params.cell_bytes), dtype=np.uint8)
            c[:, params.record_bytes:] = 0  # a cell holds nothing past its record
            return c.tobytes()

        tables = [server.table(cells(int(rng.integers(1, d + 1)))) for _ in range(3)]
        seats = [server.seat(_random_key(gpu, rng, d)) for _ in range(2)]
        count = 7
        tl = [tables[int(i)] for i in rng.integers(0, 3, count)]
        sl = [seats[int(i)] for i in rng.integers(0, 2, count)]
        queries = [_random_query(gpu, rng, d) for _ in range(count)]
        masks, bodies = server.respond_batch(tl, sl, queries)
        want = [gpu.Response(masks[k], bodies[k]).to_bytes(params) for k in range(count)]
        assert [len(w) for w in want] == [server.response_bytes()] * count
        assert server.respond_batch_wire(tl, sl, queries) == want
        size = server.response_bytes()
        stride = (size + 3) // 4 * 4 + 8
        dev = torch.device("cuda:0")
        d_seeds = torch.from_numpy(np.frombuffer(b"".join(q.seed for q in queries), np.uint8).copy()).to(dev)
        d_bodies = torch.from_numpy(np.stack([q.body for q in queries]).astype(np.uint32).view(np.int32)).to(dev)
        d_scratch = torch.zeros((count, d), dtype=torch.int32, device=dev)
        d_wire = torch.full((count, stride), 0xAB, dtype=torch.uint8, device=dev)
        server.respond_batch_device_wire(tl, sl, d_seeds.data_ptr(), d_bodies.data_ptr(), d_scratch.data_ptr(),
                                         d_wire.data_ptr(), stride)
        torch.cuda.synchronize()
        got = d_wire.cpu().numpy()
        for k in range(count):
            assert got[k, :size].tobytes() == want[k], f"query {k}"
            assert (got[k, (size + 3) // 4 * 4:] == 0xAB).all()  # nothing written past the padded length
        with pytest.raises(gpu.PirError, match="multiple of 4"):
            server.respond_batch_device_wire(tl, sl, d_seeds.data_ptr(), d_bodies.data_ptr(), d_scratch.data_ptr(),
                                             d_wire.data_ptr(), size + 1)
        for x in tables + seats:
            x.close()


def test_device_launches_on_two_streams_with_their_own_job_lists(gpu):
    """Device launches alternating between two streams, each naming its own tables and keys (so
    each uploads its own job list) with no wait between them: more launches than the server's
    ring of job lists, a launch that repeats its predecessor's handles on the other stream and one
    that repeats them on the same stream. Every
    launch equals the host-buffer path's answer to the same queries."""
    import torch

    rng = np.random.default_rng(8)
    dev = torch.device("cuda:0")
    L, n, launches = 24, 2048, 8
    with gpu.SmServer(gpu.ring2048_cell256()) as server:
        tables = [server.table(oracle.sm_kat_cells(500 + i, 2048).tobytes()) for i in range(3)]
        keys = [gpu.PackingKey(2, bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                               rng.integers(0, SM_Q, n, dtype=np.uint32), rng.integers(0, SM_P, n, dtype=np.uint32))
                for _ in range(3)]
        seats = [server.seat(k) for k in keys]
        # launch i runs on stream i % 2, the last one on the stream of the one before it; launch 2
        # repeats launch 1's handles on the other stream, the last repeats its predecessor's on
        # the same stream (the reused job list)
        on = [i % 2 for i in range(launches - 1)] + [(launches - 2) % 2]
        picks = []
        for i in range(launches):
            if i in (2, launches - 1):
                picks.append(picks[i - 1])
                continue
            picks.append(([tables[j] for j in rng.integers(0, 3, L)], [seats[j] for j in rng.integers(0, 3, L)]))
        queries = [[gpu.Query(bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                              rng.integers(0, SM_Q, n, dtype=np.uint32)) for _ in range(L)] for _ in range(launches)]
        d_seeds = [torch.from_numpy(np.frombuffer(b"".join(q.seed for q in qs), np.uint8).copy()).to(dev)
                   for qs in queries]
        d_bodies = [torch.from_numpy(np.stack([q.body for q in qs]).astype(np.uint32).view(np.int32)).to(dev)
                    for qs in queries]
        masks = [torch.empty((L, n), dtype=torch.int32, device=dev) for _ in range(launches)]
        bodies = [torch.empty((L, 2 * 246), dtype=torch.int32, device=dev) for _ in range(launches)]
        streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        torch.cuda.synchronize()
        print(f"sm: {launches} device launches on two streams", flush=True)
        for i in range(launches):
            server.respond_batch_device(picks[i][0], picks[i][1], d_seeds[i].data_ptr(), d_bodies[i].data_ptr(),
                                        masks[i].data_ptr(), bodies[i].data_ptr(), stream=streams[on[i]].cuda_stream)
        torch.cuda.synchronize()
        for i in range(launches):
            want_m, want_b = server.respond_batch(picks[i][0], picks[i][1], queries[i])
            assert np.array_equal(masks[i].cpu().numpy().view(np.uint32), want_m), i
            assert np.array_equal(bodies[i].cpu().numpy().view(np.uint32), want_b), i
        for s in seats:
            s.close()
        for t in tables:
            t.close()
