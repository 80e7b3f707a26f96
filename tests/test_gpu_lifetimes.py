"""Object lifetimes against launches in flight, and the store-less shard semantics.

* A shard destroyed right after device launches naming it were enqueued (no wait in between):
  the destroy waits for those launches (their fence events), which are answered normally; a
  launch enqueued afterwards gets status -1 for the destroyed shard and answers the others.
  Seats and workspaces destroyed behind launches likewise wait for them, not for the device.
* A seat's device memory goes back to the context's seat pool when the seat is destroyed
  (after the launches enqueued before), and a seat created afterwards takes it: that seat
  answers under its own key, and the launches behind the destroy under the old one.
* A shard uploaded without precompute_packing is packed through a store the respond builds on
  the way; after an edit it is packed from the edited columns (as the reference packs a
  store-less shard, src/pir/respond.rs:1022-1041), not refused. A shard whose store came from
  precompute_packing is refused after an edit (tests/shard_packing_store.rs:328-368).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.Q
N, GAMMA = 2048, 256


def test_destroy_waits_for_launches_in_flight(gpu):
    import torch

    rng = np.random.default_rng(31)
    dev = torch.device("cuda:0")
    w_seed = bytes(range(3, 35))
    ctx = gpu.Context(gpu.params_d2048(), w_seed, GAMMA)
    cols = [rng.integers(0, 1 << 16, (GAMMA, N), dtype=np.uint64) for _ in range(4)]
    shards = [ctx.upload_columns(i, c) for i, c in enumerate(cols)]
    for sh in shards:
        sh.precompute_packing()
    y = rng.integers(0, Q, (3, N), dtype=np.uint64)
    seat = ctx.seat(y)
    B, K = 96, 4
    batches = []
    for k in range(K):
        ids = (np.arange(B) % 4).astype(np.int32)
        seeds = rng.integers(0, 256, (B, 32), dtype=np.uint8)
        bs = rng.integers(0, Q, (B, N), dtype=np.uint64)
        want = ctx.respond_seeded_batch(seat, ids, [s.tobytes() for s in seeds], bs)
        batches.append((torch.from_numpy(ids).to(dev), torch.from_numpy(seeds).to(dev),
                        torch.from_numpy(bs.view(np.int64)).to(dev), want))
    wss = [gpu.Workspace(ctx, B) for _ in range(K)]
    outs = [(torch.zeros((B, N), dtype=torch.int64, device=dev), torch.zeros((B, GAMMA), dtype=torch.int64, device=dev))
            for _ in range(K)]
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    print("lifetimes: 4 launches in flight, then destroy shard 2", flush=True)
    for k in range(K):
        d_ids, d_seeds, d_b, _ = batches[k]
        ctx.respond_batch_device(seat, wss[k], d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), B,
                                 outs[k][0].data_ptr(), outs[k][1].data_ptr(), stream=stream.cuda_stream)
    shards[2].close()  # waits for the four launches; nothing synchronised before it
    ctx.shards.pop(2)
    for k in range(K):
        assert (wss[k].status(B) == 0).all(), k
        assert np.array_equal(outs[k][0].cpu().numpy().view(np.uint64), batches[k][3][0]), k
        assert np.array_equal(outs[k][1].cpu().numpy().view(np.uint64), batches[k][3][1]), k
    # a launch after the destroy: shard 2's queries refused on the device, the rest answered
    d_ids, d_seeds, d_b, want = batches[0]
    a, b = outs[0]
    a.zero_()
    ctx.respond_batch_device(seat, wss[0], d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), B, a.data_ptr(),
                             b.data_ptr(), stream=stream.cuda_stream)
    stream.synchronize()
    st = wss[0].status(B)
    ids = d_ids.cpu().numpy()
    assert (st[ids == 2] == -1).all() and (st[ids != 2] == 0).all()
    got = a.cpu().numpy().view(np.uint64)
    assert np.array_equal(got[ids != 2], want[0][ids != 2])
    with pytest.raises(gpu.PirError, match="not found at dense position"):
        ctx.respond_seeded_batch(seat, [2], [bytes(32)], rng.integers(0, Q, (1, N), dtype=np.uint64))
    # a seat and workspaces destroyed right behind launches that use them
    d_ids, d_seeds, d_b, want = batches[1]
    ctx.respond_batch_device(seat, wss[1], d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), B,
                             outs[1][0].data_ptr(), outs[1][1].data_ptr(), stream=stream.cuda_stream)
    seat.close()
    st = wss[1].status(B)
    ids = d_ids.cpu().numpy()
    assert (st[ids == 2] == -1).all() and (st[ids != 2] == 0).all()
    assert np.array_equal(outs[1][0].cpu().numpy().view(np.uint64)[ids != 2], want[0][ids != 2])
    for w in wss:
        w.close()
    ctx.close()


def test_store_less_shard_is_packed_from_its_edited_columns(gpu):
    rng = np.random.default_rng(32)
    n, gamma = 2048, 16
    cols = rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64)
    y = rng.integers(0, Q, (3, n), dtype=np.uint64)
    b = rng.integers(0, Q, n, dtype=np.uint64)
    qseed, w_seed = bytes(range(32)), bytes(range(50, 82))
    p = oracle.Pack(n, gamma)
    a = oracle.expand_seed(qseed, n)
    with gpu.Context(gpu.params_d2048(), w_seed, gamma) as ctx:
        lazy = ctx.upload_columns(0, cols)          # no precompute_packing
        built = ctx.upload_columns(1, cols)
        built.precompute_packing()
        got = ctx.respond_seeded_inspiring(y, 0, qseed, b)
        want = p.respond(w_seed, a, b, cols, y)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1][:gamma])
        edited = cols.copy()
        edited[5, 17] ^= 1
        lazy.write_value(17, 5, int(edited[5, 17]))
        built.write_value(17, 5, int(edited[5, 17]))
        got = ctx.respond_seeded_inspiring(y, 0, qseed, b)
        want = p.respond(w_seed, a, b, edited, y)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1][:gamma])
        with pytest.raises(gpu.PirError) as e:
            ctx.respond_seeded_inspiring(y, 1, qseed, b)
        assert e.value.code == gpu.ERR_STORE


def test_seat_memory_reused_behind_launches(gpu):
    import torch

    rng = np.random.default_rng(33)
    dev = torch.device("cuda:0")
    gamma = 16
    w_seed = bytes(range(7, 39))
    with gpu.Context(gpu.params_d2048(), w_seed, gamma) as ctx:
        cols = [rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64) for _ in range(2)]
        for i, c in enumerate(cols):
            ctx.upload_columns(i, c).precompute_packing()
        ya, yb = (rng.integers(0, Q, (3, N), dtype=np.uint64) for _ in range(2))
        B = 32
        ids = (np.arange(B) % 2).astype(np.int32)
        seeds = rng.integers(0, 256, (B, 32), dtype=np.uint8)
        bs = rng.integers(0, Q, (B, N), dtype=np.uint64)
        seat_a = ctx.seat(ya)
        want_a = ctx.respond_seeded_batch(seat_a, ids, [s.tobytes() for s in seeds], bs)
        ws = gpu.Workspace(ctx, B)
        a = torch.zeros((B, N), dtype=torch.int64, device=dev)
        b = torch.zeros((B, gamma), dtype=torch.int64, device=dev)
        stream = torch.cuda.Stream()
        torch.cuda.synchronize()
        print("lifetimes: seat destroyed behind a launch, a new seat takes its memory", flush=True)
        d_ids, d_seeds = torch.from_numpy(ids).to(dev), torch.from_numpy(seeds).to(dev)
        d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
        torch.cuda.synchronize()
        ctx.respond_batch_device(seat_a, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), B, a.data_ptr(),
                                 b.data_ptr(), stream=stream.cuda_stream)
        seat_a.close()  # waits for the launch; the block goes to the pool
        seat_b = ctx.seat(yb)
        stream.synchronize()
        assert (ws.status(B) == 0).all()
        assert np.array_equal(a.cpu().numpy().view(np.uint64), want_a[0])
        assert np.array_equal(b.cpu().numpy().view(np.uint64), want_a[1])
        got = ctx.respond_seeded_batch(seat_b, ids[:2], [s.tobytes() for s in seeds[:2]], bs[:2])
        p = oracle.Pack(N, gamma)
        for k in range(2):
            wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k].tobytes(), N), bs[k], cols[ids[k]], yb)
            assert np.array_equal(got[0][k], wa) and np.array_equal(got[1][k], wb[:gamma]), k
        seat_b.close()
        ws.close()


def test_context_without_crs_builds_stores_then_binds_one(gpu):
    """ShardData::precompute_packing needs PackParams and the NttContext only (src/pir/setup.rs:
    335-340): a context created without a CRS takes shards and builds their stores, refuses every
    respond and a server, and after inspire_ctx_attach_crs answers as a context created with the
    CRS. Binding the same seed again is a no-op; another seed is refused."""
    rng = np.random.default_rng(34)
    gamma = 16
    w_seed = bytes(range(9, 41))
    cols = rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64)
    y = rng.integers(0, Q, (3, N), dtype=np.uint64)
    b = rng.integers(0, Q, (2, N), dtype=np.uint64)
    seeds = [bytes(range(k, k + 32)) for k in range(2)]
    with gpu.Context(gpu.params_d2048(), None, gamma) as ctx:
        ctx.upload_columns(0, cols).precompute_packing()
        with pytest.raises(gpu.PirError, match="context has no CRS"):
            gpu.Server(ctx, max_launch=8)
        seat = ctx.seat(y)
        with pytest.raises(gpu.PirError, match="context has no CRS"):
            ctx.respond_seeded_batch(seat, [0, 0], seeds, b)
        ctx.attach_crs(w_seed)
        ctx.attach_crs(w_seed)
        with pytest.raises(gpu.PirError, match="already bound to another CRS"):
            ctx.attach_crs(bytes(32))
        got = ctx.respond_seeded_batch(seat, [0, 0], seeds, b)
        p = oracle.Pack(N, gamma)
        for k in range(2):
            wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k], N), b[k], cols, y)
            assert np.array_equal(got[0][k], wa) and np.array_equal(got[1][k], wb[:gamma]), k
        seat.close()
