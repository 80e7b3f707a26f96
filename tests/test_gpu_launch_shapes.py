"""The oracle at the shapes the product runs, and concurrent device launches.

* A 512-query launch over 128 shards (the batched bench's launch: 8 batches of 64 merged, the
  2^18 x 512 B database) at every launch shape, three queries per CU included: the whole
  launch equal across the shapes, every status 0, and 8 sampled queries equal to the oracle's
  per-query path (no store: packing_offline from the columns, inspiring2.rs:1014-1161), so the
  device's store path is checked against the reference's other path.
* Four workspaces driven on four streams with no wait between their launches (the concurrency
  include/inspire_gpu.h promises for inspire_respond_batch_device), each batch equal to the
  host-buffer path's answer.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.Q
Q36 = oracle.Q36
N, GAMMA, ENTRY = 2048, 256, 512


def _records(seed, shards):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, N * ENTRY, dtype=np.uint8) for _ in range(shards)]


def _columns(rec):
    return rec.view("<u2").reshape(N, GAMMA).T.astype(np.uint64)


@pytest.fixture(scope="module")
def db128(gpu):
    w_seed = bytes(range(9, 41))
    recs = _records(1818, 128)
    rng = np.random.default_rng(1819)
    y = rng.integers(0, Q, (3, N), dtype=np.uint64)
    ctx = gpu.Context(gpu.params_d2048(), w_seed, GAMMA)
    print("launch shapes: building 128 shards of 2048 x 512 B", flush=True)
    for s, r in enumerate(recs):
        ctx.upload_records(s, r.tobytes(), ENTRY).precompute_packing()
    seat = ctx.seat(y)
    yield ctx, seat, recs, y, w_seed
    seat.close()
    ctx.close()


def test_512_query_launch_over_128_shards_at_every_shape(gpu, db128):
    import torch

    ctx, seat, recs, y, w_seed = db128
    L = 512
    rng = np.random.default_rng(512)
    ids = rng.integers(0, 128, L).astype(np.int32)
    seeds = rng.integers(0, 256, (L, 32), dtype=np.uint8)
    bs = rng.integers(0, Q, (L, N), dtype=np.uint64)
    dev = torch.device("cuda:0")
    d_ids = torch.from_numpy(ids).to(dev)
    d_seeds = torch.from_numpy(seeds).to(dev)
    d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
    ws = gpu.Workspace(ctx, L)
    outs = {}
    for qpc in (3, 2, 1):
        print(f"launch shapes: 512 queries at {qpc} per CU", flush=True)
        ctx.set_queries_per_cu(qpc)
        a = torch.zeros((L, N), dtype=torch.int64, device=dev)
        b = torch.zeros((L, GAMMA), dtype=torch.int64, device=dev)
        ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), L, a.data_ptr(),
                                 b.data_ptr())
        torch.cuda.synchronize()
        assert (ws.status(L) == 0).all()
        outs[qpc] = (a.cpu().numpy().view(np.uint64), b.cpu().numpy().view(np.uint64))
    for qpc in (2, 1):
        assert np.array_equal(outs[qpc][0], outs[3][0]) and np.array_equal(outs[qpc][1], outs[3][1]), qpc
    # 8 sampled queries of the three-per-CU launch against the oracle's per-query path
    p = oracle.Pack(N, GAMMA)
    for k in np.random.default_rng(8).choice(L, 8, replace=False):
        sid = int(ids[k])
        print(f"launch shapes: query {k} on shard {sid} against the oracle", flush=True)
        wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k].tobytes(), N), bs[k], _columns(recs[sid]), y)
        assert np.array_equal(outs[3][0][k], wa), k
        assert np.array_equal(outs[3][1][k], wb[:GAMMA]), k
    ws.close()
    ctx.set_queries_per_cu(1)


def test_four_workspaces_on_four_streams(gpu, db128):
    import torch

    ctx, seat, recs, y, w_seed = db128
    B, K = 64, 4
    rng = np.random.default_rng(44)
    dev = torch.device("cuda:0")
    streams = [torch.cuda.Stream() for _ in range(K)]  # default priority
    wss = [gpu.Workspace(ctx, B) for _ in range(K)]
    batches, outs = [], []
    for k in range(K):
        ids = rng.integers(0, 128, B).astype(np.int32)
        seeds = rng.integers(0, 256, (B, 32), dtype=np.uint8)
        bs = rng.integers(0, Q, (B, N), dtype=np.uint64)
        batches.append((ids, seeds, bs))
        outs.append((torch.zeros((B, N), dtype=torch.int64, device=dev),
                     torch.zeros((B, GAMMA), dtype=torch.int64, device=dev),
                     torch.from_numpy(ids).to(dev), torch.from_numpy(seeds).to(dev),
                     torch.from_numpy(bs.view(np.int64)).to(dev)))
    torch.cuda.synchronize()
    print("four streams: launching 4 x 64 queries, no waits between them", flush=True)
    for rep in range(2):  # twice, so each workspace is reused while the others run
        for k in range(K):
            a, b, d_ids, d_seeds, d_b = outs[k]
            ctx.respond_batch_device(seat, wss[k], d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), B,
                                     a.data_ptr(), b.data_ptr(), stream=streams[k].cuda_stream,
                                     switch_modulus=Q36 if k % 2 else 0)
    torch.cuda.synchronize()
    for k in range(K):
        assert (wss[k].status(B) == 0).all(), k
        ids, seeds, bs = batches[k]
        wa, wb = ctx.respond_seeded_batch(seat, ids, [s.tobytes() for s in seeds], bs,
                                          switch_modulus=Q36 if k % 2 else 0)
        assert np.array_equal(outs[k][0].cpu().numpy().view(np.uint64), wa), k
        assert np.array_equal(outs[k][1].cpu().numpy().view(np.uint64), wb), k
    for w in wss:
        w.close()
