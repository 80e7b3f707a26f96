"""The respond pipeline captured into a hipGraph, and the launch-shape autotuner.

Each replay runs over fixed device buffers whose inputs are rewritten between replays by
eager torch work; every replay is checked against the eager device path on the same inputs
(which the parity tests pin to the CPU oracle) and one query against the oracle directly.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.Q
Q36 = 68_718_428_161


def _setup(gpu, rng, n=2048, gamma=256, shards=3):
    w_seed = bytes(range(32))
    y = rng.integers(0, Q, (3, n), dtype=np.uint64)
    cols = [rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64) for _ in range(shards)]
    ctx = gpu.Context(gpu.params_d2048(), w_seed, gamma)
    for i, c in enumerate(cols):
        ctx.upload_columns(i, c).precompute_packing()
    return ctx, ctx.seat(y), cols, y, w_seed


def _batch(rng, count, shards, n):
    ids = rng.integers(0, shards, count).astype(np.int32)
    seeds = rng.integers(0, 256, (count, 32), dtype=np.uint8)
    bs = rng.integers(0, Q, (count, n), dtype=np.uint64)
    return ids, seeds, bs


@pytest.mark.parametrize("switch_modulus", [0, Q36])
def test_graph_replays_equal_eager(gpu, switch_modulus):
    import torch

    n, gamma, count = 2048, 256, 8
    rng = np.random.default_rng(77 + switch_modulus % 5)
    dev = torch.device("cuda:0")
    ctx, seat, cols, y, w_seed = _setup(gpu, rng)
    ws = gpu.Workspace(ctx, count)
    d_ids = torch.zeros(count, dtype=torch.int32, device=dev)
    d_seeds = torch.zeros((count, 32), dtype=torch.uint8, device=dev)
    d_b = torch.zeros((count, n), dtype=torch.int64, device=dev)
    d_a_out = torch.zeros((count, n), dtype=torch.int64, device=dev)
    d_b_out = torch.zeros((count, gamma), dtype=torch.int64, device=dev)
    e_a = torch.zeros_like(d_a_out)
    e_b = torch.zeros_like(d_b_out)
    stream = torch.cuda.current_stream().cuda_stream
    graph = gpu.RespondGraph(ctx, seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                             d_a_out.data_ptr(), d_b_out.data_ptr(), switch_modulus=switch_modulus)
    try:
        for rep in range(4):
            ids, seeds, bs = _batch(rng, count, len(cols), n)
            print(f"graph replay {rep}: switch {switch_modulus}, shards {ids.tolist()}", flush=True)
            d_ids.copy_(torch.from_numpy(ids))
            d_seeds.copy_(torch.from_numpy(seeds))
            d_b.copy_(torch.from_numpy(bs.view(np.int64)))
            d_a_out.fill_(-1)
            d_b_out.fill_(-1)
            torch.cuda.synchronize()
            graph.launch(stream)
            torch.cuda.synchronize()
            got_a, got_b = d_a_out.cpu().numpy().view(np.uint64), d_b_out.cpu().numpy().view(np.uint64)
            # the eager path on the same inputs, into separate outputs
            ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                     e_a.data_ptr(), e_b.data_ptr(), stream=stream, switch_modulus=switch_modulus)
            torch.cuda.synchronize()
            assert (ws.status(count) == 0).all()
            assert np.array_equal(got_a, e_a.cpu().numpy().view(np.uint64)), f"replay {rep}: a differs"
            assert np.array_equal(got_b, e_b.cpu().numpy().view(np.uint64)), f"replay {rep}: b differs"
            if rep == 1:
                k = 5
                p = oracle.Pack(n, gamma)
                sid = int(ids[k])
                wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k].tobytes(), n), bs[k], cols[sid], y,
                                   p.store_build(cols[sid]))
                wb = wb[:gamma]
                if switch_modulus:
                    wa, wb = oracle.mod_switch(wa), oracle.mod_switch(wb)
                assert np.array_equal(got_a[k], wa) and np.array_equal(got_b[k], wb)
    finally:
        graph.close()
        ws.close()
        seat.close()
        ctx.close()


def test_autotune_keeps_results(gpu):
    import torch

    n, gamma, count = 2048, 256, 64
    rng = np.random.default_rng(5)
    dev = torch.device("cuda:0")
    ctx, seat, cols, y, _ = _setup(gpu, rng)
    ws = gpu.Workspace(ctx, count)
    ids, seeds, bs = _batch(rng, count, len(cols), n)
    d_ids = torch.from_numpy(ids).to(dev)
    d_seeds = torch.from_numpy(seeds).to(dev)
    d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
    outs = []
    try:
        for qpc in (1, 2, 3):
            ctx.set_queries_per_cu(qpc)
            a = torch.zeros((count, n), dtype=torch.int64, device=dev)
            b = torch.zeros((count, gamma), dtype=torch.int64, device=dev)
            ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                     a.data_ptr(), b.data_ptr())
            torch.cuda.synchronize()
            outs.append((a.cpu().numpy(), b.cpu().numpy()))
        for o in outs[1:]:
            assert np.array_equal(outs[0][0], o[0]) and np.array_equal(outs[0][1], o[1])
        r = gpu.autotune(ctx, seat, ws, count)
        print(f"autotune at {count} queries: {r}", flush=True)
        assert r["queries_per_cu"] in (1, 2, 3) and min(r["ms_qpc1"], r["ms_qpc2"], r["ms_qpc3"]) > 0
        a = torch.zeros((count, n), dtype=torch.int64, device=dev)
        b = torch.zeros((count, gamma), dtype=torch.int64, device=dev)
        ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                 a.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(a.cpu().numpy(), outs[0][0]) and np.array_equal(b.cpu().numpy(), outs[0][1])
        with pytest.raises(gpu.PirError, match="does not fit"):
            gpu.autotune(ctx, seat, ws, count + 1)
    finally:
        ws.close()
        seat.close()
        ctx.close()


def test_scan_splits_keep_results_and_autotune_picks_one(gpu):
    """The DB scan's row blocks per shard (8, 16, 32) change no byte: device launches at each, the
    no-alloc path at 32 (its scratch holds fewer queries then) and a workspace created before the
    change all equal the launch at 16; the scan autotuner keeps one of the three, and a query
    after tuning equals the oracle."""
    import torch

    n, gamma, count = 2048, 256, 64
    rng = np.random.default_rng(6)
    dev = torch.device("cuda:0")
    ctx, seat, cols, y, w_seed = _setup(gpu, rng)
    ws = gpu.Workspace(ctx, count)
    ids, seeds, bs = _batch(rng, count, len(cols), n)
    d_ids = torch.from_numpy(ids).to(dev)
    d_seeds = torch.from_numpy(seeds).to(dev)
    d_b = torch.from_numpy(bs.view(np.int64)).to(dev)

    def launch():
        a = torch.zeros((count, n), dtype=torch.int64, device=dev)
        b = torch.zeros((count, gamma), dtype=torch.int64, device=dev)
        ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                 a.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        assert (ws.status(count) == 0).all()
        return a.cpu().numpy(), b.cpu().numpy()

    try:
        want = launch()
        for k in (8, 32, 16):
            print(f"scan splits {k}", flush=True)
            ctx.set_scan_splits(k)
            got = launch()
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), k
        ctx.set_scan_splits(32)
        a = torch.zeros((count, n), dtype=torch.int64, device=dev)
        b = torch.zeros((count, gamma), dtype=torch.int64, device=dev)
        st = torch.zeros(count, dtype=torch.int32, device=dev)
        gpu.respond_batch_device_noalloc(ctx, seat, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                         a.data_ptr(), b.data_ptr(), st.data_ptr())
        torch.cuda.synchronize()
        assert (st.cpu().numpy() == 0).all()
        assert np.array_equal(a.cpu().numpy(), want[0]) and np.array_equal(b.cpu().numpy(), want[1])
        with pytest.raises(gpu.PirError, match="8, 16 or 32"):
            ctx.set_scan_splits(12)
        r = gpu.autotune_scan(ctx, seat, ws, count)
        print(f"scan autotune at {count} queries: {r}", flush=True)
        assert r["scan_splits"] in (8, 16, 32)
        assert min(r["scan_ms_ks8"], r["scan_ms_ks16"], r["scan_ms_ks32"]) > 0
        got = launch()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        p = oracle.Pack(n, gamma)
        k = 3
        wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k].tobytes(), n), bs[k], cols[ids[k]], y)
        assert np.array_equal(got[0][k].view(np.uint64), wa) and np.array_equal(got[1][k].view(np.uint64), wb[:gamma])
    finally:
        ws.close()
        seat.close()
        ctx.close()
