"""Shard-grouped launches against ungrouped ones, bit for bit.

A device launch of up to 4096 queries sorts its queries by shard on the device (k_group), scans
each distinct shard's records once for all of its queries (k_scan_grouped) and runs the packing
chains in shard order, a slice per XCD. inspire_respond_batch_device_noalloc keeps the per-query
scan and launch order, so it is the reference here; a few queries also go to the oracle.
Cases: long runs of one shard (more queries than one pass of the grouped scan takes), ids that
name no shard or a shard without a store (status -1 in both paths), every shard hit once, and a
launch past the grouping limit.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.Q
N = 2048


def _run_both(gpu, torch, ctx, seat, ids, gamma, seed, switch_modulus=0):
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    count = len(ids)
    seeds = rng.integers(0, 256, (count, 32), dtype=np.uint8)
    bs = rng.integers(0, Q, (count, N), dtype=np.uint64)
    d_ids = torch.from_numpy(np.asarray(ids, dtype=np.int32)).to(dev)
    d_seeds = torch.from_numpy(seeds).to(dev)
    d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
    outs = []
    for noalloc in (False, True):
        a = torch.zeros((count, N), dtype=torch.int64, device=dev)
        b = torch.zeros((count, gamma), dtype=torch.int64, device=dev)
        if noalloc:
            st = torch.zeros(count, dtype=torch.int32, device=dev)
            gpu.respond_batch_device_noalloc(ctx, seat, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                             a.data_ptr(), b.data_ptr(), st.data_ptr(),
                                             switch_modulus=switch_modulus)
            torch.cuda.synchronize()
            status = st.cpu().numpy()
        else:
            ws = gpu.Workspace(ctx, count)
            ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                     a.data_ptr(), b.data_ptr(), switch_modulus=switch_modulus)
            torch.cuda.synchronize()
            status = ws.status(count)
            ws.close()
        outs.append((a.cpu().numpy().view(np.uint64), b.cpu().numpy().view(np.uint64), status))
    return outs, seeds, bs


def test_grouped_equals_ungrouped_on_skewed_and_invalid_batches(gpu):
    import torch

    gamma = 256
    rng = np.random.default_rng(71)
    w_seed = bytes(range(1, 33))
    cols = [rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64) for _ in range(4)]
    y = rng.integers(0, Q, (3, N), dtype=np.uint64)
    with gpu.Context(gpu.params_d2048(), w_seed, gamma) as ctx:
        for i in range(3):
            ctx.upload_columns(i, cols[i]).precompute_packing()
        ctx.upload_columns(3, cols[3])  # no store: the device path refuses it
        seat = ctx.seat(y)
        # runs of about 80 per shard in shuffled order, 5 queries naming shard 3 and 4 past the table
        ids = np.concatenate([np.repeat([0, 1, 2], 80), [3] * 5, [9, 100, 7, 2 ** 31 - 1]])
        ids = ids[np.random.default_rng(5).permutation(len(ids))]
        print(f"grouping: {len(ids)} queries over 3 stored shards with invalid ids", flush=True)
        ((ga, gb, gs), (ua, ub, us)), seeds, bs = _run_both(gpu, torch, ctx, seat, ids, gamma, 72, oracle.Q36)
        ok = ids < 3
        assert (gs[ok] == 0).all() and (gs[~ok] == -1).all()
        assert np.array_equal(gs, us)
        assert np.array_equal(ga[ok], ua[ok]) and np.array_equal(gb[ok], ub[ok])
        ((ga, gb, gs), _), seeds, bs = _run_both(gpu, torch, ctx, seat, ids, gamma, 73)
        p = oracle.Pack(N, gamma)
        for k in np.flatnonzero(ok)[:3]:
            wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k].tobytes(), N), bs[k], cols[ids[k]], y)
            assert np.array_equal(ga[k], wa) and np.array_equal(gb[k], wb[:gamma]), k
        seat.close()


@pytest.mark.parametrize("count", [37, 512])
def test_grouped_equals_ungrouped_one_query_per_shard_and_mixed(gpu, count):
    import torch

    gamma = 16
    rng = np.random.default_rng(80 + count)
    with gpu.Context(gpu.params_d2048(), bytes(32), gamma) as ctx:
        for i in range(64):
            ctx.upload_columns(i, rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64)).precompute_packing()
        seat = ctx.seat(rng.integers(0, Q, (3, N), dtype=np.uint64))
        ids = rng.integers(0, 64, count)
        ids[:min(count, 64)] = np.arange(min(count, 64))
        print(f"grouping: {count} queries over 64 shards", flush=True)
        ((ga, gb, gs), (ua, ub, us)), _, _ = _run_both(gpu, torch, ctx, seat, ids, gamma, 90 + count)
        assert (gs == 0).all() and (us == 0).all()
        assert np.array_equal(ga, ua) and np.array_equal(gb, ub)
        seat.close()


def test_past_the_grouping_limit(gpu):
    import torch

    gamma = 16
    rng = np.random.default_rng(91)
    with gpu.Context(gpu.params_d2048(), bytes(32), gamma) as ctx:
        for i in range(8):
            ctx.upload_columns(i, rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64)).precompute_packing()
        seat = ctx.seat(rng.integers(0, Q, (3, N), dtype=np.uint64))
        dev = torch.device("cuda:0")
        count = 4097  # one past the grouping sort: scanned query by query
        ids = rng.integers(0, 8, count).astype(np.int32)
        seeds = rng.integers(0, 256, (count, 32), dtype=np.uint8)
        bs = rng.integers(0, Q, (count, N), dtype=np.uint64)
        d_ids, d_seeds = torch.from_numpy(ids).to(dev), torch.from_numpy(seeds).to(dev)
        d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
        outs = []
        print("grouping: 4097 queries (ungrouped) against the first 4096 (grouped)", flush=True)
        for c in (count, count - 1):
            ws = gpu.Workspace(ctx, c)
            a = torch.zeros((c, N), dtype=torch.int64, device=dev)
            b = torch.zeros((c, gamma), dtype=torch.int64, device=dev)
            ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), c, a.data_ptr(),
                                     b.data_ptr())
            torch.cuda.synchronize()
            assert (ws.status(c) == 0).all()
            outs.append((a.cpu().numpy(), b.cpu().numpy()))
            ws.close()
        assert np.array_equal(outs[0][0][:count - 1], outs[1][0]) and np.array_equal(outs[0][1][:count - 1], outs[1][1])
        seat.close()


@pytest.mark.parametrize("gamma", [256, 512, 1024])
def test_wide_shards_runs_of_every_length(gpu, gamma):
    """Wide shards (gamma >= 256, one 512 to 2048-byte row per wave of the grouped scan) are
    scanned in tasks of up to two queries of one shard. Runs of 1 to 9 queries per shard cover
    odd and even runs and their cuts into tasks; one shard's column edited after its store was
    built is refused for exactly its queries. At gamma 256 every query is compared with the
    ungrouped no-alloc path (which serves up to gamma 256), and at every width queries from runs
    of 1, 4, 5 and 8 with the oracle."""
    import torch

    rng = np.random.default_rng(200 + gamma)
    w_seed = bytes(range(40, 72))
    shards = 10
    cols = [rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64) for _ in range(shards)]
    y = rng.integers(0, Q, (3, N), dtype=np.uint64)
    with gpu.Context(gpu.params_d2048(), w_seed, gamma) as ctx:
        handles = [ctx.upload_columns(i, cols[i]) for i in range(shards)]
        for h in handles:
            h.precompute_packing()
        handles[9].write_value(1000, gamma - 1, int(cols[9][gamma - 1, 1000]) ^ 3)
        seat = ctx.seat(y)
        ids = np.concatenate([np.full(k, k, dtype=np.int64) for k in range(1, 10)] + [np.full(3, 9)])
        ids = ids[np.random.default_rng(gamma).permutation(len(ids))]
        edited = ids == 9
        print(f"grouping: gamma {gamma}, runs of 1-9 queries per shard, {len(ids)} queries", flush=True)
        if gamma == 256:
            ((ga, gb, gs), (ua, ub, us)), _, _ = _run_both(gpu, torch, ctx, seat, ids, gamma, 300, oracle.Q36)
            assert (us[edited] == gamma).all() and (us[~edited] == 0).all()
            assert np.array_equal(gs, us)
            assert np.array_equal(ga[~edited], ua[~edited]) and np.array_equal(gb[~edited], ub[~edited])
        dev = torch.device("cuda:0")
        count = len(ids)
        seeds = rng.integers(0, 256, (count, 32), dtype=np.uint8)
        bs = rng.integers(0, Q, (count, N), dtype=np.uint64)
        bs[:, 5] = 0  # a zero b word: its reversed word is zero too (brev[n - 5])
        ws = gpu.Workspace(ctx, count)
        a = torch.zeros((count, N), dtype=torch.int64, device=dev)
        b = torch.zeros((count, gamma), dtype=torch.int64, device=dev)
        d_ids = torch.from_numpy(ids.astype(np.int32)).to(dev)
        d_seeds = torch.from_numpy(seeds).to(dev)
        d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
        ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count, a.data_ptr(),
                                 b.data_ptr())
        torch.cuda.synchronize()
        status = ws.status(count)
        ws.close()
        assert (status[edited] == gamma).all() and (status[~edited] == 0).all()
        ga, gb = a.cpu().numpy().view(np.uint64), b.cpu().numpy().view(np.uint64)
        p = oracle.Pack(N, gamma)
        for run in (1, 4, 5, 8):
            k = int(np.flatnonzero(ids == run)[-1])
            wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k].tobytes(), N), bs[k], cols[ids[k]], y)
            assert np.array_equal(ga[k], wa) and np.array_equal(gb[k], wb[:gamma]), (run, k)
        seat.close()
