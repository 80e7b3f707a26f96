"""inspire_respond_batch_device_noalloc (the entry point that allocates nothing) against
inspire_respond_batch_device (the same kernels over a workspace), bit for bit: two calls in a
row on one context, the second smaller, then a second context of another width on the same
GPU (which shares the library's scratch array), and the size refusal."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.Q
N = 2048


def _ctx(gpu, gamma, shards, seed):
    rng = np.random.default_rng(seed)
    ctx = gpu.Context(gpu.params_d2048(), bytes(rng.integers(0, 256, 32, dtype=np.uint8)), gamma)
    for s in range(shards):
        ctx.upload_columns(s, rng.integers(0, 1 << 16, (gamma, N), dtype=np.uint64)).precompute_packing()
    return ctx, ctx.seat(rng.integers(0, Q, (3, N), dtype=np.uint64))


def _compare(gpu, torch, ctx, seat, count, shards, gamma, seed, switch_modulus):
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    d_ids = torch.from_numpy(rng.integers(0, shards, count).astype(np.int32)).to(dev)
    d_seeds = torch.from_numpy(rng.integers(0, 256, (count, 32), dtype=np.uint8)).to(dev)
    d_b = torch.from_numpy(rng.integers(0, Q, (count, N), dtype=np.uint64).view(np.int64)).to(dev)
    outs = []
    for noalloc in (True, False):
        a = torch.full((count, N), -1, dtype=torch.int64, device=dev)
        b = torch.full((count, gamma), -1, dtype=torch.int64, device=dev)
        st = torch.full((count,), 7, dtype=torch.int32, device=dev)
        if noalloc:
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
        outs.append((a.cpu().numpy(), b.cpu().numpy(), status))
    (a0, b0, s0), (a1, b1, s1) = outs
    assert (s0 == 0).all() and (s1 == 0).all()
    assert np.array_equal(a0, a1) and np.array_equal(b0, b1)


def test_noalloc_equals_the_workspace_path(gpu):
    import torch

    gamma, shards = 256, 6
    print("no-alloc: gamma 256, 256 then 37 queries", flush=True)
    ctx, seat = _ctx(gpu, gamma, shards, 5)
    try:
        _compare(gpu, torch, ctx, seat, 256, shards, gamma, 50, oracle.Q36)  # the most the scratch holds
        _compare(gpu, torch, ctx, seat, 37, shards, gamma, 51, 0)
        with pytest.raises(gpu.PirError, match="scan scratch"):
            gpu.respond_batch_device_noalloc(ctx, seat, 1, 1, 1, 257, 1, 1, 1)
    finally:
        seat.close()
        ctx.close()
    print("no-alloc: gamma 16, second context on the same GPU", flush=True)
    ctx, seat = _ctx(gpu, 16, 3, 6)
    try:
        _compare(gpu, torch, ctx, seat, 300, 3, 16, 52, oracle.Q36)
        _compare(gpu, torch, ctx, seat, 5, 3, 16, 53, 0)
    finally:
        seat.close()
        ctx.close()
