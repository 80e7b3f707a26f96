"""Response bytes packed on the device: the 36-bit switched form against the host codec and
the oracle, the device-buffer entry point's stride and padding, and its refusals."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.Q
Q36 = oracle.Q36


def _ctx(gpu, rng, n, gamma, shards=2):
    w_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    cols = [rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64) for _ in range(shards)]
    ctx = gpu.Context(gpu.params_d2048(ring_dim=n), w_seed, gamma)
    for i, c in enumerate(cols):
        ctx.upload_columns(i, c).precompute_packing()
    y = rng.integers(0, Q, (3, n), dtype=np.uint64)
    return ctx, cols, y, w_seed


@pytest.mark.parametrize("n,gamma", [(2048, 256), (2048, 16), (256, 1), (256, 2), (256, 128)])
@pytest.mark.parametrize("switch", [0, Q36])
def test_wire_bytes_equal_host_codec_and_oracle(gpu, n, gamma, switch):
    rng = np.random.default_rng(n + gamma + (switch & 7))
    ctx, cols, y, w_seed = _ctx(gpu, rng, n, gamma)
    with ctx:
        seat = ctx.seat(y)
        count = 5
        ids = rng.integers(0, len(cols), count)
        seeds = [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in range(count)]
        bs = rng.integers(0, Q, (count, n), dtype=np.uint64)
        print(f"wire n {n} gamma {gamma} switch {switch}", flush=True)
        wires = ctx.respond_seeded_batch_wire(seat, ids, seeds, bs, switch_modulus=switch)
        A, B = ctx.respond_seeded_batch(seat, ids, seeds, bs, switch_modulus=switch)
        seat.close()
        size = ctx.response_bytes(switch)
    if (n, gamma, switch) == (2048, 256, Q36):
        assert size == 10446
    if (n, gamma, switch) == (2048, 256, 0):
        assert size == 17358
    p = oracle.Pack(n, gamma)
    for k in range(count):
        assert len(wires[k]) == size
        assert wires[k] == gpu.encode_response(A[k], B[k], switch or Q)
        if k == 2:
            wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k], n), bs[k], cols[ids[k]], y)
            wb = wb[:gamma]
            if switch:
                wa, wb = oracle.mod_switch(wa), oracle.mod_switch(wb)
            assert wires[k] == oracle.encode_response(wa, wb, gamma, switch or Q)


def test_device_buffer_stride_padding_and_refusals(gpu):
    import torch

    n, gamma, count = 2048, 256, 6
    rng = np.random.default_rng(9)
    dev = torch.device("cuda:0")
    ctx, cols, y, _ = _ctx(gpu, rng, n, gamma)
    with ctx:
        seat = ctx.seat(y)
        ws = gpu.Workspace(ctx, count)
        ids = rng.integers(0, len(cols), count).astype(np.int32)
        seeds = rng.integers(0, 256, (count, 32), dtype=np.uint8)
        bs = rng.integers(0, Q, (count, n), dtype=np.uint64)
        d_ids = torch.from_numpy(ids).to(dev)
        d_seeds = torch.from_numpy(seeds).to(dev)
        d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
        size = ctx.response_bytes(Q36)
        stride = ((size + 3) // 4) * 4 + 16
        d_wire = torch.full((count * stride,), 0xAB, dtype=torch.uint8, device=dev)
        ctx.respond_batch_device_wire(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                      d_wire.data_ptr(), stride, switch_modulus=Q36)
        torch.cuda.synchronize()
        got = d_wire.cpu().numpy().reshape(count, stride)
        want = ctx.respond_seeded_batch_wire(seat, ids, [s.tobytes() for s in seeds], bs, switch_modulus=Q36)
        r4 = ((size + 3) // 4) * 4
        for k in range(count):
            assert got[k, :size].tobytes() == want[k]
            assert (got[k, size:r4] == 0).all()          # the last word's pad bytes
            assert (got[k, r4:] == 0xAB).all()           # nothing past the rounded length
        for bad in (size, r4 - 4, r4 + 2):
            with pytest.raises(gpu.PirError, match="multiple of 4"):
                ctx.respond_batch_device_wire(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                              d_wire.data_ptr(), bad, switch_modulus=Q36)
        with pytest.raises(gpu.PirError, match="aligned"):
            ctx.respond_batch_device_wire(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                          d_wire.data_ptr() + 2, stride, switch_modulus=Q36)
        torch.cuda.synchronize()
        ws.close()
        seat.close()
