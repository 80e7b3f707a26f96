"""Gadget-key rings 512 and 1024 on the GPU (InspireParams with ring_dim 512 or 1024 at the d2048
prime and gadget: Params::validate accepts every power-of-two ring whose 2d divides q - 1,
src/params.rs:535-580), one workgroup of n/4 threads per query.

The reference commits no response vector at these rings, so the device is checked against the
oracle's restatement (the code the d256 and d2048 vectors pin, with the ring size a parameter),
exact and approximate gadget, unswitched and at 36 bits, on the host-buffer and device-buffer
paths, and through the client: responses of a packing-key cell built at the ring in the
fixture's draw order decode to the planted records.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.Q


@pytest.mark.parametrize("n", [512, 1024])
@pytest.mark.parametrize("gamma,approx", [(16, False), (128, True), (256, False)])
def test_ring_seeded_matches_oracle(gpu, n, gamma, approx):
    import torch

    rng = np.random.default_rng(n + gamma + approx)
    rows = 2 if approx else 3
    cols = rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64)
    y = rng.integers(0, Q, (rows, n), dtype=np.uint64)
    count = 5
    ids = [0, 1, 0, 1, 1]
    bs = rng.integers(0, Q, (count, n), dtype=np.uint64)
    seeds = [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in range(count)]
    w_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    cols1 = rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64)
    p = oracle.Pack(n, gamma, 3, 1 if approx else 0)
    print(f"ring {n}: gamma {gamma}, approximate {approx}", flush=True)
    with gpu.Context(gpu.params_d2048(approx, ring_dim=n), w_seed, gamma) as ctx:
        ctx.upload_columns(0, cols).precompute_packing()
        ctx.upload_columns(1, cols1).precompute_packing()
        seat = ctx.seat(y)
        a, b = ctx.respond_seeded_batch(seat, ids, seeds, bs)
        sa, sb = ctx.respond_seeded_batch(seat, ids, seeds, bs, switch_modulus=oracle.Q36)
        dev = torch.device("cuda:0")
        ws = gpu.Workspace(ctx, count)
        da = torch.zeros((count, n), dtype=torch.int64, device=dev)
        db = torch.zeros((count, gamma), dtype=torch.int64, device=dev)
        d_ids = torch.tensor(ids, dtype=torch.int32, device=dev)
        d_seeds = torch.from_numpy(np.frombuffer(b"".join(seeds), dtype=np.uint8).copy()).to(dev)
        d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
        ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), count,
                                 da.data_ptr(), db.data_ptr())
        torch.cuda.synchronize()
        assert (ws.status(count) == 0).all()
        ws.close()
        seat.close()
    assert np.array_equal(da.cpu().numpy().view(np.uint64), a) and np.array_equal(db.cpu().numpy().view(np.uint64), b)
    for k in range(count):
        c = cols if ids[k] == 0 else cols1
        wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[k], n), bs[k], c, y)
        assert np.array_equal(a[k], wa) and np.array_equal(b[k], wb[:gamma]), k
        assert np.array_equal(sa[k], oracle.mod_switch(wa)) and np.array_equal(sb[k], oracle.mod_switch(wb[:gamma]))


@pytest.mark.parametrize("n", [512, 1024])
@pytest.mark.parametrize("approximate", [False, True], ids=["exact", "approximate"])
def test_ring_responses_decode_to_the_planted_records(gpu, n, approximate):
    entry = 32
    c = oracle.Cell(n, entry, 0x5249_4e47 + n, approximate=approximate)
    indices = [0, 3, n - 1, n, 2 * n + 1, 3 * n - 1]
    queries = [c.query(i) for i in indices]
    print(f"ring {n} decode cell: approximate {approximate}, gamma {c.gamma}", flush=True)
    with gpu.Context(gpu.params_d2048(approximate=approximate, ring_dim=n), c.w_seed, c.gamma) as ctx:
        for shard in range(3):
            ctx.upload_records(shard, c.db[shard * n * entry:(shard + 1) * n * entry].tobytes(),
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
