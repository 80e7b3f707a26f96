"""GPU parity: the HIP respond path against the reference's golden and the CPU oracle.

Integer path, so every comparison is bit-exact.
"""
import numpy as np
import pytest

import oracle
import wire

pytestmark = pytest.mark.gpu

Q = oracle.Q


def _random_case(n, gamma, rows, seed):
    rng = np.random.default_rng(seed)
    cols = rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64)
    y = rng.integers(0, Q, (rows, n), dtype=np.uint64)
    b = rng.integers(0, Q, n, dtype=np.uint64)
    qseed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    w_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    return cols, y, b, qseed, w_seed


def test_d256_reference_fixture_byte_identity(gpu):
    # tests/respond_byte_identity_kat.rs:108-145, through the C ABI
    fx = wire.load_d256_fixture()
    n, g = fx["params"]["ring_dim"], fx["num_columns"]
    with gpu.Context(gpu.params_d2048(ring_dim=n), fx["w_seed"], g) as ctx:
        ctx.upload_columns(0, fx["columns"])
        a_hat, b_prefix = ctx.respond_inspiring(fx["y_body"], 0, fx["a"], fx["b"])
        resp = gpu.encode_response(a_hat, b_prefix)
        assert oracle.fnv1a(resp) == wire.D256_GOLDEN_FNV
        assert resp == oracle.encode_response(*oracle.Pack(n, g).respond(
            fx["w_seed"], fx["a"], fx["b"], fx["columns"], fx["y_body"]), g)


@pytest.mark.parametrize("gamma,approx", [(16, False), (256, False), (16, True), (256, True), (512, False),
                                          (1, False), (2, True), (1024, False)])
def test_d2048_seeded_matches_oracle(gpu, gamma, approx):
    n = 2048
    rows = 2 if approx else 3
    cols, y, b, qseed, w_seed = _random_case(n, gamma, rows, seed=gamma + 7 * approx)
    p = oracle.Pack(n, gamma, 3, 1 if approx else 0)
    a = oracle.expand_seed(qseed, n)
    want_a, want_b = p.respond(w_seed, a, b, cols, y, p.store_build(cols))
    with gpu.Context(gpu.params_d2048(approx), w_seed, gamma) as ctx:
        sh = ctx.upload_columns(0, cols)
        sh.precompute_packing()
        got_a, got_b = ctx.respond_seeded_inspiring(y, 0, qseed, b)
        assert np.array_equal(got_a, want_a)
        assert np.array_equal(got_b, want_b[:gamma])
        # 36-bit modulus switch of the same response (mod_switch.rs:148-175)
        sa, sb = ctx.respond_seeded_inspiring(y, 0, qseed, b, switch_modulus=oracle.Q36)
        assert np.array_equal(sa, oracle.mod_switch(want_a))
        assert np.array_equal(sb, oracle.mod_switch(want_b[:gamma]))
        if gamma == 256:
            assert len(gpu.encode_response(sa, sb, oracle.Q36)) == 10446


def test_batch_equals_single_queries(gpu):
    n, gamma = 2048, 256
    rng = np.random.default_rng(11)
    w_seed = bytes(range(32))
    y = rng.integers(0, Q, (3, n), dtype=np.uint64)
    shards = [rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64) for _ in range(3)]
    with gpu.Context(gpu.params_d2048(), w_seed, gamma) as ctx:
        for i, c in enumerate(shards):
            ctx.upload_columns(i, c).precompute_packing()
        seat = ctx.seat(y)
        ids = [2, 0, 2, 1, 0]
        seeds = [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in ids]
        bs = rng.integers(0, Q, (len(ids), n), dtype=np.uint64)
        A, B = ctx.respond_seeded_batch(seat, ids, seeds, bs)
        for k, sid in enumerate(ids):
            a1, b1 = ctx.respond_seeded_batch(seat, [sid], [seeds[k]], bs[k:k + 1])
            assert np.array_equal(A[k], a1[0]) and np.array_equal(B[k], b1[0])
        # and the oracle for one of them
        p = oracle.Pack(n, gamma)
        wa, wb = p.respond(w_seed, oracle.expand_seed(seeds[3], n), bs[3], shards[1], y, p.store_build(shards[1]))
        assert np.array_equal(A[3], wa) and np.array_equal(B[3], wb[:gamma])
        seat.close()


def test_records_upload_equals_columns_upload(gpu):
    n, gamma, entry = 2048, 16, 31  # odd entry: the last column holds one byte
    rng = np.random.default_rng(5)
    recs = rng.integers(0, 256, n * entry, dtype=np.uint8).tobytes()
    cols = np.zeros((gamma, n), dtype=np.uint64)
    for m in range(n):
        e = recs[m * entry:(m + 1) * entry]
        for j in range(gamma):
            o = 2 * j
            cols[j, m] = e[o] | (e[o + 1] << 8) if o + 1 < entry else e[o]
    y = rng.integers(0, Q, (3, n), dtype=np.uint64)
    b = rng.integers(0, Q, n, dtype=np.uint64)
    with gpu.Context(gpu.params_d2048(), bytes(32), gamma) as ctx:
        ctx.upload_records(0, recs, entry)
        ctx.upload_columns(1, cols)
        r0 = ctx.respond_seeded_inspiring(y, 0, bytes(32), b)
        r1 = ctx.respond_seeded_inspiring(y, 1, bytes(32), b)
        assert np.array_equal(r0[0], r1[0]) and np.array_equal(r0[1], r1[1])


def test_edited_shard_is_refused_before_packing(gpu):
    # tests/shard_packing_store.rs:328-368: a column edited after the store was built
    n, gamma = 2048, 16
    cols, y, b, qseed, w_seed = _random_case(n, gamma, 3, seed=3)
    with gpu.Context(gpu.params_d2048(), w_seed, gamma) as ctx:
        sh = ctx.upload_columns(0, cols)
        sh.precompute_packing()
        ctx.respond_seeded_inspiring(y, 0, qseed, b)
        sh.write_value(17, 5, int(cols[5, 17]) ^ 1)
        with pytest.raises(gpu.PirError) as e:
            ctx.respond_seeded_inspiring(y, 0, qseed, b)
        assert e.value.code == gpu.ERR_STORE
        assert "column 5 is not the column the store was built from" in str(e.value)
        sh.write_value(17, 5, int(cols[5, 17]))  # restored: served again
        ctx.respond_seeded_inspiring(y, 0, qseed, b)


def test_refusals(gpu):
    n = 2048
    with pytest.raises(gpu.PirError, match="not a legal InspiRING width"):
        gpu.Context(gpu.params_d2048(), bytes(32), 24)
    with gpu.Context(gpu.params_d2048(), bytes(32), 16) as ctx:
        with pytest.raises(gpu.PirError, match="column-count mismatch"):
            ctx.upload_columns(0, np.zeros((8, n), dtype=np.uint64))
        ctx.upload_columns(0, np.zeros((16, n), dtype=np.uint64))
        with pytest.raises(gpu.PirError, match="key rows"):
            ctx.seat(np.zeros((2, n), dtype=np.uint64))
        with pytest.raises(gpu.PirError, match="gamma"):
            ctx.seat(np.zeros((3, n), dtype=np.uint64), num_to_pack=32)
        seat = ctx.seat(np.zeros((3, n), dtype=np.uint64))
        with pytest.raises(gpu.PirError, match="shard 4 not found"):
            ctx.respond_seeded_batch(seat, [4], [bytes(32)], np.zeros((1, n), dtype=np.uint64))
        bad = np.zeros((1, n), dtype=np.uint64)
        bad[0, 9] = Q
        with pytest.raises(gpu.PirError, match="not canonical"):
            ctx.respond_seeded_batch(seat, [0], [bytes(32)], bad)
        seat.close()
        # sizes past INSPIRE_MAX_BATCH are refused before anything is allocated
        with pytest.raises(gpu.PirError, match="at most 4194304 queries"):
            gpu.Workspace(ctx, (1 << 22) + 1)
        with pytest.raises(gpu.PirError, match="1 to 4194304 queries"):
            gpu.Server(ctx, max_launch=(1 << 22) + 1)


def test_empty_batch(gpu):
    """respond_seeded_batch over no queries answers an empty list, after resolving its key
    (respond.rs:1189-1202): a seat of another context is refused even then."""
    import torch

    n, gamma = 2048, 16
    with gpu.Context(gpu.params_d2048(), bytes(32), gamma) as ctx, \
            gpu.Context(gpu.params_d2048(), bytes(32), gamma) as other:
        ctx.upload_columns(0, np.zeros((gamma, n), dtype=np.uint64))
        seat = ctx.seat(np.zeros((3, n), dtype=np.uint64))
        foreign = other.seat(np.zeros((3, n), dtype=np.uint64))
        a, b = ctx.respond_seeded_batch(seat, [], [], np.zeros((0, n), dtype=np.uint64))
        assert a.shape == (0, n) and b.shape == (0, gamma)
        with pytest.raises(gpu.PirError, match="another context"):
            ctx.respond_seeded_batch(foreign, [], [], np.zeros((0, n), dtype=np.uint64))
        ws = gpu.Workspace(ctx, 4)
        dummy = torch.zeros(8, dtype=torch.int64, device="cuda:0")
        ctx.respond_batch_device(seat, ws, dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), 0,
                                 dummy.data_ptr(), dummy.data_ptr())
        torch.cuda.synchronize()
        with pytest.raises(gpu.PirError, match="another context"):
            ctx.respond_batch_device(foreign, ws, dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), 0,
                                     dummy.data_ptr(), dummy.data_ptr())
        ws.close()
        foreign.close()
        seat.close()


def test_kernel_trace_keeps_results(gpu):
    """Launches traced by the workspace (events around the scan and chain kernels, read after
    the loop) answer exactly as untraced ones; only the armed number of launches is recorded."""
    import torch

    n, gamma, L = 2048, 16, 8
    rng = np.random.default_rng(5)
    dev = torch.device("cuda:0")
    with gpu.Context(gpu.params_d2048(), bytes(32), gamma) as ctx:
        ctx.upload_columns(0, rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64)).precompute_packing()
        seat = ctx.seat(rng.integers(0, gpu.Q, (3, n), dtype=np.uint64))
        ws = gpu.Workspace(ctx, L)
        ids = torch.zeros(L, dtype=torch.int32, device=dev)
        seeds = torch.from_numpy(rng.integers(0, 256, (L, 32), dtype=np.uint8)).to(dev)
        b = torch.from_numpy(rng.integers(0, gpu.Q, (L, n), dtype=np.uint64).view(np.int64)).to(dev)
        outs = []
        for traced in (False, True):
            a_out = torch.empty((L, n), dtype=torch.int64, device=dev)
            b_out = torch.empty((L, gamma), dtype=torch.int64, device=dev)
            if traced:
                ws.trace(2)
            for _ in range(3):
                ctx.respond_batch_device(seat, ws, ids.data_ptr(), seeds.data_ptr(), b.data_ptr(), L,
                                         a_out.data_ptr(), b_out.data_ptr())
            if traced:
                group, scan, chain = ws.kernel_times()
                assert len(group) == len(scan) == len(chain) == 2
                assert (group > 0).all() and (scan > 0).all() and (chain > 0).all()
            torch.cuda.synchronize()
            assert (ws.status(L) == 0).all()
            outs.append((a_out.cpu(), b_out.cpu()))
        assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
        with pytest.raises(gpu.PirError, match="at most 1048576 launches"):
            ws.trace((1 << 20) + 1)
        ws.close()
        seat.close()


def test_device_ids_without_a_store_are_refused(gpu):
    """The ids of a device-buffer batch are read on the device: a query naming no shard with a
    store (an id past the shard table, or a shard whose store was never built) gets status -1
    and nothing reads that shard; the other queries of the batch are answered as the host
    path answers them."""
    import torch

    n, gamma = 2048, 16
    rng = np.random.default_rng(8)
    dev = torch.device("cuda:0")
    with gpu.Context(gpu.params_d2048(), bytes(32), gamma) as ctx:
        ctx.upload_columns(0, rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64)).precompute_packing()
        ctx.upload_columns(1, rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64))
        seat = ctx.seat(rng.integers(0, gpu.Q, (3, n), dtype=np.uint64))
        ids = [0, 1, 7, 0]
        seeds = rng.integers(0, 256, (len(ids), 32), dtype=np.uint8)
        bs = rng.integers(0, gpu.Q, (len(ids), n), dtype=np.uint64)
        ws = gpu.Workspace(ctx, len(ids))
        d_ids = torch.tensor(ids, dtype=torch.int32, device=dev)
        d_seeds = torch.from_numpy(seeds).to(dev)
        d_b = torch.from_numpy(bs.view(np.int64)).to(dev)
        a_out = torch.zeros((len(ids), n), dtype=torch.int64, device=dev)
        b_out = torch.zeros((len(ids), gamma), dtype=torch.int64, device=dev)
        print("device batch over shard ids", ids, flush=True)
        ctx.respond_batch_device(seat, ws, d_ids.data_ptr(), d_seeds.data_ptr(), d_b.data_ptr(), len(ids),
                                 a_out.data_ptr(), b_out.data_ptr())
        torch.cuda.synchronize()
        assert list(ws.status(len(ids))) == [0, -1, -1, 0]
        A, B = ctx.respond_seeded_batch(seat, [0, 0], [seeds[0].tobytes(), seeds[3].tobytes()], bs[[0, 3]])
        got_a, got_b = a_out.cpu().numpy().view(np.uint64), b_out.cpu().numpy().view(np.uint64)
        assert np.array_equal(got_a[[0, 3]], A) and np.array_equal(got_b[[0, 3]], B)
        ws.close()
        seat.close()

