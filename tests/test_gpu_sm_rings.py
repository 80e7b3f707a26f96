"""The special-modulus kernels at the rings Params::validate accepts beyond 64..2048
(src/special_modulus/mod.rs:100-104: powers of two in [16, 2^16]): 16 and 32 (two and four
threads per query), 4096 and 8192 (512 and 1024 threads per query, 55 and 110 KiB of LDS),
each with NTT-friendly primes of the served sizes (1 mod 2d), against the oracle with every
nibble -8 (the largest fold) and with random cells, and through the client: every response
of a query decodes to its record. Rings above 8192 are refused by the server.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SM_Q, SM_P = oracle.SM_Q, oracle.SM_P
CASES = [
    (16, 8, SM_Q, SM_P),
    (32, 16, SM_Q, SM_P),
    (4096, 512, SM_Q, 2_147_377_153),
    (8192, 1024, 536_690_689, 2_147_352_577),
]


@pytest.mark.parametrize("d,gamma,q,p", CASES)
def test_ring_matches_the_oracle_and_decodes(gpu, d, gamma, q, p):
    cell = gamma // 2
    params = gpu.SmParams(d, gamma, q, p, cell, cell, 6.4, 2)
    ring = oracle.SmRing(d=d, gamma=gamma, q=q, p=p, cell_bytes=cell, record_bytes=cell)
    rng = np.random.default_rng(d + gamma)
    worst = np.full(d * cell, 0x88, dtype=np.uint8)
    rand = rng.integers(0, 256, d * cell, dtype=np.uint8)
    cells = [worst, rand]
    clients = [oracle.SmClient(ring, bytes([k + 21] * 32)) for k in range(2)]
    print(f"sm ring {d}, gamma {gamma}", flush=True)
    with gpu.SmServer(params) as server:
        tables = [server.table(c.tobytes()) for c in cells]
        seats = [server.seat(gpu.PackingKey.from_bytes(params, c.key_bytes())) for c in clients]
        picks = []
        for j in range(6):
            ti, ci, row = j % 2, (j // 2) % 2, int(rng.integers(0, d))
            seed, body = clients[ci].query(row)
            picks.append((ti, ci, row, gpu.Query(seed, body)))
        picks.append((0, 1, 0, gpu.Query(bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                                         rng.integers(0, q, d, dtype=np.uint32))))
        masks, bodies = server.respond_batch([tables[t] for t, _, _, _ in picks],
                                             [seats[c] for _, c, _, _ in picks], [x[3] for x in picks])
        wires = server.respond_batch_wire([tables[t] for t, _, _, _ in picks[:2]],
                                          [seats[c] for _, c, _, _ in picks[:2]], [x[3] for x in picks[:2]])
        otables = [ring.table(c.tobytes()) for c in cells]
        oseats = []
        for c in clients:
            seed, kq, kp = c.packing_key()
            oseats.append(ring.seat(kq, kp, seed))
        for j, (ti, ci, row, qy) in enumerate(picks):
            wm, wb = ring.respond(otables[ti], oseats[ci], qy.seed, qy.body)
            assert np.array_equal(masks[j], wm), (d, j)
            assert np.array_equal(bodies[j], wb), (d, j)
            if j < 6:
                want = bytes(cells[ti][row * cell:(row + 1) * cell])
                assert clients[ci].decode(masks[j], bodies[j]) == want, (d, j)
            if j < 2:
                assert wires[j] == gpu.Response(masks[j], bodies[j]).to_bytes(params), (d, j)
        for x in tables + seats:
            x.close()


def test_rings_past_8192_are_refused(gpu):
    # a set Params::validate accepts (primes 1 mod 2^15), refused by the server with its message
    params = gpu.SmParams(16384, 512, 536_641_537, 2_147_352_577, 256, 256, 6.4, 2)
    params.validate()
    with pytest.raises(gpu.PirError, match=r"outside the \[16, 8192\] the gfx950 kernels run"):
        gpu.SmServer(params)
