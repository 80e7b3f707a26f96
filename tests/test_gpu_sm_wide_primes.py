"""The special-modulus kernels at ciphertext primes above the served one.

The server accepts every prime q below 2^31 that Params::validate accepts
(src/special_modulus/mod.rs:100-165). Up to 2^32 / 6 (715,827,882) the recursion keeps residues
mod q lazy below 6q; above it the kernels' WQ instantiation keeps them canonical, as it does mod
p. The served q (536,813,569) sits just under 2^29; these cases run the kernels at

* 536,903,681, the first NTT-friendly prime (1 mod 4096) above 2^29, where a product of a
  centred body coefficient (up to q/2) and a nibble of -8 no longer fits 32 bits, so the body
  fold (server.rs:398-410, i64 sums) must widen before it multiplies;
* 715,825,153, the largest prime 1 mod 4096 of the lazy instantiation, where the lazy sums of
  the recursion come within 17,000 of 2^32;
* 716,259,329, the first prime 1 mod 2^14 above 2^32 / 6, the smallest of the canonical
  instantiation;
* 2,147,352,577, the largest prime 1 mod 2^14 below the served special prime 2,147,389,441 (and
  below 2^31), where sums of two residues come within 2^22 of 2^32;

at ring 2048 (gamma 512, the served cells) and at ring 64, with every nibble -8 (the largest
fold) and with random cells, against the oracle (i64 and i128 arithmetic, oracle/sm_arith.h),
and through the client: every response decodes to its record.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SM_P = oracle.SM_P
WIDE = [536_903_681, 715_825_153, 716_259_329, 2_147_352_577]


def _params(gpu, d, gamma, q):
    cell = gamma // 2
    record = 246 if d == 2048 else cell
    return gpu.SmParams(d, gamma, q, SM_P, cell, record, 6.4, 2)


@pytest.mark.parametrize("q", WIDE)
@pytest.mark.parametrize("d,gamma", [(2048, 512), (64, 8)])
def test_wide_prime_matches_the_oracle_and_decodes(gpu, q, d, gamma):
    params = _params(gpu, d, gamma, q)
    cell, record = params.cell_bytes, params.record_bytes
    ring = oracle.SmRing(d=d, gamma=gamma, q=q, cell_bytes=cell, record_bytes=record)
    rng = np.random.default_rng(q % 1000 + d)
    worst = np.zeros((d, cell), dtype=np.uint8)
    worst[:, :record] = 0x88  # every nibble of every record -8
    worst = worst.reshape(-1)
    rand = oracle.sm_kat_cells(int(rng.integers(1, 1 << 30)), d, cell_bytes=cell, record_bytes=record)
    cells = [worst, rand]
    clients = [oracle.SmClient(ring, bytes([k + 11] * 32)) for k in range(2)]
    print(f"wide prime {q}: ring {d}, gamma {gamma}", flush=True)
    with gpu.SmServer(params) as server:
        tables = [server.table(c.tobytes()) for c in cells]
        seats = [server.seat(gpu.PackingKey.from_bytes(params, c.key_bytes())) for c in clients]
        picks = []
        for j in range(6):
            ti, ci, row = j % 2, (j // 2) % 2, int(rng.integers(0, d))
            seed, body = clients[ci].query(row)
            picks.append((ti, ci, row, gpu.Query(seed, body)))
        # a random body as well: centred values spread over the whole of (-q/2, q/2]
        picks.append((0, 1, 0, gpu.Query(bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                                         rng.integers(0, q, d, dtype=np.uint32))))
        masks, bodies = server.respond_batch([tables[t] for t, _, _, _ in picks],
                                             [seats[c] for _, c, _, _ in picks], [p[3] for p in picks])
        otables = [ring.table(c.tobytes()) for c in cells]
        oseats = [ring.seat(*_key_parts(c)) for c in clients]
        for j, (ti, ci, row, qy) in enumerate(picks):
            wm, wb = ring.respond(otables[ti], oseats[ci], qy.seed, qy.body)
            assert np.array_equal(masks[j], wm), (q, d, j)
            assert np.array_equal(bodies[j], wb), (q, d, j)
            if j < 6:
                want = bytes(cells[ti][row * cell:row * cell + record])
                assert clients[ci].decode(masks[j], bodies[j]) == want, (q, d, j)
        for x in tables + seats:
            x.close()


def _key_parts(client):
    seed, kq, kp = client.packing_key()
    return kq, kp, seed
