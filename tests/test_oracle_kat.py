"""The oracle's client side and seeded respond against the reference's recorded digests.

tests/golden/client_packing_key_kat.txt holds the key bytes and the sampler position after
them (tests/client_packing_key_kat.rs:20-50); tests/golden/packing_seat_response_kat.txt
holds the session-key and response digests of the packing-key cells
(tests/packing_seat_byte_identity_kat.rs:48-146, cells built as
tests/packing_key_fixture/mod.rs:86-147). Both files are data copied from the reference's
tests/fixtures. Only the single-modulus lines are replayed: the two-prime CRT sets are not
on the path these kernels implement.
"""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = oracle.load_packing_seat_kat(os.path.join(HERE, "golden", "packing_seat_response_kat.txt"))
TAGS = {"d2048": 2048, "d256-single-l3": 256}
CELLS = sorted(k for k in KAT if k[0] in TAGS)


def _key_lines():
    out = []
    for line in open(os.path.join(HERE, "golden", "client_packing_key_kat.txt")):
        f = line.split()
        kv = dict(x.split("=", 1) for x in f[1:])
        if kv["set"] == "single":
            out.append((int(kv["gamma"]), int(kv["seed"], 16), int(kv["len"]), int(kv["fnv"], 16), int(kv["next"])))
    return out


@pytest.mark.parametrize("gamma,seed,length,fnv,nxt", _key_lines())
def test_client_packing_key_kat(gamma, seed, length, fnv, nxt):
    kb, got_next = oracle.client_key_kat(2048, gamma, seed)
    assert (len(kb), oracle.fnv1a(kb), got_next) == (length, fnv, nxt)


def test_every_single_modulus_cell_is_replayed():
    assert len(CELLS) == 12
    assert sum(len(KAT[c]["resp"]) for c in CELLS) == 42 + 25


@pytest.mark.parametrize("cell", CELLS, ids=lambda c: f"{c[0]}-e{c[1]}-{c[2]:#x}")
def test_session_key_digest(cell):
    tag, entry, seed = cell
    c = oracle.Cell(TAGS[tag], entry, seed)
    kb = c.key_bytes()
    assert (len(kb), oracle.fnv1a(kb)) == KAT[cell]["key"]


@pytest.mark.parametrize("cell", CELLS, ids=lambda c: f"{c[0]}-e{c[1]}-{c[2]:#x}")
def test_response_digests(cell):
    """Every recorded response, alternating the per-query and the shard-store packing (the
    reference records the same line for both)."""
    tag, entry, seed = cell
    n = TAGS[tag]
    c = oracle.Cell(n, entry, seed)
    p = oracle.Pack(n, c.gamma)
    stores = {}
    for k, (index, (length, fnv)) in enumerate(sorted(KAT[cell]["resp"].items())):
        shard, qseed, b = c.query(index)
        cols = c.columns(shard)
        sums = None
        if k % 2:
            sums = stores.setdefault(shard, p.store_build(cols))
        a_hat, b_out = p.respond(c.w_seed, oracle.expand_seed(qseed, n), b, cols, c.y_body, sums)
        wire = oracle.encode_response(a_hat, b_out[:c.gamma], c.gamma)
        assert (len(wire), oracle.fnv1a(wire)) == (length, fnv), f"index {index}"


def test_planted_mistakes_miss_the_digests():
    """A foreign row, a foreign seat or a skipped sampler draw serves other bytes
    (tests/packing_seat_byte_identity_kat.rs:228-274)."""
    cell = ("d256-single-l3", 32, 0x534D4120)
    n = 256
    rec = KAT[cell]["resp"]
    c = oracle.Cell(n, 32, cell[2])
    p = oracle.Pack(n, c.gamma)
    shard, qseed, b = c.query(0)
    a = oracle.expand_seed(qseed, n)

    def digest(y, row_b, cols):
        ah, bo = p.respond(c.w_seed, a, row_b, cols, y)
        return oracle.fnv1a(oracle.encode_response(ah, bo[:c.gamma], c.gamma))

    assert digest(c.y_body, b, c.columns(shard)) == rec[0][1]
    assert digest(c.y_body, b, c.columns(1)) != rec[0][1]               # another shard's rows
    other = oracle.Cell(n, 32, cell[2] + 1)
    assert digest(other.y_body, b, c.columns(shard)) != rec[0][1]       # another client's seat
    b2 = b.copy()
    b2[0] = (int(b2[0]) + 1) % oracle.Q                                 # one error draw off by one
    assert digest(c.y_body, b2, c.columns(shard)) != rec[0][1]


def test_keys_satisfy_the_key_equation():
    """y_body[r] + s * w_mask[r] - 2^(20 r) tau_g(s) is a small error (the reference's own
    check, tests/client_packing_key_kat.rs:74-105)."""
    n, gamma = 256, 16
    c = oracle.Cell(n, 2 * gamma, 0x1234)
    pack = oracle.Pack(n, gamma)
    w = pack.w_mask(c.w_seed)
    tau = oracle.apply_automorphism(c.secret, 2 * n // gamma + 1)
    Q = oracle.Q
    for r in range(3):
        sw = oracle.negacyclic_mul_naive(c.secret, w[r])
        res = [(int(y) + int(x) - (int(t) << (20 * r))) % Q for y, x, t in zip(c.y_body[r], sw, tau)]
        centred = [v if v <= Q // 2 else v - Q for v in res]
        assert max(abs(v) for v in centred) < 64
