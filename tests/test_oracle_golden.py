"""Pins the CPU oracle to the reference's own goldens and to closed forms.

The oracle is the checker every GPU parity test trusts, so each part of it is tied
here to something that fixes it independently of the oracle's own formulas.
"""
import numpy as np
import pytest

import oracle
import wire

Q = oracle.Q


def test_gaussian_stream_hashes():
    # src/math/gaussian.rs:334-351: frozen FNV-1a of 200,000 seeded samples
    golden = [
        (3.2, 0, 0xD6A6E62F77396BD4),
        (3.2, 12345, 0x270FDB99E4274779),
        (6.4, 1, 0x220CF0EA2DC8E0E5),
        (1.0, 42, 0x5A0AB704AB0775A2),
        (10.0, 7, 0x503DF0167FFD586D),
    ]
    for sigma, seed, want in golden:
        assert oracle.gauss_stream_hash(sigma, seed, 200_000) == want, (sigma, seed)
    assert oracle.gauss_centered_hash_seed7(50_000, Q) == 0x56FDCCED8A630609


def test_chacha20_rfc7539_block():
    # RFC 7539 section 2.3.2 test vector: key 00..1f, counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00.
    # In the 64-bit counter / 64-bit stream layout, words 12-13 = (1, 0x09000000), words 14-15 = (0x4a000000, 0).
    key = np.frombuffer(bytes(range(32)), dtype="<u4")
    out = oracle.chacha_block(key, 1 | (0x09000000 << 32), 0x4A000000)
    want = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204, 0x4E6CD4C3,
            0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]
    assert list(out) == want


@pytest.mark.parametrize("n", [16, 256, 512, 1024, 2048, 4096])
def test_ntt_is_negacyclic_convolution(n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, Q, n, dtype=np.uint64)
    b = rng.integers(0, Q, n, dtype=np.uint64)
    ntt = oracle.Ntt(n)
    fa, fb = ntt.forward(a), ntt.forward(b)
    prod = np.array([ntt.mont_mul(int(x), int(y)) for x, y in zip(fa, fb)], dtype=np.uint64)
    got = ntt.inverse(prod)
    if n <= 256:
        want = oracle.negacyclic_mul_naive(a, b)
    else:  # exact big-integer schoolbook on a sparse operand keeps this fast
        b = np.zeros(n, dtype=np.uint64)
        b[3], b[n - 1] = 5, Q - 2
        fb = ntt.forward(b)
        got = ntt.inverse(np.array([ntt.mont_mul(int(x), int(y)) for x, y in zip(fa, fb)], dtype=np.uint64))
        want = [0] * n
        for i in range(n):
            for j, c in ((3, 5), (n - 1, Q - 2)):
                k = i + j
                v = int(a[i]) * c
                if k >= n:
                    want[k - n] = (want[k - n] - v) % Q
                else:
                    want[k] = (want[k] + v) % Q
        want = np.array(want, dtype=np.uint64)
    assert np.array_equal(got, want)
    assert np.array_equal(ntt.inverse(ntt.forward(a)), a)


def test_montgomery_equals_plain_product():
    ntt = oracle.Ntt(256)
    rng = np.random.default_rng(1)
    rinv = pow(2 ** 64, -1, Q)
    for _ in range(200):
        a, b = (int(x) for x in rng.integers(0, Q, 2, dtype=np.uint64))
        assert ntt.mont_mul(a, b) == a * b * rinv % Q


@pytest.mark.parametrize("n,gamma", [(256, 16), (512, 64), (1024, 128), (2048, 256), (4096, 256)])
def test_automorph_tables_invariant_and_closed_form(n, gamma):
    p = oracle.Pack(n, gamma)
    ntt = oracle.Ntt(n)
    rng = np.random.default_rng(7)
    f = rng.integers(0, Q, n, dtype=np.uint64)
    logn = n.bit_length() - 1

    def brv(x):
        return int(format(x, f"0{logn}b")[::-1], 2)

    for t in (3, 2 * n // gamma + 1, 2 * n - 1):
        tab = p.automorph_table(t)
        # tau_t(f)(X) = f(X^t) in coefficient form
        tf = [0] * n
        for i in range(n):
            e = (i * t) % (2 * n)
            if e < n:
                tf[e] = (tf[e] + int(f[i])) % Q
            else:
                tf[e - n] = (tf[e - n] - int(f[i])) % Q
        assert np.array_equal(ntt.forward(np.array(tf, dtype=np.uint64)), ntt.forward(f)[tab])
        closed = [brv(((t * (2 * brv(j) + 1)) % (2 * n)) >> 1) for j in range(n)]
        assert list(tab) == closed


def test_respond_byte_identity_d256_golden():
    # tests/respond_byte_identity_kat.rs:108-145 on the reference's own fixture
    fx = wire.load_d256_fixture()
    p = oracle.Pack(fx["params"]["ring_dim"], fx["num_columns"])
    a_hat, b = p.respond(fx["w_seed"], fx["a"], fx["b"], fx["columns"], fx["y_body"])
    resp = oracle.encode_response(a_hat, b, fx["num_columns"])
    assert oracle.fnv1a(resp) == wire.D256_GOLDEN_FNV


def test_planted_mistakes_break_the_golden():
    # a dropped term or a wrong index must fail the pin: perturb inputs the way such
    # mistakes would and check the golden moves
    fx = wire.load_d256_fixture()
    p = oracle.Pack(256, 16)
    y = fx["y_body"].copy()
    y[2] = 0  # a dropped digit row
    a_hat, b = p.respond(fx["w_seed"], fx["a"], fx["b"], fx["columns"], y)
    assert oracle.fnv1a(oracle.encode_response(a_hat, b, 16)) != wire.D256_GOLDEN_FNV
    cols = fx["columns"][::-1].copy()  # transposed column order
    a_hat, b = p.respond(fx["w_seed"], fx["a"], fx["b"], cols, fx["y_body"])
    assert oracle.fnv1a(oracle.encode_response(a_hat, b, 16)) != wire.D256_GOLDEN_FNV


def test_store_path_equals_per_query_path():
    # the store's step-1 fold is linear in the columns (shard_store.rs:1-16;
    # tests/shard_packing_store.rs:163-208 pins the same equality)
    fx = wire.load_d256_fixture()
    p = oracle.Pack(256, 16)
    S = p.store_build(fx["columns"])
    x = p.respond(fx["w_seed"], fx["a"], fx["b"], fx["columns"], fx["y_body"])
    y = p.respond(fx["w_seed"], fx["a"], fx["b"], fx["columns"], fx["y_body"], S)
    assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


@pytest.mark.parametrize("n", [512, 1024, 4096])
def test_store_path_equals_per_query_path_at_other_rings(n):
    # InspireParams::d4096 (params.rs:392-436) and the rings 512 and 1024 the same prime admits
    # (2n divides q - 1, params.rs:535-580): the same invariant on each ring, for both gadgets
    rng = np.random.default_rng(n)
    gamma = 16
    cols = rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64)
    b = rng.integers(0, Q, n, dtype=np.uint64)
    a = oracle.expand_seed(bytes(range(32)), n)
    for dropped in (0, 1):
        p = oracle.Pack(n, gamma, 3, dropped)
        y = rng.integers(0, Q, (3 - dropped, n), dtype=np.uint64)
        x = p.respond(bytes(32), a, b, cols, y)
        z = p.respond(bytes(32), a, b, cols, y, p.store_build(cols))
        assert np.array_equal(x[0], z[0]) and np.array_equal(x[1], z[1])


def test_fold_b_is_coefficient_zero_of_products():
    n = 256
    rng = np.random.default_rng(3)
    p = oracle.Pack(n, 16)
    cols = rng.integers(0, 1 << 16, (16, n), dtype=np.uint64)
    b = rng.integers(0, Q, n, dtype=np.uint64)
    got = p.fold_b(cols, b)
    for j in (0, 7, 15):
        prod = oracle.negacyclic_mul_naive(b, cols[j])
        assert got[j] == prod[0]


def test_mod_switch_closed_form():
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.integers(0, Q, 1000, dtype=np.uint64),
                           np.array([0, 1, Q // 2, Q // 2 + 1, Q - 1], dtype=np.uint64)])
    got = oracle.mod_switch(vals)
    for v, g in zip(vals, got):
        assert int(g) == ((int(v) * oracle.Q36 + Q // 2) // Q) % oracle.Q36


def test_response_wire_lengths():
    # README.md:66-77: 512-byte records respond in 17,358 B, 10,446 B mod-switched;
    # packing_seat_response_kat.txt lines for 32-byte records: 15,558 B
    n = 2048
    z = np.zeros(n, dtype=np.uint64)
    assert len(oracle.encode_response(z, z, 256)) == 17358
    assert len(oracle.encode_response(z, z, 256, oracle.Q36)) == 10446
    assert len(oracle.encode_response(z, z, 16)) == 15558


def test_seeded_expansion_uses_gen_range():
    a = oracle.expand_seed(bytes(range(32)), 64)
    assert (a < Q).all() and len(set(a.tolist())) == 64
