"""Pins the special-modulus oracle (oracle/special.c) and the host pieces of the special-
modulus path: the reference's served-set wire digests, closed forms and brute force for
the arithmetic, decode of every record, the rotated-frame recursion the gfx950 kernel runs,
the wire codecs, and the parameter refusals."""
import numpy as np
import pytest

import oracle
import inspire_gpu as g

SM_Q, SM_P = oracle.SM_Q, oracle.SM_P

# tests/special_modulus_kat.rs:380-384: FNV-1a of the key, query and response bytes of
# Client::from_seed([42; 32]), cells from ChaCha20Rng::seed_from_u64(42), query(1234)
PINNED = [0x84FE10759FE96D5F, 0xD889F6A134B7E313, 0xEF10B409C5474E4B]


@pytest.fixture(scope="module")
def served():
    return oracle.SmRing(cell_bytes=256, record_bytes=246)


# respond_portable and respond_avx512 (server.rs:306-330): the same respond_impl, the second
# compiled for AVX-512F/BW/DQ/VL
FORMS = ["portable", pytest.param("avx512", marks=pytest.mark.skipif(
    not oracle.sm_avx512_available(), reason="this CPU lacks AVX-512F/BW/DQ/VL"))]


@pytest.mark.parametrize("form", FORMS)
def test_served_wire_bytes_are_pinned(served, form):
    client = oracle.SmClient(served, bytes([42] * 32))
    cells = oracle.sm_kat_cells(42, 2048)
    table = served.table(cells)
    seed, kq, kp = client.packing_key()
    seat = served.seat(kq, kp, seed)
    key = client.key_bytes()
    qseed, qbody = client.query(1234)
    mask, body = served.respond(table, seat, qseed, qbody, avx512=form == "avx512")
    query = qseed + oracle.sm_pack(qbody, SM_Q)
    response = served.response_bytes(mask, body)
    assert (len(key), len(query), len(response)) == (15_394, 7_456, 9_208)
    assert [oracle.fnv1a(key), oracle.fnv1a(query), oracle.fnv1a(response)] == PINNED
    assert client.decode(mask, body) == bytes(cells[1234 * 256:1234 * 256 + 246])


@pytest.mark.skipif(not oracle.sm_avx512_available(), reason="this CPU lacks AVX-512F/BW/DQ/VL")
@pytest.mark.parametrize("d,gamma", [(2048, 512), (256, 32), (64, 8)])
def test_avx512_build_equals_the_portable_build(d, gamma):
    ring = oracle.SmRing(cell_bytes=256, record_bytes=246) if d == 2048 else oracle.SmRing(d=d, gamma=gamma)
    rng = np.random.default_rng(d)
    rows, cb = min(d, 300), ring.cell_bytes
    cells = np.frombuffer(rng.bytes(rows * cb), dtype=np.uint8).reshape(rows, cb).copy()
    cells[:, ring.record_bytes:] = 0
    table = ring.table(cells.tobytes())
    seat = ring.seat(rng.integers(0, SM_Q, d, dtype=np.uint32), rng.integers(0, SM_P, d, dtype=np.uint32),
                     rng.bytes(32))
    for _ in range(2):
        seed, body = rng.bytes(32), rng.integers(0, SM_Q, d, dtype=np.uint32)
        m0, b0 = ring.respond(table, seat, seed, body)
        m1, b1 = ring.respond(table, seat, seed, body, avx512=True)
        assert np.array_equal(m0, m1) and np.array_equal(b0, b1)


def _naive_negacyclic(a, b, m):
    n = len(a)
    out = [0] * n
    for i in range(n):
        for j in range(n):
            v = int(a[i]) * int(b[j])
            if i + j < n:
                out[i + j] += v
            else:
                out[i + j - n] -= v
    return np.array([x % m for x in out], dtype=np.uint32)


def test_transform_multiplies_negacyclically():
    ring = oracle.SmRing(d=64, gamma=8)
    rng = np.random.default_rng(3)
    for m, fwd in ((SM_Q, ring.ntt_q), (SM_P, ring.ntt_p)):
        a = rng.integers(0, m, 64, dtype=np.uint32)
        b = rng.integers(0, m, 64, dtype=np.uint32)
        fa, fb = fwd(a), fwd(b)
        prod = ((fa.astype(object) * fb.astype(object)) % m).astype(np.uint32)
        assert np.array_equal(fwd(prod, inverse=True), _naive_negacyclic(a, b, m))
        assert np.array_equal(fwd(fa, inverse=True), a)


@pytest.mark.parametrize("d,q,p", [(16, SM_Q, SM_P), (4096, SM_Q, 2_147_377_153), (8192, 536_690_689, 2_147_352_577)])
def test_transform_multiplies_negacyclically_at_the_other_rings(d, q, p):
    # rings 16 .. 8192 the kernels also run, with primes 1 mod 2d; an exact product with a
    # sparse operand (three terms, one wrapping past X^d) keeps the big rings fast
    ring = oracle.SmRing(d=d, gamma=8, q=q, p=p, tables=False)
    rng = np.random.default_rng(d)
    for m, fwd in ((q, ring.ntt_q), (p, ring.ntt_p)):
        a = rng.integers(0, m, d, dtype=np.uint32)
        b = np.zeros(d, dtype=np.uint32)
        b[0], b[1], b[d - 1] = 3, m - 1, 7
        fa, fb = fwd(a), fwd(b)
        prod = ((fa.astype(object) * fb.astype(object)) % m).astype(np.uint32)
        want = [0] * d
        for i in range(d):
            for j, c in ((0, 3), (1, m - 1), (d - 1, 7)):
                k, v = i + j, int(a[i]) * c
                if k >= d:
                    want[k - d] = (want[k - d] - v) % m
                else:
                    want[k] = (want[k] + v) % m
        assert np.array_equal(fwd(prod, inverse=True), np.array(want, dtype=np.uint32))
        assert np.array_equal(fwd(fa, inverse=True), a)


def _brv(x, bits):
    r = np.zeros_like(x)
    for i in range(bits):
        r |= ((x >> i) & 1) << (bits - 1 - i)
    return r


def test_automorphism_tables_have_the_closed_form(served):
    # slot j evaluates at psi^(2 brv(j) + 1): tau_t reads slot brv((t (2 brv(j) + 1) mod 2n) >> 1),
    # the form the kernels compute instead of storing tables
    d, logn = 2048, 11
    tabs = served.tables()
    j = np.arange(d, dtype=np.int64)
    t = 1
    for i in range(512):
        want = _brv(((t * (2 * _brv(j, logn) + 1)) % (2 * d)) >> 1, logn)
        assert np.array_equal(tabs[i].astype(np.int64), want), i
        t = t * 9 % (2 * d)


def test_automorphism_tables_match_the_coefficient_map():
    # arith.rs:428-452: NTT(tau_t f) = NTT(f)[table_t]
    ring = oracle.SmRing(d=64, gamma=8)
    tabs = ring.tables()
    rng = np.random.default_rng(5)
    f = rng.integers(0, SM_Q, 64, dtype=np.uint32)
    t = 1
    for i in range(8):
        tau = np.zeros(64, dtype=np.int64)
        for k, c in enumerate(f):
            e = t * k % 128
            if e < 64:
                tau[e] = int(c)
            else:
                tau[e - 64] = (SM_Q - int(c)) % SM_Q
        assert np.array_equal(ring.ntt_q(f)[tabs[i]], ring.ntt_q(tau.astype(np.uint32)))
        t = t * 17 % 128


def _crt(a, b, q, p):
    return a + q * ((b - a) % p * pow(q, -1, p) % p)


def test_lifts_and_mod_down_against_exact_arithmetic(served):
    # mod.rs:436-476: each lift boundary and each step of the conditional subtractions
    q, p = SM_Q, SM_P
    cen = lambda x, m: x - m if x > m // 2 else x
    rs = [0, 1, p // 2 - 1, p // 2, p // 2 + 1, p - 2, p - 1]
    for k in range(1, 5):
        rs += [k * q - 1, k * q, k * q + 1]
    cs = [0, 1, q // 2 - 1, q // 2, q // 2 + 1, q - 2, q - 1]
    rng = np.random.default_rng(306)
    partners = [int(x) for x in rng.integers(0, q, 8)]
    for r in rs:
        assert served.p_to_q_centred(r) == cen(r, p) % q, r
        for xq in partners + cs:
            x = _crt(xq, r, q, p)
            assert served.mod_down(xq, r) == (x + (p - 1) // 2) // p % q, (xq, r)
    for c in cs:
        assert served.q_to_p_centred(c) == cen(c, q) % p, c


def test_every_record_decodes(served):
    # tests/special_modulus_kat.rs:296-336: a short shard; rows past it read as zero records
    rows = 1500
    cells = oracle.sm_kat_cells(11, rows)
    table = served.table(cells)
    client = oracle.SmClient(served, bytes([5] * 32))
    seed, kq, kp = client.packing_key()
    seat = served.seat(kq, kp, seed)
    for row in (0, 1, 777, rows - 1, rows, 2047):
        qseed, qbody = client.query(row)
        mask, body = served.respond(table, seat, qseed, qbody)
        want = bytes(cells[row * 256:row * 256 + 246]) if row < rows else bytes(246)
        assert client.decode(mask, body) == want, row


def _frame_model(ring, table, seat, qseed, qbody):
    """The recursion the gfx950 kernel runs: state V_i = tau_i^-1(R_i), step-1 rows
    gamma^-1 NTT(P_i) (no gather), one tau_g rotation of (V, B_q, B_p) per step, and
    accumulators by Horner. Exact integers, no lazy ranges."""
    d, gamma, q, p = ring.d, ring.gamma, ring.q, ring.p
    R = 1 << 32
    std = lambda x, m: (x.astype(object) * pow(R, -1, m)) % m
    tabs = ring.tables()
    step1, values, rows = table
    sq, sp, wq, wp = [std(x, m) for x, m in zip(seat, (q, p, q, p))]
    # S'_i = tau_i^-1(S_i) = tau_{gamma-i}(S_i)
    S = [std(step1[i], q)[tabs[(gamma - i) % gamma]] for i in range(gamma)]
    a = ring.ntt_q(ring.query_mask(qseed)).astype(object)
    lift_qp = lambda c: np.array([int(x) + (p - q) if x > q // 2 else int(x) for x in c], dtype=np.uint32)
    lift_pq = lambda r: np.array([(int(x) - p) % q if x > p // 2 else int(x) % q for x in r], dtype=np.uint32)
    tp_of = lambda y: ring.ntt_p(lift_qp(ring.ntt_q(np.array(y, dtype=np.uint32), True))).astype(object)
    rq_of = lambda c: ring.ntt_q(lift_pq(ring.ntt_p(np.array(c, dtype=np.uint32), True))).astype(object)
    pinv = pow(p, -1, q)
    V = a * S[gamma - 1] % q
    bq = np.zeros(d, dtype=object)
    bp = np.zeros(d, dtype=object)
    rot = tabs[1]
    for i in range(gamma - 2, -1, -1):
        Y, bq, bp = V[rot], bq[rot], bp[rot]
        tp = tp_of(Y)
        bp = (bp + tp * sp) % p
        rq = rq_of(tp * wp % p)
        V = (a * S[i] + (Y * wq - rq) % q * pinv) % q
        bq = (bq + Y * sq) % q
    mask = ring.ntt_q(np.array(V, dtype=np.uint32), True)
    accq = ring.ntt_q(np.array(bq, dtype=np.uint32), True)
    accp = ring.ntt_p(np.array(bp, dtype=np.uint32), True)
    cen = lambda x: int(x) - q if x > q // 2 else int(x)
    fold = np.zeros(ring.nb, dtype=np.int64)
    for m in range(rows):
        c = cen(qbody[0]) if m == 0 else -cen(qbody[d - m])
        fold += c * values[m, :ring.nb].astype(np.int64)
    body = [(ring.mod_down(int(accq[k]), int(accp[k])) + int(fold[k]) % q) % q for k in range(ring.nb)]
    return mask, np.array(body, dtype=np.uint32)


@pytest.mark.parametrize("d,gamma", [(16, 8), (32, 16), (64, 8), (256, 32)])
def test_rotated_frame_recursion_equals_the_reference_recursion(d, gamma):
    ring = oracle.SmRing(d=d, gamma=gamma)
    rng = np.random.default_rng(d + gamma)
    cells = rng.integers(0, 256, d * ring.cell_bytes, dtype=np.uint8).tobytes()
    table = ring.table(cells)
    seat = ring.seat(rng.integers(0, SM_Q, d, dtype=np.uint32), rng.integers(0, SM_P, d, dtype=np.uint32), bytes(32))
    qseed = bytes(rng.integers(0, 256, 32, dtype=np.uint8))
    qbody = rng.integers(0, SM_Q, d, dtype=np.uint32)
    want = ring.respond(table, seat, qseed, qbody)
    got = _frame_model(ring, table, seat, qseed, qbody)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_wire_codecs_round_trip_and_refuse():
    # src/special_modulus/wire.rs:144-215
    params = g.ring2048_cell256()
    d = params.ring_dim
    key = g.PackingKey(2, bytes([9] * 32), np.full(d, SM_Q - 1), np.full(d, SM_P - 1))
    query = g.Query(bytes([7] * 32), np.arange(d))
    resp = g.Response(np.full(d, SM_Q - 1), np.ones(492))
    kb, qb, rb = key.to_bytes(params), query.to_bytes(params), resp.to_bytes(params)
    assert (len(kb), len(qb), len(rb)) == (params.key_bytes(), params.query_bytes(), params.response_bytes())
    k2 = g.PackingKey.from_bytes(params, kb)
    assert np.array_equal(k2.q, key.q) and np.array_equal(k2.p, key.p) and k2.mask_seed == key.mask_seed
    assert np.array_equal(g.Query.from_bytes(params, qb).body, query.body)
    assert g.Response.from_bytes(params, rb) == resp
    with pytest.raises(g.PirError, match="15393 bytes"):
        g.PackingKey.from_bytes(params, kb[1:])
    assert kb[:2] == (2).to_bytes(2, "big") and kb[2:34] == bytes([9] * 32)
    for version in (0, 1, 3):
        other = bytearray(kb)
        other[:2] = version.to_bytes(2, "big")
        with pytest.raises(g.PirError, match="key-mask version"):
            g.PackingKey.from_bytes(params, bytes(other))
    with pytest.raises(g.PirError):
        g.Query.from_bytes(params, qb + b"\0")
    # a value at the modulus, and a set padding bit (7424 * 8 = 2048 * 29 leaves none for the
    # mask; the 492-coefficient body leaves 4 spare bits)
    bad = bytearray(rb)
    bad[:4] = (SM_Q).to_bytes(4, "little")
    with pytest.raises(g.PirError, match="not canonical"):
        g.Response.from_bytes(params, bytes(bad))
    bad = bytearray(rb)
    bad[-1] |= 0x80
    with pytest.raises(g.PirError, match="unused high bits"):
        g.Response.from_bytes(params, bytes(bad))
    # a relabelled set that names the retired version refuses by name
    retired = g.ring2048_cell256()
    retired.key_mask_version = 1
    v1 = bytearray(kb)
    v1[:2] = (1).to_bytes(2, "big")
    with pytest.raises(g.PirError, match="per-key-seed version"):
        g.PackingKey.from_bytes(retired, bytes(v1))


def test_codec_matches_the_oracle_packer():
    rng = np.random.default_rng(9)
    for m, count in ((SM_Q, 2048), (SM_P, 2048), (SM_Q, 492), (SM_Q, 7)):
        v = rng.integers(0, m, count, dtype=np.uint32)
        assert g.pack_tight(v, m) == oracle.sm_pack(v, m)
        assert np.array_equal(g.unpack_tight(g.pack_tight(v, m), m, count), v)


def test_invalid_sets_are_refused():
    # src/special_modulus/mod.rs:411-434, through the C ABI's Params::validate
    edits = {
        "ring": ("ring_dim", 3000),
        "mask version 0": ("key_mask_version", 0),
        "shared-mask version": ("key_mask_version", 1),
        "unknown mask version": ("key_mask_version", 3),
        "width": ("gamma", 2048),
        "cell": ("cell_bytes", 255),
        "record": ("record_bytes", 257),
        "composite": ("q", 4097 * 4097),
        "sigma": ("sigma", float("nan")),
    }
    g.ring2048_cell256().validate()
    for name, (field, value) in edits.items():
        p = g.ring2048_cell256()
        setattr(p, field, value)
        with pytest.raises(g.PirError):
            p.validate()
    p = g.ring2048_cell256()
    p.q, p.special_prime = p.special_prime, p.q
    with pytest.raises(g.PirError, match="not above"):
        p.validate()
    # a ring the reference accepts but the kernels do not run (primes 1 mod 2^15): refused at
    # server creation, before any device call
    p = g.SmParams(16384, 512, 536_641_537, 2_147_352_577, 256, 256, 6.4, 2)
    p.validate()
    with pytest.raises(g.PirError, match=r"outside the \[16, 8192\] the gfx950 kernels run"):
        g.SmServer(p)
    # ciphertext primes above 2^32 / 6 up to 2^31 are accepted (the kernels keep residues mod q
    # canonical there); primes at or above 2^31 are not (Params::validate)
    p = g.ring2048_cell256()
    p.q = 715_829_249
    p.validate()
    p.q = 2_147_352_577
    p.validate()
    p.q, p.special_prime = 2_147_352_577, 2_147_483_659
    with pytest.raises(g.PirError, match="not a prime below 2"):
        p.validate()


@pytest.mark.parametrize("q", [716_259_329, 2_147_352_577])
def test_oracle_at_primes_above_the_lazy_bound_decodes(q):
    # src/special_modulus/mod.rs:124-133 accepts ciphertext primes up to 2^31; the oracle's
    # u64/i128 arithmetic follows the reference's at every one of them: responses to queries at
    # the first and last rows, over random cells and over cells whose every nibble is -8 (the
    # largest fold), decode to their records through the client
    d, gamma = 64, 8
    cell = record = gamma // 2
    ring = oracle.SmRing(d=d, gamma=gamma, q=q, cell_bytes=cell, record_bytes=record)
    worst = np.full(d * cell, 0x88, dtype=np.uint8)
    for cells in (oracle.sm_kat_cells(q % 997, d, cell_bytes=cell, record_bytes=record), worst):
        client = oracle.SmClient(ring, bytes([q % 251] * 32))
        seed, kq, kp = client.packing_key()
        table, seat = ring.table(cells.tobytes()), ring.seat(kq, kp, seed)
        for row in (0, d - 1):
            qs, body = client.query(row)
            mask, out = ring.respond(table, seat, qs, body)
            assert client.decode(mask, out) == bytes(cells[row * cell:(row + 1) * cell])
