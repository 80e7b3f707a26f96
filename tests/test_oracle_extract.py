"""The client's decode (PIR.Extract) restated in the oracle, pinned by the reference's own
known answers and closed forms, then used to decode the oracle's responses.

* production_post_switch_error_kat_at_45_36_and_33_bits (src/pir/mod_switch.rs:1070-1114):
  four recorded post-switch errors. They fix the client draw order, the exact-gadget respond at
  d2048 with 512-byte records, the switch at four target moduli, the secret-key switch and the
  phase; a wrong coefficient anywhere turns an error of hundreds into one near q'/2.
* the noise gate's verdicts and margins (src/pir/mod_switch.rs:36-47, 1020-1025, 1109-1124);
* reconstruct_entry's test vector (src/pir/extract.rs:499-513);
* PlaintextRounding against the exact quotient (src/rlwe/enc.rs:14-44, 555-);
* decode correctness of exact and approximate-gadget cells
  (tests/approximate_packing_gadget.rs:435-461, 463-494).
"""
import numpy as np
import pytest

import oracle

N = 2048


@pytest.fixture(scope="module")
def kat():
    k = oracle.post_switch_error_kat_inputs()
    p = oracle.Pack(k["n"], k["gamma"])
    a = oracle.expand_seed(k["seed"], k["n"])
    wa, wb = p.respond(k["w_seed"], a, k["b"], k["columns"], k["y_body"])
    k["a_hat"], k["b_prefix"] = wa, wb[:k["gamma"]]
    return k


def test_post_switch_error_kat_at_45_36_and_33_bits(kat):
    measured = {}
    for modulus, expected in oracle.POST_SWITCH_ERROR_KAT:
        sa = oracle.mod_switch(kat["a_hat"], oracle.Q, modulus)
        sb = oracle.mod_switch(kat["b_prefix"], oracle.Q, modulus)
        err = oracle.served_post_switch_error(kat["secret"], sa, sb, modulus, kat["record"])
        assert err == expected, f"target {modulus}"
        assert err < modulus // (2 * oracle.P)
        measured[modulus] = err
    # the served constant keeps nine bits under the half step; the largest 36-bit prime not four
    assert measured[oracle.Q36] << 9 < oracle.Q36 // (2 * oracle.P)
    assert not measured[68_719_464_449] << 4 < 68_719_464_449 // (2 * oracle.P)


def test_the_kat_response_decodes_at_every_implemented_modulus(kat):
    assert oracle.extract_inspiring(kat["secret"], kat["a_hat"], kat["b_prefix"], kat["entry"]) == kat["record"]
    for modulus in oracle.IMPLEMENTED_TARGETS:
        sa = oracle.mod_switch(kat["a_hat"], oracle.Q, modulus)
        sb = oracle.mod_switch(kat["b_prefix"], oracle.Q, modulus)
        got = oracle.extract_inspiring(kat["secret"], sa, sb, kat["entry"], retained=kat["gamma"], modulus=modulus)
        assert got == kat["record"]


def test_a_wrong_response_coefficient_does_not_decode(kat):
    """A planted mistake in the response (one coefficient of a off by a third of q, one of b
    off by one plaintext step) moves the error far past the recorded one."""
    for which in ("a", "b"):
        a, b = kat["a_hat"].copy(), kat["b_prefix"].copy()
        if which == "a":
            a[5] = (int(a[5]) + oracle.Q // 3) % oracle.Q
        else:
            b[3] = (int(b[3]) + (oracle.Q // oracle.P)) % oracle.Q
        sa, sb = oracle.mod_switch(a, oracle.Q, oracle.Q36), oracle.mod_switch(b, oracle.Q, oracle.Q36)
        err = oracle.served_post_switch_error(kat["secret"], sa, sb, oracle.Q36, kat["record"])
        assert err > 100 * 652


def test_noise_gate_verdicts_and_margins():
    # production_45bit_target_passes_noise_gate (mod_switch.rs:1020-1025)
    oracle.check_mod_switch_noise_budget(N, oracle.MOD_SWITCH_TARGET_45BIT)
    # the served 36-bit rung clears the gate 1.70x, the largest 36-bit NTT prime 1.45x,
    # the 33-bit prime fails at 0.57x (mod_switch.rs:41-46, 1062-1067)
    assert round(oracle.check_mod_switch_noise_budget(N, oracle.Q36), 2) == 1.70
    assert round(oracle.check_mod_switch_noise_budget(N, 68_719_464_449), 2) == 1.45
    with pytest.raises(ValueError, match="noise-budget violation"):
        oracle.check_mod_switch_noise_budget(N, 8_589_905_921)
    assert oracle.Q36 % oracle.P == 33 and 68_719_464_449 % oracle.P == 53_266
    with pytest.raises(ValueError, match="widening"):
        oracle.check_mod_switch_noise_budget(N, oracle.Q + 2 * N)
    with pytest.raises(ValueError, match="NTT-friendly"):
        oracle.check_mod_switch_noise_budget(N, oracle.Q36 + 2)
    with pytest.raises(ValueError, match="plaintext modulus"):
        oracle.check_mod_switch_noise_budget(N, 4097)


def test_reconstruct_entry_vector():
    vals = np.array([i * 256 + i for i in range(16)], np.uint64)
    e = oracle.reconstruct_entry(vals, 32)
    assert len(e) == 32
    for i in range(16):
        assert e[2 * i] | (e[2 * i + 1] << 8) == int(vals[i]) & 0xFFFF
    # an odd record keeps only the low byte of its last column
    assert oracle.reconstruct_entry(np.array([0x1234, 0xABCD], np.uint64), 3) == bytes([0x34, 0x12, 0xCD])


def test_plaintext_rounding_is_the_exact_quotient():
    rng = np.random.default_rng(3)
    for modulus in (oracle.Q, oracle.Q36, oracle.MOD_SWITCH_TARGET_45BIT):
        delta = modulus // oracle.P
        vals = [0, 1, delta // 2 - 1, delta // 2, delta - 1, delta, modulus - 1, modulus - delta // 2,
                modulus - delta // 2 - 1] + [int(v) for v in rng.integers(0, modulus, 200, dtype=np.uint64)]
        for v in vals:
            assert oracle.round_plaintext(v, delta, oracle.P) == ((v + delta // 2) // delta) % oracle.P


def test_phase_against_the_schoolbook_product():
    """phase = a*s + b mod (X^n + 1) by O(n^2) big integers, for a small secret (the i128
    path) and a full-width one (the per-product path)."""
    n, q = 64, oracle.Q36
    rng = np.random.default_rng(9)
    a = rng.integers(0, q, n, dtype=np.uint64)
    b = rng.integers(0, q, 20, dtype=np.uint64)
    for s in (np.array([int(v) % q for v in rng.integers(-20, 21, n)], np.uint64),
              rng.integers(0, q, n, dtype=np.uint64)):
        want = []
        for i in range(20):
            acc = 0
            for j in range(n):
                k = i - j
                if k >= 0:
                    acc += int(a[j]) * int(s[k])
                else:
                    acc -= int(a[j]) * int(s[n + k])
            want.append((acc + int(b[i])) % q)
        assert [int(v) for v in oracle.phase(s, a, b, 20, q)] == want


def test_switch_secret_is_the_centred_lift():
    q_new = oracle.Q36
    vals = np.array([0, 1, 5, oracle.Q - 1, oracle.Q - 7, oracle.Q // 2, oracle.Q // 2 + 1], np.uint64)
    got = [int(v) for v in oracle.switch_secret(vals, q_new)]
    want = []
    for v in vals.tolist():
        c = v - oracle.Q if v > oracle.Q // 2 else v
        want.append(c % q_new)
    assert got == want


@pytest.mark.parametrize("approximate", [False, True], ids=["exact", "approximate"])
def test_cell_responses_decode_to_their_records(approximate):
    """cell.assert_decodes (tests/packing_key_fixture/mod.rs:148-166) over the oracle's own
    responses, per query and through the shard store, at d2048 (a bounded sample of
    production_cells_decode_under_the_approximate_gadget's widths and indices)."""
    d = N
    for entry in (32, 512):
        c = oracle.Cell(d, entry, 0x4150_5244, approximate=approximate)
        p = oracle.Pack(d, c.gamma, 3, 1 if approximate else 0)
        stores = {}
        for index in (0, d - 1, 2 * d + 17):
            shard, seed, b = c.query(index)
            cols = c.columns(shard)
            if shard not in stores:
                stores[shard] = p.store_build(cols)
            a = oracle.expand_seed(seed, d)
            wa, wb = p.respond(c.w_seed, a, b, cols, c.y_body)
            sa, sb = p.respond(c.w_seed, a, b, cols, c.y_body, stores[shard])
            assert np.array_equal(wa, sa) and np.array_equal(wb, sb)
            rec = c.db[index * entry:(index + 1) * entry].tobytes()
            assert oracle.extract_inspiring(c.secret, wa, wb[:c.gamma], entry) == rec, f"entry {entry} index {index}"


def test_served_switch_decodes_through_the_wire_form():
    """production_cells_decode_after_the_served_switch (tests/approximate_packing_gadget.rs:463-494):
    switch to 36 bits, serialize, parse, decode."""
    c = oracle.Cell(N, 256, 0x4150_5357, approximate=True)
    p = oracle.Pack(N, c.gamma, 3, 1)
    for index in (3, 2047, 4096):
        shard, seed, b = c.query(index)
        wa, wb = p.respond(c.w_seed, oracle.expand_seed(seed, N), b, c.columns(shard), c.y_body)
        sa, sb = oracle.mod_switch(wa), oracle.mod_switch(wb[:c.gamma])
        wire = oracle.encode_response(sa, sb, c.gamma, oracle.Q36)
        assert len(wire) == 10_446 - (256 - c.gamma) * 36 // 8
        r = oracle.parse_response(wire)
        assert r["modulus"] == oracle.Q36 and r["retained"] == c.gamma
        assert np.array_equal(r["a"], sa) and np.array_equal(r["b_prefix"], sb)
        got = oracle.extract_inspiring(c.secret, r["a"], r["b_prefix"], 256, r["retained"], r["modulus"])
        assert got == c.db[index * 256:(index + 1) * 256].tobytes()


def test_extract_refusals(kat):
    with pytest.raises(ValueError, match="retains"):
        oracle.extract_inspiring(kat["secret"], kat["a_hat"], kat["b_prefix"], kat["entry"], retained=255)
    with pytest.raises(ValueError, match="not an implemented"):
        oracle.extract_inspiring(kat["secret"], kat["a_hat"], kat["b_prefix"], kat["entry"], modulus=8_589_905_921)
    wire = bytearray(oracle.encode_response(kat["a_hat"], kat["b_prefix"], kat["gamma"]))
    assert oracle.parse_response(bytes(wire))["modulus"] == oracle.Q
    bad = bytearray(wire)
    bad[0] = 2
    with pytest.raises(ValueError, match="not a packed"):
        oracle.parse_response(bytes(bad))
    with pytest.raises(ValueError):
        oracle.parse_response(bytes(wire[:-1]))


@pytest.mark.parametrize("d", [512, 1024])
@pytest.mark.parametrize("approximate", [False, True], ids=["exact", "approximate"])
def test_cell_responses_decode_at_rings_512_and_1024(d, approximate):
    """The same decode check over the oracle's responses at the rings 512 and 1024 the d2048 prime
    and gadget admit, per query and through the shard store, unswitched and at 36 bits."""
    entry = 32
    c = oracle.Cell(d, entry, 0x5249_4e47 + d, approximate=approximate)
    p = oracle.Pack(d, c.gamma, 3, 1 if approximate else 0)
    for index in (0, d - 1, 2 * d + 5):
        shard, seed, b = c.query(index)
        cols = c.columns(shard)
        a = oracle.expand_seed(seed, d)
        wa, wb = p.respond(c.w_seed, a, b, cols, c.y_body)
        sa, sb = p.respond(c.w_seed, a, b, cols, c.y_body, p.store_build(cols))
        assert np.array_equal(wa, sa) and np.array_equal(wb, sb)
        rec = c.db[index * entry:(index + 1) * entry].tobytes()
        assert oracle.extract_inspiring(c.secret, wa, wb[:c.gamma], entry) == rec, index
        got = oracle.extract_inspiring(c.secret, oracle.mod_switch(wa), oracle.mod_switch(wb[:c.gamma]), entry,
                                       c.gamma, oracle.Q36)
        assert got == rec, index
