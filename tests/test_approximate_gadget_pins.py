"""Pins of the approximate-gadget respond path (d2048_approximate_gadget) that do not rest on
decode correctness alone.

The reference commits no response vector for this gadget (tests/approximate_packing_gadget.rs
pins decode correctness and sizes), so the two places where its bytes differ from the exact
gadget's are pinned here by what the reference states they must satisfy, in checks that a
plausible mistake fails:

* the rounded digit split (gadget_decompose_rounded / rounded_digit_rows,
  src/inspiring/inspiring2.rs:1167-1225): digits of round(c / 2^20), half up, whose top row
  keeps the rounding carry. Checked on the kernels' own digit code (csrc/modq.hpp,
  gadget_digits, compiled for the host) and on the oracle's, at edge coefficients: a
  truncating split, a half-down tie or a carry dropped from the top row fails.
* the mask stream (kept_mask_rows, inspiring2.rs:835-855): an approximate key expands on the
  stream named by its width, "so keys of one secret at two widths, or under both gadgets,
  never share a mask", and row r is the mask the expansion gives digit 1 + r. A stream-0 or
  width-free expansion, or a key that keeps the dropped digit's row, fails.

CPU only: the device code is the host build of the same header, and the GPU tests check the
device responses against the oracle these pins hold.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "inspire-rs_amd", "libinspire_selftest.so")
Q = oracle.Q
Z = 1 << 20
HALF = 1 << 19


@pytest.fixture(scope="module")
def st():
    if not os.path.exists(LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.dirname(LIB), "libinspire_selftest.so"], check=True)
    L = ctypes.CDLL(LIB)
    u64 = ctypes.c_uint64
    L.st_gadget_digits.argtypes = [u64, ctypes.c_int, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    return L


def device_digits(st, c, exact):
    d0, d1 = ctypes.c_uint64(), ctypes.c_uint64()
    st.st_gadget_digits(c, 1 if exact else 0, ctypes.byref(d0), ctypes.byref(d1))
    return d0.value, d1.value


def edge_coefficients():
    """Rounding boundaries, ties, carries into the top digit and the ends of [0, q)."""
    cs = {0, 1, HALF - 1, HALF, HALF + 1, Z - 1, Z, Z + HALF - 1, Z + HALF, Q - 1, Q - 2, Q // 2,
          (1 << 60) - HALF - 1, (1 << 60) - HALF, (1 << 60) - (1 << 14)}
    for k in (1, 2, 3, 1000, (1 << 20) - 1, 1 << 20, (1 << 39) + 5, (1 << 40) - 1):
        for off in (-HALF - 1, -HALF, -HALF + 1, -1, 0, 1, HALF - 1, HALF, HALF + 1):
            cs.add(k * Z + off)
    # carries out of the low digit: round(c / 2^20) = 2^20 k exactly
    for k in (1, 7, (1 << 20) - 1):
        for off in (-HALF - 1, -HALF, HALF - 1):
            cs.add(k * (1 << 40) + off)
    rng = np.random.default_rng(20)
    cs.update(int(x) for x in rng.integers(0, Q, 400, dtype=np.uint64))
    return sorted(c for c in cs if 0 <= c < Q)


def check_rounded(c, d0, d1):
    """The rounded split's contract: c = 2^20 (d0 + 2^20 d1) + e, e in [-2^19, 2^19) (round half
    up), the low digit below the base, the top digit at most the base (it keeps the carry)."""
    e = c - Z * (d0 + Z * d1)
    assert -HALF <= e < HALF, f"c={c}: residual {e} is not the round-half-up residual"
    assert 0 <= d0 < Z, f"c={c}: low digit {d0}"
    assert 0 <= d1 <= Z, f"c={c}: top digit {d1}"


def test_device_rounded_digits_satisfy_the_rounding_contract(st):
    for c in edge_coefficients():
        check_rounded(c, *device_digits(st, c, exact=False))


def test_device_rounded_digits_at_ties_and_carries(st):
    # ties round up: 2^19 -> 1, 2^20 k + 2^19 -> k + 1
    assert device_digits(st, HALF, False) == (1, 0)
    assert device_digits(st, HALF - 1, False) == (0, 0)
    assert device_digits(st, 5 * Z + HALF, False) == (6, 0)
    # a carry out of the low digit into the top one
    assert device_digits(st, (1 << 40) - HALF, False) == (0, 1)
    # the top digit keeps the rounding carry: it equals the base near q
    assert device_digits(st, (1 << 60) - HALF, False) == (0, Z)
    assert device_digits(st, Q - 1, False) == (0, Z)


def test_device_exact_digits_are_the_low_two_base_digits(st):
    for c in edge_coefficients():
        d0, d1 = device_digits(st, c, exact=True)
        assert (d0, d1) == (c % Z, (c >> 20) % Z), f"c={c}"
        assert c == d0 + Z * d1 + (1 << 40) * (c >> 40)


def test_oracle_rounded_digits_satisfy_the_rounding_contract(st):
    n = 256
    p = oracle.Pack(n, 16, 3, 1)
    cs = edge_coefficients()
    cs += [0] * (-len(cs) % n)
    for s in range(0, len(cs), n):
        chunk = np.array(cs[s:s + n], dtype=np.uint64)
        d = p.decompose(chunk)
        assert d.shape == (2, n)
        for j in range(n):
            c = int(chunk[j])
            check_rounded(c, int(d[0, j]), int(d[1, j]))
            assert (int(d[0, j]), int(d[1, j])) == device_digits(st, c, exact=False)


def test_planted_truncation_and_half_down_fail_the_contract():
    """The contract has teeth: splits a plausible mistake would produce violate it."""
    def truncating(c):
        r = c >> 20
        return r % Z, r >> 20

    def half_down(c):
        r = (c + HALF - 1) >> 20
        return r % Z, r >> 20

    def top_masked(c):
        r = (c + HALF) >> 20
        return r % Z, (r >> 20) % Z

    for bad in (truncating, half_down, top_masked):
        failed = False
        for c in edge_coefficients():
            try:
                check_rounded(c, *bad(c))
            except AssertionError:
                failed = True
                break
        assert failed, bad.__name__


def mask_rows_from_stream(w_seed, stream, n, first_row, rows):
    """Rows first_row.. of expand_mask(seed, stream) straight from ChaCha20 blocks (the block
    function pinned by RFC 7539 in test_oracle_golden): row k is draws k n .. (k+1) n - 1, each
    draw the two keystream words 2d, 2d+1, low first, reduced mod q (inspiring2.rs:866-887)."""
    key = np.frombuffer(w_seed, dtype="<u4")
    out = np.zeros((rows, n), dtype=np.uint64)
    for r in range(rows):
        row = first_row + r
        for blk in range(n // 8):
            w = oracle.chacha_block(key, row * (n // 8) + blk, stream).astype(np.uint64)
            for u in range(8):
                out[r, blk * 8 + u] = int(w[2 * u] | (w[2 * u + 1] << np.uint64(32))) % Q
    return out


@pytest.mark.parametrize("n,gamma", [(256, 16), (2048, 256)])
def test_approximate_mask_rows_are_digits_one_and_two_of_the_width_stream(n, gamma):
    w_seed = bytes((7 * i + 3) & 0xFF for i in range(32))
    got = oracle.Pack(n, gamma, 3, 1).w_mask(w_seed)
    assert got.shape == (2, n)
    assert np.array_equal(got, mask_rows_from_stream(w_seed, gamma, n, 1, 2))
    exact = oracle.Pack(n, gamma, 3, 0).w_mask(w_seed)
    assert np.array_equal(exact, mask_rows_from_stream(w_seed, 0, n, 0, 3))


@pytest.mark.parametrize("n,gammas", [(256, (2, 16, 64, 128)), (2048, (16, 256, 512))])
def test_no_two_keys_share_a_mask(n, gammas):
    """kept_mask_rows' reason (inspiring2.rs:835-838): keys of one CRS seed at two widths, or under
    both gadgets, never share a mask row."""
    w_seed = bytes(range(100, 132))
    rows = {("exact", g): oracle.Pack(n, g, 3, 0).w_mask(w_seed) for g in gammas}
    rows.update({("approx", g): oracle.Pack(n, g, 3, 1).w_mask(w_seed) for g in gammas})
    # an approximate row never equals any exact row or another width's approximate row
    exact_rows = {m[r].tobytes() for (kind, _), m in rows.items() if kind == "exact" for r in range(3)}
    for g in gammas:
        for r in range(2):
            row = rows[("approx", g)][r].tobytes()
            assert row not in exact_rows, f"approximate width {g} row {r} is an exact mask row"
            for g2 in gammas:
                if g2 != g:
                    assert row not in {rows[("approx", g2)][k].tobytes() for k in range(2)}, (g, g2, r)


def test_planted_stream_mistakes_share_a_mask():
    """A stream-0 approximate expansion shares the exact key's rows 1..2, and a width-free one
    (any fixed stream) shares rows across widths: both fail the separation above."""
    n, w_seed = 256, bytes(range(32))
    exact = oracle.Pack(n, 16, 3, 0).w_mask(w_seed)
    stream0 = mask_rows_from_stream(w_seed, 0, n, 1, 2)
    assert np.array_equal(stream0, exact[1:3])
    fixed_a = mask_rows_from_stream(w_seed, 1, n, 1, 2)  # the same stream whatever the width
    fixed_b = mask_rows_from_stream(w_seed, 1, n, 1, 2)
    assert np.array_equal(fixed_a, fixed_b)
    # the real expansion at two widths does not
    assert not np.array_equal(oracle.Pack(n, 16, 3, 1).w_mask(w_seed), oracle.Pack(n, 32, 3, 1).w_mask(w_seed))
