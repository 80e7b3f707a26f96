"""The device arithmetic headers (modq.hpp, chacha.hpp), compiled for the host, against
big-integer arithmetic and the oracle. Runs on the CPU: it checks the exact code the
kernels inline, not a restatement of it."""
import ctypes
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "inspire-rs_amd", "libinspire_selftest.so")
Q = oracle.Q


@pytest.fixture(scope="module")
def st():
    if not os.path.exists(LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.dirname(LIB), "libinspire_selftest.so"], check=True)
    L = ctypes.CDLL(LIB)
    u64, u32 = ctypes.c_uint64, ctypes.c_uint32
    L.st_mulmod.argtypes = [u64, u64]
    L.st_mulmod.restype = u64
    L.st_reduce120.argtypes = [u64, u64]
    L.st_reduce120.restype = u64
    L.st_mul_wide.argtypes = [u64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.st_acc_reduce.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, u64]
    L.st_acc_reduce.restype = u64
    L.st_mod_switch.argtypes = [u64, u64]
    L.st_mod_switch.restype = u64
    L.st_mod_p61.argtypes = [u64, u64]
    L.st_mod_p61.restype = u64
    L.st_automorph_slot.argtypes = [u32, u32, ctypes.c_int]
    L.st_automorph_slot.restype = u32
    L.st_chacha_block.argtypes = [ctypes.c_char_p, u64, u64, ctypes.c_void_p]
    L.st_shoup_companion.argtypes = [u64]
    L.st_shoup_companion.restype = u64
    L.st_shoup.argtypes = [u64, u64]
    L.st_shoup.restype = u64
    L.st_uniform_q_draw.argtypes = [u64, ctypes.POINTER(u64)]
    L.st_shoup_lazy.argtypes = [u64, u64]
    L.st_shoup_lazy.restype = u64
    return L


EDGES = [0, 1, 2, Q - 1, Q - 2, Q // 2, (1 << 32) - 1, 1 << 32, (1 << 59) + 12345, (1 << 60) - (1 << 14)]


def test_mulmod_edges_and_random(st):
    rng = np.random.default_rng(0)
    vals = EDGES + [int(x) for x in rng.integers(0, Q, 300, dtype=np.uint64)]
    for a in vals:
        for b in vals[:40]:
            assert st.st_mulmod(a, b) == a * b % Q


def test_mul_wide_full_64bit(st):
    rng = np.random.default_rng(1)
    hi, lo = ctypes.c_uint64(), ctypes.c_uint64()
    for a, b in [(2**64 - 1, 2**64 - 1), (2**64 - 1, Q)] + [tuple(int(x) for x in rng.integers(0, 2**63, 2, dtype=np.uint64)) for _ in range(200)]:
        st.st_mul_wide(a, b, ctypes.byref(hi), ctypes.byref(lo))
        assert (hi.value << 64) | lo.value == a * b


def test_reduce120_range(st):
    rng = np.random.default_rng(2)
    for _ in range(2000):
        x = int(rng.integers(0, 2**60, dtype=np.uint64)) * int(rng.integers(0, 2**60, dtype=np.uint64))
        assert st.st_reduce120(x >> 64, x & (2**64 - 1)) == x % Q
    top = (Q - 1) * (Q - 1)
    assert st.st_reduce120(top >> 64, top & (2**64 - 1)) == top % Q


def test_acc128_sums_of_products(st):
    rng = np.random.default_rng(3)
    for count in (1, 3, 4, 7, 16):
        a = rng.integers(0, Q, count, dtype=np.uint64)
        b = rng.integers(0, Q, count, dtype=np.uint64)
        init = int(rng.integers(0, Q, dtype=np.uint64))
        want = (init + sum(int(x) * int(y) for x, y in zip(a, b))) % Q
        assert st.st_acc_reduce(a.ctypes.data, b.ctypes.data, count, init) == want
    a = np.full(16, Q - 1, dtype=np.uint64)
    assert st.st_acc_reduce(a.ctypes.data, a.ctypes.data, 16, Q - 1) == (Q - 1 + 16 * (Q - 1) ** 2) % Q


@pytest.mark.parametrize("qn", [oracle.Q36, oracle.MOD_SWITCH_TARGET_45BIT, 68_719_464_449, 8_589_905_921,
                                (1 << 46) - 1])
def test_mod_switch_matches_oracle(st, qn):
    """Every implemented target, two other NTT primes and the largest modulus the device
    accepts, against the oracle's u128 switch and exact big integers."""
    rng = np.random.default_rng(4)
    vals = np.array(EDGES[:6] + [int(x) for x in rng.integers(0, Q, 2000, dtype=np.uint64)], dtype=np.uint64)
    want = oracle.mod_switch(vals, oracle.Q, qn)
    assert [int(w) for w in want] == [((int(v) * qn + Q // 2) // Q) % qn for v in vals]
    assert [st.st_mod_switch(int(v), qn) for v in vals] == [int(w) for w in want]


def test_mod_p61(st):
    p = 2**61 - 1
    rng = np.random.default_rng(5)
    for _ in range(1000):
        hi = int(rng.integers(0, 2**58, dtype=np.uint64))
        lo = int(rng.integers(0, 2**63, dtype=np.uint64)) * 2 + 1
        assert st.st_mod_p61(hi, lo) == ((hi << 64) | lo) % p


@pytest.mark.parametrize("n,gamma", [(256, 16), (2048, 256)])
def test_automorph_slot_matches_reference_tables(st, n, gamma):
    p = oracle.Pack(n, gamma)
    logn = n.bit_length() - 1
    g = 2 * n // gamma + 1
    for t in (1, g, pow(g, 7, 2 * n), 2 * n - 1):
        tab = p.automorph_table(t)
        assert [st.st_automorph_slot(t, j, logn) for j in range(n)] == list(tab)


def test_chacha_and_uniform_draws_match_oracle(st):
    seed = bytes(range(7, 39))
    out = np.zeros(16, dtype=np.uint32)
    want = oracle.expand_seed(seed, 64)
    got = []
    for blk in range(8):
        st.st_chacha_block(seed, blk, 0, out.ctypes.data)
        for u in range(8):
            v = int(out[2 * u]) | (int(out[2 * u + 1]) << 32)
            r = ctypes.c_uint64()
            assert st.st_uniform_q_draw(v, ctypes.byref(r)) == 1
            got.append(r.value)
    assert got == [int(x) for x in want]
    # the rejection zone edge: lo = v*q mod 2^64 must be <= (q << 4) - 1
    r = ctypes.c_uint64()
    zone = ((Q << 4) - 1) & (2**64 - 1)
    qinv = pow(Q, -1, 2**64)
    v_reject = ((zone + 1) * qinv) % 2**64  # v*q = zone+1 (mod 2^64): just outside
    assert st.st_uniform_q_draw(v_reject, ctypes.byref(r)) == 0
    v_accept = (zone * qinv) % 2**64
    assert st.st_uniform_q_draw(v_accept, ctypes.byref(r)) == 1


def test_shoup_companion_and_product_range(st):
    rng = np.random.default_rng(6)
    ws = EDGES + [int(x) for x in rng.integers(0, Q, 500, dtype=np.uint64)]
    for w in ws:
        w %= Q
        assert st.st_shoup_companion(w) == (w << 64) // Q
    # a*w - floor(a*w'/2^64)*q lies in [0, 2q) for every a < 2^64 and is = a*w (mod q)
    for a in [0, 1, Q - 1, 2 * Q - 1, 4 * Q - 1, 2**64 - 1] + [int(x) for x in rng.integers(0, 2**63, 300, dtype=np.uint64)]:
        w = int(rng.integers(0, Q, dtype=np.uint64))
        r = st.st_shoup(a, w)
        assert r < 2 * Q and r % Q == a * w % Q


def test_lazy_shoup_range_and_residue(st):
    """The packing chain's product: a*w (mod q) in [0, 4q) for every a < 2^64, including the
    operands that maximise the dropped low partial products and middle carries."""
    rng = np.random.default_rng(7)
    M32 = (1 << 32) - 1
    worst_a = [2**64 - 1, (M32 << 32) | M32, (1 << 63) - 1, 8 * Q - 1, 16 * Q - 1, 4 * Q - 1, M32, M32 << 32]
    ws = [Q - 1, Q - 2, 1, 0, Q // 2, (1 << 59) + 7] + [int(x) for x in rng.integers(0, Q, 200, dtype=np.uint64)]
    seen_above_2q = False
    for a in worst_a + [int(x) for x in rng.integers(0, 2**64 - 1, 400, dtype=np.uint64)]:
        for w in ws[:60]:
            r = st.st_shoup_lazy(a, w)
            assert r < 4 * Q and r % Q == a * w % Q, (a, w, r)
            seen_above_2q |= r >= 2 * Q
    # search for operands whose estimate is 2 short: the bound is reached, not loose
    short2 = 0
    for _ in range(20000):
        a = int(rng.integers(0, 2**64 - 1, dtype=np.uint64))
        w = int(rng.integers(0, Q, dtype=np.uint64))
        wp = (w << 64) // Q
        a0, a1, p0, p1 = a & M32, a >> 32, wp & M32, wp >> 32
        est = a1 * p1 + ((a1 * p0) >> 32) + ((a0 * p1) >> 32)
        short2 += (a * wp >> 64) - est == 2
        r = st.st_shoup_lazy(a, w)
        assert r < 4 * Q and r % Q == a * w % Q
    assert short2 > 0 and seen_above_2q
