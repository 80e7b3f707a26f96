"""The kernels' serial rejection replays, compiled for the host, against the oracle's samplers.

The parallel expansions in the kernels take one ChaCha20 block per lane and fall back to a
serial replay of the whole stream when any lane draws a word outside the acceptance zone
(gen_range(0..q) for a query's a, src/math/poly.rs:657-685; expand_uniform for the
special-modulus query mask and key mask, src/special_modulus/mod.rs:216-237, 313-330). At the
reference's zones a rejection has probability 2^-46 (gen_range) or below 2^-34 per draw, so no
seeded input reaches the replay. Here the replay code itself (replay_gen_range_q and
replay_expand_uniform in csrc/chacha.hpp) runs with narrowed zones that reject often, against
the oracle's serial loops run with the same zones; and at the reference's zones it must give
the oracle's ordinary expansion.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "inspire-rs_amd", "libinspire_selftest.so")
Q = oracle.Q
ZONE_Q = (Q << 4) - 1  # gen_range's acceptance bound for q (lz(q) = 4)
KEY_MASK_STREAM = int.from_bytes(b"key mask", "big")
SM_Q, SM_P = 536813569, 2147389441


@pytest.fixture(scope="module")
def st():
    if not os.path.exists(LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.dirname(LIB), "libinspire_selftest.so"], check=True)
    L = ctypes.CDLL(LIB)
    u64, u32 = ctypes.c_uint64, ctypes.c_uint32
    L.st_replay_gen_range_q.argtypes = [ctypes.c_char_p, ctypes.c_int, u64, ctypes.c_void_p]
    L.st_replay_expand_uniform.argtypes = [ctypes.c_char_p, u64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p]
    L.st_reduce64.argtypes = [u64, u32]
    L.st_reduce64.restype = u32
    return L


def replay_q(st, seed, n, zone):
    out = np.zeros(n, dtype=np.uint64)
    st.st_replay_gen_range_q(seed, n, zone, out.ctypes.data)
    return out


def replay_uniform(st, seed, stream, moduli, zones, count):
    m = np.array(moduli, dtype=np.uint32)
    z = np.array(zones, dtype=np.uint64)
    out = np.zeros((len(m), count), dtype=np.uint32)
    st.st_replay_expand_uniform(seed, stream, len(m), count, m.ctypes.data, z.ctypes.data, out.ctypes.data)
    return out


SEEDS = [bytes(32), bytes(range(32)), bytes((13 * i + 5) & 0xFF for i in range(32))]


@pytest.mark.parametrize("seed", SEEDS)
def test_gen_range_replay_at_the_reference_zone_is_the_expansion(st, seed):
    for n in (256, 2048):
        assert np.array_equal(replay_q(st, seed, n, ZONE_Q), oracle.expand_seed(seed, n))


@pytest.mark.parametrize("shift", [1, 2, 5])
@pytest.mark.parametrize("seed", SEEDS)
def test_gen_range_replay_with_rejections_matches_the_serial_sampler(st, seed, shift):
    # the zone narrowed by 2^shift rejects about 1 - 2^-shift of the draws
    zone = ZONE_Q >> shift
    n = 2048
    got = replay_q(st, seed, n, zone)
    assert np.array_equal(got, oracle.expand_seed_zone(seed, n, zone))
    assert (got < Q).all()
    # the rejections moved the stream: the plain expansion differs
    assert not np.array_equal(got, oracle.expand_seed(seed, n))


def sm_zone(m):
    return ((1 << 64) - 1) // m * m


@pytest.mark.parametrize("seed", SEEDS)
def test_expand_uniform_replay_at_the_reference_zones(st, seed):
    # the query mask (stream 0, q only) and the key mask (stream "key mask", q limb then p limb)
    got = replay_uniform(st, seed, 0, [SM_Q], [sm_zone(SM_Q)], 2048)
    assert np.array_equal(got, oracle.sm_expand_uniform_zone(seed, 0, [SM_Q], [sm_zone(SM_Q)], 2048))
    assert np.array_equal(got[0], oracle.SmRing().query_mask(seed))
    got = replay_uniform(st, seed, KEY_MASK_STREAM, [SM_Q, SM_P], [sm_zone(SM_Q), sm_zone(SM_P)], 2048)
    want = oracle.sm_expand_uniform_zone(seed, KEY_MASK_STREAM, [SM_Q, SM_P], [sm_zone(SM_Q), sm_zone(SM_P)], 2048)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seed", SEEDS)
@pytest.mark.parametrize("stream,moduli", [(0, [SM_Q]), (KEY_MASK_STREAM, [SM_Q, SM_P]),
                                           (KEY_MASK_STREAM, [536903681, SM_P])])
def test_expand_uniform_replay_with_rejections_matches_the_serial_sampler(st, seed, stream, moduli):
    # zones of about 3/8 and 5/8 of the word range: most draws of one limb, and some of the
    # other, are rejected, so the limbs' boundary in the stream moves too
    zones = [(((1 << 64) * 3 // 8) // m * m) if l == 0 else ((1 << 64) * 5 // 8) // m * m
             for l, m in enumerate(moduli)]
    for count in (64, 2048):
        got = replay_uniform(st, seed, stream, moduli, zones, count)
        want = oracle.sm_expand_uniform_zone(seed, stream, moduli, zones, count)
        assert np.array_equal(got, want), (count, moduli)
        for l, m in enumerate(moduli):
            assert (got[l] < m).all()


def test_reduce64_is_the_remainder(st):
    rng = np.random.default_rng(3)
    words = [0, 1, (1 << 64) - 1, (1 << 63), (1 << 32) - 1, 1 << 32]
    words += [int(x) for x in rng.integers(0, 1 << 63, 500, dtype=np.uint64)]
    words += [int(x) | (1 << 63) for x in rng.integers(0, 1 << 63, 500, dtype=np.uint64)]
    for m in (SM_Q, SM_P, 536903681, 715827457, 2147483647, 12289, 3):
        words_m = words + [k * m + r for k in (1, (1 << 32) + 7, ((1 << 64) - 1) // m) for r in (0, 1, m - 1)
                           if k * m + r < (1 << 64)]
        for u in words_m:
            assert st.st_reduce64(u, m) == u % m, (u, m)
