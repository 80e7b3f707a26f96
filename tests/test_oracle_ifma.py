"""The oracle's AVX-512 IFMA path (the reference's simd-packing-offline feature, which the CPU
baseline times beside the scalar path) gives the scalar path's values.

* the 8-wide split-52 Montgomery multiply-accumulate equals the scalar Montgomery product
  (solinas_montgomery_kat.rs's property: Solinas REDC == classical) on edge values around
  the 52-bit split and q, and on random residues, contiguous and through a table;
* whole responds with the switch on equal those with it off, on the store and no-store paths,
  and the d256 fixture still hits the reference's golden digest.
Skipped on hosts without AVX-512 IFMA.
"""
import numpy as np
import pytest

import oracle
import wire

Q = oracle.Q

pytestmark = pytest.mark.skipif(not oracle.ifma_available(), reason="host has no AVX-512 IFMA")


def _scalar(ntt, acc, a, b):
    return np.array([(int(x) + ntt.mont_mul(int(y), int(z))) % Q for x, y, z in zip(acc, a, b)], np.uint64)


def test_mont_mul_acc_equals_scalar():
    ntt = oracle.Ntt(8)
    qinv = (-pow(Q, -1, 1 << 64)) % (1 << 64)
    edges = [0, 1, 2, (1 << 52) - 1, 1 << 52, (1 << 52) + 1, (1 << 59), Q - 1, Q - 2, Q // 2, (1 << 60) - (1 << 14)]
    rng = np.random.default_rng(8)
    vals = np.array(edges + [int(v) for v in rng.integers(0, Q, 4096 - len(edges), dtype=np.uint64)], np.uint64)
    a = vals
    b = np.roll(vals, 3)
    acc = np.roll(vals, 7)
    got = oracle.ifma_mont_mul_acc(acc, a, b, qinv)
    want = np.array([(int(x) + int(y) * int(z) * pow(2, -64, Q)) % Q for x, y, z in zip(acc, a, b)], np.uint64)
    assert np.array_equal(got, want)
    assert np.array_equal(got[:64], _scalar(ntt, acc[:64], a[:64], b[:64]))
    # every edge against every edge
    e = np.array(edges, np.uint64)
    aa, bb = np.repeat(e, len(e)), np.tile(e, len(e))
    pad = (-len(aa)) % 8
    aa, bb = np.concatenate([aa, np.zeros(pad, np.uint64)]), np.concatenate([bb, np.zeros(pad, np.uint64)])
    got = oracle.ifma_mont_mul_acc(np.zeros_like(aa), aa, bb, qinv)
    assert [int(v) for v in got] == [int(x) * int(y) * pow(2, -64, Q) % Q for x, y in zip(aa, bb)]


def test_gather_equals_scalar():
    qinv = (-pow(Q, -1, 1 << 64)) % (1 << 64)
    rng = np.random.default_rng(9)
    n = 2048
    w = rng.integers(0, Q, n, dtype=np.uint64)
    a = rng.integers(0, Q, n, dtype=np.uint64)
    acc = rng.integers(0, Q, n, dtype=np.uint64)
    tab = oracle.Pack(n, 256).automorph_table(2 * n // 256 + 1)
    got = oracle.ifma_mont_mul_acc(acc, a, w, qinv, tab=tab)
    want = [(int(x) + int(y) * int(w[t]) * pow(2, -64, Q)) % Q for x, y, t in zip(acc, a, tab)]
    assert [int(v) for v in got] == want


@pytest.mark.parametrize("n,gamma,dropped", [(256, 16, 0), (2048, 256, 0), (2048, 256, 1), (2048, 16, 0)])
def test_respond_with_the_simd_path_equals_scalar(n, gamma, dropped):
    rng = np.random.default_rng(n + gamma + dropped)
    rows = 3 - dropped
    cols = rng.integers(0, 1 << 16, (gamma, n), dtype=np.uint64)
    y = rng.integers(0, Q, (rows, n), dtype=np.uint64)
    b = rng.integers(0, Q, n, dtype=np.uint64)
    a = oracle.expand_seed(rng.integers(0, 256, 32, dtype=np.uint8).tobytes(), n)
    w_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    scalar = oracle.Pack(n, gamma, 3, dropped)
    simd = oracle.Pack(n, gamma, 3, dropped)
    assert simd.set_simd(True)
    store = scalar.store_build(cols)
    assert np.array_equal(simd.store_build(cols), store)
    for sums in (store, None):
        sa, sb = scalar.respond(w_seed, a, b, cols, y, sums)
        va, vb = simd.respond(w_seed, a, b, cols, y, sums)
        assert np.array_equal(sa, va) and np.array_equal(sb, vb)


def test_d256_fixture_golden_with_the_simd_path():
    fx = wire.load_d256_fixture()
    n, g = fx["params"]["ring_dim"], fx["num_columns"]
    p = oracle.Pack(n, g)
    assert p.set_simd(True)
    resp = oracle.encode_response(*p.respond(fx["w_seed"], fx["a"], fx["b"], fx["columns"], fx["y_body"]), g)
    assert oracle.fnv1a(resp) == wire.D256_GOLDEN_FNV
