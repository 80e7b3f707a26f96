"""Test helper: decode the reference's bincode fixture respond_byte_identity_d256.bin.

The fixture is `bincode::serialize(&(&crs, &encoded_db, &query))` with bincode 1.3
fixint little-endian encoding (tests/respond_byte_identity_kat.rs:113-132). Layouts
followed: InspireParams positional tuple (src/params.rs:226-258), ServerCrs
(src/pir/setup.rs:75-96), KeySwitchingMatrix (src/ks/setup.rs:13-20), GadgetVector
(src/rgsw/types.rs:25-32), tight Poly wire form (src/math/poly.rs:142-150, 286-310),
EncodedDatabase/ShardData (src/pir/setup.rs:271-289), ShardConfig (src/params.rs:1130-1146),
ClientQuery (src/pir/query.rs:103-117), ClientPackingKeys (src/inspiring/inspiring2.rs:684-700).
"""
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class _Reader:
    def __init__(self, data):
        self.b = data
        self.pos = 0

    def raw(self, n):
        v = self.b[self.pos:self.pos + n]
        if len(v) != n:
            raise ValueError("fixture truncated")
        self.pos += n
        return v

    def u8(self):
        return self.raw(1)[0]

    def u32(self):
        return struct.unpack("<I", self.raw(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.raw(8))[0]

    def f64(self):
        return struct.unpack("<d", self.raw(8))[0]

    def vec(self, f):
        return [f() for _ in range(self.u64())]

    def poly(self):
        nbytes = self.u64()
        data = self.raw(nbytes)
        moduli = self.vec(self.u64)
        q = self.u64()
        dim = self.u64()
        self.u64()  # crt_q0_inv_mod_q1
        is_ntt = self.u8()
        bits = (max(moduli) - 1).bit_length()
        x = int.from_bytes(data, "little")
        mask = (1 << bits) - 1
        coeffs = np.array([(x >> (i * bits)) & mask for i in range(dim * len(moduli))], dtype=np.uint64)
        return {"coeffs": coeffs, "moduli": moduli, "q": q, "dim": dim, "is_ntt": is_ntt}

    def gadget(self):
        return {"base": self.u64(), "len": self.u64(), "q": self.u64()}

    def rlwe(self):
        return {"a": self.poly(), "b": self.poly()}


def load_d256_fixture():
    """Return the decoded (crs, db, query) essentials of the d256 fixture."""
    with open(os.path.join(GOLDEN, "respond_byte_identity_d256.bin"), "rb") as f:
        r = _Reader(f.read())
    first = r.u64()
    approx = first == (1 << 64) - 1
    ring_dim = r.u64() if approx else first
    params = {
        "ring_dim": ring_dim, "q": r.u64(), "crt_moduli": r.vec(r.u64), "p": r.u64(),
        "sigma": r.f64(), "gadget_base": r.u64(), "query_gadget_len": r.u64(),
        "packing_gadget_len": r.u64(), "label": r.u32(), "approximate": approx,
    }
    r.vec(lambda: {"rows": r.vec(r.rlwe), "gadget": r.gadget()})  # galois keys (unused)
    r.gadget()  # rgsw gadget
    w_seed = r.raw(32)
    r.raw(32)  # v_seed
    num_columns = r.u64()
    shards = r.vec(lambda: {"id": r.u32(), "polys": r.vec(r.poly)})
    config = (r.u64(), r.u64(), r.u64())
    shard_id = r.u32()
    rows = r.vec(r.rlwe)
    r.gadget()
    r.u32()  # packing_mode
    keys = None
    if r.u8() == 1:
        keys = {"y_body": r.vec(r.poly), "z_body": r.vec(r.poly), "full_key": r.u8(), "num_to_pack": r.u64()}
    if r.u8() == 1:
        r.u64()
    if r.pos != len(r.b):
        raise ValueError("fixture has trailing bytes")
    columns = np.stack([p["coeffs"] for p in shards[shard_id]["polys"]])
    return {
        "params": params,
        "w_seed": w_seed,
        "num_columns": num_columns,
        "columns": columns,
        "config": config,
        "shard_id": shard_id,
        "a": rows[0]["a"]["coeffs"],
        "b": rows[0]["b"]["coeffs"],
        "y_body": np.stack([p["coeffs"] for p in keys["y_body"]]),
    }


D256_GOLDEN_FNV = 0x796A80CA60D31A52  # tests/respond_byte_identity_kat.rs:136
