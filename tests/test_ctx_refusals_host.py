"""InspireParams sets the reference accepts (InspireParams::validate, src/params.rs:535-640) that
the gfx950 kernels do not run are refused at context creation with the library's message. The
checks come before any device call, so these tests run on the host: a Rust caller routes these
sets to the reference's own CPU code."""
import pytest

import inspire_gpu as g

# the crate's two-prime CRT sets (src/params.rs:36, :1103; tests/packing_key_fixture/mod.rs:29),
# replayed by the packing-seat KAT's "upstream2", "thirty2" and "adaptive2" lines
CRT_SETS = {
    "upstream2": (268_369_921, 249_561_089),
    "thirty2": (1_073_479_681, 1_073_692_673),
    "adaptive2": (67_043_329, 132_120_577),
}


def test_rings_outside_256_to_4096_are_refused():
    for n in (16, 128, 8192, 16384):
        with pytest.raises(g.PirError, match=rf"ring_dim {n} has no gfx950 kernel here; built ring "
                                             r"dimensions are 256, 512, 1024, 2048 and 4096"):
            g.Context(g.params_d2048(ring_dim=n), bytes(32), 16)


@pytest.mark.parametrize("name", sorted(CRT_SETS))
def test_crt_moduli_are_refused(name):
    m0, m1 = CRT_SETS[name]
    p = g.params_d2048(ring_dim=256)
    p.q = m0 * m1
    with pytest.raises(g.PirError, match=rf"modulus {m0 * m1} is not the single 60-bit prime"):
        g.Context(p, bytes(32), 16)


def test_other_single_primes_and_gadgets_are_refused():
    # 2^61 - 1 is prime but not the kernels' 2^60 - 2^14 + 1
    p = g.params_d2048()
    p.q = (1 << 61) - 1
    with pytest.raises(g.PirError, match="is not the single 60-bit prime 2\\^60 - 2\\^14 \\+ 1"):
        g.Context(p, bytes(32), 16)
    for base, length in ((1 << 15, 4), (1 << 20, 4), (1 << 30, 2)):
        p = g.params_d2048()
        p.gadget_base, p.packing_gadget_len = base, length
        with pytest.raises(g.PirError, match=rf"packing gadget \(base {base}, len {length}\) unsupported"):
            g.Context(p, bytes(32), 16)


def test_widths_that_are_not_inspiring_widths_are_refused():
    for n, gamma in ((2048, 24), (2048, 2048), (256, 256), (4096, 2048), (512, 0)):
        with pytest.raises(g.PirError, match="not a legal InspiRING width"):
            g.Context(g.params_d2048(ring_dim=n), bytes(32), gamma)
