"""Pins the oracle's curve arithmetic (CPU only): generators on their curves, the G2 twist
constant 3/(9+u), r*G = infinity, Jacobian formulas against textbook affine addition, and the
fixed-base setup multiplication against double-and-add."""
import numpy as np

from snarkrs_amd.wire import R_MOD, SplitMix64, fr_array, g1_array, g1_ints, g2_ints
from tests import pyref


def test_generators_are_on_their_curves(orc):
    g1 = g1_ints(orc.generator(1))[0]
    g2 = g2_ints(orc.generator(2))[0]
    assert g1 == pyref.G1_GEN
    assert g2 == pyref.G2_GEN
    assert pyref.on_curve(g1) and pyref.on_curve(g2)
    assert orc.on_curve(1, orc.generator(1)) and orc.on_curve(2, orc.generator(2))
    # An off-curve point is refused (the oracle's own b constants are the ones used).
    bad = g1_array([(1, 3)])
    assert not orc.on_curve(1, bad)


def test_group_order_is_r(orc):
    for group, gen in ((1, pyref.G1_GEN), (2, pyref.G2_GEN)):
        g = orc.generator(group)
        # (r-1) G = -G, so r G = infinity.
        got = orc.scalar_mul(group, g, fr_array([R_MOD - 1]))
        want = pyref.neg(gen)
        ints = g1_ints(got) if group == 1 else g2_ints(got)
        assert ints[0] == want
        s = orc.point_add(group, got, g)
        assert not s.any()  # -G + G = infinity


def test_scalar_mul_and_add_match_affine_textbook(orc):
    rng = SplitMix64(21)
    for group, gen in ((1, pyref.G1_GEN), (2, pyref.G2_GEN)):
        ks = [1, 2, 3, 5, rng.next_fr(), rng.next_fr()]
        pts = np.stack([orc.generator(group)] * len(ks))
        got = orc.scalar_mul(group, pts, fr_array(ks))
        ints = g1_ints(got) if group == 1 else g2_ints(got)
        for k, p in zip(ks, ints):
            assert p == pyref.mul(gen, k)
            assert pyref.on_curve(p)
        # P + P (the doubling branch), P + Q, P + (-P).
        a, b = got[4:5], got[5:6]
        conv = g1_ints if group == 1 else g2_ints
        assert conv(orc.point_add(group, a, a))[0] == pyref.mul(gen, 2 * ks[4])
        assert conv(orc.point_add(group, a, b))[0] == pyref.mul(gen, ks[4] + ks[5])
        neg = orc.scalar_mul(group, a, fr_array([R_MOD - 1]))
        assert not orc.point_add(group, a, neg).any()
        inf = np.zeros_like(a)
        assert (orc.point_add(group, a, inf) == a).all()
        assert (orc.point_add(group, inf, a) == a).all()


def test_fixed_base_matches_double_and_add(orc):
    rng = SplitMix64(4)
    ks = [0, 1, 2, 255, 256, R_MOD - 1] + [rng.next_fr() for _ in range(40)]
    k = fr_array(ks)
    for group in (1, 2):
        g = orc.generator(group)
        fb = orc.fixed_base(group, k)
        dd = orc.scalar_mul(group, np.stack([g] * len(ks)), k)
        assert (fb == dd).all()
