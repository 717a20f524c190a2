"""Pins the oracle's Pippenger (CPU only) against the naive sum, the closed form
(sum k_i s_i) * G for bases k_i * G, every window width, and the adversarial scalar and base
classes of crates/cuda/tests/adversarial_gpu.rs:220-281."""
import numpy as np
import pytest

from snarkrs_amd.wire import R_MOD, SplitMix64, fr_array, g1_array
from tests import pyref


def _bases(orc, group, ks):
    return orc.fixed_base(group, fr_array(ks))


@pytest.mark.parametrize("group", [1, 2])
@pytest.mark.parametrize("n", [0, 1, 2, 7, 100, 1000])
def test_pippenger_matches_naive_and_closed_form(orc, group, n):
    rng = SplitMix64(100 + n)
    ks = [rng.next_fr() for _ in range(n)]
    ss = pyref.mixed_scalars(rng, n)
    bases = _bases(orc, group, ks) if n else np.zeros((0, 64 if group == 1 else 128), np.uint8)
    sc = fr_array(ss) if n else np.zeros((0, 32), np.uint8)
    pip = orc.msm(group, bases, sc)
    naive = orc.msm(group, bases, sc, naive=True)
    closed = orc.fixed_base(group, fr_array([sum(k * s for k, s in zip(ks, ss)) % R_MOD]))[0]
    assert (pip == naive).all()
    assert (pip == closed).all()


def test_every_window_width_matches_naive(orc):
    rng = SplitMix64(9)
    n = 300
    ks = [rng.next_fr() for _ in range(n)]
    ss = [rng.next_fr() for _ in range(n)]
    bases = _bases(orc, 1, ks)
    sc = fr_array(ss)
    want = orc.msm(1, bases, sc, naive=True)
    for c in range(3, 17):
        assert (orc.msm(1, bases, sc, c=c) == want).all(), c


@pytest.mark.parametrize("group", [1, 2])
def test_adversarial_scalars_and_colliding_bases(orc, group):
    n = 257
    gen = pyref.G1_GEN if group == 1 else pyref.G2_GEN
    pts = pyref.colliding_bases(n, gen)
    bases = g1_array(pts) if group == 1 else pyref.g2_array(pts)
    for name, ss in pyref.adversarial_scalars(n).items():
        sc = fr_array(ss)
        got = orc.msm(group, bases, sc)
        # Closed form through the textbook affine arithmetic: sum s_i P_i.
        want_pt = None
        for s, p in zip(ss, pts):
            want_pt = pyref.add(want_pt, pyref.mul(p, s)) if p is not None else want_pt
        want = g1_array([want_pt]) if group == 1 else pyref.g2_array([want_pt])
        assert (got == want[0]).all(), name
