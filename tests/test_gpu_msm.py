"""Device MSMs (G1 and G2) against the oracle, bit for bit: random mixed scalars at several
sizes, every window width, the adversarial scalar classes over colliding / infinity bases
(crates/cuda/tests/adversarial_gpu.rs:220-281), repeated calls (dirty scratch), and the
closed form (sum k_i s_i) * G at the 2^20 standalone size."""
import numpy as np
import pytest

from snarkrs_amd import native
from snarkrs_amd.wire import R_MOD, SplitMix64, fr_array, g1_array
from tests import pyref

pytestmark = pytest.mark.gpu


def _bases(dev, group, ks):
    return dev.fixed_base(group, fr_array(ks))


@pytest.mark.parametrize("group", [1, 2])
@pytest.mark.parametrize("n", [1, 2, 7, 100, 1000, 1 << 14])
def test_random_matches_oracle(dev, orc, group, n):
    rng = SplitMix64(500 + n + group)
    ks = [rng.next_fr() for _ in range(n)]
    ss = pyref.mixed_scalars(rng, n)
    bases = _bases(dev, group, ks)
    sc = fr_array(ss)
    assert (dev.msm(group, bases, sc) == orc.msm(group, bases, sc)).all()


def test_empty_is_infinity(dev):
    out = dev.msm(1, np.zeros((0, 64), np.uint8), np.zeros((0, 32), np.uint8))
    assert not out.any()


def test_length_mismatch_is_refused(dev):
    from snarkrs_amd.native import BackendError

    with pytest.raises(BackendError):
        dev.msm(1, np.zeros((3, 64), np.uint8), np.zeros((2, 32), np.uint8))


@pytest.mark.parametrize("group", [1, 2])
def test_adversarial_classes(dev, orc, group):
    n = 4099
    gen = pyref.G1_GEN if group == 1 else pyref.G2_GEN
    pts = pyref.colliding_bases(n, gen)
    bases = g1_array(pts) if group == 1 else pyref.g2_array(pts)
    for name, ss in pyref.adversarial_scalars(n).items():
        sc = fr_array(ss)
        got = dev.msm(group, bases, sc)
        want = orc.msm(group, bases, sc)
        assert (got == want).all(), name
        # Same call again: scratch left dirty by the previous class must not leak.
        assert (dev.msm(group, bases, sc) == want).all(), name + " (repeat)"


def test_every_window_width(dev, orc):
    from snarkrs_amd.devmem import DeviceArray

    rng = SplitMix64(77)
    n = 3000
    ks = [rng.next_fr() for _ in range(n)]
    ss = pyref.mixed_scalars(rng, n)
    bases = _bases(dev, 1, ks)
    sc = fr_array(ss)
    want = orc.msm(1, bases, sc)
    db, ds = DeviceArray(bases), DeviceArray(sc)
    for c in range(2, 17):
        got = dev.msm_device(1, db.addr, ds.addr, n, c)
        assert (got == want).all(), c
    with pytest.raises(native.BackendError):
        dev.msm_device(1, db.addr, ds.addr, n, 17)


@pytest.mark.parametrize("group,n", [(1, 32768), (1, 32769), (1, 100_001), (2, 40_000)])
def test_sort_chunk_boundaries_closed_form(dev, group, n):
    # The counting sort works in chunks of 32768 scalars: exact multiples, one past, several
    # chunks; bases k_i G so the MSM is (sum k_i s_i) G.
    rng = SplitMix64(31 + n)
    ks = [rng.next_fr() for _ in range(n)]
    ss = pyref.mixed_scalars(rng, n)
    bases = _bases(dev, group, ks)
    got = dev.msm(group, bases, fr_array(ss))
    total = sum(k * s for k, s in zip(ks, ss)) % R_MOD
    assert (got == dev.fixed_base(group, fr_array([total]))[0]).all()


def test_standalone_2_20_closed_form(dev, orc):
    # The standalone G1 workload: bases k_i G, 70% uniform scalars, 30% in {0, 1}.
    n = 1 << 20
    rng = SplitMix64(2020)
    ks = [rng.next_fr() for _ in range(n)]
    ss = pyref.mixed_scalars(rng, n)
    bases = _bases(dev, 1, ks)
    got = dev.msm(1, bases, fr_array(ss))
    total = sum(k * s for k, s in zip(ks, ss)) % R_MOD
    assert (got == orc.fixed_base(1, fr_array([total]))[0]).all()


def test_g2_2_16_closed_form(dev, orc):
    n = 1 << 16
    rng = SplitMix64(2218)
    ks = [rng.next_fr() for _ in range(n)]
    ss = pyref.mixed_scalars(rng, n)
    bases = _bases(dev, 2, ks)
    got = dev.msm(2, bases, fr_array(ss))
    total = sum(k * s for k, s in zip(ks, ss)) % R_MOD
    assert (got == orc.fixed_base(2, fr_array([total]))[0]).all()
