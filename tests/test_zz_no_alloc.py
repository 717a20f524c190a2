"""g16hip_msm_device_noalloc (every piece of its scratch in the library's static device array)
against g16hip_msm_g1_device / g16hip_msm_g2_device through the C ABI, bit for bit: two calls
in a row, the second smaller (the array keeps the first call's contents), G1 and G2, a fixed
width, the largest sizes the array holds at the library's width (2^17 G1, 2^16 G2), and the
refusal of an MSM whose scratch exceeds the array. Named to be collected last."""
import numpy as np
import pytest

from snarkrs_amd import native
from snarkrs_amd.devmem import DeviceArray
from snarkrs_amd.synth import scalars

pytestmark = pytest.mark.gpu


def _inputs(dev, group, n, seed):
    bases = dev.fixed_base(group, scalars(seed, n, 0))
    return DeviceArray(bases), DeviceArray(scalars(seed + 1, n, 300))


@pytest.mark.parametrize("group,sizes", [(1, [1 << 16, 5000]), (2, [1 << 15, 777])])
def test_noalloc_msm_matches_the_allocating_entry_point(dev, group, sizes):
    for k, n in enumerate(sizes):
        d_b, d_s = _inputs(dev, group, n, 900 + 10 * group + k)
        want = dev.msm_device(group, d_b.addr, d_s.addr, n)
        assert (dev.msm_device_noalloc(group, d_b.addr, d_s.addr, n) == want).all(), (group, n)
        assert (dev.msm_device_noalloc(group, d_b.addr, d_s.addr, n, 9) ==
                dev.msm_device(group, d_b.addr, d_s.addr, n, 9)).all(), (group, n)
        del d_b, d_s


@pytest.mark.parametrize("group,n", [(1, 1 << 17), (2, 1 << 16)])
def test_noalloc_largest_sizes_the_array_holds(dev, group, n):
    d_b, d_s = _inputs(dev, group, n, 970 + group)
    assert (dev.msm_device_noalloc(group, d_b.addr, d_s.addr, n) ==
            dev.msm_device(group, d_b.addr, d_s.addr, n)).all()


def test_noalloc_refuses_what_the_array_cannot_hold(dev):
    n = 1 << 17
    d_s = DeviceArray(scalars(991, n, 300))
    d_b = DeviceArray(np.zeros((n, 128), np.uint8))
    with pytest.raises(native.DeviceMemory, match="static scratch"):
        dev.msm_device_noalloc(2, d_b.addr, d_s.addr, n)
    assert not dev.msm_device_noalloc(1, d_b.addr, d_s.addr, 0).any()
