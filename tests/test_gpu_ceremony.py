"""The ceremony's MSM seam (crates/cuda/src/ceremony.rs:65-312) on the device: runs of at most
max_run points summed on the host give the one-call MSM and the oracle's, bit for bit, and
the closed form (sum k_i s_i) G; a run over the memory limit is halved (down to 2^10
points) before anything is allocated, and one that cannot fit at all is refused."""
import numpy as np
import pytest

from snarkrs_amd import native
from snarkrs_amd.native import CeremonyMsm
from snarkrs_amd.wire import R_MOD, SplitMix64, fr_array

pytestmark = pytest.mark.gpu


def _inputs(dev, group, n, seed):
    rng = SplitMix64(seed)
    ks = [rng.next_fr() for _ in range(n)]
    ss = [rng.next_fr() if i % 3 else i % 2 for i in range(n)]
    return ks, ss, dev.fixed_base(group, fr_array(ks)), fr_array(ss)


@pytest.mark.parametrize("group", [1, 2])
def test_runs_sum_to_the_whole_msm(dev, orc, group):
    n = 5000
    ks, ss, bases, sc = _inputs(dev, group, n, 900 + group)
    cer = CeremonyMsm(dev, max_run=1024)
    got = cer._msm(group, bases, sc)
    assert cer.last_runs() == 5
    assert (got == dev.msm(group, bases, sc)).all()
    assert (got == orc.msm(group, bases, sc)).all()
    k = sum(a * b for a, b in zip(ks, ss)) % R_MOD
    assert (got == dev.fixed_base(group, fr_array([k]))[0]).all()
    st = cer.stats()
    assert st["calls"] == 1 and st["runs"] == 5 and st["points"] == n and st["retries"] == 0


def test_a_run_over_the_memory_limit_is_halved(dev, orc):
    n = 8192
    ks, ss, bases, sc = _inputs(dev, 1, n, 77)
    limit = CeremonyMsm.scratch_bytes(1, 4096)
    assert CeremonyMsm.scratch_bytes(1, 8192) > limit
    cer = CeremonyMsm(dev, max_run=8192, vram_limit=limit)
    got = cer.msm_g1(bases, sc)
    assert cer.last_runs() == 2
    assert (got == orc.msm(1, bases, sc)).all()
    tiny = CeremonyMsm(dev, max_run=8192, vram_limit=CeremonyMsm.scratch_bytes(1, 1024) - 1)
    with pytest.raises(native.DeviceMemory):
        tiny.msm_g1(bases, sc)


def test_empty_and_mismatched_calls(dev):
    cer = CeremonyMsm(dev)
    assert not cer.msm_g2(np.zeros((0, 128), np.uint8), np.zeros((0, 32), np.uint8)).any()
    assert cer.last_runs() == 0
    with pytest.raises(native.BackendError):
        cer.msm_g1(np.zeros((3, 64), np.uint8), np.zeros((2, 32), np.uint8))
