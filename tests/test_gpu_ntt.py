"""Device NTTs against the oracle, bit for bit: forward, inverse and the prover's coset
pipeline (iNTT -> x shift^j -> NTT), every size up to 2^14 (one, two and three LDS passes),
and 2^16, 2^18 and the 2^21 standalone
workload through a round trip and the oracle's coset pipeline."""
import numpy as np
import pytest

from snarkrs_amd.wire import SplitMix64, fr_array

pytestmark = pytest.mark.gpu


def _rand(n, seed):
    rng = SplitMix64(seed)
    return fr_array([rng.next_fr() for _ in range(n)])


@pytest.mark.parametrize("log_n", list(range(0, 15)))
def test_small_sizes_match_oracle(dev, orc, log_n):
    a = _rand(1 << log_n, 1000 + log_n)
    assert (dev.ntt(a, 0) == orc.ntt(a)).all(), "forward"
    assert (dev.ntt(a, 1) == orc.ntt(a, inverse=True)).all(), "inverse"
    assert (dev.ntt(a, 2) == orc.coset_pipeline(a)).all(), "coset pipeline"


def _fast_rand(n, seed):
    # Uniform-enough Fr values without a Python loop: 254-bit limbs, top limb below r's.
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 2**32, size=(n, 8), dtype=np.uint64).astype(np.uint32)
    a[:, 7] %= 0x30644E72
    return a.view(np.uint8).reshape(n, 32).copy()


@pytest.mark.parametrize("log_n", [16, 18, 21])
def test_large_sizes(dev, orc, log_n):
    a = _fast_rand(1 << log_n, log_n)
    f = dev.ntt(a, 0)
    assert (dev.ntt(f, 1) == a).all(), "round trip"
    assert (dev.ntt(a, 2) == orc.coset_pipeline(a)).all(), "coset pipeline vs oracle"
    if log_n <= 18:
        assert (f == orc.ntt(a)).all()
