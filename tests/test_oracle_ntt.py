"""Pins the oracle's NTT (CPU only) with the closed-form checks of the reference's NTT tests
(crates/ntt/src/lib.rs:1244-1380): the naive DFT, impulses, round trips, cyclic convolution,
and the coset pipeline as polynomial evaluation on the odd coset."""
import numpy as np
import pytest

from snarkrs_amd.wire import R_MOD, SplitMix64, fr_array, fr_ints


def _root(log_n):
    return pow(pow(5, (R_MOD - 1) >> 28, R_MOD), 1 << (28 - log_n), R_MOD)


def _dft(x, w):
    n = len(x)
    return [sum(x[j] * pow(w, j * k, R_MOD) for j in range(n)) % R_MOD for k in range(n)]


@pytest.mark.parametrize("log_n", range(0, 7))
def test_matches_the_naive_dft(orc, log_n):
    n = 1 << log_n
    rng = SplitMix64(log_n)
    x = [rng.next_fr() for _ in range(n)]
    w = _root(log_n)
    assert fr_ints(orc.ntt(fr_array(x))) == _dft(x, w)
    inv = [v * pow(n, -1, R_MOD) % R_MOD for v in _dft(x, pow(w, -1, R_MOD))]
    assert fr_ints(orc.ntt(fr_array(x), inverse=True)) == inv


def test_impulses_above_small_sizes(orc):
    log_n = 12
    n = 1 << log_n
    w = _root(log_n)
    for j in (0, 1, 5, n - 1):
        x = [0] * n
        x[j] = 1
        got = fr_ints(orc.ntt(fr_array(x)))
        wj = pow(w, j, R_MOD)
        assert got[:50] == [pow(wj, k, R_MOD) for k in range(50)]
        assert got[-3:] == [pow(wj, k, R_MOD) for k in range(n - 3, n)]


def test_round_trip(orc):
    rng = SplitMix64(77)
    for log_n in (8, 12, 14):
        a = fr_array([rng.next_fr() for _ in range(1 << log_n)])
        assert (orc.ntt(orc.ntt(a), inverse=True) == a).all()


def test_pointwise_product_is_cyclic_convolution(orc):
    rng = SplitMix64(8)
    n = 32
    a = [rng.next_fr() for _ in range(n)]
    b = [rng.next_fr() for _ in range(n)]
    fa, fb = fr_ints(orc.ntt(fr_array(a))), fr_ints(orc.ntt(fr_array(b)))
    prod = fr_array([x * y % R_MOD for x, y in zip(fa, fb)])
    got = fr_ints(orc.ntt(prod, inverse=True))
    want = [sum(a[i] * b[(k - i) % n] for i in range(n)) % R_MOD for k in range(n)]
    assert got == want


@pytest.mark.parametrize("log_n", [0, 1, 3, 6])
def test_coset_pipeline_evaluates_on_the_odd_coset(orc, log_n):
    # Evaluations on the domain -> the same polynomial evaluated at shift * w^k, with
    # shift the primitive 2n-th root (crates/field/src/lib.rs:124-130).
    n = 1 << log_n
    rng = SplitMix64(31 + log_n)
    ev = [rng.next_fr() for _ in range(n)]
    w = _root(log_n)
    coeffs = [v * pow(n, -1, R_MOD) % R_MOD for v in _dft(ev, pow(w, -1, R_MOD))]
    shift = _root(log_n + 1)
    want = [sum(c * pow(shift * pow(w, k, R_MOD) % R_MOD, j, R_MOD) for j, c in enumerate(coeffs))
            % R_MOD for k in range(n)]
    assert fr_ints(orc.coset_pipeline(fr_array(ev))) == want
