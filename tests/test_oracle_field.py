"""Pins the oracle's field arithmetic, domain and digit recoding (CPU only).

* Fr/Fq/Fq2 ops against Python big integers on the reference's edge operands
  (crates/cuda/tests/field_gpu.rs:200-244) and seeded random pairs;
* the ffjavascript domain roots (crates/groth16/src/prove.rs:549-563);
* the coset-shift identities (crates/groth16/src/cpu.rs:915-927);
* window_size KATs (crates/msm/src/lib.rs:1969-1970) and signed-digit reconstruction
  (crates/msm/src/lib.rs:1974-2062).
"""
import numpy as np
import pytest

from snarkrs_amd.wire import MONT_R, Q_MOD, R_MOD, SplitMix64, int_to_le, le_to_int
from tests.pyref import edge_limbs, limbs_to_value

MODS = {0: R_MOD, 1: Q_MOD}


def _limb_array(limbs):
    return np.stack([np.frombuffer(int_to_le(x), dtype=np.uint8) for x in limbs])


def _vals(arr, p):
    return [limbs_to_value(le_to_int(row.tobytes()), p) for row in arr.reshape(-1, 32)]


def _operands(p, n_random=200, seed=7):
    e = edge_limbs(p)
    a = [x for x in e for _ in e]
    b = [y for _ in e for y in e]
    rng = SplitMix64(seed)
    for _ in range(n_random):
        a.append(rng.next_fr() % p * 3 % p)
        b.append(rng.next_u64() * rng.next_u64() % p)
    return a, b


@pytest.mark.parametrize("field", [0, 1])
def test_base_field_ops_match_big_integers(orc, field):
    p = MODS[field]
    a_l, b_l = _operands(p)
    A, B = _limb_array(a_l), _limb_array(b_l)
    av, bv = _vals(A, p), _vals(B, p)
    checks = {
        0: lambda x, y: (x + y) % p,
        1: lambda x, y: (x - y) % p,
        2: lambda x, y: x * y % p,
        3: lambda x, y: x * x % p,
        4: lambda x, y: pow(x, -1, p) if x else 0,
        5: lambda x, y: (-x) % p,
    }
    for op, f in checks.items():
        got = _vals(orc.field_op(field, op, A, B), p)
        want = [f(x, y) for x, y in zip(av, bv)]
        assert got == want, f"field {field} op {op}"
        # Outputs are fully reduced limbs.
        raw = [le_to_int(r.tobytes()) for r in orc.field_op(field, op, A, B)]
        assert all(x < p for x in raw)


@pytest.mark.parametrize("field", [0, 1])
def test_montgomery_conversion_round_trip(orc, field):
    p = MODS[field]
    rng = SplitMix64(11)
    xs = [0, 1, 2, p - 1] + [rng.next_fr() % p for _ in range(64)]
    std = _limb_array(xs)
    mont = orc.field_op(field, 6, std)
    assert [le_to_int(r.tobytes()) for r in mont] == [x * MONT_R % p for x in xs]
    back = orc.field_op(field, 7, mont)
    assert [le_to_int(r.tobytes()) for r in back] == xs


def test_fq2_ops_match_big_integers(orc):
    from tests.pyref import fq2_add, fq2_inv, fq2_mul, fq2_sub

    e = edge_limbs(Q_MOD)
    rng = SplitMix64(5)
    pairs = [((x, y), (y, x)) for x in e for y in e[:6]]
    pairs += [((rng.next_fr() % Q_MOD, rng.next_u64()), (rng.next_u64(), rng.next_fr() % Q_MOD))
              for _ in range(100)]

    def arr(vals):
        return np.stack([np.concatenate([np.frombuffer(int_to_le(c), dtype=np.uint8) for c in v])
                         for v in vals])

    A = arr([p[0] for p in pairs])
    B = arr([p[1] for p in pairs])

    def val(a):
        return [(limbs_to_value(le_to_int(r[:32].tobytes()), Q_MOD),
                 limbs_to_value(le_to_int(r[32:].tobytes()), Q_MOD)) for r in a.reshape(-1, 64)]

    av, bv = val(A), val(B)
    assert val(orc.field_op(2, 0, A, B)) == [fq2_add(x, y) for x, y in zip(av, bv)]
    assert val(orc.field_op(2, 1, A, B)) == [fq2_sub(x, y) for x, y in zip(av, bv)]
    assert val(orc.field_op(2, 2, A, B)) == [fq2_mul(x, y) for x, y in zip(av, bv)]
    assert val(orc.field_op(2, 3, A)) == [fq2_mul(x, x) for x in av]
    assert val(orc.field_op(2, 4, A)) == [fq2_inv(x) if x != (0, 0) else (0, 0) for x in av]


def test_domain_roots_match_ffjavascript(orc):
    # crates/groth16/src/prove.rs:549-563: values printed from ffjavascript's bn128 Fr.w.
    def root(log_n):
        return limbs_to_value(le_to_int(orc.domain_root(log_n).tobytes()), R_MOD)

    assert root(28) == 19103219067921713944291392827692070036145651957329286315305642004821462161904
    assert root(4) == 14940766826517323942636479241147756311199852622225275649687664389641784935947
    # A primitive root of exactly the stated order, from 5^((r-1)/2^28) (arkworks' generator).
    assert root(28) == pow(5, (R_MOD - 1) >> 28, R_MOD)
    for log_n in range(0, 29):
        w = root(log_n)
        assert pow(w, 1 << log_n, R_MOD) == 1
        if log_n:
            assert pow(w, 1 << (log_n - 1), R_MOD) == R_MOD - 1


def test_coset_shift_is_the_odd_half_of_the_2n_th_roots(orc):
    # crates/groth16/src/cpu.rs:915-927.
    for log_n in range(0, 12):
        n = 1 << log_n
        shift = limbs_to_value(le_to_int(orc.domain_root(log_n + 1).tobytes()), R_MOD)
        w = limbs_to_value(le_to_int(orc.domain_root(log_n).tobytes()), R_MOD)
        assert shift * shift % R_MOD == w
        assert pow(shift, n, R_MOD) == R_MOD - 1


def test_window_size_kats(orc):
    # crates/msm/src/lib.rs:1969-1970.
    assert orc.window_size(1 << 16) == 13
    assert orc.window_size(1 << 18) == 15
    # The model itself, recomputed: argmin over c of ceil(255/c) * (n + 3 * 2^(c-1)).
    for n in [1, 10, 1000, 734003, 1 << 20, 1 << 21, 1 << 24]:
        best = min(range(3, 17), key=lambda c: (-(-255 // c) * (n + 3 * (1 << (c - 1))), c))
        assert orc.window_size(n) == best, n


def test_signed_digits_reconstruct_the_scalar(orc):
    rng = SplitMix64(3)
    scalars = [0, 1, 2, R_MOD - 1, R_MOD - 2, (1 << 253), (1 << 254) - 1 - (1 << 200)]
    scalars += [rng.next_fr() for _ in range(20)]
    for c in range(3, 17):
        W = -(-255 // c)
        for k in scalars:
            digits = [orc.signed_digit(k, i, c) for i in range(W)]
            assert all(-(1 << (c - 1)) <= d <= (1 << (c - 1)) for d in digits)
            assert sum(d << (c * i) for i, d in enumerate(digits)) == k, (c, k)
