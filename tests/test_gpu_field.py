"""Device field arithmetic (bn254.cuh) against the oracle, bit for bit, on the reference's
edge operands (crates/cuda/tests/field_gpu.rs:200-252) plus seeded random pairs, through the
C ABI's field probe; and the device fixed-base multiplication against the oracle's."""
import numpy as np
import pytest

from snarkrs_amd.wire import Q_MOD, R_MOD, SplitMix64, fr_array, int_to_le
from tests.pyref import edge_limbs

pytestmark = pytest.mark.gpu

MODS = {0: R_MOD, 1: Q_MOD}


def _limbs(xs):
    return np.stack([np.frombuffer(int_to_le(x), dtype=np.uint8) for x in xs])


def _pairs(p, seed, n_random=4000):
    e = edge_limbs(p)
    a = [x for x in e for _ in e]
    b = [y for _ in e for y in e]
    rng = SplitMix64(seed)
    for _ in range(n_random):
        a.append(rng.next_fr() % p)
        b.append((rng.next_fr() * 7 + rng.next_u64()) % p)
    return _limbs(a), _limbs(b)


@pytest.mark.parametrize("field", [0, 1])
@pytest.mark.parametrize("op", range(8))
def test_fp_ops_match_oracle(dev, orc, field, op):
    A, B = _pairs(MODS[field], seed=10 + field)
    if op == 4:  # inversion is slow on one thread per element; edges + a few hundred
        A, B = A[:600], B[:600]
    got = dev.field_probe(field, op, A, B)
    want = orc.field_op(field, op, A, B)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, f"field {field} op {op}: first mismatch at {bad[:5]}"


@pytest.mark.parametrize("op", range(8))
def test_fq2_ops_match_oracle(dev, orc, op):
    A, B = _pairs(Q_MOD, seed=3)
    A2 = np.concatenate([A[: len(A) // 2 * 2].reshape(-1, 64)], axis=0)
    B2 = np.concatenate([B[: len(B) // 2 * 2].reshape(-1, 64)], axis=0)
    if op == 4:
        A2, B2 = A2[:300], B2[:300]
    got = dev.field_probe(2, op, A2, B2)
    want = orc.field_op(2, op, A2, B2)
    assert (got == want).all(), f"fq2 op {op}"


@pytest.mark.parametrize("op", range(8))
def test_msm_form_fq_ops_match_python(dev, op):
    # Field 3: Fq in Montgomery form with R' = 2^261 (the MSM kernels' nine-limb product).
    # Closed forms with Python integers: a value x is stored as x R' mod q.
    A, B = _pairs(Q_MOD, seed=33)
    if op == 4:
        A, B = A[:600], B[:600]
    got = dev.field_probe(3, op, A, B)
    a = [int.from_bytes(bytes(x), "little") for x in A]
    b = [int.from_bytes(bytes(x), "little") for x in B]
    q, Rp = Q_MOD, 1 << 261
    Ri = pow(Rp, -1, q)
    f = {0: lambda x, y: (x + y) % q,
         1: lambda x, y: (x - y) % q,
         2: lambda x, y: x * y * Ri % q,
         3: lambda x, y: x * x * Ri % q,
         4: lambda x, y: (Rp * Rp * pow(x, -1, q)) % q if x else 0,
         5: lambda x, y: (-x) % q,
         6: lambda x, y: x * Rp % q,
         7: lambda x, y: x * Ri % q}[op]
    want = _limbs([f(x, y) for x, y in zip(a, b)])
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, f"op {op}: first mismatch at {bad[:5]}"


@pytest.mark.parametrize("group", [1, 2])
def test_fixed_base_matches_oracle(dev, orc, group):
    rng = SplitMix64(40 + group)
    ks = [0, 1, 2, R_MOD - 1] + [rng.next_fr() for _ in range(2000)]
    k = fr_array(ks)
    assert (dev.fixed_base(group, k) == orc.fixed_base(group, k)).all()
