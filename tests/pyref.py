"""Pure-Python BN254 arithmetic with big integers: the closed forms the oracle (and through it
the HIP path) is pinned against. Textbook affine formulas, slow, for small cases only."""
from __future__ import annotations

import numpy as np

from snarkrs_amd.wire import MONT_R, Q_MOD, R_MOD, SplitMix64, int_to_le

# ---------------------------------------------------------------------------------------
# fields
# ---------------------------------------------------------------------------------------


def edge_limbs(p: int) -> list[int]:
    """Stored (Montgomery) limb patterns at the edges of the carry chains, the list of the
    reference's crates/cuda/tests/field_gpu.rs:200-244 (the same list serves Fr and Fq)."""
    R = MONT_R % p
    two_254_mod = (1 << 254) - p
    return [
        0,
        R,  # one
        p - R,  # minus one
        pow(2, 512, p),  # the element R
        1,
        0xFFFFFFFF,
        p - 1,
        p - 2,
        (p - 1) // 2,
        (p + 1) // 2,
        1 << 253,
        (1 << 253) + 1,
        two_254_mod,
        (0x30644E71FFFFFFFF << 192) | ((1 << 192) - 1),
        (1 << 224) - 1,
        int("2aaaaaaaaaaaaaaa" + "aaaaaaaaaaaaaaaa" * 3, 16),
    ]


def limbs_to_value(limbs: int, p: int) -> int:
    return limbs * pow(MONT_R, -1, p) % p


def fq2_mul(a, b):
    (a0, a1), (b0, b1) = a, b
    return ((a0 * b0 - a1 * b1) % Q_MOD, (a0 * b1 + a1 * b0) % Q_MOD)


def fq2_add(a, b):
    return ((a[0] + b[0]) % Q_MOD, (a[1] + b[1]) % Q_MOD)


def fq2_sub(a, b):
    return ((a[0] - b[0]) % Q_MOD, (a[1] - b[1]) % Q_MOD)


def fq2_inv(a):
    t = pow(a[0] * a[0] + a[1] * a[1], -1, Q_MOD)
    return (a[0] * t % Q_MOD, (-a[1]) * t % Q_MOD)


# ---------------------------------------------------------------------------------------
# curves (affine, None = infinity)
# ---------------------------------------------------------------------------------------
G1_GEN = (1, 2)
G2_GEN = (
    (10857046999023057135944570762232829481370756359578518086990519993285655852781,
     11559732032986387107991004021392285783925812861821192530917403151452391805634),
    (8495653923123431417604973247489272438418190587263600148770280649306958101930,
     4082367875863433681332203403145435568316851327593401208105741076214120093531),
)
B1 = 3
B2 = fq2_mul((3, 0), fq2_inv((9, 1)))


class _F1:
    @staticmethod
    def add(a, b): return (a + b) % Q_MOD
    @staticmethod
    def sub(a, b): return (a - b) % Q_MOD
    @staticmethod
    def mul(a, b): return a * b % Q_MOD
    @staticmethod
    def inv(a): return pow(a, -1, Q_MOD)
    @staticmethod
    def small(k): return k % Q_MOD
    zero = 0


class _F2:
    add = staticmethod(fq2_add)
    sub = staticmethod(fq2_sub)
    mul = staticmethod(fq2_mul)
    inv = staticmethod(fq2_inv)
    @staticmethod
    def small(k): return (k % Q_MOD, 0)
    zero = (0, 0)


def _field(p):
    return _F2 if isinstance(p[0], tuple) else _F1


def on_curve(p) -> bool:
    if p is None:
        return True
    F = _field(p)
    x, y = p
    b = B2 if F is _F2 else B1
    return F.mul(y, y) == F.add(F.mul(F.mul(x, x), x), b)


def neg(p):
    if p is None:
        return None
    F = _field(p)
    return (p[0], F.sub(F.zero, p[1]))


def add(p, q):
    if p is None:
        return q
    if q is None:
        return p
    F = _field(p)
    if p[0] == q[0]:
        if p[1] == q[1] and p[1] != F.zero:
            lam = F.mul(F.mul(F.small(3), F.mul(p[0], p[0])), F.inv(F.mul(F.small(2), p[1])))
        else:
            return None
    else:
        lam = F.mul(F.sub(q[1], p[1]), F.inv(F.sub(q[0], p[0])))
    x3 = F.sub(F.sub(F.mul(lam, lam), p[0]), q[0])
    y3 = F.sub(F.mul(lam, F.sub(p[0], x3)), p[1])
    return (x3, y3)


def mul(p, k: int):
    k %= R_MOD
    acc = None
    while k:
        if k & 1:
            acc = add(acc, p)
        p = add(p, p)
        k >>= 1
    return acc


def colliding_bases(n: int, gen):
    """Seven points cycled, every fifth negated and every eleventh at infinity
    (crates/cuda/tests/adversarial_gpu.rs:220-247)."""
    pts = [mul(gen, k) for k in range(1, 8)]
    out = []
    for i in range(n):
        if i % 11 == 4:
            out.append(None)
        elif i % 5 == 2:
            out.append(neg(pts[i % 7]))
        else:
            out.append(pts[i % 7])
    return out


def adversarial_scalars(n: int) -> dict[str, list[int]]:
    """crates/cuda/tests/adversarial_gpu.rs:249-281, plus top-of-field values (:829)."""
    a = (-0x7F3A19C5D02B6E41) % R_MOD
    b = 0x123456789ABCDEF1 * 0x0FEDCBA987654321 % R_MOD
    return {
        "one value": [a] * n,
        "a value and its negation": [a if i % 2 == 0 else R_MOD - a for i in range(n)],
        "two values": [a if i % 3 == 0 else b for i in range(n)],
        "single bits": [pow(2, i % 254, R_MOD) for i in range(n)],
        "small": [i % 9 for i in range(n)],
        "zeros": [0] * n,
        "ones": [1] * n,
        "top of field": [R_MOD - 1 - (i % 5) for i in range(n)],
    }


def random_fr(rng: SplitMix64, n: int) -> list[int]:
    return [rng.next_fr() for _ in range(n)]


def mixed_scalars(rng: SplitMix64, n: int, zero_one_permille: int = 300) -> list[int]:
    """The standalone-MSM workload's scalars: uniform Fr, except a share in {0, 1}."""
    out = []
    for _ in range(n):
        if rng.next_u64() % 1000 < zero_one_permille:
            out.append(rng.next_u64() & 1)
        else:
            out.append(rng.next_fr())
    return out


def g2_array(points) -> np.ndarray:
    from snarkrs_amd.wire import to_mont

    out = np.zeros((len(points), 128), dtype=np.uint8)
    for i, p in enumerate(points):
        if p is None:
            continue
        (x0, x1), (y0, y1) = p
        for j, v in enumerate((x0, x1, y0, y1)):
            out[i, 32 * j:32 * (j + 1)] = np.frombuffer(int_to_le(to_mont(v, Q_MOD)), dtype=np.uint8)
    return out
