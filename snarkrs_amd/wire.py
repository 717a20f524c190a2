"""The packed wire format that crosses the C ABI, and Python-int conversions for it.

Mirrors crates/gpu-layout/src/lib.rs:75-185 of the reference:

* a field element is 8 little-endian u32 limbs (32 bytes); ``PackedFr``/``PackedFq`` hold
  the **Montgomery** representative ``x * 2^256 mod p`` (lib.rs:30-47), a ``PackedScalar``
  the plain integer;
* ``Fq2`` is ``c0`` then ``c1`` (64 bytes);
* a G1 affine point is ``x`` then ``y`` (64 bytes), G2 ``x`` then ``y`` as Fq2 (128 bytes);
  the all-zero encoding is the point at infinity (lib.rs:136-167).

Arrays of elements are numpy ``uint8`` arrays of shape ``(n, 32)`` / ``(n, 64)`` /
``(n, 128)``; anything C-contiguous with those byte strides is accepted by the bindings.
"""
from __future__ import annotations

import numpy as np

R_MOD = 21888242871839275222246405745257275088548364400416034343698204186575808495617
Q_MOD = 21888242871839275222246405745257275088696311157297823662689037894645226208583
MONT_R = 1 << 256

FR_BYTES = 32
G1_BYTES = 64
G2_BYTES = 128


def int_to_le(x: int, nbytes: int = 32) -> bytes:
    return int(x).to_bytes(nbytes, "little")


def le_to_int(b) -> int:
    return int.from_bytes(bytes(b), "little")


def to_mont(x: int, p: int = R_MOD) -> int:
    return (x % p) * MONT_R % p


def from_mont(x: int, p: int = R_MOD) -> int:
    return x * pow(MONT_R, -1, p) % p


def fr_array(values, mont: bool = True) -> np.ndarray:
    """Python ints -> (n, 32) uint8, Montgomery (``PackedFr``) or plain (``PackedScalar``)."""
    out = np.zeros((len(values), 32), dtype=np.uint8)
    for i, v in enumerate(values):
        v = int(v) % R_MOD
        out[i] = np.frombuffer(int_to_le(to_mont(v) if mont else v), dtype=np.uint8)
    return out


def fr_ints(arr: np.ndarray, mont: bool = True) -> list[int]:
    arr = np.ascontiguousarray(arr, dtype=np.uint8).reshape(-1, 32)
    vals = [le_to_int(row.tobytes()) for row in arr]
    return [from_mont(v) for v in vals] if mont else vals


def fq_ints(arr: np.ndarray, mont: bool = True) -> list[int]:
    arr = np.ascontiguousarray(arr, dtype=np.uint8).reshape(-1, 32)
    vals = [le_to_int(row.tobytes()) for row in arr]
    return [from_mont(v, Q_MOD) for v in vals] if mont else vals


def g1_ints(arr: np.ndarray):
    """(n, 64) packed affine G1 -> list of (x, y) ints, ``None`` for infinity."""
    arr = np.ascontiguousarray(arr, dtype=np.uint8).reshape(-1, 64)
    out = []
    for row in arr:
        if not row.any():
            out.append(None)
            continue
        x, y = fq_ints(row.reshape(2, 32))
        out.append((x, y))
    return out


def g2_ints(arr: np.ndarray):
    """(n, 128) packed affine G2 -> list of ((x0, x1), (y0, y1)) ints, ``None`` for infinity."""
    arr = np.ascontiguousarray(arr, dtype=np.uint8).reshape(-1, 128)
    out = []
    for row in arr:
        if not row.any():
            out.append(None)
            continue
        x0, x1, y0, y1 = fq_ints(row.reshape(4, 32))
        out.append(((x0, x1), (y0, y1)))
    return out


def g1_array(points) -> np.ndarray:
    out = np.zeros((len(points), 64), dtype=np.uint8)
    for i, p in enumerate(points):
        if p is None:
            continue
        out[i, :32] = np.frombuffer(int_to_le(to_mont(p[0], Q_MOD)), dtype=np.uint8)
        out[i, 32:] = np.frombuffer(int_to_le(to_mont(p[1], Q_MOD)), dtype=np.uint8)
    return out


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014), the seeded generator behind every synthetic
    input; the reference's test RNG (crates/gpu-layout/src/lib.rs:702-727)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.MASK

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def next_fr(self) -> int:
        """Uniform in [0, r): 256 random bits, little-endian, reduced mod r."""
        x = 0
        for i in range(4):
            x |= self.next_u64() << (64 * i)
        return x % R_MOD
