"""Decoding of gnark v0.14 BN254 files for the verifier's known-answer test (test code only).

gnark writes encoder points as canonical big-endian coordinates, x then y, G2 coordinates
A1 before A0; the top two bits of the first byte are a flag (00 uncompressed, 01 compressed
infinity, 10 / 11 compressed with the lexicographically smaller / larger y), as
crates/gnark/src/lib.rs:1-17 and :159-300 describe. The verifying key is
``VerifyingKey.WriteTo`` (crates/gnark/src/vk.rs:1-12), the proof ``Proof.WriteRawTo``
(crates/gnark/src/proof.rs:1-8), the public inputs 32-byte big-endian scalars.

Compressed points need square roots: in Fq, q = 3 mod 4, so sqrt(a) = a^((q+1)/4); in Fq2,
Algorithm 9 of Adj and Rodriguez-Henriquez, "Square root computation over even extension
fields" (IEEE TC 2014), for q = 3 mod 4.
"""
from __future__ import annotations

import numpy as np

from snarkrs_amd.wire import Q_MOD, R_MOD, fr_array, g1_array, int_to_le, to_mont

Q = Q_MOD
XI = (9, 1)
B_G1 = 3


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def f2_add(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def f2_pow(a, e):
    acc = (1, 0)
    while e:
        if e & 1:
            acc = f2_mul(acc, a)
        a = f2_mul(a, a)
        e >>= 1
    return acc


def f2_inv(a):
    d = pow(a[0] * a[0] + a[1] * a[1], -1, Q)
    return (a[0] * d % Q, -a[1] * d % Q)


B_G2 = f2_mul((3, 0), f2_inv(XI))  # the twist's b = 3 / (9 + u)


def fq_sqrt(a):
    y = pow(a, (Q + 1) // 4, Q)
    return y if y * y % Q == a % Q else None


def fq2_sqrt(a):
    a1 = f2_pow(a, (Q - 3) // 4)
    alpha = f2_mul(f2_mul(a1, a1), a)
    a0 = f2_mul((alpha[0], -alpha[1] % Q), alpha)  # alpha^q alpha
    if a0 == (Q - 1, 0):
        return None
    x0 = f2_mul(a1, a)
    if alpha == (Q - 1, 0):
        x = f2_mul((0, 1), x0)
    else:
        b = f2_pow(f2_add((1, 0), alpha), (Q - 1) // 2)
        x = f2_mul(b, x0)
    return x if f2_mul(x, x) == (a[0] % Q, a[1] % Q) else None


def _largest(y: int) -> bool:
    return y > (Q - 1) // 2


def _largest2(y) -> bool:
    return _largest(y[1]) if y[1] else _largest(y[0])


def g1_from_gnark(b: bytes):
    """One encoder G1 point -> ((x, y) ints or None for infinity, bytes consumed)."""
    flag = b[0] & 0xC0
    if flag == 0x00:
        raw = b[:64]
        if not any(raw):
            return None, 64
        return (int.from_bytes(raw[:32], "big"), int.from_bytes(raw[32:64], "big")), 64
    x = int.from_bytes(bytes([b[0] & 0x3F]) + b[1:32], "big")
    if flag == 0x40:
        return None, 32
    y = fq_sqrt((x * x * x + B_G1) % Q)
    assert y is not None, "not the abscissa of a curve point"
    if _largest(y) != (flag == 0xC0):
        y = (-y) % Q
    return (x, y), 32


def g2_from_gnark(b: bytes):
    """One encoder G2 point -> (((x0, x1), (y0, y1)) or None, bytes consumed)."""
    flag = b[0] & 0xC0
    if flag == 0x00:
        raw = b[:128]
        if not any(raw):
            return None, 128
        c = [int.from_bytes(raw[32 * i:32 * i + 32], "big") for i in range(4)]
        return ((c[1], c[0]), (c[3], c[2])), 128
    x1 = int.from_bytes(bytes([b[0] & 0x3F]) + b[1:32], "big")
    x0 = int.from_bytes(b[32:64], "big")
    if flag == 0x40:
        return None, 64
    x = (x0, x1)
    y = fq2_sqrt(f2_add(f2_mul(f2_mul(x, x), x), B_G2))
    assert y is not None, "not the abscissa of a curve point"
    if _largest2(y) != (flag == 0xC0):
        y = ((-y[0]) % Q, (-y[1]) % Q)
    return (x, y), 64


def g2_array(points) -> np.ndarray:
    out = np.zeros((len(points), 128), np.uint8)
    for i, p in enumerate(points):
        if p is None:
            continue
        (x0, x1), (y0, y1) = p
        out[i] = np.frombuffer(b"".join(int_to_le(to_mont(v, Q)) for v in (x0, x1, y0, y1)),
                               np.uint8)
    return out


def read_vk(data: bytes):
    """vk.bin -> dict of wire-encoded points (alpha_g1, beta_g1, beta_g2, gamma_g2, delta_g1,
    delta_g2, ic)."""
    pos = 0
    out = {}
    for name, g in (("alpha_g1", 1), ("beta_g1", 1), ("beta_g2", 2), ("gamma_g2", 2),
                    ("delta_g1", 1), ("delta_g2", 2)):
        p, n = (g1_from_gnark if g == 1 else g2_from_gnark)(data[pos:])
        out[name] = (g1_array if g == 1 else g2_array)([p])[0]
        pos += n
    count = int.from_bytes(data[pos:pos + 4], "big")
    pos += 4
    ic = []
    for _ in range(count):
        p, n = g1_from_gnark(data[pos:])
        ic.append(p)
        pos += n
    out["ic"] = g1_array(ic)
    committed = int.from_bytes(data[pos:pos + 4], "big")
    commitments = int.from_bytes(data[pos + 4:pos + 8], "big")
    assert committed == 0 and commitments == 0 and pos + 8 == len(data)
    return out


def read_proof(data: bytes) -> np.ndarray:
    """proof.bin (WriteRawTo, no commitments) -> the 256-byte wire proof A | B | C."""
    a, n1 = g1_from_gnark(data)
    b, n2 = g2_from_gnark(data[n1:])
    c, n3 = g1_from_gnark(data[n1 + n2:])
    pos = n1 + n2 + n3
    assert int.from_bytes(data[pos:pos + 4], "big") == 0  # no commitments
    pok, _ = g1_from_gnark(data[pos + 4:])
    assert pok is None
    return np.concatenate([g1_array([a])[0], g2_array([b])[0], g1_array([c])[0]])


def read_public(data: bytes) -> np.ndarray:
    """public.bin -> Fr Montgomery (n, 32)."""
    vals = [int.from_bytes(data[i:i + 32], "big") for i in range(0, len(data), 32)]
    assert all(v < R_MOD for v in vals)
    return fr_array(vals)


class _Reader:
    def __init__(self, data: bytes):
        self.d, self.p = data, 0

    def take(self, n: int) -> bytes:
        assert self.p + n <= len(self.d), "truncated"
        b = self.d[self.p:self.p + n]
        self.p += n
        return b

    def u64_le(self) -> int:
        return int.from_bytes(self.take(8), "little")

    def u64_be(self) -> int:
        return int.from_bytes(self.take(8), "big")

    def u32_be(self) -> int:
        return int.from_bytes(self.take(4), "big")

    def g1(self):
        p, n = g1_from_gnark(self.d[self.p:])
        self.p += n
        return g1_array([p])[0]

    def g2(self):
        p, n = g2_from_gnark(self.d[self.p:])
        self.p += n
        return g2_array([p])[0]

    def dumped(self, stride: int) -> np.ndarray:
        """unsafe.WriteSlice: u64 LE length, then the elements' memory (Montgomery limbs,
        little-endian, G2 coordinates A0 before A1: the boundary's own encoding)."""
        n = self.u64_le()
        return np.frombuffer(self.take(n * stride), np.uint8).reshape(n, stride).copy()


MARKER = (0xDEADBEEF).to_bytes(8, "little")


def read_pk_dump(data: bytes) -> dict:
    """pk.dump (ProvingKey.WriteDump, crates/gnark/src/pk.rs:1-20) -> header values and the
    queries re-expanded to one base per wire (infinity where gnark leaves a wire out)."""
    r = _Reader(data)
    assert r.take(8) == MARKER
    card = r.u64_be()
    dom = [int.from_bytes(r.take(32), "big") for _ in range(5)]
    assert r.take(1)[0] in (0, 1)
    out = {"domain_size": card, "card_inv": dom[0], "generator": dom[1],
           "generator_inv": dom[2]}
    for name in ("alpha_g1", "beta_g1", "delta_g1"):
        out[name] = r.g1()
    for name in ("beta_g2", "delta_g2"):
        out[name] = r.g2()
    n_wires, inf_a, inf_b = r.u64_be(), r.u64_be(), r.u64_be()
    flags_a = np.frombuffer(r.take(n_wires), np.uint8)
    flags_b = np.frombuffer(r.take(n_wires), np.uint8)
    assert int(flags_a.sum()) == inf_a and int(flags_b.sum()) == inf_b
    assert r.u32_be() == 0  # no commitments
    g1_a, g1_b, g1_z, g1_k = r.dumped(64), r.dumped(64), r.dumped(64), r.dumped(64)
    g2_b = r.dumped(128)
    assert r.p == len(data)

    def expand(pts, flags, width):
        full = np.zeros((n_wires, width), np.uint8)
        full[flags == 0] = pts
        return full

    out.update(n_wires=n_wires, a_query=expand(g1_a, flags_a, 64),
               b_g1_query=expand(g1_b, flags_b, 64), b_g2_query=expand(g2_b, flags_b, 128),
               z=g1_z, k=g1_k)
    assert len(g1_z) == card - 1
    return out


def read_witness_dump(data: bytes):
    """witness.bin (crates/gnark/src/witness.rs:1-12): W (every wire, W[0] = 1), A.w, B.w."""
    r = _Reader(data)
    assert r.take(8) == MARKER
    w, aw, bw = r.dumped(32), r.dumped(32), r.dumped(32)
    assert r.p == len(data)
    return w, aw, bw


def toy_proving_key(pk: dict, vk: dict):
    """The toy's gnark key as a g16hip proving key (gnark's H basis, no matrices)."""
    from snarkrs_amd.key import H_BASIS_GNARK, ProvingKey
    n_pub = len(vk["ic"]) - 1
    return ProvingKey(
        n_vars=pk["n_wires"], n_public=n_pub, domain_size=pk["domain_size"],
        alpha_g1=pk["alpha_g1"], beta_g1=pk["beta_g1"], beta_g2=pk["beta_g2"],
        gamma_g2=vk["gamma_g2"], delta_g1=pk["delta_g1"], delta_g2=pk["delta_g2"],
        ic=np.ascontiguousarray(vk["ic"]), row_ptr=None, signal=None, value_id=None,
        values=None, a_query=pk["a_query"], b_g1_query=pk["b_g1_query"],
        b_g2_query=pk["b_g2_query"], l_query=np.ascontiguousarray(pk["k"]),
        h_query=np.ascontiguousarray(pk["z"]), h_basis=H_BASIS_GNARK)
