"""Groth16 verification and snarkjs' JSON files, over libg16hip's host verifier.

Mirrors ``snarkrs_groth16::verify`` (crates/groth16/src/verify.rs:49-175) and
``snarkrs_groth16::json`` (crates/groth16/src/json.rs:97-243), plus
``VerifyingKey::from_json_str`` (crates/formats/src/lib.rs:1371-1430). The arithmetic (BN254
optimal ate pairing, point and subgroup checks, decimal conversion, the JSON reader) is the
library's C++; nothing here needs a GPU. Points and scalars use the boundary's wire encoding:
Montgomery limbs, all-zero = infinity (G1 64 bytes, G2 128 bytes, Fr 32 bytes).

Errors are the ``VerifyError`` family of :mod:`snarkrs_amd.native` (``PublicInputCount``,
``EmptyVerifyingKey``, ``PairingFailed``, ``InvalidProof``, ``UnsafeKey``) and ``JsonError``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from .native import ProofOut, VkStruct, _c8, _check, _p8, load

G1_BYTES, G2_BYTES, FR_BYTES = 64, 128, 32


@dataclass
class VerifyingKey:
    """``VerifyingKey`` (crates/formats/src/lib.rs:353-360) in the wire encoding."""

    alpha_g1: np.ndarray
    beta_g2: np.ndarray
    gamma_g2: np.ndarray
    delta_g2: np.ndarray
    ic: np.ndarray  # (n_public + 1, 64)

    @property
    def n_public(self) -> int:
        return len(self.ic) - 1

    @staticmethod
    def from_bytes(b) -> "VerifyingKey":
        """alpha_g1 | beta_g2 | gamma_g2 | delta_g2 | ic, as g16hip_circuit_vk writes it."""
        b = _c8(b).reshape(-1)
        if len(b) < 448 or (len(b) - 448) % 64:
            raise ValueError(f"{len(b)} bytes is not a verifying key")
        return VerifyingKey(b[0:64].copy(), b[64:192].copy(), b[192:320].copy(),
                            b[320:448].copy(), b[448:].reshape(-1, 64).copy())

    def to_bytes(self) -> np.ndarray:
        return np.concatenate([_c8(self.alpha_g1).reshape(-1), _c8(self.beta_g2).reshape(-1),
                               _c8(self.gamma_g2).reshape(-1), _c8(self.delta_g2).reshape(-1),
                               _c8(self.ic).reshape(-1)])

    @staticmethod
    def from_proving_key(pk) -> "VerifyingKey":
        """The verifying half of a :class:`snarkrs_amd.key.ProvingKey`."""
        return VerifyingKey(_c8(pk.alpha_g1).reshape(-1), _c8(pk.beta_g2).reshape(-1),
                            _c8(pk.gamma_g2).reshape(-1), _c8(pk.delta_g2).reshape(-1),
                            _c8(pk.ic).reshape(-1, 64))

    @staticmethod
    def from_json_str(text: str, check_structure: bool = True) -> "VerifyingKey":
        """snarkjs' verification_key.json (``from_json_str`` / ``from_json_str_unchecked``)."""
        lib = load()
        raw = text.encode()
        npub = ctypes.c_uint32()
        _check(lib.g16hip_vk_from_json(raw, len(raw), int(check_structure), None, 0,
                                       ctypes.byref(npub)))
        out = np.zeros(448 + 64 * (npub.value + 1), np.uint8)
        _check(lib.g16hip_vk_from_json(raw, len(raw), int(check_structure), _p8(out), len(out),
                                       ctypes.byref(npub)))
        return VerifyingKey.from_bytes(out)

    def _struct(self):
        parts = [_c8(self.alpha_g1).reshape(-1), _c8(self.beta_g2).reshape(-1),
                 _c8(self.gamma_g2).reshape(-1), _c8(self.delta_g2).reshape(-1),
                 _c8(self.ic).reshape(-1)]
        for p, n in zip(parts, (64, 128, 128, 128)):
            if len(p) != n:
                raise ValueError("verifying-key point of the wrong size")
        if len(parts[4]) % 64:
            raise ValueError("IC is not a whole number of G1 points")
        st = VkStruct(_p8(parts[0]), _p8(parts[1]), _p8(parts[2]), _p8(parts[3]),
                      _p8(parts[4]) if len(parts[4]) else None, len(parts[4]) // 64)
        return st, parts  # keep the arrays alive for the call


def _proof(proof) -> ProofOut:
    b = _c8(proof).reshape(-1)
    if len(b) != 256:
        raise ValueError(f"a proof is 256 bytes (A | B | C), got {len(b)}")
    return ProofOut.from_buffer_copy(b.tobytes())


def _public(public):
    p = _c8(public).reshape(-1, FR_BYTES) if len(public) else np.zeros((0, 32), np.uint8)
    return p, (_p8(p) if len(p) else None), len(p)


def _verify(fn, vk: VerifyingKey, public, proof) -> None:
    st, keep = vk._struct()
    p, pp, n = _public(public)
    pr = _proof(proof)
    _check(fn(ctypes.byref(st), pp, n, ctypes.byref(pr)))
    del keep, p


def verify(vk: VerifyingKey, public, proof) -> None:
    """``verify`` (crates/groth16/src/verify.rs:49-52): raises a ``VerifyError`` unless the
    key is well formed, the proof's points are valid and the pairing equation holds."""
    _verify(load().g16hip_verify, vk, public, proof)


def verify_allowing_unblinded(vk: VerifyingKey, public, proof) -> None:
    """``verify_allowing_unblinded`` (verify.rs:66-76): accepts gamma_g2 == delta_g2."""
    _verify(load().g16hip_verify_allowing_unblinded, vk, public, proof)


def verify_unchecked(vk: VerifyingKey, public, proof) -> None:
    """``verify_unchecked`` (verify.rs:104-134): the bare pairing check."""
    _verify(load().g16hip_verify_unchecked, vk, public, proof)


def pairing_check(g1, g2) -> bool:
    """Whether prod_i e(g1[i], g2[i]) = 1 (``Bn254::multi_pairing(..).is_zero()``)."""
    a = _c8(g1).reshape(-1, G1_BYTES)
    b = _c8(g2).reshape(-1, G2_BYTES)
    if len(a) != len(b):
        raise ValueError("as many G1 as G2 points")
    out = ctypes.c_int()
    _check(load().g16hip_pairing_check(_p8(a) if len(a) else None, _p8(b) if len(b) else None,
                                       len(a), ctypes.byref(out)))
    return bool(out.value)


def pairing(g1, g2) -> np.ndarray:
    """e(g1, g2), 384 bytes: Fq12 coordinates c0.(a0, a1, a2), c1.(a0, a1, a2), Fq2 c0 then c1."""
    out = np.zeros(384, np.uint8)
    _check(load().g16hip_pairing(_p8(_c8(g1).reshape(-1)), _p8(_c8(g2).reshape(-1)), _p8(out)))
    return out


def point_valid(group: int, point) -> bool:
    """The point decodes below the modulus, is on the curve and (G2) in the subgroup."""
    out = ctypes.c_int()
    _check(load().g16hip_point_check(group, _p8(_c8(point).reshape(-1)), ctypes.byref(out)))
    return bool(out.value)


def _text(fn, *args) -> str:
    n = ctypes.c_size_t()
    _check(fn(*args, None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _check(fn(*args, buf, len(buf), ctypes.byref(n)))
    return buf.value.decode()


def proof_to_string(proof) -> str:
    """``proof_to_string`` (crates/groth16/src/json.rs:97-108): snarkjs' proof.json."""
    pr = _proof(proof)
    return _text(load().g16hip_proof_to_json, ctypes.byref(pr))


def public_to_string(public) -> str:
    """``public_to_string`` (json.rs:111-114): snarkjs' public.json."""
    p, pp, n = _public(public)
    return _text(load().g16hip_public_to_json, pp, n)


def proof_from_str(text: str) -> np.ndarray:
    """``proof_from_str`` (json.rs:230-236): a proof (A | B | C, 256 bytes)."""
    raw = text.encode()
    out = ProofOut()
    _check(load().g16hip_proof_from_json(raw, len(raw), ctypes.byref(out)))
    return out.as_array()


def public_from_str(text: str) -> np.ndarray:
    """``public_from_str`` (json.rs:238-243): the public signals, Fr Montgomery (n, 32)."""
    lib = load()
    raw = text.encode()
    n = ctypes.c_size_t()
    _check(lib.g16hip_public_from_json(raw, len(raw), None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, 32), np.uint8)
    if n.value:
        _check(lib.g16hip_public_from_json(raw, len(raw), _p8(out), n.value, ctypes.byref(n)))
    return out
