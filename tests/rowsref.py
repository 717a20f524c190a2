"""Test helpers for the rows path: a gnark-basis copy of a synthetic key, and A.w / B.w rows.

gnark's H basis (``HBasis::GnarkBitReversed``, crates/formats/src/lib.rs:66-89) is
``domain_size - 1`` bases ``[tau^bitrev(j) t(tau) / delta]_1``, t(x) = x^n - 1; a synthetic
key knows tau and delta, so the basis is computed from the toxic waste. A gnark key carries
no constraint matrices: the copy drops them, so it can only prove from rows.
"""
import dataclasses

import numpy as np

from snarkrs_amd.key import H_BASIS_GNARK
from snarkrs_amd.wire import R_MOD, fr_array, fr_ints


def bitrev(j: int, bits: int) -> int:
    return int(format(j, f"0{bits}b")[::-1], 2) if bits else 0


def gnark_basis_scalars(toxic, n: int) -> list:
    tau, _, _, _, delta = fr_ints(toxic)
    bits = n.bit_length() - 1
    t = (pow(tau, n, R_MOD) - 1) * pow(delta, -1, R_MOD) % R_MOD
    return [pow(tau, bitrev(j, bits), R_MOD) * t % R_MOD for j in range(n - 1)]


def gnark_key(key, toxic, fixed_base, keep_matrices: bool = False):
    """``key`` with gnark's H basis (points by ``fixed_base(group, k)``)."""
    z = fixed_base(1, fr_array(gnark_basis_scalars(toxic, key.domain_size)))
    kw = dict(h_query=np.ascontiguousarray(z), h_basis=H_BASIS_GNARK)
    if not keep_matrices:
        kw.update(row_ptr=None, signal=None, value_id=None, values=None)
    return dataclasses.replace(key, **kw)


def rows(orc, key, witness, trim: bool = True):
    """A.w and B.w per row from the key's matrices (the oracle's gather); with ``trim`` the
    zero rows past the last nonzero one are dropped, as a solver hands over one row per
    constraint."""
    a = orc.gather(key, 0, witness)
    b = orc.gather(key, 1, witness)
    if trim:
        nz = np.nonzero(a.any(axis=1) | b.any(axis=1))[0]
        m = int(nz[-1]) + 1 if len(nz) else 0
        a, b = a[:m], b[:m]
    return np.ascontiguousarray(a), np.ascontiguousarray(b)
