"""Host-side proving key container and its C-ABI view (``g16hip_key``).

The fields follow the reference's ``ProvingKey`` (crates/formats/src/lib.rs:35-75) with
section 4 in CSR form (``Coefficients``, crates/formats/src/lib.rs:118-150). Arrays are
numpy, in the packed wire format of :mod:`snarkrs_amd.wire`.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)


H_BASIS_SNARKJS, H_BASIS_GNARK = 0, 1


class KeyStruct(ctypes.Structure):
    """``g16hip_key`` (include/g16hip.h)."""

    _fields_ = [
        ("n_vars", ctypes.c_uint32),
        ("n_public", ctypes.c_uint32),
        ("domain_size", ctypes.c_uint32),
        ("n_values", ctypes.c_uint32),
        ("alpha_g1", _u8p),
        ("beta_g1", _u8p),
        ("beta_g2", _u8p),
        ("gamma_g2", _u8p),
        ("delta_g1", _u8p),
        ("delta_g2", _u8p),
        ("ic", _u8p),
        ("row_ptr", _u32p * 2),
        ("signal", _u32p * 2),
        ("value_id", _u32p * 2),
        ("values", _u8p),
        ("a_query", _u8p),
        ("b_g1_query", _u8p),
        ("b_g2_query", _u8p),
        ("l_query", _u8p),
        ("h_query", _u8p),
        ("h_basis", ctypes.c_uint32),
    ]


def _p8(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] and a.dtype == np.uint8, (a.dtype, a.flags)
    return a.ctypes.data_as(_u8p)


def _p32(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] and a.dtype == np.uint32
    return a.ctypes.data_as(_u32p)


@dataclass
class ProvingKey:
    n_vars: int
    n_public: int
    domain_size: int
    alpha_g1: np.ndarray
    beta_g1: np.ndarray
    beta_g2: np.ndarray
    gamma_g2: np.ndarray
    delta_g1: np.ndarray
    delta_g2: np.ndarray
    ic: np.ndarray
    row_ptr: list  # [A, B] uint32 (domain_size + 1)
    signal: list
    value_id: list
    values: np.ndarray  # (n_values, 32) uint8, Fr Montgomery
    a_query: np.ndarray
    b_g1_query: np.ndarray
    b_g2_query: np.ndarray
    l_query: np.ndarray
    h_query: np.ndarray
    # What h_query is a basis for (HBasis, crates/formats/src/lib.rs:66-89): H_BASIS_SNARKJS
    # (domain_size bases, the odd coset) or H_BASIS_GNARK (domain_size - 1 bases
    # [tau^bitrev(j) t(tau) / delta]_1, H's coefficients in bit-reversed order).
    h_basis: int = 0
    extra: dict = field(default_factory=dict)

    def struct(self) -> KeyStruct:
        """A ``g16hip_key`` view; the arrays must outlive its use."""
        s = KeyStruct()
        s.n_vars = self.n_vars
        s.n_public = self.n_public
        s.domain_size = self.domain_size
        s.n_values = 0 if self.values is None else len(self.values)
        s.alpha_g1 = _p8(self.alpha_g1)
        s.beta_g1 = _p8(self.beta_g1)
        s.beta_g2 = _p8(self.beta_g2)
        s.gamma_g2 = _p8(self.gamma_g2)
        s.delta_g1 = _p8(self.delta_g1)
        s.delta_g2 = _p8(self.delta_g2)
        s.ic = _p8(self.ic)
        if self.row_ptr is not None:  # a key without matrices proves from A.w, B.w rows only
            for m in range(2):
                s.row_ptr[m] = _p32(self.row_ptr[m])
                s.signal[m] = _p32(self.signal[m])
                s.value_id[m] = _p32(self.value_id[m])
        if self.values is not None:
            s.values = _p8(self.values)
        s.a_query = _p8(self.a_query)
        s.b_g1_query = _p8(self.b_g1_query)
        s.b_g2_query = _p8(self.b_g2_query)
        s.l_query = _p8(self.l_query)
        s.h_query = _p8(self.h_query)
        s.h_basis = self.h_basis
        return s

    @property
    def nbytes(self) -> int:
        tot = 0
        for v in self.__dict__.values():
            if isinstance(v, np.ndarray):
                tot += v.nbytes
            elif isinstance(v, list):
                tot += sum(x.nbytes for x in v)
        return tot
