"""snarkjs .zkey / .wtns files (include/g16zkey.h) to and from :class:`ProvingKey`."""
from __future__ import annotations

import ctypes

import numpy as np

from .key import KeyStruct, ProvingKey
from .native import _check, load

_u8p = ctypes.POINTER(ctypes.c_uint8)


def _declare(lib):
    if getattr(lib, "_zkey_declared", False):
        return
    lib.g16zkey_load.restype = ctypes.c_int
    lib.g16zkey_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.g16zkey_key.restype = ctypes.POINTER(KeyStruct)
    lib.g16zkey_key.argtypes = [ctypes.c_void_p]
    lib.g16zkey_free.argtypes = [ctypes.c_void_p]
    lib.g16zkey_write.restype = ctypes.c_int
    lib.g16zkey_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(KeyStruct)]
    lib.g16wtns_load.restype = ctypes.c_int
    lib.g16wtns_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(_u8p),
                                 ctypes.POINTER(ctypes.c_size_t)]
    lib.g16wtns_free.argtypes = [_u8p]
    lib.g16wtns_write.restype = ctypes.c_int
    lib.g16wtns_write.argtypes = [ctypes.c_char_p, _u8p, ctypes.c_size_t]
    lib._zkey_declared = True


def _copy(ptr, nbytes, dtype=np.uint8, shape=None):
    if nbytes == 0:
        return np.zeros(0 if shape is None else shape, dtype=dtype)
    buf = ctypes.string_at(ctypes.cast(ptr, ctypes.c_void_p), nbytes)
    a = np.frombuffer(buf, dtype=dtype).copy()
    return a.reshape(shape) if shape is not None else a


def load_zkey(path: str) -> ProvingKey:
    """ProvingKey::load (crates/formats/src/lib.rs:430): parse, validate, CSR-sort section 4."""
    lib = load()
    _declare(lib)
    h = ctypes.c_void_p()
    _check(lib.g16zkey_load(str(path).encode(), ctypes.byref(h)))
    try:
        k = lib.g16zkey_key(h).contents
        nv, npub, n = k.n_vars, k.n_public, k.domain_size
        rp = [_copy(k.row_ptr[m], 4 * (n + 1), np.uint32) for m in range(2)]
        nnz = [int(rp[m][-1]) for m in range(2)]
        return ProvingKey(
            n_vars=nv, n_public=npub, domain_size=n,
            alpha_g1=_copy(k.alpha_g1, 64), beta_g1=_copy(k.beta_g1, 64),
            beta_g2=_copy(k.beta_g2, 128), gamma_g2=_copy(k.gamma_g2, 128),
            delta_g1=_copy(k.delta_g1, 64), delta_g2=_copy(k.delta_g2, 128),
            ic=_copy(k.ic, 64 * (npub + 1), shape=(npub + 1, 64)),
            row_ptr=rp,
            signal=[_copy(k.signal[m], 4 * nnz[m], np.uint32) for m in range(2)],
            value_id=[_copy(k.value_id[m], 4 * nnz[m], np.uint32) for m in range(2)],
            values=_copy(k.values, 32 * k.n_values, shape=(k.n_values, 32)),
            a_query=_copy(k.a_query, 64 * nv, shape=(nv, 64)),
            b_g1_query=_copy(k.b_g1_query, 64 * nv, shape=(nv, 64)),
            b_g2_query=_copy(k.b_g2_query, 128 * nv, shape=(nv, 128)),
            l_query=_copy(k.l_query, 64 * (nv - npub - 1), shape=(nv - npub - 1, 64)),
            h_query=_copy(k.h_query, 64 * n, shape=(n, 64)),
        )
    finally:
        lib.g16zkey_free(h)


def write_zkey(path: str, key: ProvingKey) -> None:
    lib = load()
    _declare(lib)
    st = key.struct()
    _check(lib.g16zkey_write(str(path).encode(), ctypes.byref(st)))


def load_wtns(path: str) -> np.ndarray:
    """Witness::load (crates/formats/src/wtns.rs:17-63), returned as Fr Montgomery (n, 32)."""
    lib = load()
    _declare(lib)
    p = _u8p()
    n = ctypes.c_size_t()
    _check(lib.g16wtns_load(str(path).encode(), ctypes.byref(p), ctypes.byref(n)))
    try:
        return _copy(p, 32 * n.value, shape=(n.value, 32))
    finally:
        lib.g16wtns_free(p)


def write_wtns(path: str, witness: np.ndarray) -> None:
    lib = load()
    _declare(lib)
    w = np.ascontiguousarray(witness, dtype=np.uint8).reshape(-1, 32)
    _check(lib.g16wtns_write(str(path).encode(), w.ctypes.data_as(_u8p), len(w)))
