"""ctypes binding of libg16hip.so (include/g16hip.h).

This is the only way the package reaches the GPU. There is no fallback: if the library is
missing or cannot be loaded, every entry point raises :class:`NativeUnavailable`.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

from .key import KeyStruct, ProvingKey

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libg16hip.so")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p

# g16hip_status
(OK, WITNESS_LENGTH, BACKEND, DEVICE, DEVICE_MEMORY, CONSTANT_WIRE, ZERO_BLINDER, SELF_VERIFY,
 PUBLIC_COUNT, EMPTY_KEY, PAIRING_FAILED, INVALID_PROOF, UNSAFE_KEY, BAD_JSON) = range(14)


class NativeUnavailable(RuntimeError):
    """libg16hip.so is not built or cannot be loaded; there is no CPU fallback."""


class ProveError(RuntimeError):
    """``ProveError`` (crates/groth16/src/lib.rs:79-154); ``status`` is the g16hip_status."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status

    def is_device_fault(self) -> bool:
        return self.status == DEVICE


class WitnessLength(ProveError):
    pass


class BackendError(ProveError):
    pass


class DeviceFault(ProveError):
    pass


class DeviceMemory(ProveError):
    pass


class ConstantWire(ProveError):
    pass


class ZeroBlinder(ProveError):
    pass


class SelfVerify(ProveError):
    """The proof failed the check of its result (``ProveError::SelfVerify``)."""


class VerifyError(ProveError):
    """``VerifyError`` (crates/groth16/src/verify.rs:16-31)."""


class PublicInputCount(VerifyError):
    pass


class EmptyVerifyingKey(VerifyError):
    pass


class PairingFailed(VerifyError):
    pass


class InvalidProof(VerifyError):
    pass


class UnsafeKey(VerifyError):
    pass


class JsonError(ProveError):
    """A malformed proof.json / public.json / verification_key.json
    (crates/groth16/src/json.rs:39-41, ``ZkeyError::BadJson``)."""


_ERRORS = {WITNESS_LENGTH: WitnessLength, BACKEND: BackendError, DEVICE: DeviceFault,
           DEVICE_MEMORY: DeviceMemory, CONSTANT_WIRE: ConstantWire, ZERO_BLINDER: ZeroBlinder,
           SELF_VERIFY: SelfVerify, PUBLIC_COUNT: PublicInputCount, EMPTY_KEY: EmptyVerifyingKey,
           PAIRING_FAILED: PairingFailed, INVALID_PROOF: InvalidProof, UNSAFE_KEY: UnsafeKey,
           BAD_JSON: JsonError}


class Timings(ctypes.Structure):
    """``g16hip_timings`` = StageTimings (crates/groth16/src/lib.rs:322-368), microseconds."""

    _fields_ = [(n, ctypes.c_uint64) for n in
                ("gather_us", "ntt_us", "pointwise_us", "msm_us", "assemble_us", "verify_us")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class MsmStats(ctypes.Structure):
    """``g16hip_msm_stats``: phase split of the last standalone MSM (milliseconds)."""

    _fields_ = [("sort_ms", ctypes.c_float), ("accumulate_ms", ctypes.c_float),
                ("merge_ms", ctypes.c_float), ("reduce_ms", ctypes.c_float),
                ("total_ms", ctypes.c_float), ("entries", ctypes.c_uint64),
                ("c", ctypes.c_uint32), ("windows", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class KernelStats(ctypes.Structure):
    """``g16hip_kernel_stats``: running totals over bucket-accumulation launches."""

    _fields_ = [("launches", ctypes.c_uint64 * 2), ("entries", ctypes.c_uint64 * 2),
                ("accumulate_ms", ctypes.c_double * 2)]


class BatchItem(ctypes.Structure):
    """``g16hip_batch_item``."""

    _fields_ = [("circuit", _vp), ("witness", _u8p), ("n_witness", ctypes.c_size_t),
                ("r", _u8p), ("s", _u8p)]


class MsmOut(ctypes.Structure):
    _fields_ = [("a_g1", ctypes.c_uint8 * 64), ("b_g2", ctypes.c_uint8 * 128),
                ("b_g1", ctypes.c_uint8 * 64), ("l_g1", ctypes.c_uint8 * 64),
                ("h_g1", ctypes.c_uint8 * 64)]

    def as_array(self) -> np.ndarray:
        return np.frombuffer(bytes(self), dtype=np.uint8).copy()


class ProofOut(ctypes.Structure):
    _fields_ = [("a", ctypes.c_uint8 * 64), ("b", ctypes.c_uint8 * 128), ("c", ctypes.c_uint8 * 64)]

    def as_array(self) -> np.ndarray:
        return np.frombuffer(bytes(self), dtype=np.uint8).copy()


_lib = None
_lib_lock = threading.Lock()


class CeremonyStats(ctypes.Structure):
    """``g16hip_ceremony_stats``: where a ceremony handle's calls went."""
    _fields_ = [("calls", ctypes.c_uint64), ("runs", ctypes.c_uint64), ("points", ctypes.c_uint64),
                ("retries", ctypes.c_uint64), ("ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class VkStruct(ctypes.Structure):
    """``g16hip_vk`` = VerifyingKey (crates/formats/src/lib.rs:353-360)."""
    _fields_ = [("alpha_g1", _u8p), ("beta_g2", _u8p), ("gamma_g2", _u8p), ("delta_g2", _u8p),
                ("ic", _u8p), ("n_ic", ctypes.c_uint32)]


def load() -> ctypes.CDLL:
    """Loads libg16hip.so once; raises NativeUnavailable if it is not there."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NativeUnavailable(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C snarkrs_amd/csrc)")
        try:
            lib = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise NativeUnavailable(f"cannot load {LIB_PATH}: {e}") from e
        _declare(lib)
        _lib = lib
        return lib


EXPORTS = {
    # name: (restype, argtypes)
    "g16hip_last_error": (ctypes.c_char_p, []),
    "g16hip_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "g16hip_close": (None, [_vp]),
    "g16hip_mem_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(ctypes.c_size_t)]),
    "g16hip_set_vram_limit": (ctypes.c_int, [_vp, ctypes.c_size_t]),
    "g16hip_device_budget": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "g16hip_key_resident_bytes": (ctypes.c_int, [ctypes.POINTER(KeyStruct), ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_size_t)]),
    "g16hip_prepare": (ctypes.c_int, [_vp, ctypes.POINTER(KeyStruct), ctypes.POINTER(_vp)]),
    "g16hip_circuit_release": (None, [_vp]),
    "g16hip_circuit_dims": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_uint32)] * 3),
    "g16hip_circuit_resident_bytes": (ctypes.c_size_t, [_vp]),
    "g16hip_compute_h": (ctypes.c_int, [_vp, _u8p, ctypes.c_size_t, ctypes.POINTER(_vp),
                                        ctypes.POINTER(Timings)]),
    "g16hip_h_to_host": (ctypes.c_int, [_vp, _vp, _u8p]),
    "g16hip_h_release": (None, [_vp]),
    "g16hip_msms": (ctypes.c_int, [_vp, _u8p, ctypes.c_size_t, _vp, ctypes.POINTER(MsmOut),
                                   ctypes.POINTER(Timings)]),
    "g16hip_h_and_msms": (ctypes.c_int, [_vp, _u8p, ctypes.c_size_t, ctypes.POINTER(MsmOut),
                                         ctypes.POINTER(Timings)]),
    "g16hip_h_and_msms_device": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t,
                                                ctypes.POINTER(MsmOut), ctypes.POINTER(Timings)]),
    "g16hip_prove_with_blinders": (ctypes.c_int, [_vp, _u8p, ctypes.c_size_t, _u8p, _u8p,
                                                  ctypes.POINTER(ProofOut), ctypes.POINTER(Timings)]),
    "g16hip_prove_with_blinders_device": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _u8p, _u8p,
                                                         ctypes.POINTER(ProofOut),
                                                         ctypes.POINTER(Timings)]),
    "g16hip_circuit_add_slots": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_uint32)]),
    "g16hip_set_graph_replay": (None, [ctypes.c_int]),
    "g16hip_prove_batch": (ctypes.c_int, [ctypes.POINTER(BatchItem), ctypes.c_size_t,
                                          ctypes.POINTER(ProofOut), ctypes.POINTER(ctypes.c_int)]),
    "g16hip_set_batch_cap": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "g16hip_batch_autotune": (ctypes.c_int, [_vp, _u8p, ctypes.c_size_t, ctypes.c_uint32,
                                             ctypes.POINTER(ctypes.c_uint32),
                                             ctypes.POINTER(ctypes.c_double), ctypes.c_size_t,
                                             ctypes.POINTER(ctypes.c_size_t),
                                             ctypes.POINTER(ctypes.c_uint32)]),
    "g16hip_server_open": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_uint32,
                                          ctypes.POINTER(_vp)]),
    "g16hip_server_submit": (ctypes.c_int, [_vp, _vp, _u8p, ctypes.c_size_t, _u8p, _u8p,
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "g16hip_server_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ProofOut),
                                          ctypes.POINTER(ctypes.c_double)]),
    "g16hip_server_close": (None, [_vp]),
    "g16hip_msm_g1": (ctypes.c_int, [_vp, _u8p, _u8p, ctypes.c_size_t, _u8p]),
    "g16hip_msm_g2": (ctypes.c_int, [_vp, _u8p, _u8p, ctypes.c_size_t, _u8p]),
    "g16hip_msm_g1_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _u8p]),
    "g16hip_msm_g2_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _u8p]),
    "g16hip_msm_device_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_size_t,
                                                 ctypes.c_int, _u8p]),
    "g16hip_msm_last_stats": (ctypes.c_int, [_vp, ctypes.POINTER(MsmStats)]),
    "g16hip_kernel_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(KernelStats)]),
    "g16hip_kernel_stats_reset": (ctypes.c_int, [_vp]),
    "g16hip_ntt": (ctypes.c_int, [_vp, _u8p, ctypes.c_int, ctypes.c_int]),
    "g16hip_ntt_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int]),
    "g16hip_field_probe": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _u8p, _u8p, _u8p,
                                          ctypes.c_size_t]),
    "g16hip_valu_probe": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "g16hip_fixed_base": (ctypes.c_int, [_vp, ctypes.c_int, _u8p, ctypes.c_size_t, _u8p]),
    "g16hip_ceremony_open": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_size_t,
                                            ctypes.POINTER(_vp)]),
    "g16hip_ceremony_close": (None, [_vp]),
    "g16hip_ceremony_msm": (ctypes.c_int, [_vp, ctypes.c_int, _u8p, _u8p, ctypes.c_size_t, _u8p]),
    "g16hip_set_constant_work": (ctypes.c_int, [_vp, ctypes.c_int]),
    "g16hip_msm_scratch_bytes": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t)]),
    "g16hip_ceremony_last_runs": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "g16hip_ceremony_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(CeremonyStats)]),
    "g16hip_h_and_msms_from_rows": (ctypes.c_int, [_vp, _u8p, ctypes.c_size_t, _u8p, _u8p,
                                                   ctypes.c_size_t, ctypes.POINTER(MsmOut),
                                                   ctypes.POINTER(Timings)]),
    "g16hip_prove_from_rows": (ctypes.c_int, [_vp, _u8p, ctypes.c_size_t, _u8p, _u8p,
                                              ctypes.c_size_t, _u8p, _u8p,
                                              ctypes.POINTER(ProofOut), ctypes.POINTER(Timings)]),
    "g16hip_prove": (ctypes.c_int, [_vp, _u8p, ctypes.c_size_t, _u8p, _u8p,
                                    ctypes.POINTER(ProofOut), ctypes.POINTER(Timings)]),
    "g16hip_circuit_vk": (ctypes.c_int, [_vp, _u8p]),
    "g16hip_verify": (ctypes.c_int, [ctypes.POINTER(VkStruct), _u8p, ctypes.c_size_t,
                                     ctypes.POINTER(ProofOut)]),
    "g16hip_verify_allowing_unblinded": (ctypes.c_int, [ctypes.POINTER(VkStruct), _u8p,
                                                        ctypes.c_size_t, ctypes.POINTER(ProofOut)]),
    "g16hip_verify_unchecked": (ctypes.c_int, [ctypes.POINTER(VkStruct), _u8p, ctypes.c_size_t,
                                               ctypes.POINTER(ProofOut)]),
    "g16hip_pairing_check": (ctypes.c_int, [_u8p, _u8p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_int)]),
    "g16hip_pairing": (ctypes.c_int, [_u8p, _u8p, _u8p]),
    "g16hip_point_check": (ctypes.c_int, [ctypes.c_int, _u8p, ctypes.POINTER(ctypes.c_int)]),
    "g16hip_proof_to_json": (ctypes.c_int, [ctypes.POINTER(ProofOut), ctypes.c_char_p,
                                            ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "g16hip_public_to_json": (ctypes.c_int, [_u8p, ctypes.c_size_t, ctypes.c_char_p,
                                             ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "g16hip_proof_from_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t,
                                              ctypes.POINTER(ProofOut)]),
    "g16hip_public_from_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, _u8p,
                                               ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "g16hip_vk_from_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, _u8p,
                                           ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32)]),
}


def _declare(lib) -> None:
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue  # reported by exported_symbols()
        fn.restype = res
        fn.argtypes = args


def exported_symbols() -> dict:
    """Which of the declared entry points the loaded library exports."""
    lib = load()
    return {name: hasattr(lib, name) for name in EXPORTS}


def _check(rc: int) -> None:
    if rc != OK:
        msg = load().g16hip_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, ProveError)(rc, msg)


def _p8(a: np.ndarray):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the ABI must be C-contiguous"
    return a.ctypes.data_as(_u8p)


def _c8(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint8)


class Device:
    """An open GPU (``g16hip_device``)."""

    def __init__(self, ordinal: int = 0):
        self.lib = load()
        h = _vp()
        _check(self.lib.g16hip_open(ordinal, ctypes.byref(h)))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.g16hip_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def mem_info(self) -> tuple[int, int]:
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        _check(self.lib.g16hip_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def set_vram_limit(self, nbytes: int) -> None:
        """``with_vram_limit`` (crates/cuda/src/backend.rs:120-130): cap what circuits
        prepared afterwards may plan against (0 = no cap)."""
        _check(self.lib.g16hip_set_vram_limit(self.h, nbytes))

    def budget(self) -> int:
        """Bytes an allocation may plan against now (free less headroom, capped)."""
        b = ctypes.c_size_t()
        _check(self.lib.g16hip_device_budget(self.h, ctypes.byref(b)))
        return b.value

    def field_probe(self, field: int, op: int, a: np.ndarray, b: np.ndarray | None = None):
        w = 64 if field == 2 else 32
        a = _c8(a).reshape(-1, w)
        b = _c8(b).reshape(-1, w) if b is not None else None
        out = np.zeros_like(a)
        _check(self.lib.g16hip_field_probe(self.h, field, op, _p8(a), _p8(b), _p8(out), len(a)))
        return out

    def valu_probe(self, kind: int) -> float:
        """kind 0: v_mad_u64_u32 per second; kind 1: Fq products per second."""
        x = ctypes.c_double()
        _check(self.lib.g16hip_valu_probe(self.h, kind, ctypes.byref(x)))
        return x.value

    def fixed_base(self, group: int, k: np.ndarray) -> np.ndarray:
        k = _c8(k).reshape(-1, 32)
        out = np.zeros((len(k), 64 if group == 1 else 128), dtype=np.uint8)
        _check(self.lib.g16hip_fixed_base(self.h, group, _p8(k), len(k), _p8(out)))
        return out

    def msm(self, group: int, bases: np.ndarray, scalars: np.ndarray) -> np.ndarray:
        """MsmBackend::try_msm_g1/g2 (crates/msm/src/lib.rs:61-81)."""
        w = 64 if group == 1 else 128
        bases = _c8(bases).reshape(-1, w)
        scalars = _c8(scalars).reshape(-1, 32)
        if len(bases) != len(scalars):
            raise BackendError(BACKEND, f"{len(bases)} bases but {len(scalars)} scalars")
        out = np.zeros(w, dtype=np.uint8)
        fn = self.lib.g16hip_msm_g1 if group == 1 else self.lib.g16hip_msm_g2
        _check(fn(self.h, _p8(bases), _p8(scalars), len(bases), _p8(out)))
        return out

    def msm_device(self, group: int, d_bases: int, d_scalars: int, n: int, c: int = 0):
        w = 64 if group == 1 else 128
        out = np.zeros(w, dtype=np.uint8)
        fn = self.lib.g16hip_msm_g1_device if group == 1 else self.lib.g16hip_msm_g2_device
        _check(fn(self.h, _vp(d_bases), _vp(d_scalars), n, c, _p8(out)))
        return out

    def msm_device_noalloc(self, group: int, d_bases: int, d_scalars: int, n: int,
                           c: int = 0) -> np.ndarray:
        """msm_device with all scratch in the library's static device array (no allocation)."""
        out = np.zeros(64 if group == 1 else 128, dtype=np.uint8)
        _check(self.lib.g16hip_msm_device_noalloc(self.h, group, _vp(d_bases), _vp(d_scalars), n,
                                                  c, _p8(out)))
        return out

    def set_constant_work(self, on: bool) -> None:
        """Work::Constant for standalone MSMs and circuits prepared after the call."""
        _check(self.lib.g16hip_set_constant_work(self.h, int(on)))

    def msm_stats(self) -> dict:
        st = MsmStats()
        _check(self.lib.g16hip_msm_last_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def kernel_stats(self) -> dict:
        st = KernelStats()
        _check(self.lib.g16hip_kernel_stats_get(self.h, ctypes.byref(st)))
        return {"launches": list(st.launches), "entries": list(st.entries),
                "accumulate_ms": list(st.accumulate_ms)}

    def kernel_stats_reset(self) -> None:
        _check(self.lib.g16hip_kernel_stats_reset(self.h))

    def ntt(self, data: np.ndarray, kind: int = 0) -> np.ndarray:
        a = _c8(data).reshape(-1, 32).copy()
        log_n = len(a).bit_length() - 1
        if len(a) != 1 << log_n:
            raise BackendError(BACKEND, "NTT length must be a power of two")
        _check(self.lib.g16hip_ntt(self.h, _p8(a), log_n, kind))
        return a

    def ntt_device(self, d_data: int, log_n: int, kind: int = 0) -> None:
        _check(self.lib.g16hip_ntt_device(self.h, _vp(d_data), log_n, kind))

    def prepare(self, key: ProvingKey) -> "Circuit":
        return Circuit(self, key)


def key_resident_bytes(key: ProvingKey, window_tables: bool = False) -> int:
    """Device bytes ``key`` and one proof slot take once prepared (``resident_bytes``,
    crates/cuda/src/backend.rs:626-643), with or without the per-window base tables."""
    st = key.struct()
    b = ctypes.c_size_t()
    _check(load().g16hip_key_resident_bytes(ctypes.byref(st), int(window_tables), ctypes.byref(b)))
    return b.value


class HandleH:
    """``HPoly::Device``: H left on the GPU by compute_h, owned by its circuit."""

    def __init__(self, circuit: "Circuit", h):
        self.circuit = circuit
        self.h = h

    def release(self) -> None:
        if self.h:
            self.circuit.lib.g16hip_h_release(self.h)
            self.h = None

    def __del__(self):
        self.release()


class Circuit:
    """A prepared circuit (``g16hip_circuit``), the PreparedCircuit of the reference
    (crates/groth16/src/lib.rs:382-455)."""

    def __init__(self, dev: Device, key: ProvingKey):
        self.dev = dev
        self.lib = dev.lib
        self.key = key
        st = key.struct()
        h = _vp()
        _check(self.lib.g16hip_prepare(dev.h, ctypes.byref(st), ctypes.byref(h)))
        self.h = h
        nv, npub, dom = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(self.lib.g16hip_circuit_dims(h, ctypes.byref(nv), ctypes.byref(npub),
                                            ctypes.byref(dom)))
        self.n_vars, self.n_public, self.domain_size = nv.value, npub.value, dom.value

    def release(self) -> None:
        if getattr(self, "h", None):
            self.lib.g16hip_circuit_release(self.h)
            self.h = None

    def __del__(self):
        self.release()

    def resident_bytes(self) -> int:
        return int(self.lib.g16hip_circuit_resident_bytes(self.h))

    def set_batch_cap(self, cap: int) -> None:
        """Proofs per wave in g16hip_prove_batch (0 = the library's default)."""
        _check(self.lib.g16hip_set_batch_cap(self.h, cap))

    def batch_autotune(self, witness: np.ndarray, max_cap: int = 32) -> dict:
        """Measures batched proofs/s per wave size on this circuit and keeps the best."""
        w = _c8(witness).reshape(-1, 32)
        caps = (ctypes.c_uint32 * 8)()
        rates = (ctypes.c_double * 8)()
        tried, best = ctypes.c_size_t(), ctypes.c_uint32()
        _check(self.lib.g16hip_batch_autotune(self.h, _p8(w), len(w), max_cap, caps, rates, 8,
                                              ctypes.byref(tried), ctypes.byref(best)))
        return {"best": best.value,
                "tried": {int(caps[i]): float(rates[i]) for i in range(min(tried.value, 8))}}

    def add_slots(self, count: int, reserve_bytes: int = 8 << 30) -> int:
        """More proofs in flight on this circuit, admitted while device memory allows."""
        total = ctypes.c_uint32()
        _check(self.lib.g16hip_circuit_add_slots(self.h, count, reserve_bytes, ctypes.byref(total)))
        return total.value

    def compute_h(self, witness: np.ndarray, t: Timings | None = None) -> HandleH:
        w = _c8(witness).reshape(-1, 32)
        h = _vp()
        _check(self.lib.g16hip_compute_h(self.h, _p8(w), len(w), ctypes.byref(h),
                                         ctypes.byref(t or Timings())))
        return HandleH(self, h)

    def h_to_host(self, hh: HandleH) -> np.ndarray:
        out = np.zeros((self.domain_size, 32), dtype=np.uint8)
        _check(self.lib.g16hip_h_to_host(self.h, hh.h, _p8(out)))
        return out

    def msms(self, witness: np.ndarray, hh: HandleH, t: Timings | None = None) -> np.ndarray:
        w = _c8(witness).reshape(-1, 32)
        out = MsmOut()
        _check(self.lib.g16hip_msms(self.h, _p8(w), len(w), hh.h, ctypes.byref(out),
                                    ctypes.byref(t or Timings())))
        return out.as_array()

    def h_and_msms(self, witness: np.ndarray, t: Timings | None = None) -> np.ndarray:
        w = _c8(witness).reshape(-1, 32)
        out = MsmOut()
        _check(self.lib.g16hip_h_and_msms(self.h, _p8(w), len(w), ctypes.byref(out),
                                          ctypes.byref(t or Timings())))
        return out.as_array()

    def h_and_msms_device(self, d_witness: int, n: int, t: Timings | None = None) -> np.ndarray:
        out = MsmOut()
        _check(self.lib.g16hip_h_and_msms_device(self.h, _vp(d_witness), n, ctypes.byref(out),
                                                 ctypes.byref(t or Timings())))
        return out.as_array()

    def prove_with_blinders(self, witness: np.ndarray, r: np.ndarray, s: np.ndarray,
                            t: Timings | None = None) -> np.ndarray:
        w = _c8(witness).reshape(-1, 32)
        out = ProofOut()
        _check(self.lib.g16hip_prove_with_blinders(self.h, _p8(w), len(w), _p8(_c8(r)),
                                                   _p8(_c8(s)), ctypes.byref(out),
                                                   ctypes.byref(t or Timings())))
        return out.as_array()

    def prove(self, witness: np.ndarray, r: np.ndarray, s: np.ndarray,
              t: Timings | None = None) -> np.ndarray:
        """``prove`` (crates/groth16/src/prove.rs:12-41): the proof, checked against this
        circuit's verifying key before it is returned (``SelfVerify`` if it fails). r and s
        are the caller's random blinders (Fr Montgomery)."""
        w = _c8(witness).reshape(-1, 32)
        out = ProofOut()
        _check(self.lib.g16hip_prove(self.h, _p8(w), len(w), _p8(_c8(r)), _p8(_c8(s)),
                                     ctypes.byref(out), ctypes.byref(t or Timings())))
        return out.as_array()

    @staticmethod
    def _rows(a_rows, b_rows):
        a = _c8(a_rows).reshape(-1, 32)
        b = _c8(b_rows).reshape(-1, 32)
        if len(a) != len(b):
            raise BackendError(BACKEND, f"A.w has {len(a)} rows and B.w has {len(b)}")
        return a, b

    def h_and_msms_from_rows(self, witness: np.ndarray, a_rows, b_rows,
                             t: Timings | None = None) -> np.ndarray:
        """``h_and_msms_from_rows`` (crates/groth16/src/lib.rs:425-444): stage 0 from the
        caller's A.w and B.w (one per constraint, Fr Montgomery)."""
        w = _c8(witness).reshape(-1, 32)
        a, b = self._rows(a_rows, b_rows)
        out = MsmOut()
        _check(self.lib.g16hip_h_and_msms_from_rows(
            self.h, _p8(w), len(w), _p8(a) if len(a) else None, _p8(b) if len(b) else None,
            len(a), ctypes.byref(out), ctypes.byref(t or Timings())))
        return out.as_array()

    def prove_from_rows(self, witness: np.ndarray, a_rows, b_rows, r: np.ndarray,
                        s: np.ndarray, t: Timings | None = None) -> np.ndarray:
        """``prove_from_rows`` (crates/groth16/src/prove.rs:93-115), checked like ``prove``."""
        w = _c8(witness).reshape(-1, 32)
        a, b = self._rows(a_rows, b_rows)
        out = ProofOut()
        _check(self.lib.g16hip_prove_from_rows(
            self.h, _p8(w), len(w), _p8(a) if len(a) else None, _p8(b) if len(b) else None,
            len(a), _p8(_c8(r)), _p8(_c8(s)), ctypes.byref(out), ctypes.byref(t or Timings())))
        return out.as_array()

    def vk_bytes(self) -> np.ndarray:
        """``key().vk`` as alpha_g1 | beta_g2 | gamma_g2 | delta_g2 | ic (wire encoding)."""
        nv, npub, dom = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(self.lib.g16hip_circuit_dims(self.h, ctypes.byref(nv), ctypes.byref(npub),
                                            ctypes.byref(dom)))
        out = np.zeros(64 + 3 * 128 + 64 * (npub.value + 1), np.uint8)
        _check(self.lib.g16hip_circuit_vk(self.h, _p8(out)))
        return out

    def prove_with_blinders_device(self, d_witness: int, n: int, r: np.ndarray, s: np.ndarray,
                                   t: Timings | None = None) -> np.ndarray:
        out = ProofOut()
        _check(self.lib.g16hip_prove_with_blinders_device(self.h, _vp(d_witness), n,
                                                          _p8(_c8(r)), _p8(_c8(s)),
                                                          ctypes.byref(out),
                                                          ctypes.byref(t or Timings())))
        return out.as_array()


class CeremonyMsm:
    """The ceremony's ``MsmBackend`` (``CudaMsmBackend``, crates/cuda/src/ceremony.rs:65-312):
    MSMs over host slices, cut into runs of at most ``max_run`` points summed on the host,
    a run refused for memory halved, a sum off the curve re-run. ``vram_limit`` (bytes, 0 =
    none) refuses a run whose scratch would exceed it before allocating."""

    MIN_RUN = 1 << 10

    def __init__(self, dev: Device, max_run: int = 1 << 21, vram_limit: int = 0):
        self.lib = dev.lib
        self.dev = dev
        self.h = _vp()
        _check(self.lib.g16hip_ceremony_open(dev.h, max_run, vram_limit, ctypes.byref(self.h)))

    @staticmethod
    def name() -> str:
        return "hip"

    def _msm(self, group: int, bases, scalars) -> np.ndarray:
        w = 64 if group == 1 else 128
        bases = _c8(bases).reshape(-1, w)
        scalars = _c8(scalars).reshape(-1, 32)
        if len(bases) != len(scalars):
            raise BackendError(BACKEND, f"{len(bases)} bases but {len(scalars)} scalars")
        out = np.zeros(w, np.uint8)
        n = len(bases)
        _check(self.lib.g16hip_ceremony_msm(self.h, group, _p8(bases) if n else None,
                                            _p8(scalars) if n else None, n, _p8(out)))
        return out

    def msm_g1(self, bases, scalars) -> np.ndarray:
        return self._msm(1, bases, scalars)

    def msm_g2(self, bases, scalars) -> np.ndarray:
        return self._msm(2, bases, scalars)

    @staticmethod
    def scratch_bytes(group: int, n: int) -> int:
        """Device bytes one run of n points takes (what ``vram_limit`` is compared with)."""
        b = ctypes.c_size_t()
        _check(load().g16hip_msm_scratch_bytes(group, n, ctypes.byref(b)))
        return b.value

    def last_runs(self) -> int:
        r = ctypes.c_uint64()
        _check(self.lib.g16hip_ceremony_last_runs(self.h, ctypes.byref(r)))
        return r.value

    def stats(self) -> dict:
        st = CeremonyStats()
        _check(self.lib.g16hip_ceremony_stats_get(self.h, ctypes.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.g16hip_ceremony_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


def set_graph_replay(on: bool) -> None:
    """Batches and the server: replay captured graphs (True) or launch eagerly (False)."""
    load().g16hip_set_graph_replay(1 if on else 0)


def prove_batch(items, statuses: bool = False):
    """items: (circuit, witness, r, s) tuples (arrays as for prove_with_blinders). Returns the
    proofs in order and raises on the first failure; with statuses=True returns (proofs,
    per-item statuses) instead, a failed item's proof left zero."""
    lib = load()
    keep = []
    arr = (BatchItem * len(items))()
    for i, (c, w, r, s) in enumerate(items):
        w, r, s = _c8(w).reshape(-1, 32), _c8(r), _c8(s)
        keep += [w, r, s]
        arr[i] = BatchItem(c.h, _p8(w), len(w), _p8(r), _p8(s))
    out = (ProofOut * len(items))()
    st = (ctypes.c_int * len(items))()
    rc = lib.g16hip_prove_batch(arr, len(items), out, st)
    proofs = [out[i].as_array() for i in range(len(items))]
    if statuses:
        return proofs, [int(st[i]) for i in range(len(items))]
    _check(rc)
    return proofs


class Server:
    """The continuous-batching serving loop (``g16hip_server``)."""

    def __init__(self, dev: Device, reserve_bytes: int = 8 << 30, lanes: int = 4):
        self.lib = dev.lib
        self._held = {}  # ticket -> witness array, alive until the ticket is waited for
        h = _vp()
        _check(self.lib.g16hip_server_open(dev.h, reserve_bytes, lanes, ctypes.byref(h)))
        self.h = h

    def submit(self, circ: Circuit, witness, r, s) -> int:
        w = _c8(witness).reshape(-1, 32)
        t = ctypes.c_uint64()
        _check(self.lib.g16hip_server_submit(self.h, circ.h, _p8(w), len(w), _p8(_c8(r)),
                                             _p8(_c8(s)), ctypes.byref(t)))
        self._held[t.value] = w
        return t.value

    def wait(self, ticket: int):
        """(proof bytes, latency in ms)."""
        out = ProofOut()
        lat = ctypes.c_double()
        try:
            _check(self.lib.g16hip_server_wait(self.h, ticket, ctypes.byref(out),
                                               ctypes.byref(lat)))
        finally:
            self._held.pop(ticket, None)
        return out.as_array(), lat.value

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.g16hip_server_close(self.h)
            self.h = None

    def __del__(self):
        self.close()
