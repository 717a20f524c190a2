"""ctypes wrapper of liboracle.so, the CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import this
module. It is the checker, never the thing measured or shipped: the product path
(``snarkrs_amd``) never imports it and fails loudly when its HIP library is missing.

See oracle/oracle.cpp for which reference file each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from snarkrs_amd.key import ProvingKey
from snarkrs_amd.wire import R_MOD, fr_array

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u8p = ctypes.POINTER(ctypes.c_uint8)


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE], check=True)


class SynthParams(ctypes.Structure):
    _fields_ = [
        ("n_constraints", ctypes.c_uint32),
        ("n_public", ctypes.c_uint32),
        ("n_private_in", ctypes.c_uint32),
        ("bool_permille", ctypes.c_uint32),
        ("input_bool_permille", ctypes.c_uint32),
        ("terms", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("setup_seed", ctypes.c_uint64),
        ("witness_seed", ctypes.c_uint64),
    ]


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p)


class Oracle:
    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            build()
        self.lib = lib = ctypes.CDLL(path)
        lib.orc_window_size.restype = ctypes.c_int
        lib.orc_window_size.argtypes = [ctypes.c_uint64]
        lib.orc_signed_digit.restype = ctypes.c_int64
        lib.orc_signed_digit.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, ctypes.c_int]
        lib.orc_synth_new.restype = ctypes.c_void_p
        lib.orc_synth_new.argtypes = [ctypes.POINTER(SynthParams)]
        lib.orc_synth_free.argtypes = [ctypes.c_void_p]
        lib.orc_synth_witness.argtypes = [ctypes.c_void_p, ctypes.c_uint64, _u8p]
        lib.orc_synth_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        lib.orc_synth_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        lib.orc_shadow.argtypes = [ctypes.c_void_p, _u8p, _u8p, _u8p]
        lib.orc_threads.restype = ctypes.c_int
        for name in ("orc_field_op", "orc_ntt", "orc_coset_pipeline", "orc_on_curve",
                     "orc_scalar_mul", "orc_point_add", "orc_fixed_base", "orc_msm",
                     "orc_gather", "orc_compute_h", "orc_prove", "orc_prove_rows"):
            getattr(lib, name).restype = ctypes.c_int
        lib.orc_field_op.argtypes = [ctypes.c_int, ctypes.c_int, _u8p, _u8p, _u8p, ctypes.c_size_t]
        lib.orc_ntt.argtypes = [_u8p, ctypes.c_int, ctypes.c_int]
        lib.orc_coset_pipeline.argtypes = [_u8p, ctypes.c_int]
        lib.orc_on_curve.argtypes = [ctypes.c_int, _u8p, ctypes.c_size_t]
        lib.orc_scalar_mul.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_size_t, _u8p]
        lib.orc_point_add.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_size_t, _u8p]
        lib.orc_fixed_base.argtypes = [ctypes.c_int, _u8p, ctypes.c_size_t, _u8p]
        lib.orc_msm.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_size_t, ctypes.c_int,
                                ctypes.c_int, _u8p]
        lib.orc_gather.argtypes = [ctypes.c_void_p, ctypes.c_int, _u8p, _u8p]
        lib.orc_compute_h.argtypes = [ctypes.c_void_p, _u8p, _u8p]
        lib.orc_prove.argtypes = [ctypes.c_void_p, _u8p, _u8p, _u8p, _u8p, _u8p]
        lib.orc_prove_rows.argtypes = [ctypes.c_void_p, _u8p, _u8p, _u8p, ctypes.c_size_t, _u8p,
                                       _u8p, _u8p, _u8p]
        lib.orc_domain_root.argtypes = [ctypes.c_int, _u8p]
        lib.orc_g1_generator.argtypes = [_u8p]
        lib.orc_g2_generator.argtypes = [_u8p]
        lib.orc_set_threads.argtypes = [ctypes.c_int]

    # -- knobs ----------------------------------------------------------------------------
    def set_threads(self, t: int) -> None:
        self.lib.orc_set_threads(t)

    def threads(self) -> int:
        return self.lib.orc_threads()

    def window_size(self, n: int) -> int:
        return self.lib.orc_window_size(n)

    def signed_digit(self, scalar_std: int, i: int, c: int) -> int:
        limbs = (ctypes.c_uint64 * 4)(*[(scalar_std >> (64 * k)) & (2**64 - 1) for k in range(4)])
        return self.lib.orc_signed_digit(limbs, i, c)

    # -- field ----------------------------------------------------------------------------
    def field_op(self, field: int, op: int, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
        width = 64 if field == 2 else 32
        a = np.ascontiguousarray(a, dtype=np.uint8).reshape(-1, width)
        out = np.zeros_like(a)
        bp = _p(np.ascontiguousarray(b, dtype=np.uint8).reshape(-1, width)) if b is not None else None
        rc = self.lib.orc_field_op(field, op, _p(a), bp, _p(out), len(a))
        assert rc == 0
        return out

    def domain_root(self, log_n: int) -> np.ndarray:
        out = np.zeros(32, dtype=np.uint8)
        self.lib.orc_domain_root(log_n, _p(out))
        return out

    # -- curve ----------------------------------------------------------------------------
    def generator(self, group: int) -> np.ndarray:
        out = np.zeros(64 if group == 1 else 128, dtype=np.uint8)
        (self.lib.orc_g1_generator if group == 1 else self.lib.orc_g2_generator)(_p(out))
        return out

    def on_curve(self, group: int, pts: np.ndarray) -> bool:
        pts = np.ascontiguousarray(pts, dtype=np.uint8)
        n = pts.size // (64 if group == 1 else 128)
        return bool(self.lib.orc_on_curve(group, _p(pts), n))

    def scalar_mul(self, group: int, pts: np.ndarray, k: np.ndarray) -> np.ndarray:
        w = 64 if group == 1 else 128
        pts = np.ascontiguousarray(pts, dtype=np.uint8).reshape(-1, w)
        k = np.ascontiguousarray(k, dtype=np.uint8).reshape(-1, 32)
        out = np.zeros_like(pts)
        self.lib.orc_scalar_mul(group, _p(pts), _p(k), len(pts), _p(out))
        return out

    def point_add(self, group: int, p: np.ndarray, q: np.ndarray) -> np.ndarray:
        w = 64 if group == 1 else 128
        p = np.ascontiguousarray(p, dtype=np.uint8).reshape(-1, w)
        q = np.ascontiguousarray(q, dtype=np.uint8).reshape(-1, w)
        out = np.zeros_like(p)
        self.lib.orc_point_add(group, _p(p), _p(q), len(p), _p(out))
        return out

    def fixed_base(self, group: int, k: np.ndarray) -> np.ndarray:
        k = np.ascontiguousarray(k, dtype=np.uint8).reshape(-1, 32)
        out = np.zeros((len(k), 64 if group == 1 else 128), dtype=np.uint8)
        self.lib.orc_fixed_base(group, _p(k), len(k), _p(out))
        return out

    def msm(self, group: int, bases: np.ndarray, scalars: np.ndarray, naive: bool = False,
            c: int = 0) -> np.ndarray:
        w = 64 if group == 1 else 128
        bases = np.ascontiguousarray(bases, dtype=np.uint8).reshape(-1, w)
        scalars = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1, 32)
        assert len(bases) == len(scalars)
        out = np.zeros(w, dtype=np.uint8)
        self.lib.orc_msm(group, _p(bases), _p(scalars), len(bases), int(naive), c, _p(out))
        return out

    # -- NTT ------------------------------------------------------------------------------
    def ntt(self, data: np.ndarray, inverse: bool = False) -> np.ndarray:
        a = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1, 32).copy()
        log_n = len(a).bit_length() - 1
        assert 1 << log_n == len(a)
        self.lib.orc_ntt(_p(a), log_n, int(inverse))
        return a

    def coset_pipeline(self, data: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1, 32).copy()
        log_n = len(a).bit_length() - 1
        self.lib.orc_coset_pipeline(_p(a), log_n)
        return a

    # -- prove ----------------------------------------------------------------------------
    def gather(self, key: ProvingKey, m: int, witness: np.ndarray) -> np.ndarray:
        s = key.struct()
        out = np.zeros((key.domain_size, 32), dtype=np.uint8)
        self.lib.orc_gather(ctypes.byref(s), m, _p(witness), _p(out))
        return out

    def compute_h(self, key: ProvingKey, witness: np.ndarray) -> np.ndarray:
        s = key.struct()
        out = np.zeros((key.domain_size, 32), dtype=np.uint8)
        rc = self.lib.orc_compute_h(ctypes.byref(s), _p(witness), _p(out))
        assert rc == 0
        return out

    def prove_rows(self, key: ProvingKey, witness: np.ndarray, a_rows: np.ndarray,
                   b_rows: np.ndarray, r: int, s: int):
        """prove_from_rows with fixed blinders: (msm outputs, proof)."""
        st = key.struct()
        a_rows = np.ascontiguousarray(a_rows, dtype=np.uint8).reshape(-1, 32)
        b_rows = np.ascontiguousarray(b_rows, dtype=np.uint8).reshape(-1, 32)
        assert len(a_rows) == len(b_rows)
        rb, sb = fr_array([r]), fr_array([s])
        msm = np.zeros(384, dtype=np.uint8)
        proof = np.zeros(256, dtype=np.uint8)
        rc = self.lib.orc_prove_rows(ctypes.byref(st), _p(np.ascontiguousarray(witness)),
                                     _p(a_rows), _p(b_rows), len(a_rows), _p(rb), _p(sb),
                                     _p(msm), _p(proof))
        if rc != 0:
            raise ValueError(f"orc_prove_rows: {rc}")
        return msm, proof

    def prove(self, key: ProvingKey, witness: np.ndarray, r: int, s: int):
        """Returns (msm outputs, proof) as 384- and 256-byte arrays."""
        st = key.struct()
        rb, sb = fr_array([r]), fr_array([s])
        msm = np.zeros(384, dtype=np.uint8)
        proof = np.zeros(256, dtype=np.uint8)
        rc = self.lib.orc_prove(ctypes.byref(st), _p(np.ascontiguousarray(witness)), _p(rb),
                                _p(sb), _p(msm), _p(proof))
        assert rc == 0
        return msm, proof


class Synth:
    """A synthetic circuit + witness + toxic waste built by the oracle (the construction is
    specified in DESIGN.md and mirrored by snarkrs_amd/csrc/synth.cpp)."""

    def __init__(self, orc: Oracle, n_constraints: int, n_public: int = 2, n_private_in: int = 8,
                 bool_permille: int = 0, input_bool_permille: int = 0, terms: int = 3,
                 seed: int = 1, setup_seed: int = 2, witness_seed: int = 3):
        self.orc = orc
        self.params = SynthParams(n_constraints, n_public, n_private_in, bool_permille,
                                  input_bool_permille, terms, seed, setup_seed, witness_seed)
        self.h = orc.lib.orc_synth_new(ctypes.byref(self.params))
        d = (ctypes.c_uint32 * 7)()
        orc.lib.orc_synth_dims(self.h, d)
        (self.n_vars, self.n_public, self.domain, self.n_constraints,
         self.nnz_a, self.nnz_b, self.nnz_c) = list(d)

    def __del__(self):
        if getattr(self, "h", None):
            self.orc.lib.orc_synth_free(self.h)
            self.h = None

    def _get(self, what: int, n: int, width: int, dtype=np.uint8) -> np.ndarray:
        if dtype == np.uint32:
            out = np.zeros(n, dtype=np.uint32)
        else:
            out = np.zeros((n, width), dtype=np.uint8)
        self.orc.lib.orc_synth_get(self.h, what, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def arrays(self) -> dict:
        n, nv = self.domain, self.n_vars
        nnz = [self.nnz_a, self.nnz_b, self.nnz_c]
        d = {
            "row_ptr": [self._get(m, n + 1, 4, np.uint32) for m in range(3)],
            "signal": [self._get(3 + m, nnz[m], 4, np.uint32) for m in range(3)],
            "values": [self._get(6 + m, nnz[m], 32) for m in range(3)],
            "witness": self._get(9, nv, 32),
            "toxic": self._get(10, 5, 32),
            "u": self._get(11, nv, 32),
            "v": self._get(12, nv, 32),
            "ic_s": self._get(13, self.n_public + 1, 32),
            "l_s": self._get(14, nv - self.n_public - 1, 32),
            "h_s": self._get(15, n, 32),
        }
        return d

    def key(self) -> tuple[ProvingKey, np.ndarray]:
        """Proving key (points by the oracle's fixed-base multiplication) and witness."""
        a = self.arrays()
        orc = self.orc
        tau, alpha, beta, gamma, delta = a["toxic"]
        vals = np.ascontiguousarray(np.concatenate([a["values"][0], a["values"][1]]))
        vid = [np.arange(self.nnz_a, dtype=np.uint32),
               np.arange(self.nnz_a, self.nnz_a + self.nnz_b, dtype=np.uint32)]
        g1 = lambda k: orc.fixed_base(1, k)
        g2 = lambda k: orc.fixed_base(2, k)
        key = ProvingKey(
            n_vars=self.n_vars, n_public=self.n_public, domain_size=self.domain,
            alpha_g1=g1(alpha)[0], beta_g1=g1(beta)[0], beta_g2=g2(beta)[0],
            gamma_g2=g2(gamma)[0], delta_g1=g1(delta)[0], delta_g2=g2(delta)[0],
            ic=g1(a["ic_s"]), row_ptr=a["row_ptr"][:2], signal=a["signal"][:2],
            value_id=vid, values=vals, a_query=g1(a["u"]), b_g1_query=g1(a["v"]),
            b_g2_query=g2(a["v"]), l_query=g1(a["l_s"]), h_query=g1(a["h_s"]),
        )
        return key, a["witness"]

    def witness(self, witness_seed: int) -> np.ndarray:
        """Another satisfying witness of the same circuit."""
        out = np.zeros((self.n_vars, 32), dtype=np.uint8)
        self.orc.lib.orc_synth_witness(self.h, witness_seed, _p(out))
        return out

    def shadow(self, r: int, s: int) -> list[int]:
        """Closed-form scalars: msm A, B, L, H, pi_a, pi_b, pi_c, verify lhs, verify rhs."""
        from snarkrs_amd.wire import fr_ints
        out = np.zeros((9, 32), dtype=np.uint8)
        self.orc.lib.orc_shadow(self.h, _p(fr_array([r])), _p(fr_array([s])), _p(out))
        return fr_ints(out)


__all__ = ["Oracle", "Synth", "build", "R_MOD"]
