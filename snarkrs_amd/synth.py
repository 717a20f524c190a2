"""Synthetic Groth16 workloads (include/g16synth.h) on the product path: the circuit, witness
and query scalars come from the host half of libg16hip, the key points from its GPU fixed-base
kernel. Used by bench.py; tests hold it bit-identical to the oracle's construction."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from .key import ProvingKey
from .native import Device, load


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


def _declare(lib):
    if getattr(lib, "_synth_declared", False):
        return
    lib.g16synth_new.restype = ctypes.c_void_p
    lib.g16synth_new.argtypes = [ctypes.POINTER(SynthParams)]
    lib.g16synth_free.argtypes = [ctypes.c_void_p]
    lib.g16synth_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
    lib.g16synth_get.restype = ctypes.c_int
    lib.g16synth_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.g16synth_witness.restype = ctypes.c_int
    lib.g16synth_witness.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint8)]
    lib.g16synth_scalars.argtypes = [ctypes.c_uint64, ctypes.c_size_t, ctypes.c_uint32,
                                     ctypes.POINTER(ctypes.c_uint8)]
    lib._synth_declared = True


@dataclass
class Workload:
    """One synthetic circuit configuration."""

    n_constraints: int
    n_public: int = 2
    n_private_in: int = 8
    bool_permille: int = 0
    input_bool_permille: int = 0
    terms: int = 3
    seed: int = 1
    setup_seed: int = 2
    witness_seed: int = 3


# The plain-prove configuration: 2^16 multiplication-chain constraints, uniform inputs.
PLAIN_2_16 = Workload(n_constraints=1 << 16, n_public=2, n_private_in=14, seed=16, setup_seed=17,
                      witness_seed=18)
# The 2^20-constraint class: 1.1M constraints (domain 2^21), about 1.2M wires, 3 nonzeros per
# A/B row, 40% of wires boolean.
LARGE_1_1M = Workload(n_constraints=1_100_000, n_public=2, n_private_in=100_000,
                      bool_permille=400, input_bool_permille=400, seed=21, setup_seed=22,
                      witness_seed=23)


def mixed_workloads(count: int, lo: float, hi: float, seed: int, base_seed: int) -> list:
    """`count` circuits whose log2 constraint count is drawn uniformly in [lo, hi] (seeded
    SplitMix64): 40% boolean witness, 3 nonzeros per A/B row, about 9% of the wires private
    inputs. The batch workload is (4, 14, 20, seed=3, base_seed=500), the serving one
    (8, 16, 22, seed=4, base_seed=700)."""
    from .wire import SplitMix64

    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        x = lo + (hi - lo) * (rng.next_u64() / 2.0**64)
        nc = int(2.0**x)
        out.append(Workload(n_constraints=nc, n_public=2, n_private_in=max(8, nc // 11),
                            bool_permille=400, input_bool_permille=400, seed=base_seed + i,
                            setup_seed=base_seed + 100 + i, witness_seed=base_seed + 200 + i))
    return out


BATCH_WORKLOADS = dict(count=4, lo=14.0, hi=20.0, seed=3, base_seed=500)
SERVE_WORKLOADS = dict(count=8, lo=16.0, hi=22.0, seed=4, base_seed=700)


class Synth:
    def __init__(self, w: Workload):
        self.lib = load()
        _declare(self.lib)
        self.w = w
        p = SynthParams(w.n_constraints, w.n_public, w.n_private_in, w.bool_permille,
                        w.input_bool_permille, w.terms, w.seed, w.setup_seed, w.witness_seed)
        self.h = self.lib.g16synth_new(ctypes.byref(p))
        if not self.h:
            raise ValueError(f"bad synthetic workload {w}")
        d = (ctypes.c_uint32 * 7)()
        self.lib.g16synth_dims(self.h, d)
        (self.n_vars, self.n_public, self.domain, self.n_constraints,
         self.nnz_a, self.nnz_b, self.nnz_c) = list(d)

    def close(self):
        if getattr(self, "h", None):
            self.lib.g16synth_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _get(self, what, n, fr=True):
        out = np.zeros((n, 32), np.uint8) if fr else np.zeros(n, np.uint32)
        rc = self.lib.g16synth_get(self.h, what, out.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0
        return out

    def arrays(self) -> dict:
        n, nv = self.domain, self.n_vars
        nnz = [self.nnz_a, self.nnz_b, self.nnz_c]
        return {
            "row_ptr": [self._get(m, n + 1, False) for m in range(3)],
            "signal": [self._get(3 + m, nnz[m], False) for m in range(3)],
            "values": [self._get(6 + m, nnz[m]) for m in range(3)],
            "witness": self._get(9, nv),
            "toxic": self._get(10, 5),
            "u": self._get(11, nv),
            "v": self._get(12, nv),
            "ic_s": self._get(13, self.n_public + 1),
            "l_s": self._get(14, nv - self.n_public - 1),
            "h_s": self._get(15, n),
        }

    def key(self, dev: Device) -> tuple[ProvingKey, np.ndarray]:
        """Proving key with points k * G from the GPU fixed-base kernel, and the witness."""
        a = self.arrays()
        tau, alpha, beta, gamma, delta = a["toxic"]
        g1 = lambda k: dev.fixed_base(1, k)
        g2 = lambda k: dev.fixed_base(2, k)
        vals = np.ascontiguousarray(np.concatenate([a["values"][0], a["values"][1]]))
        vid = [np.arange(self.nnz_a, dtype=np.uint32),
               np.arange(self.nnz_a, self.nnz_a + self.nnz_b, dtype=np.uint32)]
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
        """Another satisfying witness of the same circuit (same key)."""
        out = np.zeros((self.n_vars, 32), np.uint8)
        rc = self.lib.g16synth_witness(self.h, witness_seed,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        assert rc == 0
        return out


def scalars(seed: int, n: int, zero_one_permille: int = 300) -> np.ndarray:
    """Seeded Fr scalars (Montgomery): a share in {0, 1}, the rest uniform."""
    lib = load()
    _declare(lib)
    out = np.zeros((n, 32), np.uint8)
    lib.g16synth_scalars(seed, n, zero_one_permille, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out
