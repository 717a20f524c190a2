"""The product's synthetic workload generator (include/g16synth.h, host code) is
bit-identical to the oracle's construction: circuit, witness, toxic waste and every query
scalar (CPU only; the GPU half, k * G, is pinned in test_gpu_field.py)."""
import numpy as np
import pytest

from oracle.oracle import Synth as OracleSynth
from snarkrs_amd.synth import Synth, Workload, scalars
from snarkrs_amd.wire import SplitMix64, fr_array
from tests import pyref


@pytest.mark.parametrize("w", [
    Workload(n_constraints=1, n_public=0, n_private_in=1, seed=3, setup_seed=4),
    Workload(n_constraints=37, n_public=3, n_private_in=4, bool_permille=400,
             input_bool_permille=400, seed=5, setup_seed=6),
    Workload(n_constraints=5000, n_public=2, n_private_in=50, bool_permille=250, terms=4,
             seed=7, setup_seed=8),
])
def test_product_synth_matches_oracle(orc, w):
    mine = Synth(w)
    ref = OracleSynth(orc, w.n_constraints, w.n_public, w.n_private_in, w.bool_permille,
                      w.input_bool_permille, w.terms, w.seed, w.setup_seed)
    assert (mine.n_vars, mine.domain, mine.nnz_a, mine.nnz_b, mine.nnz_c) == \
        (ref.n_vars, ref.domain, ref.nnz_a, ref.nnz_b, ref.nnz_c)
    a, b = mine.arrays(), ref.arrays()
    for k in a:
        if isinstance(a[k], list):
            for x, y in zip(a[k], b[k]):
                assert (x == y).all(), k
        else:
            assert (a[k] == b[k]).all(), k


def test_scalar_generator_matches_python_reference():
    got = scalars(99, 500, 300)
    rng = SplitMix64(99)
    assert (got == fr_array(pyref.mixed_scalars(rng, 500, 300))).all()
