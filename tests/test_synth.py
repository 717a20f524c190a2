"""The product's synthetic workload generator (include/g16synth.h, host code) is
bit-identical to the oracle's construction: circuit, witness, toxic waste and every query
scalar (CPU only; the GPU half, k * G, is pinned in test_gpu_field.py)."""
import numpy as np
import pytest

from oracle.oracle import Synth as OracleSynth
from snarkrs_amd.synth import Synth, Workload, scalars
from snarkrs_amd.wire import R_MOD, SplitMix64, fr_array, fr_ints
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
                      w.input_bool_permille, w.terms, w.seed, w.setup_seed, w.witness_seed)
    assert (mine.n_vars, mine.domain, mine.nnz_a, mine.nnz_b, mine.nnz_c) == \
        (ref.n_vars, ref.domain, ref.nnz_a, ref.nnz_b, ref.nnz_c)
    a, b = mine.arrays(), ref.arrays()
    for k in a:
        if isinstance(a[k], list):
            for x, y in zip(a[k], b[k]):
                assert (x == y).all(), k
        else:
            assert (a[k] == b[k]).all(), k
    for ws in (0, 77):
        assert (mine.witness(ws) == ref.witness(ws)).all(), ws
    assert (mine.witness(w.witness_seed) == a["witness"]).all()


def _check_r1cs(arr, wit, n_constraints, n_public, terms):
    """(A w)(B w) = (C w) on every row, A and B rows of the constraints have `terms` entries,
    and the public tail rows are A[N + s] = [s]. Python integers, independent of both C++
    constructions."""
    w = fr_ints(wit)
    assert w[0] == 1
    rows = []
    for m in range(3):
        rp, sig, val = arr["row_ptr"][m], arr["signal"][m], fr_ints(arr["values"][m])
        dot = []
        for c in range(len(rp) - 1):
            acc = 0
            for e in range(rp[c], rp[c + 1]):
                acc += val[e] * w[sig[e]]
            dot.append(acc % R_MOD)
        rows.append(dot)
        if m < 2:
            for c in range(n_constraints):
                assert rp[c + 1] - rp[c] == terms, (m, c)
    A, B, C = rows
    for c in range(len(A)):
        assert A[c] * B[c] % R_MOD == C[c], c
    for s in range(n_public + 1):
        assert A[n_constraints + s] == w[s]


def test_every_witness_seed_satisfies_the_same_circuit():
    w = Workload(n_constraints=300, n_public=3, n_private_in=6, bool_permille=400,
                 input_bool_permille=400, terms=3, seed=11, setup_seed=12, witness_seed=13)
    syn = Synth(w)
    arr = syn.arrays()
    bits = 0
    for ws in (13, 14, 15):
        wit = arr["witness"] if ws == 13 else syn.witness(ws)
        _check_r1cs(arr, wit, w.n_constraints, w.n_public, w.terms)
        vals = fr_ints(wit)
        out = vals[1 + w.n_public + w.n_private_in:]
        bits = sum(v in (0, 1) for v in out)
    # about 40% of the constraint outputs are bits (Bernoulli, 300 draws)
    assert 80 <= bits <= 160
    assert not (syn.witness(14) == syn.witness(15)).all()


def test_even_terms_pad_with_a_zero_entry():
    w = Workload(n_constraints=64, n_public=1, n_private_in=3, bool_permille=500, terms=4,
                 seed=21, setup_seed=22, witness_seed=23)
    syn = Synth(w)
    _check_r1cs(syn.arrays(), syn.witness(5), w.n_constraints, w.n_public, w.terms)


def test_scalar_generator_matches_python_reference():
    got = scalars(99, 500, 300)
    rng = SplitMix64(99)
    assert (got == fr_array(pyref.mixed_scalars(rng, 500, 300))).all()
