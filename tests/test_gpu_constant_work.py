"""Constant work (opt in; Work::Constant, crates/cuda/src/msm.rs:143-164, the reference's
constant_work_device_accounting cases, crates/cuda/src/msm.rs:6000-6033): every scalar emits in
every window, so the entry count is n * windows whatever the scalars (all zero, all one,
alternating, sparse -1, random), and every MSM and proof is bit-identical to the default
path and to the oracle. The cases cover the complete formulas' exceptional inputs (every
scalar 1: one bucket spanning every segment, folded over several levels; alternating: the
identity and repeated points), sizes from one point (two slots for the closing scan) to ones
whose fold runs several levels before it, and a batch of constant-work proofs (proved on
slots)."""
import numpy as np
import pytest

from oracle.oracle import Synth
from snarkrs_amd import native
from snarkrs_amd.wire import R_MOD, SplitMix64, fr_array

pytestmark = pytest.mark.gpu


@pytest.fixture
def constant(dev):
    dev.set_constant_work(True)
    try:
        yield dev
    finally:
        dev.set_constant_work(False)


def _cases(n, seed):
    rng = SplitMix64(seed)
    return {
        "zeros": [0] * n,
        "ones": [1] * n,
        "alternating": [i % 2 for i in range(n)],
        "sparse -1": [R_MOD - 1 if i % 64 == 0 else 0 for i in range(n)],
        "random": [rng.next_fr() for _ in range(n)],
    }


@pytest.mark.parametrize("group", [1, 2])
@pytest.mark.parametrize("n", [1, 2, 129, 5000])
def test_entries_do_not_depend_on_the_scalars(dev, orc, constant, group, n):
    rng = SplitMix64(800 + n + group)
    bases = dev.fixed_base(group, fr_array([rng.next_fr() for _ in range(n)]))
    counts = set()
    for name, ss in _cases(n, 900 + n).items():
        sc = fr_array(ss)
        got = dev.msm(group, bases, sc)
        st = dev.msm_stats()
        counts.add(st["entries"])
        assert st["entries"] == n * st["windows"], name
        assert (got == orc.msm(group, bases, sc)).all(), name
    assert len(counts) == 1


def test_variable_work_counts_only_nonzero_digits(dev):
    n = 129
    bases = dev.fixed_base(1, fr_array(list(range(1, n + 1))))
    dev.msm(1, bases, fr_array([0] * n))
    assert dev.msm_stats()["entries"] == 0
    dev.msm(1, bases, fr_array([1] * n))
    assert dev.msm_stats()["entries"] == n


def test_constant_work_proofs_match_the_oracle(dev, orc, constant):
    syn = Synth(orc, 3000, n_public=2, n_private_in=20, bool_permille=400,
                input_bool_permille=400, seed=81, setup_seed=82)
    key, wit = syn.key()
    circ = constant.prepare(key)
    r, s = fr_array([0x1234567]), fr_array([0x7654321])
    msm_o, proof_o = orc.prove(key, wit, 0x1234567, 0x7654321)
    assert (circ.h_and_msms(wit) == msm_o).all()
    assert (circ.prove(wit, r, s) == proof_o).all()


def test_constant_work_batch_matches_single_proofs(dev, orc, constant):
    syn = Synth(orc, 2000, n_public=2, n_private_in=20, bool_permille=400,
                input_bool_permille=400, seed=83, setup_seed=84)
    key, wit = syn.key()
    circ = constant.prepare(key)
    rs = [(fr_array([0x1000 + k]), fr_array([0x2000 + k])) for k in range(3)]
    proofs = native.prove_batch([(circ, wit, r, s) for r, s in rs])
    for (r, s), p in zip(rs, proofs):
        assert (p == circ.prove_with_blinders(wit, r, s)).all()

