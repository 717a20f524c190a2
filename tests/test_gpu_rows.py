"""The rows path on the GPU (h_and_msms_from_rows, prove_from_rows: crates/groth16/src/lib.rs
:425-444, crates/groth16/src/prove.rs:93-115) in both H bases, against the oracle bit for bit:
snarkjs' basis from rows is the gather path's proof; gnark's basis (H's coefficients in
bit-reversed order against [tau^bitrev(j) t(tau)/delta]_1, converted on the device) puts the
same point into the proof. Also the refusals: the gather path on a gnark key, too many rows,
rows that are not the witness's (the check of the result)."""
import numpy as np
import pytest

from oracle.oracle import Synth
from snarkrs_amd import native
from snarkrs_amd.wire import fr_array
from tests import rowsref

pytestmark = pytest.mark.gpu

R_BLIND = 0x1234567890ABCDEF1234567890ABCDEF
S_BLIND = 0xFEDCBA0987654321FEDCBA0987654321


@pytest.mark.parametrize("nc", [1, 7, 300, 5000])
def test_rows_proofs_match_the_oracle_in_both_bases(dev, orc, nc):
    syn = Synth(orc, nc, n_public=3, n_private_in=5, bool_permille=400, input_bool_permille=400,
                seed=60 + nc, setup_seed=61)
    key, wit = syn.key()
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])
    msm_o, proof_o = orc.prove(key, wit, R_BLIND, S_BLIND)
    a, b = rowsref.rows(orc, key, wit)
    circ = dev.prepare(key)
    assert (circ.h_and_msms_from_rows(wit, a, b) == msm_o).all()
    assert (circ.prove_from_rows(wit, a, b, r, s) == proof_o).all()
    assert (circ.prove_with_blinders(wit, r, s) == proof_o).all()  # the gather path still works
    gk = rowsref.gnark_key(key, syn.arrays()["toxic"], orc.fixed_base)
    msm_g, proof_g = orc.prove_rows(gk, wit, a, b, R_BLIND, S_BLIND)
    assert (msm_g == msm_o).all() and (proof_g == proof_o).all()
    gc = dev.prepare(gk)
    assert (gc.h_and_msms_from_rows(wit, a, b) == msm_o).all()
    assert (gc.prove_from_rows(wit, a, b, r, s) == proof_o).all()
    full_a, full_b = rowsref.rows(orc, key, wit, trim=False)
    assert (gc.prove_from_rows(wit, full_a, full_b, r, s) == proof_o).all()


def test_rows_path_refusals(dev, orc):
    syn = Synth(orc, 300, n_public=2, n_private_in=4, seed=71, setup_seed=72)
    key, wit = syn.key()
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])
    a, b = rowsref.rows(orc, key, wit, trim=False)
    gk = rowsref.gnark_key(key, syn.arrays()["toxic"], orc.fixed_base)
    gc = dev.prepare(gk)
    with pytest.raises(native.BackendError, match="gnark key carries no constraint matrices"):
        gc.prove_with_blinders(wit, r, s)
    with pytest.raises(native.BackendError, match="gnark key carries no constraint matrices"):
        gc.compute_h(wit)
    extra = np.zeros((1, 32), np.uint8)
    with pytest.raises(native.BackendError, match="at most the domain size"):
        gc.prove_from_rows(wit, np.concatenate([a, extra]), np.concatenate([b, extra]), r, s)
    with pytest.raises(native.WitnessLength):
        gc.prove_from_rows(wit[:-1], a, b, r, s)
    bad = a.copy()
    bad[3] = fr_array([12345])[0]
    with pytest.raises(native.SelfVerify):
        gc.prove_from_rows(wit, bad, b, r, s)
    # The same key keeps proving after the refusals.
    assert (gc.prove_from_rows(wit, a, b, r, s) == orc.prove(key, wit, R_BLIND, S_BLIND)[1]).all()


def test_rows_on_a_gpu_built_2_16_key(dev, orc):
    from snarkrs_amd.synth import PLAIN_2_16, Synth as GpuSynth
    syn = GpuSynth(PLAIN_2_16)
    key, wit = syn.key(dev)
    toxic = syn.arrays()["toxic"]
    syn.close()
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])
    proof = dev.prepare(key).prove(wit, r, s)
    a, b = rowsref.rows(orc, key, wit)
    gk = rowsref.gnark_key(key, toxic, dev.fixed_base)
    assert (dev.prepare(gk).prove_from_rows(wit, a, b, r, s) == proof).all()


def test_the_gnark_toy_key_proves_on_the_gpu(dev, orc):
    # A key from gnark's setup, a witness and rows from gnark's solver (test-support/gnark/toy):
    # the GPU proof equals the oracle's and verifies under gnark's own verifying key.
    from snarkrs_amd.verify import VerifyingKey, verify
    from tests.test_verify import _toy, _toy_key
    vk, _, pub, _ = _toy()
    _, _, key, w, aw, bw = _toy_key()
    msm_o, proof_o = orc.prove_rows(key, w, aw, bw, R_BLIND, S_BLIND)
    circ = dev.prepare(key)
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])
    assert (circ.h_and_msms_from_rows(w, aw, bw) == msm_o).all()
    proof = circ.prove_from_rows(w, aw, bw, r, s)
    assert (proof == proof_o).all()
    verify(vk, pub, proof)
    assert (VerifyingKey.from_bytes(circ.vk_bytes()).to_bytes() == vk.to_bytes()).all()
