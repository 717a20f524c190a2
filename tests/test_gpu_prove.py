"""The full HIP prove path (stages 0-11) through the C ABI against the oracle, bit for bit,
on synthetic circuits: H element by element, the five MSM outputs, the proof with fixed
blinders, the closed-form shadow, concurrency on one prepared circuit, and the error paths
of the reference contract (crates/groth16/src/lib.rs:79-154, :285-307;
crates/groth16/src/prove.rs:272-283)."""
import threading

import numpy as np
import pytest

from oracle.oracle import Synth
from snarkrs_amd import native
from snarkrs_amd.wire import R_MOD, fr_array

pytestmark = pytest.mark.gpu

R_BLIND = 0x1234567890ABCDEF1234567890ABCDEF
S_BLIND = 0xFEDCBA0987654321FEDCBA0987654321


@pytest.fixture(scope="module")
def small(orc):
    syn = Synth(orc, 300, n_public=2, n_private_in=6, bool_permille=300,
                input_bool_permille=300, seed=5, setup_seed=6)
    key, wit = syn.key()
    return syn, key, wit


@pytest.mark.parametrize("nc,bools", [(1, 0), (7, 0), (300, 300), (5000, 400)])
def test_proof_matches_oracle_and_shadow(dev, orc, nc, bools):
    syn = Synth(orc, nc, n_public=3, n_private_in=5, bool_permille=bools,
                input_bool_permille=bools, seed=nc, setup_seed=nc + 1)
    key, wit = syn.key()
    circ = dev.prepare(key)
    hh = circ.compute_h(wit)
    assert (circ.h_to_host(hh) == orc.compute_h(key, wit)).all(), "H"
    msm_gpu = circ.msms(wit, hh)
    msm_orc, proof_orc = orc.prove(key, wit, R_BLIND, S_BLIND)
    assert (msm_gpu == msm_orc).all(), "MSM outputs"
    assert (circ.h_and_msms(wit) == msm_orc).all(), "h_and_msms"
    proof = circ.prove_with_blinders(wit, fr_array([R_BLIND]), fr_array([S_BLIND]))
    assert (proof == proof_orc).all(), "proof"
    _, _, _, _, pa, pb, pc, lhs, rhs = syn.shadow(R_BLIND, S_BLIND)
    assert lhs == rhs
    assert (proof[:64] == orc.fixed_base(1, fr_array([pa]))[0]).all()
    assert (proof[64:192] == orc.fixed_base(2, fr_array([pb]))[0]).all()
    assert (proof[192:] == orc.fixed_base(1, fr_array([pc]))[0]).all()


def test_repeated_proofs_are_identical(dev, orc, small):
    _, key, wit = small
    circ = dev.prepare(key)
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])
    first = circ.prove_with_blinders(wit, r, s)
    for _ in range(3):
        assert (circ.prove_with_blinders(wit, r, s) == first).all()


def test_concurrent_proofs_on_one_circuit(dev, orc, small):
    _, key, wit = small
    circ = dev.prepare(key)
    want = orc.prove(key, wit, R_BLIND, S_BLIND)[1]
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])
    results, errors = [], []

    def work():
        try:
            for _ in range(3):
                results.append(circ.prove_with_blinders(wit, r, s))
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    ts = [threading.Thread(target=work) for _ in range(4)]
    [t.start() for t in ts]
    [t.join() for t in ts]
    assert not errors, errors
    assert len(results) == 12 and all((p == want).all() for p in results)


def test_error_paths(dev, orc, small):
    _, key, wit = small
    circ = dev.prepare(key)
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])
    with pytest.raises(native.WitnessLength):
        circ.prove_with_blinders(wit[:-1], r, s)
    bad = wit.copy()
    bad[0] = fr_array([2])[0]
    with pytest.raises(native.ConstantWire):
        circ.prove_with_blinders(bad, r, s)
    with pytest.raises(native.ZeroBlinder):
        circ.prove_with_blinders(wit, fr_array([0]), s)
    # A device witness's w[0] is checked after its proof ran: the call still fails, and the
    # slot proves the next witness correctly.
    from snarkrs_amd.devmem import DeviceArray
    d_bad, d_good = DeviceArray(np.asarray(bad)), DeviceArray(np.asarray(wit))
    with pytest.raises(native.ConstantWire):
        circ.prove_with_blinders_device(d_bad.addr, len(bad), r, s)
    assert (circ.prove_with_blinders_device(d_good.addr, len(wit), r, s)
            == circ.prove_with_blinders(wit, r, s)).all()
    del d_bad, d_good
    other = dev.prepare(key)
    hh = other.compute_h(wit)
    with pytest.raises(native.BackendError):
        circ.msms(wit, hh)


def test_malformed_keys_are_refused(dev, orc, small):
    import copy

    _, key, _ = small
    k = copy.copy(key)
    k.row_ptr = [key.row_ptr[0].copy(), key.row_ptr[1]]
    k.row_ptr[0][0] = 1
    with pytest.raises(native.BackendError, match="starts at 1"):
        dev.prepare(k)
    k = copy.copy(key)
    k.domain_size = 3
    with pytest.raises(native.BackendError, match="power of two"):
        dev.prepare(k)
    k = copy.copy(key)
    k.n_public = key.n_vars
    with pytest.raises(native.BackendError, match="exceeds n_vars"):
        dev.prepare(k)
    k = copy.copy(key)
    k.alpha_g1 = np.zeros(64, np.uint8)
    with pytest.raises(native.BackendError, match="infinity"):
        dev.prepare(k)
