"""Pins the oracle's whole prove (stages 0-11) with the closed-form "shadow" (CPU only).

With the toxic waste known, every MSM output and every proof point is a known multiple of
the generator (computed in O(n + nnz) scalar arithmetic, no NTT and no MSM), and the Groth16
verification equation reduces to a*b = alpha*beta + sum_pub + delta*c in the exponent.
Also: the hand-built CSR key of crates/groth16/src/cpu.rs:652-684 / :851-861, and H
against the polynomial definition on a tiny circuit.
"""
import numpy as np
import pytest

from oracle.oracle import Synth
from snarkrs_amd.key import ProvingKey
from snarkrs_amd.wire import R_MOD, fr_array, fr_ints

R_BLIND = 0x1234567890ABCDEF1234567890ABCDEF
S_BLIND = 0xFEDCBA0987654321FEDCBA0987654321


def _g1(orc, scalars):
    return orc.fixed_base(1, fr_array(scalars))


def _g2(orc, scalars):
    return orc.fixed_base(2, fr_array(scalars))


@pytest.mark.parametrize("n_constraints,bool_permille", [(1, 0), (5, 0), (100, 0), (300, 400)])
def test_prove_matches_closed_form_shadow(orc, n_constraints, bool_permille):
    syn = Synth(orc, n_constraints, n_public=2, n_private_in=5, bool_permille=bool_permille,
                input_bool_permille=bool_permille, seed=n_constraints, setup_seed=99)
    key, wit = syn.key()
    msm, proof = orc.prove(key, wit, R_BLIND, S_BLIND)
    mA, mB, mL, mH, pa, pb, pc, lhs, rhs = syn.shadow(R_BLIND, S_BLIND)
    assert lhs == rhs, "the synthetic witness does not satisfy its circuit"
    assert (msm[0:64] == _g1(orc, [mA])[0]).all()
    assert (msm[64:192] == _g2(orc, [mB])[0]).all()
    assert (msm[192:256] == _g1(orc, [mB])[0]).all()
    assert (msm[256:320] == _g1(orc, [mL])[0]).all()
    assert (msm[320:384] == _g1(orc, [mH])[0]).all()
    assert (proof[0:64] == _g1(orc, [pa])[0]).all()
    assert (proof[64:192] == _g2(orc, [pb])[0]).all()
    assert (proof[192:256] == _g1(orc, [pc])[0]).all()


def test_domain_follows_snarkjs_power_rule(orc):
    # crates/ceremony/src/setup.rs:125-135: 2^(floor(log2(nC + nPub)) + 1).
    for nc, npub, want in [(1, 2, 4), (2, 2, 8), (6, 2, 16), (1 << 16, 2, 1 << 17), (100, 0, 128)]:
        syn = Synth(orc, nc, n_public=npub, n_private_in=3)
        assert syn.domain == want, (nc, npub)


def test_h_is_a_times_b_minus_c_on_the_odd_coset(orc):
    syn = Synth(orc, 6, n_public=1, n_private_in=2, seed=3)
    key, wit = syn.key()
    n = key.domain_size
    a = fr_ints(orc.gather(key, 0, wit))
    b = fr_ints(orc.gather(key, 1, wit))
    c = [x * y % R_MOD for x, y in zip(a, b)]
    w = pow(pow(5, (R_MOD - 1) >> 28, R_MOD), 1 << (28 - (n.bit_length() - 1)), R_MOD)
    shift = pow(pow(5, (R_MOD - 1) >> 28, R_MOD), 1 << (28 - n.bit_length()), R_MOD)

    def interp_eval(ev, x):
        # Lagrange interpolation on the domain, evaluated at x.
        tot = 0
        for i, v in enumerate(ev):
            wi = pow(w, i, R_MOD)
            num = (pow(x, n, R_MOD) - 1) * wi % R_MOD
            den = n * (x - wi) % R_MOD
            tot += v * num * pow(den, -1, R_MOD)
        return tot % R_MOD

    pts = [shift * pow(w, k, R_MOD) % R_MOD for k in range(n)]
    want = [(interp_eval(a, x) * interp_eval(b, x) - interp_eval(c, x)) % R_MOD for x in pts]
    assert fr_ints(orc.compute_h(key, wit)) == want


def test_gather_accepts_empty_rows(orc):
    # crates/groth16/src/cpu.rs:652-684 and :851-861: rows [0,1,1,2,2], both matrices.
    z64, z128 = np.zeros(64, np.uint8), np.zeros(128, np.uint8)
    key = ProvingKey(
        n_vars=2, n_public=0, domain_size=4, alpha_g1=z64, beta_g1=z64, beta_g2=z128,
        gamma_g2=z128, delta_g1=z64, delta_g2=z128, ic=np.zeros((1, 64), np.uint8),
        row_ptr=[np.array([0, 1, 1, 2, 2], np.uint32)] * 2,
        signal=[np.array([0, 1], np.uint32)] * 2,
        value_id=[np.zeros(2, np.uint32)] * 2, values=fr_array([1]),
        a_query=np.zeros((2, 64), np.uint8), b_g1_query=np.zeros((2, 64), np.uint8),
        b_g2_query=np.zeros((2, 128), np.uint8), l_query=np.zeros((1, 64), np.uint8),
        h_query=np.zeros((4, 64), np.uint8))
    wit = fr_array([2, 3])
    for m in (0, 1):
        assert fr_ints(orc.gather(key, m, wit)) == [2, 0, 3, 0]


def test_gather_bit_shortcuts_match_multiplication(orc):
    # crates/groth16/src/cpu.rs:887-908: witness all 0, all 1, all -1, all random.
    syn = Synth(orc, 50, n_public=1, n_private_in=3, seed=8)
    key, _ = syn.key()
    vals = fr_ints(key.values)
    for v in (0, 1, R_MOD - 1, 0x5EED5EED5EED):
        wit = fr_array([v] * key.n_vars)
        for m in (0, 1):
            rp, sig, vid = key.row_ptr[m], key.signal[m], key.value_id[m]
            want = [sum(vals[vid[k]] * v for k in range(rp[c], rp[c + 1])) % R_MOD
                    for c in range(key.domain_size)]
            assert fr_ints(orc.gather(key, m, wit)) == want


# ---- the rows path and gnark's H basis (crates/groth16/src/cpu.rs:250-286, :387-470) ----

@pytest.mark.parametrize("nc", [1, 7, 300])
def test_rows_path_gives_the_gather_paths_proof_in_both_h_bases(orc, nc):
    from tests import rowsref
    from snarkrs_amd.verify import VerifyingKey, verify
    syn = Synth(orc, nc, n_public=2, n_private_in=5, bool_permille=300, input_bool_permille=300,
                seed=40 + nc, setup_seed=41)
    key, wit = syn.key()
    msm, proof = orc.prove(key, wit, R_BLIND, S_BLIND)
    a, b = rowsref.rows(orc, key, wit)
    assert len(a) < key.domain_size or nc == 1
    # snarkjs' basis from rows (trimmed or full) = the gather path, bit for bit.
    for ra, rb in ((a, b), rowsref.rows(orc, key, wit, trim=False)):
        m2, p2 = orc.prove_rows(key, wit, ra, rb, R_BLIND, S_BLIND)
        assert (m2 == msm).all() and (p2 == proof).all()
    # gnark's basis: H's coefficients against [tau^bitrev(j) t(tau)/delta]_1 put the same
    # [H(tau) t(tau)/delta]_1 into the proof, so every output is the same point.
    gk = rowsref.gnark_key(key, syn.arrays()["toxic"], orc.fixed_base)
    assert len(gk.h_query) == key.domain_size - 1 and gk.row_ptr is None
    m3, p3 = orc.prove_rows(gk, wit, a, b, R_BLIND, S_BLIND)
    assert (m3 == msm).all() and (p3 == proof).all()
    verify(VerifyingKey.from_proving_key(gk), np.asarray(wit).reshape(-1, 32)[1:3], p3)


def test_rows_path_refusals(orc):
    from tests import rowsref
    syn = Synth(orc, 20, n_public=1, n_private_in=3, seed=9, setup_seed=10)
    key, wit = syn.key()
    a, b = rowsref.rows(orc, key, wit, trim=False)
    extra = np.zeros((1, 32), np.uint8)
    with pytest.raises(ValueError, match="-2"):
        orc.prove_rows(key, wit, np.concatenate([a, extra]), np.concatenate([b, extra]), 1, 2)
