"""The host verifier and the snarkjs JSON files (CPU only: no device is opened).

Pinned by:
* a known answer: gnark's toy proof (test-support/gnark/toy, committed under
  tests/golden/gnark_toy), written and accepted by gnark's own prover and verifier, must
  verify, and must stop verifying when any public input or proof element is changed;
* the pairing's defining properties: bilinearity (e(aP, bQ) = e(abP, Q) = e(P, abQ)),
  non-degeneracy, e(P, Q)^r = 1 through a product check, and the split-loop identity of
  crates/groth16/src/verify.rs:199-226;
* proofs from the oracle prover, whose points are known multiples of the generators, and the
  refusals of crates/groth16/src/verify.rs:80-101 and :229-275;
* json.rs's format rules (decimal normal form, c0 before c1, z = "1"), recomputed here from
  the wire bytes with Python integers.
"""
import json

import numpy as np
import pytest

from snarkrs_amd import native
from snarkrs_amd.verify import (VerifyingKey, pairing, pairing_check, point_valid,
                                proof_from_str, proof_to_string, public_from_str,
                                public_to_string, verify, verify_allowing_unblinded,
                                verify_unchecked)
from snarkrs_amd.wire import Q_MOD, R_MOD, fq_ints, fr_array, fr_ints, g1_ints, g2_ints
from tests import gnark_codec as gc

TOY = "tests/golden/gnark_toy/"


def _toy():
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    rd = lambda f: open(os.path.join(root, TOY, f), "rb").read()
    vk = gc.read_vk(rd("vk.bin"))
    v = VerifyingKey(vk["alpha_g1"], vk["beta_g2"], vk["gamma_g2"], vk["delta_g2"], vk["ic"])
    return v, vk, gc.read_public(rd("public.bin")), gc.read_proof(rd("proof.bin"))


def _neg_g1(p):
    q = np.array(p, np.uint8).reshape(64).copy()
    if not q.any():
        return q
    (x, y), = g1_ints(q)
    return gc.g1_array([(x, (-y) % Q_MOD)])[0]


# ---- the gnark known answer ------------------------------------------------------------

def test_gnark_toy_proof_verifies():
    vk, raw, pub, proof = _toy()
    assert vk.n_public == 4 and len(pub) == 4
    # The public inputs of the toy circuit: X = 5, Y, Sum = sum(3i + 7, i < 1000), Low = 7.
    vals = fr_ints(pub)
    assert vals[0] == 5 and vals[2] == sum(3 * i + 7 for i in range(1000)) and vals[3] == 7
    for k in ("alpha_g1", "beta_g1", "delta_g1"):
        assert point_valid(1, raw[k])
    for k in ("beta_g2", "gamma_g2", "delta_g2"):
        assert point_valid(2, raw[k])
    verify(vk, pub, proof)
    verify_allowing_unblinded(vk, pub, proof)
    verify_unchecked(vk, pub, proof)


@pytest.mark.parametrize("which", range(4))
def test_gnark_toy_refuses_a_changed_public_input(which):
    vk, _, pub, proof = _toy()
    vals = fr_ints(pub)
    vals[which] = (vals[which] + 1) % R_MOD
    with pytest.raises(native.PairingFailed):
        verify(vk, fr_array(vals), proof)


def test_gnark_toy_refuses_changed_proof_elements():
    vk, _, pub, proof = _toy()
    a, b, c = proof[:64], proof[64:192], proof[192:]
    for bad in (np.concatenate([c, b, a]),             # A and C swapped
                np.concatenate([_neg_g1(a), b, c]),    # -A
                np.concatenate([a, b, _neg_g1(c)])):   # -C
        with pytest.raises(native.PairingFailed):
            verify(vk, pub, bad)


def test_gnark_toy_refuses_the_wrong_public_count():
    vk, _, pub, proof = _toy()
    with pytest.raises(native.PublicInputCount, match="expected 4 public inputs, got 3"):
        verify(vk, pub[:3], proof)


# ---- the pairing ---------------------------------------------------------------------

def _g1(orc, k):
    return orc.fixed_base(1, fr_array([k]))[0]


def _g2(orc, k):
    return orc.fixed_base(2, fr_array([k]))[0]


def test_pairing_is_bilinear_and_nondegenerate(orc):
    a, b = 0x1234567890ABCDEF1122, 0xFEDCBA0987654321
    e_ab = pairing(_g1(orc, a), _g2(orc, b))
    assert (e_ab == pairing(_g1(orc, a * b), _g2(orc, 1))).all()
    assert (e_ab == pairing(_g1(orc, 1), _g2(orc, a * b))).all()
    assert (e_ab == pairing(_g1(orc, b), _g2(orc, a))).all()
    assert not (pairing(_g1(orc, 1), _g2(orc, 1)) == pairing(_g1(orc, 2), _g2(orc, 1))).all()
    # e(P, Q) != 1 but e(P, Q) e(-P, Q) = 1 and e(P, Q)^r: e(rP, Q) with rP = infinity.
    assert not pairing_check([_g1(orc, 1)], [_g2(orc, 1)])
    assert pairing_check([_g1(orc, 7), _g1(orc, R_MOD - 7)], [_g2(orc, 3), _g2(orc, 3)])
    assert pairing_check([_g1(orc, a), _g1(orc, R_MOD - a * b % R_MOD)], [_g2(orc, b), _g2(orc, 1)])
    assert pairing_check([], [])
    one = np.zeros(384, np.uint8)
    e1 = pairing(_g1(orc, 1), np.zeros(128, np.uint8))  # a pair with infinity contributes 1
    fq_one = fq_ints(e1.reshape(12, 32))
    assert fq_one == [1] + [0] * 11 and not (e1 == one).all()


def test_split_pairs_and_pairs_with_infinity(orc):
    # crates/groth16/src/verify.rs:199-226: pairs with a zero element drop out of the product.
    g1 = lambda k: _g1(orc, k)
    g2 = lambda k: _g2(orc, k)
    z1, z2 = np.zeros(64, np.uint8), np.zeros(128, np.uint8)
    pairs = [(g1(3), g2(5)), (g1(7), g2(11)), (g1(13), g2(17)), (g1(19), g2(23))]
    total = 3 * 5 + 7 * 11 + 13 * 17 + 19 * 23
    p1 = [p for p, _ in pairs] + [g1(R_MOD - total)]
    p2 = [q for _, q in pairs] + [g2(1)]
    assert pairing_check(p1, p2)
    assert not pairing_check(p1[:-1], p2[:-1])
    with_zero = [(g1(3), g2(5)), (z1, g2(11)), (g1(13), g2(17)), (g1(19), z2)]
    total = 3 * 5 + 13 * 17
    p1 = [p for p, _ in with_zero] + [g1(R_MOD - total)]
    p2 = [q for _, q in with_zero] + [g2(1)]
    assert pairing_check(p1, p2)


def test_point_checks_catch_off_curve_and_off_subgroup_points(orc):
    g = _g2(orc, 1)
    assert point_valid(2, g) and point_valid(1, _g1(orc, 1))
    # A twist point outside the prime-order subgroup: some x with x^3 + b' a square.
    for x0 in range(1, 50):
        y = gc.fq2_sqrt(gc.f2_add(gc.f2_mul(gc.f2_mul((x0, 1), (x0, 1)), (x0, 1)), gc.B_G2))
        if y is not None:
            break
    off = gc.g2_array([((x0, 1), y)])[0]
    assert not point_valid(2, off)
    bad = _g1(orc, 5).copy()
    bad[40] ^= 1
    assert not point_valid(1, bad)
    too_big = np.frombuffer(Q_MOD.to_bytes(32, "little") * 2, np.uint8).copy()
    assert not point_valid(1, too_big)


# ---- Groth16 on oracle proofs -----------------------------------------------------------

@pytest.mark.parametrize("n_constraints,n_public", [(1, 0), (20, 1), (60, 3), (300, 6)])
def test_oracle_proofs_verify(orc, n_constraints, n_public):
    from oracle.oracle import Synth
    syn = Synth(orc, n_constraints, n_public=n_public, n_private_in=4, seed=n_constraints,
                setup_seed=5)
    key, wit = syn.key()
    _, proof = orc.prove(key, wit, 0x1234567, 0x7654321)
    vk = VerifyingKey.from_proving_key(key)
    pub = np.asarray(wit).reshape(-1, 32)[1:1 + n_public]
    verify(vk, pub, proof)
    if n_public:
        vals = fr_ints(pub)
        vals[-1] = (vals[-1] + 1) % R_MOD
        with pytest.raises(native.PairingFailed):
            verify(vk, fr_array(vals), proof)
    # Malleability (crates/groth16/src/lib.rs:164-186): (A z^-1, B z, C) verifies too.
    z = 0xABCDEF
    a2 = orc.scalar_mul(1, proof[:64], fr_array([pow(z, -1, R_MOD)]))[0]
    b2 = orc.scalar_mul(2, proof[64:192], fr_array([z]))[0]
    verify(vk, pub, np.concatenate([a2, b2, proof[192:]]))


def _hand_key(orc, alpha=3, gamma=7, delta=7):
    g1 = lambda k: _g1(orc, k) if k else np.zeros(64, np.uint8)
    return VerifyingKey(g1(alpha), _g2(orc, 5), _g2(orc, gamma), _g2(orc, delta),
                        np.stack([g1(11), g1(13)]))


def test_the_unblinded_opt_in_accepts_a_forgery_and_nothing_else(orc):
    # crates/groth16/src/verify.rs:229-275: gamma = delta lets (alpha, beta, -L_bar) pass.
    vk = _hand_key(orc)
    public = fr_array([42])
    forged = np.concatenate([vk.alpha_g1, vk.beta_g2, _g1(orc, R_MOD - (11 + 42 * 13))])
    with pytest.raises(native.UnsafeKey, match="gamma_g2 equals delta_g2"):
        verify(vk, public, forged)
    verify_allowing_unblinded(vk, public, forged)
    with pytest.raises(native.UnsafeKey, match="alpha_g1 is the point at infinity"):
        verify_allowing_unblinded(_hand_key(orc, alpha=0), public, forged)
    bad = forged.copy()
    bad[:64] = 0
    with pytest.raises(native.InvalidProof, match="A is the point at infinity"):
        verify_allowing_unblinded(vk, public, bad)
    with pytest.raises(native.PairingFailed):
        verify_allowing_unblinded(vk, fr_array([43]), forged)


def test_refusals_of_keys_and_proof_elements(orc):
    vk = _hand_key(orc, gamma=9)
    pub = fr_array([1])
    good = np.concatenate([_g1(orc, 2), _g2(orc, 3), _g1(orc, 4)])
    for lo, hi, what in ((0, 64, "A"), (64, 192, "B"), (192, 256, "C")):
        bad = good.copy()
        bad[lo:hi] = 0
        with pytest.raises(native.InvalidProof, match=f"{what} is the point at infinity"):
            verify(vk, pub, bad)
    off_curve = good.copy()
    off_curve[200] ^= 1
    with pytest.raises(native.InvalidProof, match="C is not a valid G1 point"):
        verify(vk, pub, off_curve)
    for x0 in range(1, 50):
        y = gc.fq2_sqrt(gc.f2_add(gc.f2_mul(gc.f2_mul((x0, 1), (x0, 1)), (x0, 1)), gc.B_G2))
        if y is not None:
            break
    off_sub = good.copy()
    off_sub[64:192] = gc.g2_array([((x0, 1), y)])[0]
    with pytest.raises(native.InvalidProof, match="B is not a valid G2 point"):
        verify(vk, pub, off_sub)
    empty = VerifyingKey(vk.alpha_g1, vk.beta_g2, vk.gamma_g2, vk.delta_g2,
                         np.zeros((0, 64), np.uint8))
    with pytest.raises(native.UnsafeKey, match="IC is empty"):
        verify(empty, fr_array([]), good)
    # crates/groth16/src/verify.rs:186-197: the bare check reports the empty IC itself.
    with pytest.raises(native.EmptyVerifyingKey):
        verify_unchecked(empty, fr_array([]), good)
    with pytest.raises(native.EmptyVerifyingKey):
        verify_unchecked(empty, fr_array([1]), good)
    with pytest.raises(native.PairingFailed):
        verify(vk, pub, good)


# ---- snarkjs JSON ----------------------------------------------------------------------

def _expected_proof_json(proof):
    (ax, ay), = g1_ints(proof[:64])
    ((bx0, bx1), (by0, by1)), = g2_ints(proof[64:192])
    (cx, cy), = g1_ints(proof[192:])
    return ('{\n "pi_a": ["%d","%d","1"],\n "pi_b": [["%d","%d"],["%d","%d"],["1","0"]],\n'
            ' "pi_c": ["%d","%d","1"],\n "protocol": "groth16",\n "curve": "bn128"\n}\n'
            % (ax, ay, bx0, bx1, by0, by1, cx, cy))


def test_proof_and_public_json_follow_snarkjs():
    vk, _, pub, proof = _toy()
    text = proof_to_string(proof)
    assert text == _expected_proof_json(proof)
    doc = json.loads(text)
    assert doc["protocol"] == "groth16" and doc["curve"] == "bn128"
    assert (proof_from_str(text) == proof).all()
    ptext = public_to_string(pub)
    assert ptext == "[\n " + ",\n ".join('"%d"' % v for v in fr_ints(pub)) + "\n]\n"
    assert (public_from_str(ptext) == pub).all()
    assert public_to_string(fr_array([])) == "[\n \n]\n"
    # The files read back verify.
    verify(vk, public_from_str(ptext), proof_from_str(text))


def test_json_infinity_forms():
    z1, z2 = np.zeros(64, np.uint8), np.zeros(128, np.uint8)
    text = proof_to_string(np.concatenate([z1, z2, z1]))
    doc = json.loads(text)
    assert doc["pi_a"] == ["0", "1", "0"] and doc["pi_c"] == ["0", "1", "0"]
    assert doc["pi_b"] == [["0", "0"], ["1", "0"], ["0", "0"]]
    assert (proof_from_str(text) == 0).all()


def _doc(proof):
    return json.loads(proof_to_string(proof))


@pytest.mark.parametrize("mutate,match", [
    (lambda d: d["pi_a"].__setitem__(2, "2"), "z = 2, expected 1"),
    (lambda d: d["pi_a"].__setitem__(2, "0"), "z = 0 but"),
    (lambda d: d["pi_a"].__setitem__(0, str(Q_MOD)), "not less than the field modulus"),
    (lambda d: d["pi_a"].__setitem__(0, "12x"), "is not a decimal integer"),
    (lambda d: d["pi_a"].__setitem__(0, 5), "expected a decimal string"),
    (lambda d: d["pi_a"].__setitem__(1, "1"), "not on the curve"),
    (lambda d: d["pi_b"].__setitem__(2, ["1", "1"]), "z is not \\[1, 0\\]"),
    (lambda d: d["pi_b"][0].__setitem__(0, "1"), "not on the curve"),
    (lambda d: d.__setitem__("pi_c", ["1", "2"]), "expected 3 entries"),
    (lambda d: d.pop("pi_c"), "proof has no pi_c"),
    (lambda d: d.__setitem__("protocol", "plonk"), "expected groth16"),
    (lambda d: d.__setitem__("curve", "bls12381"), "expected bn128"),
])
def test_proof_json_reader_refuses(mutate, match):
    _, _, _, proof = _toy()
    d = _doc(proof)
    mutate(d)
    with pytest.raises(native.JsonError, match=match):
        proof_from_str(json.dumps(d))


def test_proof_json_reader_refuses_off_subgroup_b_and_bad_text():
    _, _, _, proof = _toy()
    for x0 in range(1, 50):
        y = gc.fq2_sqrt(gc.f2_add(gc.f2_mul(gc.f2_mul((x0, 1), (x0, 1)), (x0, 1)), gc.B_G2))
        if y is not None:
            break
    d = _doc(proof)
    d["pi_b"] = [[str(x0), "1"], [str(y[0]), str(y[1])], ["1", "0"]]
    with pytest.raises(native.JsonError, match="not in the prime-order subgroup"):
        proof_from_str(json.dumps(d))
    for text in ("", "{", '{"pi_a": [1,2,3]', "[1,2] x", '{"a": "\\x"}', "[" * 200):
        with pytest.raises(native.JsonError):
            proof_from_str(text)
    with pytest.raises(native.JsonError, match="not less than the field modulus"):
        public_from_str('["%d"]' % R_MOD)
    with pytest.raises(native.JsonError, match="flat JSON array"):
        public_from_str('{"a": 1}')
    assert fr_ints(public_from_str('["+7", "0", "007"]')) == [7, 0, 7]


def _vk_json(vk: VerifyingKey, n_public=None):
    g1 = lambda p: (["0", "1", "0"] if not np.asarray(p).any() else
                    [str(v) for v in g1_ints(p)[0]] + ["1"])

    def g2(p):
        (x0, x1), (y0, y1) = g2_ints(p)[0]
        return [[str(x0), str(x1)], [str(y0), str(y1)], ["1", "0"]]
    return json.dumps({
        "protocol": "groth16", "curve": "bn128",
        "nPublic": vk.n_public if n_public is None else n_public,
        "vk_alpha_1": g1(vk.alpha_g1), "vk_beta_2": g2(vk.beta_g2),
        "vk_gamma_2": g2(vk.gamma_g2), "vk_delta_2": g2(vk.delta_g2),
        "IC": [g1(p) for p in vk.ic],
    })


def test_verification_key_json(orc):
    vk, _, pub, proof = _toy()
    got = VerifyingKey.from_json_str(_vk_json(vk))
    assert (got.to_bytes() == vk.to_bytes()).all()
    verify(got, pub, proof)
    with pytest.raises(native.JsonError, match="IC has 5 entries, nPublic 3 implies 4"):
        VerifyingKey.from_json_str(_vk_json(vk, n_public=3))
    d = json.loads(_vk_json(vk))
    d["protocol"] = "plonk"
    with pytest.raises(native.JsonError, match="expected groth16"):
        VerifyingKey.from_json_str(json.dumps(d))
    unblinded = _hand_key(orc)
    with pytest.raises(native.UnsafeKey, match="gamma_g2 equals delta_g2"):
        VerifyingKey.from_json_str(_vk_json(unblinded))
    assert (VerifyingKey.from_json_str(_vk_json(unblinded), check_structure=False).to_bytes()
            == unblinded.to_bytes()).all()


# ---- the gnark toy key proved through the rows path (crates/gnark/tests/prove.rs) --------

def _toy_key():
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    rd = lambda f: open(os.path.join(root, TOY, f), "rb").read()
    pk = gc.read_pk_dump(rd("pk.dump"))
    vk = gc.read_vk(rd("vk.bin"))
    w, aw, bw = gc.read_witness_dump(rd("witness.bin"))
    return pk, vk, gc.toy_proving_key(pk, vk), w, aw, bw


def test_gnark_toy_dump_matches_its_verifying_key_and_domain():
    pk, vk, key, w, aw, bw = _toy_key()
    n = pk["domain_size"]
    log_n = n.bit_length() - 1
    # gnark's domain generator is the one every prover here uses for 2^12.
    assert pk["generator"] == pow(pow(5, (R_MOD - 1) >> 28, R_MOD), 1 << (28 - log_n), R_MOD)
    assert pk["card_inv"] == pow(n, -1, R_MOD)
    for k in ("alpha_g1", "beta_g1", "delta_g1", "beta_g2", "delta_g2"):
        assert (pk[k] == vk[k]).all(), k
    assert key.n_vars == len(w) == 3072 and key.n_public == 4 and len(aw) == len(bw) == 2071
    assert fr_ints(w[:1]) == [1]
    _, _, pub, _ = _toy()
    assert (w[1:5] == pub).all()


def test_the_oracle_proves_the_gnark_toy_key_from_rows(orc):
    vk, _, pub, _ = _toy()
    _, _, key, w, aw, bw = _toy_key()
    _, proof = orc.prove_rows(key, w, aw, bw, 0x1234567, 0x7654321)
    verify(vk, pub, proof)
    with pytest.raises(native.PairingFailed):
        verify(vk, fr_array([6] + fr_ints(pub)[1:]), proof)
