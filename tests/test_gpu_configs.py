"""Parity at the benchmark workloads' real sizes, through the C ABI (the pattern of
crates/cuda/tests/size_sweep_gpu.rs:140-189):

* the 1.1M-constraint prove (domain 2^21, 40% boolean witness, host witness through the
  pinned staging) against the oracle bit for bit: the five MSM outputs and the proof;
* the standalone G2 MSM at 2^18 (bases k_i G2, 30% of scalars in {0, 1}) against the closed
  form (sum k_i s_i) G2;
* the batch workload: 64 proofs with independent witnesses over 4 circuits of 2^14..2^20
  constraints through g16hip_prove_batch, each equal to the same proof made alone, each
  verified, a subset against the oracle;
* the serving workload: 8 circuits of 2^16..2^22 constraints, 32 requests through the
  serving loop, each equal to the same proof made alone, a subset against the oracle.
"""
import numpy as np
import pytest

from snarkrs_amd import native
from snarkrs_amd import synth as psynth
from snarkrs_amd.verify import VerifyingKey, verify
from snarkrs_amd.wire import R_MOD, SplitMix64, fr_array, fr_ints

pytestmark = pytest.mark.gpu

R_BLIND = 0x1234567890ABCDEF1234567890ABCDEF
S_BLIND = 0xFEDCBA0987654321FEDCBA0987654321


def _verifies(circ, wit, proof) -> bool:
    vk = VerifyingKey.from_bytes(circ.vk_bytes())
    try:
        verify(vk, np.asarray(wit).reshape(-1, 32)[1:1 + vk.n_public], proof)
    except native.VerifyError:
        return False
    return True


@pytest.mark.timeout(600)
def test_prove_1_1m_constraints_matches_oracle(dev, orc):
    wl = psynth.LARGE_1_1M
    syn = psynth.Synth(wl)
    key, wit = syn.key(dev)
    syn.close()
    assert key.domain_size == 1 << 21 and key.n_vars == 1_200_003
    circ = dev.prepare(key)
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])
    msm_gpu = circ.h_and_msms(wit)
    proof = circ.prove_with_blinders(wit, r, s)
    msm_orc, proof_orc = orc.prove(key, wit, R_BLIND, S_BLIND)
    assert (msm_gpu == msm_orc).all(), "MSM outputs (A, B2, B1, L, H)"
    assert (proof == proof_orc).all(), "proof"
    assert _verifies(circ, wit, proof)
    circ.release()


@pytest.mark.timeout(600)
def test_g2_msm_2_18_closed_form(dev, orc):
    n = 1 << 18
    ks = psynth.scalars(4343, n, 0)
    sc = psynth.scalars(2121, n, 300)
    bases = dev.fixed_base(2, ks)
    got = dev.msm(2, bases, sc)
    total = sum(k * x for k, x in zip(fr_ints(ks), fr_ints(sc))) % R_MOD
    assert (got == orc.fixed_base(2, fr_array([total]))[0]).all()


def _build(dev, which, per):
    circs = []
    for wl in psynth.mixed_workloads(**which):
        syn = psynth.Synth(wl)
        key, wit = syn.key(dev)
        wits = [wit] + [syn.witness(wl.witness_seed + 1000 * (k + 1)) for k in range(per - 1)]
        syn.close()
        circs.append((wl, key, dev.prepare(key), wits))
    return circs


def _blind(k):
    return R_BLIND + 7 * k, S_BLIND + 11 * k


@pytest.mark.timeout(900)
def test_batch_workload_64_proofs(dev, orc):
    circs = _build(dev, psynth.BATCH_WORKLOADS, 16)
    assert sorted(c[1].domain_size for c in circs) == [1 << 15, 1 << 15, 1 << 18, 1 << 19]
    for _, _, circ, _ in circs:
        circ.add_slots(3)
    items, meta = [], []
    for k in range(64):
        ci = k % 4
        _, key, circ, wits = circs[ci]
        w = wits[k // 4]
        r, s = _blind(k)
        items.append((circ, w, fr_array([r]), fr_array([s])))
        meta.append((ci, k // 4, r, s))
    got = native.prove_batch(items)
    for k, ((circ, w, r, s), g) in enumerate(zip(items, got)):
        assert (g == circ.prove_with_blinders(w, r, s)).all(), k
        assert _verifies(circ, w, g), k
    # one proof of each circuit against the oracle
    for k in range(4):
        ci, wi, r, s = meta[k]
        _, key, _, wits = circs[ci]
        assert (got[k] == orc.prove(key, wits[wi], r, s)[1]).all(), k
    for _, _, circ, _ in circs:
        circ.release()


@pytest.mark.timeout(900)
def test_serving_workload_32_requests(dev, orc):
    circs = _build(dev, psynth.SERVE_WORKLOADS, 2)
    assert max(c[1].domain_size for c in circs) == 1 << 22
    rng = SplitMix64(41)
    reqs = []
    for k in range(32):
        ci = rng.next_u64() % len(circs)
        r, s = _blind(k)
        reqs.append((ci, k % 2, r, s))
    srv = native.Server(dev, reserve_bytes=16 << 30, lanes=4)
    try:
        tickets = [srv.submit(circs[ci][2], circs[ci][3][wi], fr_array([r]), fr_array([s]))
                   for ci, wi, r, s in reqs]
        served = [srv.wait(t)[0] for t in tickets]
    finally:
        srv.close()
    for k, ((ci, wi, r, s), p) in enumerate(zip(reqs, served)):
        circ, w = circs[ci][2], circs[ci][3][wi]
        assert (p == circ.prove_with_blinders(w, fr_array([r]), fr_array([s]))).all(), k
    # the two smallest circuits' requests against the oracle
    small = sorted(range(len(circs)), key=lambda i: circs[i][0].n_constraints)[:2]
    checked = 0
    for k, (ci, wi, r, s) in enumerate(reqs):
        if ci in small and checked < 2:
            _, key, _, wits = circs[ci]
            assert (served[k] == orc.prove(key, wits[wi], r, s)[1]).all(), k
            checked += 1
    assert checked >= 1
    for _, _, circ, _ in circs:
        circ.release()
