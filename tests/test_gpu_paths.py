"""Prove paths that the small parity tests do not reach, against the oracle bit for bit:

* the heavy-bucket branch of the merged (window-table) merge: more than 512 segments of 32
  entries in one bucket, which an all-boolean witness of 40,000 constraints puts in bucket 0;
* the prove without window tables (G16HIP_WINDOW_TABLES=0, the path a device short of memory
  takes), at 300, 5,000 and 2^16 constraints, and its heavy buckets (k_merge's branch);
* a prove from a snarkjs .zkey / .wtns pair written to disk and read back by the codec
  (the reference's warm bench input path, bin/snarkrs/src/bench.rs:461-467);
* the device-memory budget: a key over it is refused as DeviceMemory with nothing allocated,
  a key at it prepares and proves (crates/cuda/src/backend.rs:722-771);
* H MSM widths above 16 over window tables (G16HIP_C_H = 17 and 20: buckets of 2^16 and 2^19,
  the 24-bit digit code, rows cut into many slices), which the width model does not pick.
"""
import numpy as np
import pytest

from oracle.oracle import Synth
from snarkrs_amd import native
from snarkrs_amd.wire import MONT_R, R_MOD, fr_array, fr_ints
from snarkrs_amd.zkey import load_wtns, load_zkey, write_wtns, write_zkey

pytestmark = pytest.mark.gpu

R_BLIND = 0x1234567890ABCDEF1234567890ABCDEF
S_BLIND = 0xFEDCBA0987654321FEDCBA0987654321
RS = (fr_array([R_BLIND]), fr_array([S_BLIND]))


def _check_against_oracle(dev, orc, key, wit):
    circ = dev.prepare(key)
    try:
        msm_orc, proof_orc = orc.prove(key, wit, R_BLIND, S_BLIND)
        assert (circ.h_and_msms(wit) == msm_orc).all(), "MSM outputs"
        assert (circ.prove_with_blinders(wit, *RS) == proof_orc).all(), "proof"
    finally:
        circ.release()


def test_merged_heavy_bucket(dev, orc):
    syn = Synth(orc, 40_000, n_public=2, n_private_in=8, bool_permille=1000,
                input_bool_permille=1000, seed=71, setup_seed=72)
    key, wit = syn.key()
    one = (MONT_R % R_MOD).to_bytes(32, "little")
    ones = int((np.asarray(wit).reshape(-1, 32) == np.frombuffer(one, np.uint8)).all(axis=1).sum())
    # each 1 is one entry of bucket 0 (window 0, digit 1): past 512 segments of 32 entries
    assert ones > 32 * 512, ones
    _check_against_oracle(dev, orc, key, wit)


@pytest.mark.parametrize("nc,bools", [(300, 300), (5000, 400), (1 << 16, 400)])
def test_prove_without_window_tables(dev, orc, monkeypatch, nc, bools):
    monkeypatch.setenv("G16HIP_WINDOW_TABLES", "0")
    syn = Synth(orc, nc, n_public=2, n_private_in=max(8, nc // 11), bool_permille=bools,
                input_bool_permille=bools, seed=80 + nc, setup_seed=81 + nc)
    key, wit = syn.key()
    _check_against_oracle(dev, orc, key, wit)


@pytest.mark.parametrize("ch", [17, 20])
def test_prove_with_wide_h_windows(dev, orc, monkeypatch, ch):
    monkeypatch.setenv("G16HIP_C_H", str(ch))
    syn = Synth(orc, 30_000, n_public=2, n_private_in=2000, bool_permille=400,
                input_bool_permille=400, seed=60 + ch, setup_seed=61 + ch)
    key, wit = syn.key()
    _check_against_oracle(dev, orc, key, wit)


def test_prove_from_zkey_and_wtns_files(dev, orc, tmp_path):
    syn = Synth(orc, 5000, n_public=3, n_private_in=300, bool_permille=400,
                input_bool_permille=400, seed=91, setup_seed=92)
    key, wit = syn.key()
    write_zkey(tmp_path / "circuit.zkey", key)
    write_wtns(tmp_path / "witness.wtns", wit)
    key_f = load_zkey(tmp_path / "circuit.zkey")
    wit_f = load_wtns(tmp_path / "witness.wtns")
    circ = dev.prepare(key_f)
    try:
        _, proof_orc = orc.prove(key, wit, R_BLIND, S_BLIND)
        assert (circ.prove_with_blinders(wit_f, *RS) == proof_orc).all()
        # and the checked prove accepts it under the file's verifying key
        assert (circ.prove(wit_f, *RS) == proof_orc).all()
    finally:
        circ.release()


def test_key_over_the_vram_budget_is_refused_before_allocating(dev, orc):
    syn = Synth(orc, 3000, n_public=2, n_private_in=100, bool_permille=300,
                input_bool_permille=300, seed=95, setup_seed=96)
    key, wit = syn.key()
    need = native.key_resident_bytes(key, window_tables=False)
    assert need > native.key_resident_bytes(key, window_tables=False) - 1 > 0
    try:
        dev.set_vram_limit(need - 1)
        free_before = dev.mem_info()[0]
        with pytest.raises(native.DeviceMemory, match=str(need)):
            dev.prepare(key)
        assert dev.mem_info()[0] == free_before, "the refusal allocated"
        dev.set_vram_limit(need)
        circ = dev.prepare(key)  # under the cap the window tables are off
        assert circ.resident_bytes() <= need
        assert (circ.prove_with_blinders(wit, *RS) == orc.prove(key, wit, R_BLIND, S_BLIND)[1]).all()
        circ.release()
    finally:
        dev.set_vram_limit(0)
