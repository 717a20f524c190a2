"""snarkjs .zkey / .wtns codec (include/g16zkey.h, host code, CPU only): a synthetic key and
witness written and read back are the same key (up to value interning) and the same witness,
the byte conventions hold (section 4 stores v R^2, the witness plain integers, points raw
Montgomery limbs), and malformed files are refused with the reference's reasons."""
import struct

import numpy as np
import pytest

from oracle.oracle import Synth
from snarkrs_amd import native
from snarkrs_amd.wire import MONT_R, R_MOD, fr_ints, le_to_int
from snarkrs_amd.zkey import load_wtns, load_zkey, write_wtns, write_zkey


@pytest.fixture(scope="module")
def sample(orc, tmp_path_factory):
    syn = Synth(orc, 300, n_public=3, n_private_in=5, bool_permille=300, seed=41, setup_seed=42)
    key, wit = syn.key()
    d = tmp_path_factory.mktemp("zkey")
    write_zkey(d / "c.zkey", key)
    write_wtns(d / "c.wtns", wit)
    return key, wit, d


def _rows(key, m):
    vals = key.values
    out = []
    for c in range(key.domain_size):
        lo, hi = key.row_ptr[m][c], key.row_ptr[m][c + 1]
        out.append([(int(key.signal[m][e]), vals[key.value_id[m][e]].tobytes()) for e in range(lo, hi)])
    return out


def test_round_trip(sample):
    key, wit, d = sample
    back = load_zkey(d / "c.zkey")
    assert (back.n_vars, back.n_public, back.domain_size) == (key.n_vars, key.n_public, key.domain_size)
    for f in ("alpha_g1", "beta_g1", "beta_g2", "gamma_g2", "delta_g1", "delta_g2", "ic",
              "a_query", "b_g1_query", "b_g2_query", "l_query", "h_query"):
        assert (getattr(back, f) == getattr(key, f)).all(), f
    for m in range(2):
        assert (back.row_ptr[m] == key.row_ptr[m]).all()
        assert _rows(back, m) == _rows(key, m)
    assert len(back.values) <= len(key.values)  # interned
    assert (load_wtns(d / "c.wtns") == wit).all()


def test_byte_conventions(sample):
    key, wit, d = sample
    raw = (d / "c.zkey").read_bytes()
    assert raw[:4] == b"zkey" and struct.unpack_from("<II", raw, 4) == (1, 10)
    # Walk to section 4 and check the first record's value is v R^2 mod r.
    off, sec = 12, {}
    while off < len(raw):
        sid, ln = struct.unpack_from("<IQ", raw, off)
        sec[sid] = (off + 12, ln)
        off += 12 + ln
    o4, _ = sec[4]
    m, c, s = struct.unpack_from("<III", raw, o4 + 4)
    stored = le_to_int(raw[o4 + 16:o4 + 48])
    e = int(key.row_ptr[m][c])
    v = fr_ints(key.values[key.value_id[m][e]:key.value_id[m][e] + 1])[0]
    assert stored == v * MONT_R * MONT_R % R_MOD
    # Witness section 2: plain integers.
    wraw = (d / "c.wtns").read_bytes()
    assert wraw[:4] == b"wtns"
    plain = [le_to_int(wraw[-32 * len(wit) + 32 * i:][:32]) for i in range(3)]
    assert plain == fr_ints(wit[:3])


def test_malformed_files_are_refused(sample, tmp_path):
    key, wit, d = sample
    raw = bytearray((d / "c.zkey").read_bytes())
    bad = tmp_path / "bad.zkey"
    bad.write_bytes(b"zkey" + bytes(raw[4:20]))
    with pytest.raises(native.BackendError):
        load_zkey(bad)
    bad.write_bytes(bytes(raw[:-10]))
    with pytest.raises(native.BackendError):
        load_zkey(bad)
    # witness[0] != 1
    w2 = wit.copy()
    w2[0] = w2[1]
    write_wtns(tmp_path / "bad.wtns", w2)
    with pytest.raises(native.BackendError, match="witness"):
        load_wtns(tmp_path / "bad.wtns")
