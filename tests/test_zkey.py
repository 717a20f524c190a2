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


def _sections(raw: bytes, magic: bytes) -> dict:
    """The binfile container (crates/formats/src/binfile.rs:465-541), decoded here with
    struct alone: magic, u32 version, u32 section count, then (u32 id, u64 length, bytes)."""
    assert raw[:4] == magic
    _, count = struct.unpack_from("<II", raw, 4)
    off, sec = 12, {}
    for _ in range(count):
        sid, ln = struct.unpack_from("<IQ", raw, off)
        assert sid not in sec
        sec[sid] = raw[off + 12:off + 12 + ln]
        off += 12 + ln
    assert off == len(raw)
    return sec


def test_every_record_decodes_independently(sample):
    """Every byte the writer emits, read back without the codec: the groth16 header, every
    section-4 coefficient (m, c, signal, v R^2 mod r) against the key's CSR rows as a multiset,
    every query point, and every witness value as a plain integer."""
    from snarkrs_amd.wire import Q_MOD

    key, wit, d = sample
    sec = _sections((d / "c.zkey").read_bytes(), b"zkey")
    assert struct.unpack("<I", sec[1]) == (1,)  # groth16
    s2 = sec[2]
    n8q, = struct.unpack_from("<I", s2, 0)
    assert n8q == 32 and le_to_int(s2[4:36]) == Q_MOD
    n8r, = struct.unpack_from("<I", s2, 36)
    assert n8r == 32 and le_to_int(s2[40:72]) == R_MOD
    nv, npub, dom = struct.unpack_from("<III", s2, 72)
    assert (nv, npub, dom) == (key.n_vars, key.n_public, key.domain_size)
    pts = s2[84:]
    want = b"".join(getattr(key, f).tobytes() for f in
                    ("alpha_g1", "beta_g1", "beta_g2", "gamma_g2", "delta_g1", "delta_g2"))
    assert pts == want and len(pts) == 3 * 64 + 3 * 128
    assert sec[3] == key.ic.tobytes()
    # section 4: u32 count, then 44-byte records; values double Montgomery (v R^2 mod r)
    s4 = sec[4]
    count, = struct.unpack_from("<I", s4, 0)
    assert len(s4) == 4 + 44 * count
    got = []
    for i in range(count):
        m, c, s = struct.unpack_from("<III", s4, 4 + 44 * i)
        got.append((m, c, s, le_to_int(s4[16 + 44 * i:48 + 44 * i])))
    vals = fr_ints(key.values)
    want = []
    for m in range(2):
        for c in range(key.domain_size):
            for e in range(int(key.row_ptr[m][c]), int(key.row_ptr[m][c + 1])):
                v = vals[int(key.value_id[m][e])]
                want.append((m, c, int(key.signal[m][e]), v * MONT_R * MONT_R % R_MOD))
    assert sorted(got) == sorted(want)
    # sections 5-9: raw Montgomery points; L covers the private wires only
    assert sec[5] == key.a_query.tobytes()
    assert sec[6] == key.b_g1_query.tobytes()
    assert sec[7] == key.b_g2_query.tobytes()
    assert sec[8] == key.l_query.tobytes() and len(sec[8]) == 64 * (nv - npub - 1)
    assert sec[9] == key.h_query.tobytes() and len(sec[9]) == 64 * dom
    # .wtns: section 1 = n8, prime, count; section 2 = plain little-endian integers
    ws = _sections((d / "c.wtns").read_bytes(), b"wtns")
    n8, = struct.unpack_from("<I", ws[1], 0)
    assert n8 == 32 and le_to_int(ws[1][4:36]) == R_MOD
    assert struct.unpack_from("<I", ws[1], 36) == (len(wit),)
    plain = [le_to_int(ws[2][32 * i:32 * i + 32]) for i in range(len(wit))]
    assert plain == fr_ints(wit) and plain[0] == 1


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
