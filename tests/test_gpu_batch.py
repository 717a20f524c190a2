"""Batched proving and the serving loop (g16hip_prove_batch, g16hip_server_*): every proof
is bit-identical to the same proof made alone (and to the oracle's), whatever else is in
flight; graph replays interleaved with eager proofs on the same circuit all agree."""
import threading

import numpy as np
import pytest

from oracle.oracle import Synth
from snarkrs_amd import native
from snarkrs_amd.wire import fr_array, fr_ints

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def circuits(dev, orc):
    out = []
    for i, nc in enumerate([40, 700, 3000]):
        syn = Synth(orc, nc, n_public=2, n_private_in=4, bool_permille=300,
                    input_bool_permille=300, seed=100 + i, setup_seed=200 + i)
        key, wit = syn.key()
        circ = dev.prepare(key)
        out.append((syn, key, wit, circ))
    return out


def _blind(i):
    return fr_array([1000 + i]), fr_array([2000 + 3 * i])


@pytest.fixture(params=[False, True], ids=["eager", "graphs"])
def replay(request):
    native.set_graph_replay(request.param)
    yield request.param
    native.set_graph_replay(False)


def test_batch_matches_single_proofs_and_oracle(dev, orc, circuits, replay):
    for _, _, _, circ in circuits:
        circ.add_slots(2)
    # independent witnesses of the same circuits
    items, want, wits = [], [], []
    for i in range(12):
        syn, key, _, circ = circuits[i % 3]
        wit = syn.witness(900 + i)
        r, s = _blind(i)
        items.append((circ, wit, r, s))
        wits.append(wit)
        want.append(circ.prove_with_blinders(wit, r, s))
    got = native.prove_batch(items)
    for i, (g, w) in enumerate(zip(got, want)):
        assert (g == w).all(), i
    # and against the oracle for a few
    for i in (0, 4, 8):
        syn, key, _, circ = circuits[i % 3]
        r, s = _blind(i)
        assert (got[i] == orc.prove(key, wits[i], fr_ints(r)[0], fr_ints(s)[0])[1]).all()


def test_graph_replays_interleaved_with_eager_proofs(dev, circuits):
    native.set_graph_replay(True)
    syn, key, wit, circ = circuits[2]
    r, s = _blind(7)
    eager = circ.prove_with_blinders(wit, r, s)
    for k in range(4):
        (g,) = native.prove_batch([(circ, wit, r, s)])
        assert (g == eager).all(), f"replay {k}"
        assert (circ.prove_with_blinders(wit, r, s) == eager).all()
    native.set_graph_replay(False)


def test_server_results_are_the_same_alone_and_in_the_mix(dev, circuits, replay):
    alone = []
    for i in range(9):
        _, _, wit, circ = circuits[i % 3]
        r, s = _blind(50 + i)
        alone.append(circ.prove_with_blinders(wit, r, s))
    srv = native.Server(dev, reserve_bytes=1 << 30, lanes=3)
    tickets = []
    lock = threading.Lock()

    def submit(i):
        _, _, wit, circ = circuits[i % 3]
        r, s = _blind(50 + i)
        t = srv.submit(circ, wit, r, s)
        with lock:
            tickets.append((i, t))

    ths = [threading.Thread(target=submit, args=(i,)) for i in range(9)]
    [t.start() for t in ths]
    [t.join() for t in ths]
    for i, t in tickets:
        proof, lat = srv.wait(t)
        assert lat > 0
        assert (proof == alone[i]).all(), i
    srv.close()


def test_batch_reports_bad_items(dev, circuits):
    _, _, wit, circ = circuits[0]
    r, s = _blind(1)
    with pytest.raises(native.WitnessLength):
        native.prove_batch([(circ, wit[:-1], r, s)])


@pytest.mark.parametrize("cap", [1, 3, 16])
def test_waves_of_every_size_match_single_proofs(dev, circuits, cap):
    # 7 witnesses of one circuit: full waves and a partial last wave (cap 3: 3 + 3 + 1)
    syn, key, _, circ = circuits[1]
    circ.set_batch_cap(cap)
    try:
        items, want = [], []
        for i in range(7):
            w = syn.witness(3100 + i)
            r, s = _blind(300 + i)
            items.append((circ, w, r, s))
            want.append(circ.prove_with_blinders(w, r, s))
        for rep in range(2):  # the wave's buffers are reused (dirty) the second time
            got = native.prove_batch(items)
            for i, (g, w) in enumerate(zip(got, want)):
                assert (g == w).all(), (cap, rep, i)
    finally:
        circ.set_batch_cap(0)


def test_bad_items_inside_a_wave_keep_their_status(dev, circuits):
    syn, key, wit, circ = circuits[2]
    circ.set_batch_cap(4)
    try:
        bad = wit.copy()
        bad[0] = fr_array([2])[0]
        items = []
        for i in range(5):
            r, s = _blind(400 + i)
            items.append((circ, bad if i == 1 else syn.witness(4100 + i),
                          fr_array([0]) if i == 3 else r, s))
        proofs, st = native.prove_batch(items, statuses=True)
        assert st[1] == native.CONSTANT_WIRE and st[3] == native.ZERO_BLINDER
        for i in (0, 2, 4):
            assert st[i] == native.OK
            c, w, r, s = items[i]
            assert (proofs[i] == circ.prove_with_blinders(w, r, s)).all(), i
        with pytest.raises(native.ConstantWire):
            native.prove_batch(items)
    finally:
        circ.set_batch_cap(0)


def test_autotune_picks_a_measured_wave_size(dev, orc, circuits):
    syn, key, wit, circ = circuits[2]
    res = circ.batch_autotune(wit, 8)
    assert res["best"] in res["tried"] and set(res["tried"]) <= {1, 2, 4, 8}
    assert all(v > 0 for v in res["tried"].values())
    items = [(circ, syn.witness(5100 + i), *_blind(500 + i)) for i in range(6)]
    got = native.prove_batch(items)
    for i, (c, w, r, s) in enumerate(items):
        assert (got[i] == orc.prove(key, w, fr_ints(r)[0], fr_ints(s)[0])[1]).all(), i
    circ.set_batch_cap(0)
