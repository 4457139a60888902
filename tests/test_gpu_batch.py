"""Mixed batches: every member's results are bit-identical alone and in the
batch (one persistent launch over a packed work list), in the persistent and
graph forms, across repeated runs and after autotuning."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ps():
    import plaquette_simulator as ps
    import plaquette_simulator_bindings as b
    if b.device_count() < 1:
        pytest.skip("no GPU")
    return ps


def say(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def members():
    from plaquette_simulator import workloads as W
    from test_gpu_parity import random_program
    out = []
    for i, (d, B, p) in enumerate([(3, 64, 0.01), (5, 100, 0.003), (3, 333, 0.02), (7, 130, 0.001),
                                   (5, 64, 0.0), (9, 200, 0.001)]):
        ops, n = W.surface_code_memory_z(d, d, p)
        out.append((n, B, 1000 + i, ops))
    rng = np.random.default_rng(4)
    for i, (n, B) in enumerate([(5, 70), (12, 128), (33, 65), (64, 190)]):
        out.append((n, B, 2000 + i, random_program(W, rng, n, 250)))
    return out


def snapshot(st):
    recs = [st.get_record(i) for i in range(st.num_records)]
    return st.read_planes(), recs


def standalone(ps, spec, runs=1):
    n, B, seed, ops = spec
    st = ps.CliffordState(n, B, seed, True)
    prog = ps.Program(st, ops)
    for _ in range(runs):
        prog.run(st, "persistent")
    st.synchronize()
    st.sync_records()
    return snapshot(st)


def batched(ps, specs, form, runs=1, autotune=False, word_tile=1):
    batch = ps.Batch()
    batch.set_word_tile(word_tile)
    states = []
    for n, B, seed, ops in specs:
        st = ps.CliffordState(n, B, seed, True)
        prog = ps.Program(st, ops)
        batch.add(st, prog)
        states.append(st)
    if autotune:
        choice = batch.autotune(1)
        say(f"[autotune] {choice}")
    for _ in range(runs):
        batch.run(form)
    out = []
    for st in states:
        st.sync_records()
        out.append(snapshot(st))
    return out, batch


def same(a, b):
    (pa, ra), (pb, rb) = a, b
    assert np.array_equal(pa, pb), "tableau differs"
    assert len(ra) == len(rb)
    for (x1, y1), (x2, y2) in zip(ra, rb):
        assert np.array_equal(x1, x2) and np.array_equal(y1, y2), "record differs"


@pytest.mark.parametrize("form", ["persistent", "graph"])
def test_batch_equals_standalone(ps, form):
    specs = members()
    say(f"[case] batch of {len(specs)} form={form}")
    got, batch = batched(ps, specs, form)
    info = batch.info()
    assert info["members"] == len(specs)
    assert info["items"] == sum((B + 63) // 64 for _, B, _, _ in specs)
    for spec, g in zip(specs, got):
        same(standalone(ps, spec), g)


@pytest.mark.parametrize("form,tile", [("persistent", 2), ("persistent", 3), ("graph", 4)])
def test_batch_word_tiles_equal_standalone(ps, form, tile):
    # several word groups per work item (members of 1..6 groups, so tiles
    # end short at a member's last group)
    specs = members()
    say(f"[case] batch of {len(specs)} form={form} word tile {tile}")
    got, batch = batched(ps, specs, form, word_tile=tile)
    assert batch.info()["word_tile"] == tile
    for spec, g in zip(specs, got):
        same(standalone(ps, spec), g)


def test_batch_repeated_runs_and_oracle(ps):
    specs = members()[:5]
    say("[case] batch graph x2")
    got, _ = batched(ps, specs, "graph", runs=2)
    for spec, g in zip(specs, got):
        same(standalone(ps, spec, runs=2), g)
    # one member against the oracle, two runs
    n, B, seed, ops = specs[1]
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops.astype(np.int32))
    orc.run_program(ops.astype(np.int32))
    planes, recs = got[1]
    st = ps.CliffordState(n, B, seed, True)
    prog = ps.Program(st, ops)
    prog.run(st, "persistent"); prog.run(st, "persistent"); st.synchronize(); st.sync_records()
    x, z, r = st.device_to_host_numpy()
    assert np.array_equal(st.read_planes(), planes)
    assert np.array_equal(x, orc.x) and np.array_equal(z, orc.z) and np.array_equal(r, orc.r)
    for i, (res, det) in enumerate(orc.records):
        assert np.array_equal(recs[i][0], res) and np.array_equal(recs[i][1], det)


def test_batch_after_autotune(ps):
    specs = members()
    say("[case] batch autotune + graph")
    got, batch = batched(ps, specs, "graph", autotune=True)
    for spec, g in zip(specs, got):
        same(standalone(ps, spec), g)


def frame_members():
    """Members the frame form takes: surface codes and random programs without
    flagged H/S/CX/CZ (resets and other flagged Paulis included)."""
    from plaquette_simulator import workloads as W
    from test_frame_cpu import frame_program
    out = [m for m in members() if m[0] in (17, 49, 97, 161)]  # the surface codes
    rng = np.random.default_rng(6)
    for i, (n, B) in enumerate([(5, 70), (12, 128), (33, 65), (64, 190)]):
        out.append((n, B, 3000 + i, frame_program(W, rng, n, 250)))
    return out


def test_batch_frame_equals_standalone(ps):
    """The frame form of a mixed batch (one noise prologue, one frame kernel
    over the packed list, members of one structure sharing a plan) equals each
    member's standalone persistent run; a reset and a second run too."""
    specs = frame_members()
    say(f"[case] frame batch of {len(specs)}")
    batch = ps.Batch()
    states = []
    for n, B, seed, ops in specs:
        st = ps.CliffordState(n, B, seed, True)
        batch.add(st, ps.Program(st, ops))
        states.append(st)
    batch.run("frame")
    for spec, st in zip(specs, states):
        st.sync_records()
        same(standalone(ps, spec), snapshot(st))
    for st in states:
        st.reset_clifford_state()
    batch.run("frame")
    for spec, st in zip(specs, states):
        st.sync_records()
        n, B, seed, ops = spec
        ref = ps.CliffordState(n, B, seed, True)
        prog = ps.Program(ref, ops)
        prog.run(ref, "persistent")
        ref.reset_clifford_state()
        prog.run(ref, "persistent")
        ref.synchronize()
        ref.sync_records()
        same(snapshot(ref), snapshot(st))


def test_batch_frame_refusals(ps):
    """A member the frame form cannot take, or one not at |0...0>, refuses the
    whole run before any member changes."""
    from plaquette_simulator import workloads as W
    specs = frame_members()[:3]
    ob = W.OpList()
    m = ob.measure(0)
    ob.flagged(W.OP_H_FLAG, 1, m)
    batch = ps.Batch()
    states = []
    for n, B, seed, ops in specs + [(4, 64, 9, ob.array())]:
        st = ps.CliffordState(n, B, seed, True)
        batch.add(st, ps.Program(st, ops))
        states.append(st)
    with pytest.raises(RuntimeError, match="frame form"):
        batch.run("frame")
    assert all(st.num_records == 0 for st in states)
    batch2 = ps.Batch()
    st2 = []
    for n, B, seed, ops in specs:
        st = ps.CliffordState(n, B, seed, True)
        batch2.add(st, ps.Program(st, ops))
        st2.append(st)
    st2[1].apply_hadamard_gate(0)
    with pytest.raises(RuntimeError, match="reset"):
        batch2.run("frame")
    assert all(st.num_records == 0 for st in st2)
