"""The frame-form compiler (host only) against the oracle, on the CPU.

``frame_plan`` compiles an op list into the stream the GPU's frame kernel
runs; ``frame_interp`` executes that stream in Python.  Final X/Z (shared),
every shot's R column and every record must equal the oracle's run of the
same ops, bit for bit.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
import frame_interp as FI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "plaquette-simulator_amd"))


@pytest.fixture(scope="module")
def b():
    import plaquette_simulator_bindings as b
    return b


@pytest.fixture(scope="module")
def W():
    from plaquette_simulator import workloads as W
    return W


def frame_program(W, rng, n, nops, pflag=0.06, pmeas=0.08, pnoise=0.08, prand=0.05):
    """Random op list without flagged H/S/CX/CZ (X/Z stay shared)."""
    ob = W.OpList()
    for _ in range(nops):
        u = rng.random()
        if u < pmeas:
            q = int(rng.integers(0, n))
            bias = float(rng.choice([0.5, 0.5, 0.0, 1.0, 0.3]))
            if ob.nrec and rng.random() < 0.2:
                ob.measure(q, bias, slot=int(rng.integers(0, ob.nrec)))  # re-use a record
            else:
                ob.measure(q, bias)
        elif u < pmeas + pflag and ob.nrec:
            code = int(rng.choice([W.OP_X_FLAG, W.OP_Y_FLAG, W.OP_Z_FLAG]))
            ob.flagged(code, int(rng.integers(0, n)), int(rng.integers(0, ob.nrec)))
        elif u < pmeas + pflag + prand:
            gate = int(rng.choice([W.OP_X, W.OP_Y, W.OP_Z]))
            ob.with_prob(gate, int(rng.integers(0, n)), float(rng.choice([0.1, 0.5, 1.0])))
        elif u < pmeas + pflag + prand + pnoise:
            ob.noise(int(rng.integers(0, n)), 0.1, 0.05, 0.2)
        else:
            code = int(rng.choice([W.OP_H, W.OP_S, W.OP_CX, W.OP_CZ, W.OP_X, W.OP_Y, W.OP_Z]))
            if code in (W.OP_CX, W.OP_CZ):
                if n > 1:
                    c, t = rng.choice(n, 2, replace=False)
                    ob.gate(code, int(c), int(t))
            else:
                ob.gate(code, int(rng.integers(0, n)))
    return ob.array()


def check(b, n, B, seed, ops):
    plan = b.frame_plan(n, ops)
    r, recs = FI.run(plan, n, B, seed, ops)
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops)
    x, z = FI.shared_xz(plan, n)
    for s in range(B):
        assert np.array_equal(orc.x[s], x), f"x differs (shot {s})"
        assert np.array_equal(orc.z[s], z), f"z differs (shot {s})"
    assert np.array_equal(orc.r, r), "r differs"
    for slot, (res, det) in enumerate(orc.records):
        assert slot in recs, f"record {slot} missing"
        assert np.array_equal(recs[slot][0], res), f"record {slot} results differ"
        assert np.array_equal(recs[slot][1], det), f"record {slot} determined differ"
    return plan


@pytest.mark.parametrize("seed,n,B", [(1, 1, 5), (2, 2, 64), (3, 5, 70), (4, 9, 40),
                                      (5, 16, 65), (6, 30, 20)])
def test_random_frame_programs(b, W, seed, n, B):
    rng = np.random.default_rng(seed)
    ops = frame_program(W, rng, n, 250)
    check(b, n, B, seed, ops)


@pytest.mark.parametrize("d,rounds,p,B", [(3, 3, 0.05, 100), (5, 2, 0.02, 64)])
def test_surface_code_frame(b, W, d, rounds, p, B):
    ops, n = W.surface_code_memory_z(d, rounds, p)
    plan = check(b, n, B, 7, ops)
    assert plan["random"] > 0 and plan["determined"] > 0


def test_chunking_and_noise_windows(b, W):
    """A long noisy program spans several stream chunks and noise windows."""
    rng = np.random.default_rng(9)
    n = 40
    ops = frame_program(W, rng, n, 6000, pflag=0.02, pmeas=0.01, pnoise=0.5, prand=0.05)
    plan = check(b, n, 30, 9, ops)
    assert plan["chunks_n"] >= 3
    ks = plan["chunks"][:, 3]
    assert (ks > 0).sum() >= 2


def test_pauli_only_program_keeps_identity(b, W):
    ob = W.OpList()
    for q in range(4):
        ob.gate(W.OP_X, q)
        ob.noise(q, 0.2, 0.1, 0.1)
    plan = check(b, 4, 64, 3, ob.array())
    assert plan["final_identity"] == 1


def test_flagged_h_is_not_frame_eligible(b, W):
    ob = W.OpList()
    m = ob.measure(0)
    ob.flagged(W.OP_H_FLAG, 1, m)
    with pytest.raises(RuntimeError, match="frame form"):
        b.frame_plan(3, ob.array())
    ob2 = W.OpList()
    ob2.with_prob(W.OP_CX, 0, 0.5, q1=1)
    with pytest.raises(RuntimeError, match="frame form"):
        b.frame_plan(3, ob2.array())


def test_gates_without_phase_flips(b, W):
    """H, CX, S on |0...0> flip no phase, so the stream is empty, but X/Z
    change (the plan's final X/Z, which the run writes into every group)."""
    ob = W.OpList()
    ob.h(0)
    ob.cx(0, 1)
    ob.s(2)
    plan = check(b, 3, 10, 1, ob.array())
    assert plan["stream_words"] == 0 and plan["chunks_n"] == 0


def test_record_pool_reuse(b, W):
    """More flag-read records live at once than the pool hands out fresh
    (128), then reuse: every value must still reach its readers."""
    rng = np.random.default_rng(21)
    n = 12
    ob = W.OpList()
    for q in range(n):
        ob.h(q)
    for _ in range(3):
        slots = [ob.measure(int(rng.integers(0, n)), float(rng.choice([0.5, 0.2]))) for _ in range(150)]
        for s in rng.permutation(slots):
            ob.flagged(int(rng.choice([W.OP_X_FLAG, W.OP_Y_FLAG, W.OP_Z_FLAG])),
                       int(rng.integers(0, n)), int(s))
            if rng.random() < 0.2:
                ob.noise(int(rng.integers(0, n)), 0.1, 0.1, 0.1)
    plan = check(b, n, 70, 21, ob.array())
    assert plan["nref"] > 128


def test_long_tasks_split(b, W):
    """A long run of updates on one qubit pair outgrows one task (256 stream
    words) and one task's noise span; the splits keep the order."""
    ob = W.OpList()
    ob.h(0)
    for i in range(700):
        ob.cx(0, 1)
        ob.s(1)
        ob.noise(0, 0.05, 0.05, 0.05)
        if i % 50 == 0:
            ob.noise(1, 0.2, 0.0, 0.1)
    ob.measure(0)
    ob.measure(1)
    check(b, 2, 130, 5, ob.array())


def test_determined_measurements_out_of_program_order(b, W):
    """Determined measurements join an earlier batch when no random step
    wrote their rows; the scratch row still ends with the program's last."""
    rng = np.random.default_rng(4)
    n = 10
    ob = W.OpList()
    for q in range(0, n, 2):
        ob.h(q)
        ob.cx(q, q + 1)
    for _ in range(6):
        for q in rng.permutation(n):
            ob.measure(int(q), float(rng.choice([0.5, 0.1])))
            if rng.random() < 0.3:
                ob.noise(int(rng.integers(0, n)), 0.1, 0.05, 0.1)
        for q in range(0, n, 2):
            ob.h(q)
            ob.cx(q, q + 1)
    plan = check(b, n, 100, 4, ob.array())
    assert plan["random"] > 0 and plan["determined"] > 0


@pytest.mark.parametrize("d,rounds,p", [(5, 5, 0.01), (7, 7, 0.001), (9, 9, 0.001)])
def test_two_part_split_point(b, W, d, rounds, p):
    """A run of one round of word groups splits at ``split_chunk``: the chunks
    before it read only noise ops below the largest ``k0 + nk`` among them,
    which covers at least half, and not all, of the ops the plan reads."""
    ops, n = W.surface_code_memory_z(d, rounds, p)
    plan = b.frame_plan(n, ops)
    c = plan["split_chunk"]
    assert 0 < c < plan["chunks_n"]
    ch = plan["chunks"].astype(np.int64)
    ends = np.where(ch[:, 3] > 0, ch[:, 2] + ch[:, 3], 0)
    split_k, total = ends[:c].max(), ends.max()
    assert 2 * split_k >= total - 1 and split_k < total
    # the first chunk from which the split point is reached
    assert 2 * ends[:c - 1].max(initial=0) < total - 1 or c == 1


def test_no_split_without_noise_or_second_chunk(b, W):
    ops, n = W.surface_code_memory_z(3, 3, 0.05)
    assert b.frame_plan(n, ops)["chunks_n"] == 1
    assert b.frame_plan(n, ops)["split_chunk"] == 0
    ob = W.OpList()
    for q in range(6):
        ob.h(q)
    for q in range(6):
        ob.measure(q)
    assert b.frame_plan(6, ob.array())["split_chunk"] == 0
