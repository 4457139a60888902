"""GPU parity of the frame form (X/Z shared by all shots, simulated once on
the host; only R runs on the GPU) against the oracle and against the
bit-plane persistent form: bit-exact tableaux and records."""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))


@pytest.fixture(scope="module")
def ps():
    import plaquette_simulator as ps
    import plaquette_simulator_bindings as b
    if b.device_count() < 1:
        pytest.skip("no GPU")
    return ps


@pytest.fixture(scope="module")
def W():
    from plaquette_simulator import workloads as W
    return W


def say(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def assert_same(st, orc):
    x, z, r = st.device_to_host_numpy()
    assert np.array_equal(x, orc.x), "x differs"
    assert np.array_equal(z, orc.z), "z differs"
    assert np.array_equal(r, orc.r), "r differs"
    for i, (res, det) in enumerate(orc.records):
        gres, gdet = st.get_record(i)
        assert np.array_equal(gres, res), f"record {i} results differ"
        assert np.array_equal(gdet, det), f"record {i} determined differ"


@pytest.mark.parametrize("seed,n,B", [(1, 1, 1), (2, 2, 63), (3, 5, 64), (4, 9, 65),
                                      (5, 20, 300), (6, 64, 129), (7, 150, 100)])
def test_random_frame_programs(ps, W, seed, n, B):
    from test_frame_cpu import frame_program
    rng = np.random.default_rng(seed)
    ops = frame_program(W, rng, n, 400)
    say(f"[frame] n={n} B={B} seed={seed} ops={len(ops)}")
    st = ps.CliffordState(n, B, seed, True)
    st.run_program(ops, "frame")
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops)
    assert_same(st, orc)


@pytest.mark.parametrize("d,rounds,p,B", [(3, 3, 0.05, 1000), (5, 5, 0.01, 200)])
def test_surface_code_frame(ps, W, d, rounds, p, B):
    ops, n = W.surface_code_memory_z(d, rounds, p)
    say(f"[frame] surface d={d} B={B}")
    st = ps.CliffordState(n, B, 7, True)
    st.run_program(ops, "frame")
    orc = O.OracleState(n, B, 7)
    orc.run_program(ops)
    assert_same(st, orc)


def test_long_program_many_chunks(ps, W):
    """Several stream chunks and noise windows (the CPU test checks the plan)."""
    from test_frame_cpu import frame_program
    rng = np.random.default_rng(9)
    n, B = 40, 200
    ops = frame_program(W, rng, n, 6000, pflag=0.02, pmeas=0.01, pnoise=0.5, prand=0.05)
    st = ps.CliffordState(n, B, 9, True)
    st.run_program(ops, "frame")
    orc = O.OracleState(n, B, 9)
    orc.run_program(ops)
    assert_same(st, orc)


def test_d9_frame_equals_persistent(ps, W):
    """The d=9, 16384-shot surface code: frame and persistent forms give the
    same planes and records, run after run.  Its 256 word groups fit one
    round of the GPU, so each frame run goes in two parts (the frame and
    record words cross a side buffer between the two launches)."""
    import plaquette_simulator_bindings as b
    wl = W.config_surface_d9()
    assert b.frame_plan(wl.num_qubits, wl.ops)["split_chunk"] > 0
    a = ps.CliffordState(wl.num_qubits, wl.batch_size, wl.seed, True)
    bst = ps.CliffordState(wl.num_qubits, wl.batch_size, wl.seed, True)
    pa = ps.Program(a, wl.ops)
    pb = ps.Program(bst, wl.ops)
    for it in range(2):
        say(f"[frame] d=9 run {it}")
        a.reset_clifford_state()
        bst.reset_clifford_state()
        pa.run(a, "persistent")
        pb.run(bst, "frame")
        assert np.array_equal(a.read_planes(), bst.read_planes()), "planes differ"
        for i in range(a.num_records):
            ra, da = a.get_record(i)
            rb, db = bst.get_record(i)
            assert np.array_equal(ra, rb) and np.array_equal(da, db), f"record {i}"


def test_frame_records_across_runs(ps, W):
    """A flag reading a record written by an earlier run (preloaded by the
    kernel), reset in between."""
    n, B = 4, 130
    a = W.OpList()
    a.h(0)
    m0 = a.measure(0)
    b_ = W.OpList()
    b_.nrec = 1
    b_.flagged(W.OP_X_FLAG, 1, m0)
    b_.flagged(W.OP_Z_FLAG, 2, m0)
    b_.h(2)
    b_.measure(1)
    b_.measure(2)
    st = ps.CliffordState(n, B, 5, True)
    orc = O.OracleState(n, B, 5)
    st.run_program(a.array(), "frame")
    orc.run_program(a.array())
    st.reset_clifford_state()
    orc.reset_clifford_state()
    st.run_program(b_.array(), "frame")
    orc.run_program(b_.array())
    assert_same(st, orc)


def test_frame_refusals_leave_state_untouched(ps, W):
    n, B = 3, 64
    st = ps.CliffordState(n, B, 1, True)
    st.apply_hadamard_gate(0)
    ob = W.OpList()
    ob.measure(0)
    with pytest.raises(RuntimeError, match="reset"):
        st.run_program(ob.array(), "frame")
    assert st.num_records == 0
    st.reset_clifford_state()
    bad = W.OpList()
    m = bad.measure(0)
    bad.flagged(W.OP_H_FLAG, 1, m)
    with pytest.raises(RuntimeError, match="frame form"):
        st.run_program(bad.array(), "frame")
    # Pauli-only programs leave X/Z at |0...0>, so runs can follow each other
    p = W.OpList()
    p.gate(W.OP_X, 1)
    p.noise(2, 0.3, 0.1, 0.1)
    st.run_program(p.array(), "frame")
    st.run_program(p.array(), "frame")
    orc = O.OracleState(n, B, 1)
    orc.run_program(p.array())
    orc.run_program(p.array())
    x, z, r = st.device_to_host_numpy()
    assert np.array_equal(r, orc.r) and np.array_equal(x, orc.x) and np.array_equal(z, orc.z)


def test_gates_without_phase_flips(ps, W):
    """An op stream with nothing for R still writes the new X/Z."""
    ob = W.OpList()
    ob.h(0)
    ob.cx(0, 1)
    ob.s(2)
    st = ps.CliffordState(3, 70, 1, True)
    st.run_program(ob.array(), "frame")
    orc = O.OracleState(3, 70, 1)
    orc.run_program(ob.array())
    assert_same(st, orc)


@pytest.mark.parametrize("n,B,nops,seed", [(300, 128, 3000, 41), (1024, 64, 4000, 42)])
def test_large_states_frame_equals_persistent(ps, W, n, B, nops, seed):
    """Large random Clifford circuits (10% measurements, so most chunks hold
    long random-measurement row lists): the frame form equals the persistent
    form, planes and records."""
    ops = W.random_clifford(n, nops, seed, meas_prob=0.1)
    say(f"[frame] large n={n} B={B} ops={len(ops)}")
    a = ps.CliffordState(n, B, seed, True)
    b = ps.CliffordState(n, B, seed, True)
    a.run_program(ops, "persistent")
    b.run_program(ops, "frame")
    assert np.array_equal(a.read_planes(), b.read_planes()), "planes differ"
    assert a.num_records == b.num_records
    for i in range(a.num_records):
        ra, da = a.get_record(i)
        rb, db = b.get_record(i)
        assert np.array_equal(ra, rb) and np.array_equal(da, db), f"record {i}"


def test_three_forms_agree_at_n2000(ps, W):
    """n=2000 (the measurement kernels keep the scratch row in LDS): the
    plain, persistent and frame forms give the same planes and records."""
    n, B, seed = 2000, 100, 43
    ops = W.random_clifford(n, 2500, seed, meas_prob=0.04)
    say(f"[forms] n={n} B={B} ops={len(ops)}")
    states = {}
    for form in ("plain", "persistent", "frame"):
        st = ps.CliffordState(n, B, seed, True)
        st.run_program(ops, form)
        states[form] = st
    ref = states["persistent"]
    planes = ref.read_planes()
    for form in ("plain", "frame"):
        st = states[form]
        assert np.array_equal(st.read_planes(), planes), f"{form}: planes differ"
        for i in range(ref.num_records):
            ra, da = ref.get_record(i)
            rb, db = st.get_record(i)
            assert np.array_equal(ra, rb) and np.array_equal(da, db), f"{form}: record {i}"


def test_sampling_state(ps, W):
    """A sampling state (phase column only) gives the records of a full state,
    alone and in a frame batch, across a reset; entry points that need the
    full tableau refuse it."""
    ops, n = W.surface_code_memory_z(5, 5, 0.01)
    B, seed = 300, 21
    full = ps.CliffordState(n, B, seed, True)
    prog_f = ps.Program(full, ops)
    samp = ps.SamplingState(n, B, seed)
    prog_s = ps.Program(samp, ops)
    for it in range(2):
        full.reset_clifford_state()
        samp.reset_clifford_state()
        prog_f.run(full, "persistent")
        prog_s.run(samp, "frame")
        assert samp.num_records == full.num_records
        for i in range(full.num_records):
            ra, da = full.get_record(i)
            rb, db = samp.get_record(i)
            assert np.array_equal(ra, rb) and np.array_equal(da, db), f"run {it} record {i}"
    for call in (lambda: samp.device_to_host_numpy(), lambda: samp.apply_hadamard_gate(0),
                 lambda: samp.measure_qubit(0), lambda: samp.read_planes(),
                 lambda: samp.run_program(ops, "persistent")):
        with pytest.raises(RuntimeError):
            call()
    # in a frame batch beside full states
    batch = ps.Batch()
    s2 = ps.SamplingState(n, B, seed)
    batch.add(s2, ps.Program(s2, ops))
    f2 = ps.CliffordState(n, 130, 5, True)
    batch.add(f2, ps.Program(f2, ops))
    batch.run("frame")
    for i in range(full.num_records):
        ra, da = full.get_record(i)
        rb, db = s2.get_record(i)
        assert np.array_equal(ra, rb) and np.array_equal(da, db), f"batch record {i}"
    with pytest.raises(RuntimeError):
        batch.run("persistent")


def test_lazy_reset_reads_and_runs(ps, W):
    """A reset writes only the R column and leaves X/Z pending: whatever
    reads or updates the full tableau next (a host copy, a checksum, a
    gate, a persistent or plain run) sees |0...0>.  A frame run leaves the
    plan's shared X/Z pending in turn, written on the same demand."""
    n, B, seed = 9, 130, 12
    rng = np.random.default_rng(seed)
    ob = W.OpList()
    for _ in range(120):
        u = rng.random()
        q = int(rng.integers(0, n))
        if u < 0.15:
            ob.measure(q, 0.5)
        elif u < 0.3:
            ob.noise(q, 0.1, 0.05, 0.1)
        elif u < 0.5:
            c, t = rng.choice(n, 2, replace=False)
            ob.cx(int(c), int(t))
        else:
            ob.gate(int(rng.choice([W.OP_H, W.OP_S, W.OP_X])), q)
    ops = ob.array()
    st = ps.CliffordState(n, B, seed, True)
    orc = O.OracleState(n, B, seed)
    ident = O.OracleState(n, B, seed)
    say("lazy reset: persistent, reset, host copy")
    st.run_program(ops, "persistent")
    orc.run_program(ops)
    st.reset_clifford_state()
    orc.reset_clifford_state()
    x, z, r = st.device_to_host_numpy()
    assert np.array_equal(x, ident.x) and np.array_equal(z, ident.z) and np.array_equal(r, ident.r)
    say("lazy reset: reset, checksum")
    st.run_program(ops, "persistent")
    orc.run_program(ops)
    st.reset_clifford_state()
    orc.reset_clifford_state()
    assert st.check_sum(0, B) == orc.check_sum(0, B)
    say("lazy reset: reset, frame run, host copy")
    st.run_program(ops, "frame")
    orc.run_program(ops)
    assert_same(st, orc)
    say("lazy reset: reset, frame run, persistent run (X/Z written on demand)")
    st.reset_clifford_state()
    orc.reset_clifford_state()
    st.run_program(ops, "frame")
    orc.run_program(ops)
    st.run_program(ops, "persistent")
    orc.run_program(ops)
    assert_same(st, orc)
    for form in ("persistent", "plain"):
        say(f"lazy reset: reset, gate, {form} run")
        st.reset_clifford_state()
        orc.reset_clifford_state()
        st.apply_hadamard_gate(3)
        orc.apply_hadamard_gate(3)
        st.run_program(ops, form)
        orc.run_program(ops)
        assert_same(st, orc)


def test_sliced_frame_run_equals_persistent(ps, W):
    """More word groups than the GPU holds at once: the frame run goes in
    slices, each slice's noise drawn on a second stream while the previous
    slice runs.  d=5 (4 groups per CU, 1024 at once) with 2^17 shots is 2048
    groups, so two slices; two runs with a reset between also check that the
    second run's noise waits for the first run's kernels."""
    ops, n = W.surface_code_memory_z(5, 3, 0.02)
    B, seed = 1 << 17, 8
    say(f"sliced frame run: n={n} B={B}")
    full = ps.CliffordState(n, B, seed, True)
    prog_f = ps.Program(full, ops)
    samp = ps.SamplingState(n, B, seed)
    prog_s = ps.Program(samp, ops)
    for it in range(2):
        full.reset_clifford_state()
        samp.reset_clifford_state()
        prog_f.run(full, "persistent")
        prog_s.run(samp, "frame")
        for i in range(full.num_records):
            ra, da = full.get_record(i)
            rb, db = samp.get_record(i)
            assert np.array_equal(ra, rb) and np.array_equal(da, db), f"run {it} record {i}"
