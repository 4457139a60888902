"""GPU parity of two orderings the reference gets from running one op at a time.

* A measurement that writes an existing record overwrites it, so of several
  measurements into one record the last in program order is what stays
  (CliffordState.hpp:608-611).  The compiled forms run a run of measurements
  that are determined in every shot in parallel, one wave each; these tests
  put several such measurements into one record, next to random ones, in
  every execution form and in batches, against the oracle.
* Kokkos::deep_copy has finished before the source or the destination
  changes again (copy constructor CliffordState.hpp:168-174, DeviceToDevice
  :268-274).  These tests change both states straight after the copy and
  compare the copy with a snapshot of the source taken before it.
"""
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
    assert st.num_records == len(orc.records)
    for i, (res, det) in enumerate(orc.records):
        gres, gdet = st.get_record(i)
        assert np.array_equal(gres, res), f"record {i} results differ"
        assert np.array_equal(gdet, det), f"record {i} determined differ"


def reuse_program(W, n, rounds):
    """Runs of consecutive measurements that reuse records.

    q1 and q3 are flipped, so they measure 1 and q0, q2 measure 0, each
    determined in every shot; q5.. are put in superposition, so their first
    measurement is random.  Each round: several determined measurements into
    one record with alternating outcomes, a new record written twice in the
    same run, random-then-determined and determined-then-random pairs, and a
    long alternating chain into one record (more measurements than waves).
    """
    b = W.OpList()
    b.gate(W.OP_X, 1)
    b.gate(W.OP_X, 3)
    for q in range(5, n):
        b.h(q)
    for rnd in range(rounds):
        s0 = b.measure(0)
        b.measure(1, slot=s0)      # 0 then 1: record s0 must read 1
        s1 = b.measure(1)
        b.measure(0, slot=s1)      # 1 then 0: record s1 must read 0
        b.measure(2, slot=s1)
        b.measure(3, slot=s1)      # ... then 1
        s2 = b.measure(5 + (rnd % (n - 5)))  # random (first round), then determined
        b.measure(3, slot=s2)
        s3 = b.measure(2)
        b.measure(6 + (rnd % (n - 6)), slot=s3)  # determined, then random/determined
        s4 = b.measure(0)
        for k in range(23):        # alternating 0/1 into one record, ends on q2
            b.measure((0, 1, 2, 3)[k % 4], slot=s4)
        b.measure(1, slot=s0)
        # gates between rounds keep the next round's runs apart and move the
        # random qubits on
        for q in range(5, n):
            b.h(q)
        b.cx(5, n - 1)
        b.gate(W.OP_X_FLAG, 4, s2)
    return b.array()


@pytest.mark.parametrize("form", ["plain", "persistent", "graph", "frame"])
@pytest.mark.parametrize("n,B,seed", [(9, 200, 3), (12, 64, 4), (40, 130, 5)])
def test_record_reuse_in_measurement_runs(ps, W, form, n, B, seed):
    ops = reuse_program(W, n, 3)
    say(f"[case] record reuse n={n} B={B} seed={seed} form={form} ops={len(ops)}")
    st = ps.CliffordState(n, B, seed, True)
    st.run_program(ops, form)
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops.astype(np.int32))
    # last writers: s0 q1 (outcome 1), s1 q3 (1), the chain into s4 q2 (0)
    assert orc.records[0][0].all() and orc.records[1][0].all() and not orc.records[4][0].any()
    assert orc.records[4][1].all()
    assert_same(st, orc)


@pytest.mark.parametrize("form", ["persistent", "graph", "frame"])
def test_record_reuse_in_batches(ps, W, form):
    members = [(9, 200, 3), (12, 64, 4), (40, 130, 5), (20, 1000, 6)]
    b = ps.Batch()
    keep = []
    for i, (n, B, seed) in enumerate(members):
        ops = reuse_program(W, n, 2 + i % 2)
        st = ps.CliffordState(n, B, seed, True)
        prog = ps.Program(st, ops)
        b.add(st, prog)
        keep.append((st, prog, ops, n, B, seed))
    say(f"[case] record reuse batch form={form}")
    b.run(form)
    for st, _, ops, n, B, seed in keep:
        st.sync_records()  # a batch run updates the records behind the state's back
        orc = O.OracleState(n, B, seed)
        orc.run_program(ops.astype(np.int32))
        assert_same(st, orc)


def test_batch_refuses_a_program_twice(ps, W):
    ops = reuse_program(W, 9, 1)
    st1 = ps.CliffordState(9, 64, 1, True)
    st2 = ps.CliffordState(9, 64, 2, True)
    prog = ps.Program(st1, ops)
    b = ps.Batch()
    b.add(st1, prog)
    with pytest.raises(RuntimeError, match="program already in the batch"):
        b.add(st2, prog)


def _scrambled(ps, W, n, B, seed, nops=400):
    rng = np.random.default_rng(seed)
    b = W.OpList()
    for _ in range(nops):
        u = rng.random()
        if u < 0.04:
            b.measure(int(rng.integers(0, n)))
        elif u < 0.5:
            c, t = rng.choice(n, 2, replace=False)
            b.cx(int(c), int(t))
        else:
            b.gate(int(rng.choice([W.OP_H, W.OP_S, W.OP_Y])), int(rng.integers(0, n)))
    ops = b.array()
    st = ps.CliffordState(n, B, seed, True)
    st.run_program(ops, "persistent")
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops.astype(np.int32))
    return st, orc


def _mutate(obj, n):
    for q in range(n):
        obj.apply_hadamard_gate(q)
        obj.apply_control_not_gate(q, (q + 1) % n)
        obj.apply_phase_gate(q)


def test_copy_then_change_source_against_oracle(ps, W):
    n, B = 24, 700
    say(f"[case] copy ctor then source changes n={n} B={B}")
    st, orc = _scrambled(ps, W, n, B, 41)
    x0, z0, r0 = orc.x.copy(), orc.z.copy(), orc.r.copy()
    cp = st.copy()
    _mutate(st, n)
    st.measure_qubit(0)
    _mutate(orc, n)
    orc.measure_qubit(0)
    x, z, r = cp.device_to_host_numpy()
    assert np.array_equal(x, x0) and np.array_equal(z, z0) and np.array_equal(r, r0)
    assert_same(st, orc)


def test_device_to_device_then_change_both(ps, W):
    n, B = 24, 700
    say(f"[case] device_to_device then both change n={n} B={B}")
    src, orc = _scrambled(ps, W, n, B, 42)
    dst, _ = _scrambled(ps, W, n, B, 43)   # busy with its own queued work
    snap = O.OracleState.from_flat(orc.x.ravel(), orc.z.ravel(), orc.r.ravel(), n, B, 42)
    dst.device_to_device(src)
    _mutate(src, n)
    _mutate(orc, n)
    dst.apply_hadamard_gate(3)              # dst's own work comes after the copy
    snap.apply_hadamard_gate(3)
    x, z, r = dst.device_to_host_numpy()
    assert np.array_equal(x, snap.x) and np.array_equal(z, snap.z) and np.array_equal(r, snap.r)
    assert_same(src, orc)


def test_copy_of_large_state_then_source_changes(ps, W):
    """The d=9 tableau (214 MB, 16384 shots): the copy takes tens of
    microseconds, and gate launches on the source follow at once."""
    w = W.config_surface_d9()
    n = w.num_qubits
    say(f"[case] copy d=9 tableau n={n} B={w.batch_size}")
    st = ps.CliffordState(n, w.batch_size, w.seed, True)
    st.run_program(w.ops, "persistent")
    before = st.read_planes()
    cp = st.copy()
    for q in range(0, n, 2):
        st.apply_hadamard_gate(q)
        st.apply_control_not_gate(q, (q + 1) % n)
    dst = ps.CliffordState(n, w.batch_size, 5, True)
    dst.device_to_device(cp)
    for q in range(n):
        cp.apply_phase_gate(q)
    assert np.array_equal(dst.read_planes(), before)
    assert not np.array_equal(st.read_planes(), before)


def test_program_runs_on_two_states_are_ordered(ps, W):
    """One program run on two states back to back (its noise planes serve
    both runs): each state equals the oracle running the program alone."""
    ops, n = W.surface_code_memory_z(5, 5, 0.01)
    B = 4096
    a = ps.CliffordState(n, B, 9, True)
    b = ps.CliffordState(n, B, 10, True)
    prog = ps.Program(a, ops)
    say(f"[case] one program on two states n={n} B={B}")
    prog.run(a, "persistent")
    prog.run(b, "persistent")
    for st, seed in ((a, 9), (b, 10)):
        ref = ps.CliffordState(n, B, seed, True)
        ref.run_program(ops, "persistent")
        assert np.array_equal(st.read_planes(), ref.read_planes())
    for st, seed in ((a, 9), (b, 10)):
        orc = O.OracleState(n, 130, seed)
        orc.run_program(ops.astype(np.int32))
        x, z, r = st.device_to_host_numpy()
        assert np.array_equal(x[:130], orc.x) and np.array_equal(r[:130], orc.r)
