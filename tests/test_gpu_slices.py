"""Word groups split between workgroups ("slices") of one cooperative launch.

A persistent run with k workgroups per 64-shot word group must leave the same
tableau words and records as the run with one workgroup per group, and as the
CPU oracle: random circuits (random measurements with deep rowsums, pivot
groups, determined chains) and a noisy surface code (noise planes, batches of
determined measurements, row-tiled gate segments), in both kernel builds
(up to 192 qubits, and above)."""
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


def run_sliced(ps, n, B, seed, ops, slices):
    st = ps.CliffordState(n, B, seed, True)
    prog = ps.Program(st, ops)
    prog.set_slices(slices)
    prog.run(st, "persistent")
    st.sync_records()
    used, failed = prog.slice_status(st)
    return st, used, failed


def check_against_oracle(st, n, B, seed, ops):
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops.astype(np.int32))
    x, z, r = st.device_to_host_numpy()
    assert np.array_equal(x, orc.x), "x differs"
    assert np.array_equal(z, orc.z), "z differs"
    assert np.array_equal(r, orc.r), "r differs"
    assert st.num_records == len(orc.records)
    for i, (res, det) in enumerate(orc.records):
        gres, gdet = st.get_record(i)
        assert np.array_equal(gres, res), f"record {i} results differ"
        assert np.array_equal(gdet, det), f"record {i} determined differ"


def random_ops(ps, n, nops, seed, meas_prob):
    from plaquette_simulator import workloads as W
    return W.random_clifford(n, nops, seed, meas_prob=meas_prob)


@pytest.mark.parametrize("slices", [2, 3, 4, 8])
def test_random_n64_slices_match(ps, slices):
    # also the small build (up to 192 qubits), with slices asked for
    n, B, seed = 64, 256, 17
    ops = random_ops(ps, n, 2000, 5, 0.1)
    say(f"[case] random n64 B{B}: slices {slices} vs 1")
    a, used, failed = run_sliced(ps, n, B, seed, ops, slices)
    assert (used, failed) == (slices, 0)
    b, used1, _ = run_sliced(ps, n, B, seed, ops, 1)
    assert used1 == 1
    assert a.compare(b) == (0, 0)
    check_against_oracle(a, n, B, seed, ops)


def test_random_n520_auto_slices(ps):
    # 520 qubits, 2 word groups: the automatic choice splits each group
    n, B, seed = 520, 128, 3
    ops = random_ops(ps, n, 3000, 9, 0.1)
    say(f"[case] random n520 B{B}: automatic slices vs 1")
    a, used, failed = run_sliced(ps, n, B, seed, ops, 0)
    say(f"[slices] {used}")
    assert used == 2 and failed == 0
    b, _, _ = run_sliced(ps, n, B, seed, ops, 1)
    assert a.compare(b) == (0, 0)
    check_against_oracle(a, n, B, seed, ops)


def test_random_n300_slices(ps):
    # the large build below the automatic threshold, slices asked for
    n, B, seed = 300, 128, 3
    ops = random_ops(ps, n, 3000, 9, 0.1)
    say(f"[case] random n300 B{B}: slices 16 vs 1")
    a, used, failed = run_sliced(ps, n, B, seed, ops, 16)
    assert (used, failed) == (16, 0)
    b, used0, _ = run_sliced(ps, n, B, seed, ops, 0)
    assert used0 == 1
    assert a.compare(b) == (0, 0)
    check_against_oracle(a, n, B, seed, ops)


def test_measure_heavy_n200_slices(ps):
    # every other op a measurement: many barriers and pivot groups per step
    n, B, seed = 200, 192, 21
    ops = random_ops(ps, n, 1500, 4, 0.5)
    say(f"[case] measure-heavy n200 B{B}: slices 4 vs 1")
    a, used, failed = run_sliced(ps, n, B, seed, ops, 4)
    assert (used, failed) == (4, 0)
    b, _, _ = run_sliced(ps, n, B, seed, ops, 1)
    assert a.compare(b) == (0, 0)
    check_against_oracle(a, n, B, seed, ops)


@pytest.mark.parametrize("slices", [2, 5])
def test_surface_code_slices(ps, slices):
    from plaquette_simulator import workloads as W
    ops, n = W.surface_code_memory_z(5, 5, 0.01)
    B, seed = 320, 8
    say(f"[case] surface d5 B{B}: slices {slices} vs 1")
    a, used, failed = run_sliced(ps, n, B, seed, ops, slices)
    assert (used, failed) == (slices, 0)
    b, _, _ = run_sliced(ps, n, B, seed, ops, 1)
    assert a.compare(b) == (0, 0)
    check_against_oracle(a, n, B, seed, ops)


def test_sliced_rerun_and_graph(ps):
    # a sliced program run twice on one state (barrier counters restart),
    # then the graph form (one workgroup per group) continues it
    from plaquette_simulator import workloads as W
    ops, n = W.surface_code_memory_z(3, 3, 0.02)
    B, seed = 128, 12
    a = ps.CliffordState(n, B, seed, True)
    b = ps.CliffordState(n, B, seed, True)
    pa = ps.Program(a, ops)
    pa.set_slices(4)
    pb = ps.Program(b, ops)
    pb.set_slices(1)
    for form in ("persistent", "persistent", "graph"):
        say(f"[case] rerun: {form}")
        pa.run(a, form)
        pb.run(b, form)
    a.sync_records()
    b.sync_records()
    assert pa.slice_status(a)[1] == 0
    assert a.compare(b) == (0, 0)
    orc = O.OracleState(n, B, seed)
    for _ in range(3):
        orc.run_program(ops.astype(np.int32))
    x, z, r = a.device_to_host_numpy()
    assert np.array_equal(x, orc.x) and np.array_equal(z, orc.z) and np.array_equal(r, orc.r)


def test_slices_that_do_not_fit_are_refused(ps):
    n, B = 64, 64 * 64
    ops = random_ops(ps, n, 100, 1, 0.1)
    st = ps.CliffordState(n, B, 1, True)
    prog = ps.Program(st, ops)
    prog.set_slices(1000)  # 64 groups x 1000 workgroups cannot be resident
    with pytest.raises(RuntimeError):
        prog.run(st, "persistent")


def test_barrier_give_up_is_reported_and_the_run_ends(ps):
    # a one-poll wait makes a slice that arrives first give up at once: the
    # run must still end, and the status must say its results are not sound
    n, B, seed = 520, 128, 5
    ops = random_ops(ps, n, 2000, 13, 0.1)
    say(f"[case] random n520 B{B}: slices 2, barriers poll once")
    st = ps.CliffordState(n, B, seed, True)
    prog = ps.Program(st, ops)
    prog.set_slices(2)
    prog.set_slice_wait(1)
    prog.run(st, "persistent")
    st.synchronize()
    used, failed = prog.slice_status(st)
    assert used == 2 and failed == 1
    # a fresh program with the default wait on a fresh state is sound again
    a, used, failed = run_sliced(ps, n, B, seed, ops, 2)
    assert (used, failed) == (2, 0)
    check_against_oracle(a, n, B, seed, ops)
