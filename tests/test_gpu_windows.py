"""Gate segments cut into qubit windows.

A gate segment over more qubits than a row tile can hold is cut into windows
of fewer qubits, each its own row-tiled segment; every qubit's ops keep their
program order, so the windows must leave the oracle's tableau words and
records, bit-exact.  Small window sizes (set process-wide for programs
created after the call) force the cut on small circuits; a d=13 surface code
(337 qubits) is cut at the default size."""
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
    yield ps
    b.set_window_qubits(0)


def say(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def check(st, orc):
    st.sync_records()
    x, z, r = st.device_to_host_numpy()
    assert np.array_equal(x, orc.x), "x differs"
    assert np.array_equal(z, orc.z), "z differs"
    assert np.array_equal(r, orc.r), "r differs"
    assert st.num_records == len(orc.records)
    for i, (res, det) in enumerate(orc.records):
        gres, gdet = st.get_record(i)
        assert np.array_equal(gres, res), f"record {i} results differ"
        assert np.array_equal(gdet, det), f"record {i} determined differ"


def segments(ps, n, B, seed, ops, cap):
    import plaquette_simulator_bindings as b
    b.set_window_qubits(cap)
    st = ps.CliffordState(n, B, seed, True)
    prog = ps.Program(st, ops)
    b.set_window_qubits(0)
    return st, prog


def cases():
    from plaquette_simulator import workloads as W
    sops, sn = W.surface_code_memory_z(5, 3, 0.02)
    rops = W.random_clifford(64, 3000, 21, meas_prob=0.02)
    return [("surface d5", sn, 192, 4, sops), ("random n64", 64, 128, 6, rops)]


@pytest.mark.parametrize("cap", [6, 16, 40])
def test_windows_match_oracle(ps, cap):
    for name, n, B, seed, ops in cases():
        _, ref = segments(ps, n, B, seed, ops, 0xFFFFFFFF)
        for form in ("persistent", "graph"):
            say(f"[case] {name} windows of {cap} qubits, {form} x2")
            st, prog = segments(ps, n, B, seed, ops, cap)
            assert prog.info()["segments"] > ref.info()["segments"]
            prog.run(st, form)
            prog.run(st, form)
            orc = O.OracleState(n, B, seed)
            orc.run_program(ops.astype(np.int32))
            orc.run_program(ops.astype(np.int32))
            check(st, orc)


def test_windows_in_a_batch_and_sliced(ps):
    name, n, B, seed, ops = cases()[0]
    say(f"[case] {name} windows of 12 qubits: batch member and 3 slices")
    a, pa = segments(ps, n, B, seed, ops, 12)
    c, pc = segments(ps, n, B, seed, ops, 12)
    batch = ps.Batch()
    batch.add(a, pa)
    batch.run("persistent")
    pc.set_slices(3)
    pc.run(c, "persistent")
    assert pc.slice_status(c) == (3, 0)
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops.astype(np.int32))
    check(a, orc)
    check(c, orc)


def every_op_program(n, seed):
    """Layers of every op kind: gates, Paulis, record-flagged and randomly
    flagged gates (one and two qubits), noise fused and alone, measurements."""
    from plaquette_simulator import workloads as W
    rng = np.random.default_rng(seed)
    b = W.OpList()
    for q in range(n):
        b.h(q)
    for _ in range(4):
        for _ in range(3 * n):
            k = int(rng.integers(0, 7))
            q, t = (int(x) for x in rng.choice(n, 2, replace=False))
            if k == 0:
                b.cx(q, t)
            elif k == 1:
                b.cz(q, t)
            elif k == 2:
                b.gate(int(rng.choice([W.OP_S, W.OP_X, W.OP_Y, W.OP_Z, W.OP_H])), q)
            elif k == 3 and b.nrec:
                b.flagged(int(rng.choice([W.OP_H_FLAG, W.OP_S_FLAG, W.OP_X_FLAG, W.OP_Y_FLAG,
                                          W.OP_Z_FLAG])), q, int(rng.integers(0, b.nrec)))
            elif k == 4 and b.nrec:
                b.flagged(int(rng.choice([W.OP_CX_FLAG, W.OP_CZ_FLAG])), q,
                          int(rng.integers(0, b.nrec)), t)
            elif k == 5:
                b.with_prob(int(rng.choice([W.OP_H, W.OP_S, W.OP_CX])), q, 0.3, t)
            else:
                b.noise(q, 0.05, 0.02, 0.05)
        for q in rng.choice(n, n // 4, replace=False):
            b.measure(int(q))
    return b.array()


@pytest.mark.parametrize("cap", [5, 23])
def test_windows_every_op_kind(ps, cap):
    n, B, seed = 40, 192, 13
    ops = every_op_program(n, 2)
    say(f"[case] every op kind n{n}, windows of {cap} qubits, persistent and graph")
    for form in ("persistent", "graph"):
        st, prog = segments(ps, n, B, seed, ops, cap)
        prog.run(st, form)
        orc = O.OracleState(n, B, seed)
        orc.run_program(ops.astype(np.int32))
        check(st, orc)


def test_surface_d13_cut_at_default_size(ps):
    from plaquette_simulator import workloads as W
    ops, n = W.surface_code_memory_z(13, 2, 0.005)
    B, seed = 64, 3
    say(f"[case] surface d13 (n={n}) default windows")
    st, prog = segments(ps, n, B, seed, ops, 0)
    _, whole = segments(ps, n, B, seed, ops, 0xFFFFFFFF)
    assert prog.info()["segments"] > whole.info()["segments"]
    prog.run(st, "persistent")
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops.astype(np.int32))
    check(st, orc)
