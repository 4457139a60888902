"""GPU parity: the gfx950 engine against the CPU oracle, bit-exact.

Every comparison covers the full X/Z/R tableau (all 2n+1 rows, scratch row
included) and every result and determined bit.  Coins and noise are keyed by
(seed, sequence number, global shot index), so the oracle run on a slice of
shots (shot0 offset) reproduces exactly those shots of a large GPU run.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KATS = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")))
FORMS = ("plain", "persistent", "graph")


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


def assert_same(st, orc, shots=None):
    x, z, r = st.device_to_host_numpy()
    sl = slice(None) if shots is None else shots
    assert np.array_equal(x[sl], orc.x), "x differs"
    assert np.array_equal(z[sl], orc.z), "z differs"
    assert np.array_equal(r[sl], orc.r), "r differs"
    assert st.num_records >= len(orc.records)
    for i, (res, det) in enumerate(orc.records):
        gres, gdet = st.get_record(i)
        assert np.array_equal(gres[sl], res), f"record {i} results differ"
        assert np.array_equal(gdet[sl], det), f"record {i} determined differ"


def run_both(ps, n, B, seed, ops, form):
    say(f"[case] n={n} B={B} seed={seed} ops={len(ops)} form={form}")
    st = ps.CliffordState(n, B, seed, True)
    st.run_program(ops, form)
    orc = O.OracleState(n, B, seed)
    orc.run_program(ops.astype(np.int32))
    return st, orc


# ---------------- reference known answers on the GPU ----------------

def test_init_tableau(ps):
    k = KATS["init_n3"]
    st = ps.CliffordState(3, 2, 0, True)
    x, z, r = st.device_to_host()
    x = np.array(x).reshape(2, 7, 3); z = np.array(z).reshape(2, 7, 3); r = np.array(r).reshape(2, 7)
    for b in range(2):
        assert x[b].tolist() == k["x_rows"] and z[b].tolist() == k["z_rows"]
        assert r[b].tolist() == k["r_rows"]
    assert st.get_num_qubits() == 3 and st.get_batch_size() == 2 and st.get_seed() == 0
    assert st.get_kokkos_backend().startswith("HIP(")


def test_checksum_kat(ps):
    k = KATS["checksum_5q"]
    st = ps.CliffordState.from_host(k["x"], k["z"], k["r"], 5, 1)
    assert st.check_sum() == sum(k["x"]) + sum(k["z"]) + sum(k["r"])


def test_row_product_sign_kat(ps):
    k = KATS["row_product_sign_5q"]
    for i, j, e in k["expected"]:
        st = ps.CliffordState.from_host(k["x"], k["z"], k["r"], 5, 1)
        assert int(st.row_product_sign(i, j)[0]) == e, (i, j)


def test_row_multiply_kat(ps):
    k = KATS["row_multiply_5q"]
    for i, j, e in k["expected_checksum"]:
        st = ps.CliffordState.from_host(k["x"], k["z"], k["r"], 5, 1)
        st.row_multiply(i, j)
        assert st.check_sum() == e, (i, j)


def test_measure_determined_kat(ps):
    k = KATS["measure_determined_5q"]
    for q in range(5):
        st = ps.CliffordState.from_host(k["x"], k["z"], k["r"], 5, 1)
        assert int(st.measure_determined(q)[0]) == k["expected"][q]


def test_measure_random_forced(ps):
    k = KATS["measure_random_forced"]
    for bias, expected in k["cases"]:
        st = ps.CliffordState(2, 1, -1, True)
        st.apply_hadamard_gate(1); st.apply_phase_gate(1)
        st.apply_hadamard_gate(0); st.apply_control_not_gate(0, 1)
        st.measure_qubit(1, bias, None)
        assert st.check_sum() == expected


# ---------------- the reference's behavioural tests (API shape) ----------------

def test_measure_and_epr(ps):
    st = ps.CliffordState(1, 1, -1, True)
    st.apply_hadamard_gate(0); st.apply_phase_gate(0); st.apply_phase_gate(0); st.apply_hadamard_gate(0)
    st.measure_qubit(0, 0.5, None)
    assert st.get_measurement(0, 0) == (1, 1)
    assert st.get_measurement(1, 0) is None and st.get_measurement(0, 1) is None
    for seed in range(6):
        st = ps.CliffordState(2, 1, seed, True)
        st.apply_hadamard_gate(0); st.apply_control_not_gate(0, 1)
        st.measure_qubit(0); st.measure_qubit(1)
        a, b = st.get_measurement(0, 0), st.get_measurement(1, 0)
        assert a[1] == 0 and b[1] == 1 and a[0] == b[0]


def test_s_state_distillation_batched(ps):
    """Test_CliffordState.hpp:674-741 with B=100 and flag-conditioned gates."""
    phasors = [[0], [1], [2], [0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
    B = 100
    st = ps.CliffordState(5, B, 1231232132, True)
    orc = O.OracleState(5, B, 1231232132)
    anc = 4
    for ph in phasors:
        for s in (st, orc):
            s.apply_hadamard_gate(anc)
            for k in ph:
                s.apply_control_not_gate(anc, k)
            s.apply_hadamard_gate(anc); s.apply_phase_gate(anc); s.apply_hadamard_gate(anc)
        mid = st.measure_qubit(anc)
        omid = orc.measure_qubit(anc)
        assert all(st.get_measurement(mid, b)[1] == 0 for b in range(B))
        flag = orc.records[omid][0]
        for k in ph + [anc]:
            st.apply_hadamard_gate_with_apply_flag(k, mid)
            st.apply_phase_gate_with_apply_flag(k, mid)
            st.apply_phase_gate_with_apply_flag(k, mid)
            st.apply_hadamard_gate_with_apply_flag(k, mid)
            orc.apply_hadamard_gate_with_apply_flag(k, flag)
            orc.apply_phase_gate_with_apply_flag(k, flag)
            orc.apply_phase_gate_with_apply_flag(k, flag)
            orc.apply_hadamard_gate_with_apply_flag(k, flag)
    for k in range(3):
        mid = st.measure_qubit(k); orc.measure_qubit(k)
        assert all(st.get_measurement(mid, b) == (0, 1) for b in range(B))
    st.apply_phase_gate(3); st.apply_hadamard_gate(3); orc.apply_phase_gate(3); orc.apply_hadamard_gate(3)
    mid = st.measure_qubit(3); orc.measure_qubit(3)
    assert all(st.get_measurement(mid, b) == (1, 1) for b in range(B))
    assert_same(st, orc)


def test_record_reuse_and_errors(ps):
    st = ps.CliffordState(3, 70, 5, True)
    st.apply_hadamard_gate(0)
    m0 = st.measure_qubit(0)
    first = [st.get_measurement(0, b) for b in range(70)]
    assert all(d == 0 for _, d in first)
    # re-measuring the collapsed qubit into the same record: determined, same
    # outcome, and the host copy is refreshed (not the stale first copy)
    assert st.measure_qubit(0, 0.5, m0) == m0
    assert st.num_records == 1
    assert [st.get_measurement(0, b) for b in range(70)] == [(r, 1) for r, _ in first]
    with pytest.raises(RuntimeError):
        st.measure_qubit(0, 0.5, 5)
    with pytest.raises(RuntimeError):
        st.apply_hadamard_gate(3)
    with pytest.raises(RuntimeError):
        st.apply_control_not_gate(1, 1)


# ---------------- random circuits, all forms, ragged batches ----------------

def random_program(W, rng, n, nops, pflag=0.08, pmeas=0.08, pnoise=0.05, flag2=False,
                   prand=0.0):
    """Random op list.  flag2 adds flagged CX/CZ to the flagged ops; prand is
    the share of gates applied with a probability (random flags)."""
    b = W.OpList()
    codes = [W.OP_H_FLAG, W.OP_S_FLAG, W.OP_X_FLAG, W.OP_Y_FLAG, W.OP_Z_FLAG]
    if flag2 and n > 1:
        codes += [W.OP_CX_FLAG, W.OP_CZ_FLAG]
    for _ in range(nops):
        u = rng.random()
        if u < pmeas:
            b.measure(int(rng.integers(0, n)), float(rng.choice([0.5, 0.5, 0.0, 1.0, 0.3])))
        elif u < pmeas + pflag and b.nrec:
            code = int(rng.choice(codes))
            c, t = rng.choice(n, 2, replace=False) if n > 1 else (0, 0)
            b.flagged(code, int(c), int(rng.integers(0, b.nrec)), int(t))
        elif u < pmeas + pflag + prand:
            gate = int(rng.choice([W.OP_H, W.OP_S, W.OP_X, W.OP_Y, W.OP_Z] +
                                  ([W.OP_CX, W.OP_CZ] if n > 1 else [])))
            c, t = rng.choice(n, 2, replace=False) if n > 1 else (0, 0)
            b.with_prob(gate, int(c), float(rng.choice([0.0, 0.1, 0.5, 1.0])), int(t))
        elif u < pmeas + pflag + prand + pnoise:
            b.noise(int(rng.integers(0, n)), 0.1, 0.05, 0.2)
        else:
            code = int(rng.choice([W.OP_H, W.OP_S, W.OP_CX, W.OP_CZ, W.OP_X, W.OP_Y, W.OP_Z]))
            if code in (W.OP_CX, W.OP_CZ) and n > 1:
                c, t = rng.choice(n, 2, replace=False)
                b.gate(code, int(c), int(t))
            elif code not in (W.OP_CX, W.OP_CZ):
                b.gate(code, int(rng.integers(0, n)))
    return b.array()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("seed,n,B", [(1, 1, 1), (2, 2, 63), (3, 5, 64), (4, 7, 65),
                                      (5, 12, 130), (6, 20, 300), (7, 33, 1000), (8, 64, 129)])
def test_random_programs(ps, W, form, seed, n, B):
    rng = np.random.default_rng(seed)
    ops = random_program(W, rng, n, 300)
    st, orc = run_both(ps, n, B, seed, ops, form)
    assert_same(st, orc)


@pytest.mark.parametrize("form", ["persistent", "graph"])
@pytest.mark.parametrize("seed,n,B,nops,pmeas", [(11, 5, 200, 1500, 0.001), (12, 40, 300, 2000, 0.01),
                                                 (13, 64, 129, 1200, 0.004), (14, 200, 64, 1500, 0.003),
                                                 (15, 600, 100, 2500, 0.002),
                                                 (16, 300, 70, 2000, 0.002)])
def test_long_gate_segments(ps, W, form, seed, n, B, nops, pmeas):
    """Long runs of gates (flags, noise, CZ included) between measurements:
    the compiler marks them row-tiled, segments over 448 instructions are
    staged in several chunks, and n=300 and n=600 leave room for so few rows
    per tile that one wave holds several ops (lanes with different ops)."""
    rng = np.random.default_rng(seed)
    ops = random_program(W, rng, n, nops, pflag=0.03, pmeas=pmeas, pnoise=0.1)
    if not any(ops[:, 0] == W.OP_M):
        ops = np.vstack([np.array([[W.OP_M, 0, 0, 0, O.coin_threshold(0.5), 0, 0, 0]], np.uint32), ops])
    st0 = ps.CliffordState(n, B, seed, True)
    info = ps.Program(st0, ops).info()
    say(f"[layout] {info}")
    assert info["tiled_segments"] >= 1
    st, orc = run_both(ps, n, B, seed, ops, form)
    assert_same(st, orc)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("seed,n,B", [(21, 2, 64), (22, 6, 100), (23, 17, 200), (24, 40, 65)])
def test_flagged_two_qubit_and_random_flags(ps, W, form, seed, n, B):
    """Flagged CX/CZ and gates applied with a probability (random flags that
    share the noise sequence with NOISE ops), mixed with everything else."""
    rng = np.random.default_rng(seed)
    ops = random_program(W, rng, n, 400, pflag=0.1, flag2=True, prand=0.1)
    st, orc = run_both(ps, n, B, seed, ops, form)
    assert_same(st, orc)


@pytest.mark.parametrize("form", ["persistent", "graph"])
@pytest.mark.parametrize("seed,n,B", [(31, 64, 130), (32, 300, 70)])
def test_long_gate_segments_flagged(ps, W, form, seed, n, B):
    """Row-tiled segments carrying flagged CX/CZ and random flags, with one op
    per wave (n=64) and with several ops per wave (n=300)."""
    rng = np.random.default_rng(seed)
    ops = random_program(W, rng, n, 2000, pflag=0.05, pmeas=0.002, pnoise=0.1, flag2=True,
                         prand=0.05)
    ops = np.vstack([np.array([[W.OP_M, 0, 0, 0, O.coin_threshold(0.5), 0, 0, 0]], np.uint32), ops])
    st0 = ps.CliffordState(n, B, seed, True)
    info = ps.Program(st0, ops).info()
    say(f"[layout] {info}")
    assert info["tiled_segments"] >= 1
    st, orc = run_both(ps, n, B, seed, ops, form)
    assert_same(st, orc)


def test_flag_and_prob_methods(ps, W):
    """The per-gate API: flagged CX/CZ/Paulis (record id or host flags) and the
    *_with_prob methods, against the oracle's program interpreter."""
    n, B = 5, 300
    st = ps.CliffordState(n, B, 23, True)
    orc = O.OracleState(n, B, 23)
    for q in range(n):
        st.apply_hadamard_gate(q); orc.apply_hadamard_gate(q)
    st.apply_control_not_gate(0, 1); orc.apply_control_not_gate(0, 1)
    m = st.measure_qubit(2)
    orc.measure_qubit(2)
    host = (np.arange(B) % 5 < 2).astype(np.int32)
    st.apply_control_not_gate_with_apply_flag(3, 4, m)
    orc.apply_control_not_gate_with_apply_flag(3, 4, orc.records[m][0])
    st.apply_control_phase_gate_with_apply_flag(1, 3, host.tolist())
    orc.apply_control_phase_gate_with_apply_flag(1, 3, host)
    st.apply_control_not_gate_with_apply_flag(4, 0, host.tolist())
    orc.apply_control_not_gate_with_apply_flag(4, 0, host)
    st.apply_control_phase_gate_with_apply_flag(2, 0, m)
    orc.apply_control_phase_gate_with_apply_flag(2, 0, orc.records[m][0])
    st.apply_pauli_x_gate_with_apply_flag(0, m)
    st.apply_pauli_y_gate_with_apply_flag(1, host.tolist())
    st.apply_pauli_z_gate_with_apply_flag(2, m)
    orc.apply_pauli_flag(0, orc.records[m][0], None)
    orc.apply_pauli_flag(1, host, host)
    orc.apply_pauli_flag(2, None, orc.records[m][0])
    rows = []
    for k, (name, gate, q0, q1) in enumerate([
            ("apply_hadamard_gate_with_prob", W.OP_H, 0, 0),
            ("apply_phase_gate_with_prob", W.OP_S, 1, 0),
            ("apply_pauli_x_gate_with_prob", W.OP_X, 2, 0),
            ("apply_pauli_y_gate_with_prob", W.OP_Y, 3, 0),
            ("apply_pauli_z_gate_with_prob", W.OP_Z, 4, 0),
            ("apply_control_not_gate_with_prob", W.OP_CX, 0, 3),
            ("apply_control_phase_gate_with_prob", W.OP_CZ, 4, 1)]):
        p = (0.25, 0.5, 1.0, 0.0, 0.75, 0.4, 0.6)[k]
        if q1 or gate in (W.OP_CX, W.OP_CZ):
            getattr(st, name)(q0, q1, p)
        else:
            getattr(st, name)(q0, p)
        rows.append([W.FLAGGED[gate], q0, q1, W.FLAG_RANDOM, W.coin_threshold(p), 0, 0, 0])
    orc.run_program(np.array(rows, np.int64))
    for q in range(n):
        st.measure_qubit(q); orc.measure_qubit(q)
    assert_same(st, orc)


@pytest.mark.parametrize("n,B", [(10, 256), (400, 128)])
def test_divergent_pivots(ps, W, n, B):
    """Per-shot flagged H makes X/Z differ across the 64 lanes of a word, so
    one measurement sees many pivot groups and mixed random/determined.  At
    n=10 a random measurement has a thread per tableau row; at n=400 (800
    rows, more than a workgroup's threads) each thread walks several rows."""
    b = W.OpList()
    for q in range(n):
        b.h(q)
    m = b.measure(0)
    for k in range(4):
        b.flagged(W.OP_H_FLAG, k + 1, m)
        b.cx(k + 1, k + 5)
        m = b.measure(k + 2)
        b.flagged(W.OP_S_FLAG, (k + 3) % n, m)
    for q in range(min(n, 40)):
        b.measure(q)
    ops = b.array()
    for form in FORMS:
        st, orc = run_both(ps, n, B, 99, ops, form)
        assert_same(st, orc)


def test_coin_replay(ps):
    n, B = 6, 200
    rng = np.random.default_rng(3)
    st = ps.CliffordState(n, B, 1, True)
    orc = O.OracleState(n, B, 1)
    for q in range(n):
        st.apply_hadamard_gate(q); orc.apply_hadamard_gate(q)
    st.apply_control_not_gate(0, 1); orc.apply_control_not_gate(0, 1)
    for q in range(n):
        coins = rng.integers(0, 2, B).astype(np.int32)
        st.measure_qubit(q, 0.5, None, coins.tolist())
        orc.measure_qubit(q, coins=coins)
    assert_same(st, orc)


def test_pauli_noise_and_conditional_paulis(ps):
    n, B = 4, 500
    st = ps.CliffordState(n, B, 17, True)
    orc = O.OracleState(n, B, 17)
    for q in range(n):
        st.apply_hadamard_gate(q); orc.apply_hadamard_gate(q)
    st.apply_pauli_noise(1, 0.1, 0.2, 0.3)
    from plaquette_simulator.workloads import noise_thresholds
    t0, t1, t2 = noise_thresholds(0.1, 0.2, 0.3)
    orc.run_program(np.array([[O.ORC_NOISE, 1, 0, 0, t0, t1, t2, 0]], np.uint32).astype(np.int32))
    st.measure_qubit(0); orc.measure_qubit(0)
    for kind, code in (("X", O.ORC_X_FLAG), ("Y", O.ORC_Y_FLAG), ("Z", O.ORC_Z_FLAG)):
        st.apply_pauli_with_apply_flag(kind, 2, 0)
        orc.run_program(np.array([[code, 2, 0, 0, 0, 0, 0, 0]], np.int32))
    st.apply_hadamard_gate_with_apply_flag(3, (np.arange(B) % 3 == 0).astype(int).tolist())
    orc.apply_hadamard_gate_with_apply_flag(3, (np.arange(B) % 3 == 0).astype(np.int32))
    assert_same(st, orc)


def test_reset_checksum_and_io_roundtrip(ps):
    n, B = 9, 150
    rng = np.random.default_rng(8)
    x = rng.integers(0, 2, B * (2 * n + 1) * n).tolist()
    z = rng.integers(0, 2, B * (2 * n + 1) * n).tolist()
    r = rng.integers(0, 2, B * (2 * n + 1)).tolist()
    st = ps.CliffordState.from_host(x, z, r, n, B)
    gx, gz, gr = st.device_to_host()
    assert gx == x and gz == z and gr == r
    orc = O.OracleState.from_flat(x, z, r, n, B)
    assert st.check_sum() == orc.check_sum()
    assert st.check_sum(10, 77) == orc.check_sum(10, 77)
    cp = st.copy()
    assert cp.device_to_host() == (x, z, r)
    st.reset_clifford_state()
    fresh = O.OracleState(n, B)
    x2, z2, r2 = st.device_to_host_numpy()
    assert np.array_equal(x2, fresh.x) and np.array_equal(z2, fresh.z) and np.array_equal(r2, fresh.r)


# ---------------- the measured workloads ----------------

@pytest.mark.parametrize("form", FORMS)
def test_random_n64_full(ps, W, form):
    w = W.config_random_n64()
    st, orc = run_both(ps, w.num_qubits, w.batch_size, w.seed, w.ops, form)
    assert_same(st, orc)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("d,p", [(3, 0.01), (5, 0.003)])
def test_surface_code_small(ps, W, form, d, p):
    ops, n = W.surface_code_memory_z(d, d, p)
    st, orc = run_both(ps, n, 200, 7, ops, form)
    assert_same(st, orc)


def test_surface_d9_full_size(ps, W):
    """16384 shots on the GPU (persistent and graph forms); the oracle
    recomputes two shot slices exactly; plain form equals persistent form on
    the whole bit-plane tableau."""
    w = W.config_surface_d9()
    say(f"[case] surface d9 B={w.batch_size} persistent")
    st = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
    st.run_program(w.ops, "persistent")
    for shot0, nb in ((0, 192), (8192 + 64, 100)):
        orc = O.OracleState(w.num_qubits, nb, w.seed, shot0=shot0)
        orc.run_program(w.ops.astype(np.int32))
        assert_same(st, orc, shots=slice(shot0, shot0 + nb))
    planes = st.read_planes()
    for form in ("plain", "graph"):
        say(f"[case] surface d9 B={w.batch_size} {form}")
        st2 = ps.CliffordState(w.num_qubits, w.batch_size, w.seed, True)
        st2.run_program(w.ops, form)
        assert np.array_equal(st2.read_planes(), planes)
        for i in range(0, w.measurements, 97):
            assert np.array_equal(st2.get_record(i)[0], st.get_record(i)[0])


def test_graph_replays_continue_sequences(ps, W):
    """A program run three times (graph form, eager work in between) equals
    the oracle running it three times: coin/noise streams advance per run."""
    ops, n = W.surface_code_memory_z(3, 2, 0.02)
    B = 130
    st = ps.CliffordState(n, B, 3, True)
    prog = ps.Program(st, ops)
    orc = O.OracleState(n, B, 3)
    for rep in range(3):
        say(f"[case] graph replay {rep}")
        prog.run(st, "graph")
        st.synchronize()
        st.sync_records()
        orc.run_program(ops.astype(np.int32))
        assert_same(st, orc)
        st.apply_hadamard_gate(0); orc.apply_hadamard_gate(0)


def test_large_n_prefix(ps, W):
    """n=1024 (the large random-circuit shape) on the first 3000 ops."""
    w = W.config_random_n1024()
    ops = w.ops[:3000]
    for form in ("plain", "persistent"):
        st, orc = run_both(ps, 1024, 128, w.seed, ops, form)
        assert_same(st, orc)
