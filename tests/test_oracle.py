"""Pins for the CPU oracle (runs without a GPU).

The oracle is trusted only after these checks: the reference's own
known-answer data (tests/golden/reference_kats.json, transcribed from
Test_CliffordState.hpp), a brute-force statevector, Clifford identities and
symplectic invariants, and the reference's behavioural tests.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from statevector import StateVector

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")
KATS = json.load(open(GOLD))


# ---------------- reference known-answer data ----------------

def test_init_tableau_n3():
    """Test_CliffordState.hpp:29-64 (scratch row r = 1)."""
    k = KATS["init_n3"]
    s = O.OracleState(k["num_qubits"], k["batch_size"])
    for b in range(k["batch_size"]):
        assert s.x[b].tolist() == k["x_rows"]
        assert s.z[b].tolist() == k["z_rows"]
        assert s.r[b].tolist() == k["r_rows"]


def test_g_table():
    """Test_CliffordState.hpp:160-174."""
    k = KATS["g_table"]
    P = k["paulis"]
    for i in range(4):
        for j in range(4):
            assert O.g(P[i][0], P[i][1], P[j][0], P[j][1]) == k["expected"][i][j]
    for x1, z1, x2, z2, e in k["extra"]:
        assert O.g(x1, z1, x2, z2) == e


def test_g_is_pauli_product_phase():
    """g(P1, P2) is the exponent e with P1 P2 = i^e P3 (P3 the Pauli of
    x1^x2, z1^z2, Y = iXZ): checked against 2x2 matrices."""
    I = np.eye(2); X = np.array([[0, 1], [1, 0]]); Z = np.diag([1, -1])
    Y = np.array([[0, -1j], [1j, 0]])
    M = {(0, 0): I, (1, 0): X, (1, 1): Y, (0, 1): Z}
    for (x1, z1), A in M.items():
        for (x2, z2), Bm in M.items():
            C = M[(x1 ^ x2, z1 ^ z2)]
            prod = A @ Bm
            e = O.g(x1, z1, x2, z2)
            assert np.allclose(prod, (1j ** e) * C)


def test_checksum_5q():
    """Test_CliffordState.hpp:176-202."""
    k = KATS["checksum_5q"]
    s = O.OracleState.from_flat(k["x"], k["z"], k["r"], 5, 1)
    assert s.check_sum() == sum(k["x"]) + sum(k["z"]) + sum(k["r"])


def test_row_product_sign_all_110():
    """Test_CliffordState.hpp:204-342: the reference checks the 20 pairs
    with i, j < 5; all 110 listed pairs are checked here."""
    k = KATS["row_product_sign_5q"]
    total = sum(k["x"]) + sum(k["z"]) + sum(k["r"])
    for i, j, e in k["expected"]:
        s = O.OracleState.from_flat(k["x"], k["z"], k["r"], 5, 1)
        assert s.row_product_sign(0, i, j) == e, (i, j)
        assert s.check_sum() == total


def test_row_product_sign_odd_phase_sums_present():
    """At least one of the 110 pairs anticommutes (odd phase sum), so the
    'bit 1 of S mod 4' convention for odd S is pinned by the data."""
    k = KATS["row_product_sign_5q"]
    x = np.array(k["x"]).reshape(11, 5)
    z = np.array(k["z"]).reshape(11, 5)
    odd = 0
    for i, j, _ in k["expected"]:
        s = sum(O.g(int(x[i, c]), int(z[i, c]), int(x[j, c]), int(z[j, c])) for c in range(5))
        odd += s % 2
    assert odd > 0


def test_row_multiply_all_110():
    """Test_CliffordState.hpp:344-454."""
    k = KATS["row_multiply_5q"]
    for i, j, e in k["expected_checksum"]:
        s = O.OracleState.from_flat(k["x"], k["z"], k["r"], 5, 1)
        s.row_multiply(0, i, j)
        assert s.check_sum() == e, (i, j)


def test_measure_determined_5q():
    """Test_CliffordState.hpp:456-511."""
    k = KATS["measure_determined_5q"]
    for q in range(5):
        s = O.OracleState.from_flat(k["x"], k["z"], k["r"], 5, 1)
        assert s.measure_determined(0, q) == k["expected"][q]


def test_measure_random_forced_coins():
    """Test_CliffordState.hpp:513-542: bias 0 -> coin 0 -> checksum 9;
    bias 1 -> coin 1 -> checksum 10."""
    k = KATS["measure_random_forced"]
    for bias, expected in k["cases"]:
        s = O.OracleState(2, 1)
        s.apply_hadamard_gate(1); s.apply_phase_gate(1)
        s.apply_hadamard_gate(0); s.apply_control_not_gate(0, 1)
        mid = s.measure_qubit(1, bias)
        assert s.check_sum() == expected
        assert s.get_measurement(mid, 0) == (int(bias >= 1.0), 0)


# ---------------- counter-based draws ----------------

def test_splitmix64_published_vector():
    """SplitMix64 seeded with 0: first outputs 0xE220A8397B1DCDAF,
    0x6E789E6AA1B965F4, 0x06C45D188009454F (published test vector)."""
    G = 0x9E3779B97F4A7C15
    M = 2 ** 64
    outs = [O.mix64((k * G) % M) for k in (1, 2, 3)]
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert O.draw(0, 0, 0) == O.mix64((O.mix64(G) + G) % M)


def test_coin_threshold_matches_float_compare():
    rng = np.random.default_rng(1)
    for bias in [0.0, 1.0, 0.5, 1e-3, 0.3, 0.999, -1.0, 2.0] + list(rng.random(20)):
        T = O.coin_threshold(np.float32(bias))
        for u in [0, 1, T - 1, T, T + 1, 2 ** 23, 2 ** 24 - 1]:
            if 0 <= u < 2 ** 24:
                assert (u < T) == (np.float32(u) * np.float32(2.0 ** -24) < np.float32(bias))


# ---------------- statevector cross-check ----------------

def _random_circuit(rng, n, nops, pmeas=0.15):
    ops = []
    for _ in range(nops):
        k = rng.integers(0, 8 if n > 1 else 5)
        if rng.random() < pmeas:
            ops.append(("M", int(rng.integers(0, n))))
        elif k < 5:
            ops.append(("HSXYZ"[k], int(rng.integers(0, n))))
        else:
            c, t = rng.choice(n, 2, replace=False)
            ops.append(("CX" if k < 7 else "CZ", int(c), int(t)))
    return ops


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_statevector(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 7))
    s = O.OracleState(n, 1, seed=seed)
    sv = StateVector(n)
    for op in _random_circuit(rng, n, 60):
        name = op[0]
        if name == "M":
            q = op[1]
            p1 = sv.prob_one(q)
            coin = int(rng.integers(0, 2))
            mid = s.measure_qubit(q, coins=np.array([coin], np.int32))
            res, det = s.get_measurement(mid, 0)
            if p1 < 1e-9 or p1 > 1 - 1e-9:
                assert det == 1 and res == int(p1 > 0.5)
            else:
                assert abs(p1 - 0.5) < 1e-9
                assert det == 0 and res == coin
            sv.project(q, res)
        else:
            getattr(sv, name.lower())(*op[1:])
            {"H": s.apply_hadamard_gate, "S": s.apply_phase_gate, "X": s.apply_pauli_x_gate,
             "Y": s.apply_pauli_y_gate, "Z": s.apply_pauli_z_gate,
             "CX": s.apply_control_not_gate, "CZ": s.apply_control_phase_gate}[name](*op[1:])
    for i in range(n, 2 * n):
        v = sv.pauli_expectation_sign(s.x[0, i], s.z[0, i], s.r[0, i])
        assert abs(v - 1) < 1e-9, (i, v)


def _symplectic_ok(s, b=0):
    n = s.n
    X = s.x[b, : 2 * n].astype(int)
    Z = s.z[b, : 2 * n].astype(int)
    W = (X @ Z.T + Z @ X.T) % 2
    J = np.zeros((2 * n, 2 * n), int)
    J[np.arange(n), np.arange(n) + n] = 1
    J[np.arange(n) + n, np.arange(n)] = 1
    return np.array_equal(W, J)


@pytest.mark.parametrize("seed", range(6))
def test_symplectic_invariant(seed):
    """Destabilizer i anticommutes only with stabilizer i; all other pairs
    commute -- after every op of a random circuit."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2, 9))
    s = O.OracleState(n, 2, seed=seed)
    for op in _random_circuit(rng, n, 80):
        if op[0] == "M":
            s.measure_qubit(op[1])
        else:
            {"H": s.apply_hadamard_gate, "S": s.apply_phase_gate, "X": s.apply_pauli_x_gate,
             "Y": s.apply_pauli_y_gate, "Z": s.apply_pauli_z_gate,
             "CX": s.apply_control_not_gate, "CZ": s.apply_control_phase_gate}[op[0]](*op[1:])
        assert _symplectic_ok(s, 0) and _symplectic_ok(s, 1)


# ---------------- Clifford identities (bit-exact on the tableau) ----------------

def _scrambled(n, B, seed):
    rng = np.random.default_rng(seed)
    s = O.OracleState(n, B, seed=seed)
    for op in _random_circuit(rng, n, 40, pmeas=0.0):
        {"H": s.apply_hadamard_gate, "S": s.apply_phase_gate, "X": s.apply_pauli_x_gate,
         "Y": s.apply_pauli_y_gate, "Z": s.apply_pauli_z_gate,
         "CX": s.apply_control_not_gate, "CZ": s.apply_control_phase_gate}[op[0]](*op[1:])
    return s


def test_gate_identities_bit_exact():
    s = _scrambled(5, 3, 7)
    ref = [a.copy() for a in (s.x, s.z, s.r)]
    for q in range(5):
        s.apply_hadamard_gate(q); s.apply_hadamard_gate(q)
        for _ in range(4):
            s.apply_phase_gate(q)
    s.apply_control_not_gate(1, 3); s.apply_control_not_gate(1, 3)
    s.apply_control_phase_gate(0, 4); s.apply_control_phase_gate(0, 4)
    for a, b in zip((s.x, s.z, s.r), ref):
        assert np.array_equal(a, b)


def test_flagged_hssh_is_conditional_x_and_ss_is_conditional_z():
    """H.S.S.H with a per-shot flag equals the X rule on flagged shots and
    S.S equals the Z rule -- the identities that let per-shot Pauli noise and
    reset run as masked Paulis."""
    B = 16
    flag = (np.arange(B) % 3 == 0).astype(np.int32)
    a = _scrambled(5, B, 11)
    b = _scrambled(5, B, 11)
    for q in (0, 2):
        a.apply_hadamard_gate_with_apply_flag(q, flag)
        a.apply_phase_gate_with_apply_flag(q, flag)
        a.apply_phase_gate_with_apply_flag(q, flag)
        a.apply_hadamard_gate_with_apply_flag(q, flag)
        b.apply_pauli_flag(q, fx=flag)
    a.apply_phase_gate_with_apply_flag(3, flag)
    a.apply_phase_gate_with_apply_flag(3, flag)
    b.apply_pauli_flag(3, fz=flag)
    for u, v in zip((a.x, a.z, a.r), (b.x, b.z, b.r)):
        assert np.array_equal(u, v)
    # unconditional forms: HSSH == X, SS == Z, and X.Z == Y
    c = _scrambled(4, 2, 3); d = _scrambled(4, 2, 3)
    c.apply_hadamard_gate(1); c.apply_phase_gate(1); c.apply_phase_gate(1); c.apply_hadamard_gate(1)
    d.apply_pauli_x_gate(1)
    c.apply_pauli_x_gate(2); c.apply_pauli_z_gate(2)
    d.apply_pauli_y_gate(2)
    for u, v in zip((c.x, c.z, c.r), (d.x, d.z, d.r)):
        assert np.array_equal(u, v)


# ---------------- the reference's behavioural tests ----------------

def test_measure0_to_3_and_epr():
    """Test_CliffordState.hpp:66-158."""
    s = O.OracleState(1, 1)
    s.apply_hadamard_gate(0); s.apply_phase_gate(0); s.apply_phase_gate(0); s.apply_hadamard_gate(0)
    s.measure_qubit(0)
    assert s.get_measurement(0, 0) == (1, 1)
    s = O.OracleState(1, 1); s.measure_qubit(0)
    assert s.get_measurement(0, 0)[0] == 0
    s = O.OracleState(2, 1); s.measure_qubit(0); s.measure_qubit(1)
    assert s.get_measurement(0, 0) == (0, 1) and s.get_measurement(1, 0) == (0, 1)
    s = O.OracleState(2, 1)
    s.apply_hadamard_gate(0); s.apply_phase_gate(0); s.apply_phase_gate(0); s.apply_hadamard_gate(0)
    s.measure_qubit(0); s.measure_qubit(1)
    assert s.get_measurement(0, 0) == (1, 1) and s.get_measurement(1, 0) == (0, 1)
    for seed in range(8):
        s = O.OracleState(2, 1, seed=seed)
        s.apply_hadamard_gate(0); s.apply_control_not_gate(0, 1)
        s.measure_qubit(0); s.measure_qubit(1)
        r0, r1 = s.get_measurement(0, 0), s.get_measurement(1, 0)
        assert r0[1] == 0 and r1[1] == 1 and r0[0] == r1[0]


def test_phase_kickback():
    """Test_CliffordState.hpp:544-572."""
    for seed in range(8):
        s = O.OracleState(2, 1, seed=seed)
        s.apply_hadamard_gate(1); s.apply_phase_gate(1)
        s.apply_hadamard_gate(0); s.apply_control_not_gate(0, 1)
        s.measure_qubit(1)
        r0 = s.get_measurement(0, 0)
        assert r0[1] == 0
        if r0[0]:
            s.apply_phase_gate(0); s.apply_phase_gate(0)
        s.apply_phase_gate(0); s.apply_hadamard_gate(0)
        s.measure_qubit(0)
        assert s.get_measurement(1, 0) == (1, 1)


PHASORS = [[0], [1], [2], [0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]


def test_s_state_distillation_100_seeds():
    """Test_CliffordState.hpp:574-633."""
    for seed in range(100):
        s = O.OracleState(5, 1, seed=seed)
        anc = 4
        for ph in PHASORS:
            s.apply_hadamard_gate(anc)
            for k in ph:
                s.apply_control_not_gate(anc, k)
            s.apply_hadamard_gate(anc); s.apply_phase_gate(anc); s.apply_hadamard_gate(anc)
            mid = s.measure_qubit(anc)
            v = s.get_measurement(mid, 0)
            assert v[1] == 0
            if v[0]:
                for k in ph:
                    s.apply_hadamard_gate(k); s.apply_phase_gate(k); s.apply_phase_gate(k); s.apply_hadamard_gate(k)
                s.apply_hadamard_gate(anc); s.apply_phase_gate(anc); s.apply_phase_gate(anc); s.apply_hadamard_gate(anc)
        for k in range(3):
            mid = s.measure_qubit(k)
            assert s.get_measurement(mid, 0) == (0, 1)
        s.apply_phase_gate(3); s.apply_hadamard_gate(3)
        mid = s.measure_qubit(3)
        assert s.get_measurement(mid, 0) == (1, 1)


def test_batched_flag_tests():
    """Test_CliffordState.hpp:635-672 and :674-741 (batch 2 and 100)."""
    s = O.OracleState(2, 2)
    s.apply_hadamard_gate(1); s.apply_phase_gate(1); s.apply_hadamard_gate(0); s.apply_control_not_gate(0, 1)
    mid = s.measure_qubit(1)
    flag = s.records[mid][0]
    assert all(s.get_measurement(mid, b)[1] == 0 for b in range(2))
    s.apply_phase_gate_with_apply_flag(0, flag); s.apply_phase_gate_with_apply_flag(0, flag)
    s.apply_phase_gate(0); s.apply_hadamard_gate(0)
    mid = s.measure_qubit(0)
    assert all(s.get_measurement(mid, b) == (1, 1) for b in range(2))

    B = 100
    s = O.OracleState(5, B, seed=1231232132)
    anc = 4
    for ph in PHASORS:
        s.apply_hadamard_gate(anc)
        for k in ph:
            s.apply_control_not_gate(anc, k)
        s.apply_hadamard_gate(anc); s.apply_phase_gate(anc); s.apply_hadamard_gate(anc)
        mid = s.measure_qubit(anc)
        assert all(s.get_measurement(mid, b)[1] == 0 for b in range(B))
        flag = s.records[mid][0]
        for k in ph + [anc]:
            s.apply_hadamard_gate_with_apply_flag(k, flag)
            s.apply_phase_gate_with_apply_flag(k, flag)
            s.apply_phase_gate_with_apply_flag(k, flag)
            s.apply_hadamard_gate_with_apply_flag(k, flag)
    for k in range(3):
        mid = s.measure_qubit(k)
        assert all(s.get_measurement(mid, b) == (0, 1) for b in range(B))
    s.apply_phase_gate(3); s.apply_hadamard_gate(3)
    mid = s.measure_qubit(3)
    assert all(s.get_measurement(mid, b) == (1, 1) for b in range(B))


def test_program_interpreter_matches_method_calls():
    """orc_run_program (used as the checker of the GPU program runner) equals
    the same ops issued as individual oracle calls."""
    rng = np.random.default_rng(5)
    n, B = 6, 70
    ops = []
    a = O.OracleState(n, B, seed=99)
    nrec = 0
    for k in range(120):
        u = rng.random()
        if u < 0.1:
            q = int(rng.integers(0, n))
            ops.append([O.ORC_M, q, 0, nrec, O.coin_threshold(0.5), 0, 0, 0])
            a.measure_qubit(q, 0.5)
            nrec += 1
        elif u < 0.2 and nrec:
            q = int(rng.integers(0, n)); slot = int(rng.integers(0, nrec))
            code = [O.ORC_H_FLAG, O.ORC_S_FLAG, O.ORC_X_FLAG, O.ORC_Y_FLAG, O.ORC_Z_FLAG][k % 5]
            ops.append([code, q, 0, slot, 0, 0, 0, 0])
            f = a.records[slot][0]
            if code == O.ORC_H_FLAG:
                a.apply_hadamard_gate_with_apply_flag(q, f)
            elif code == O.ORC_S_FLAG:
                a.apply_phase_gate_with_apply_flag(q, f)
            else:
                a.apply_pauli_flag(q, f if code != O.ORC_Z_FLAG else None,
                                   f if code != O.ORC_X_FLAG else None)
        else:
            code = int(rng.integers(1, 8))
            c, t = rng.choice(n, 2, replace=False)
            ops.append([code, int(c), int(t), 0, 0, 0, 0, 0])
            a._gate(code, int(c), int(t))
    b = O.OracleState(n, B, seed=99)
    b.run_program(np.array(ops, np.int32))
    for u, v in zip((a.x, a.z, a.r), (b.x, b.z, b.r)):
        assert np.array_equal(u, v)
    for (r1, d1), (r2, d2) in zip(a.records, b.records):
        assert np.array_equal(r1, r2) and np.array_equal(d1, d2)


def test_noise_draw_rates():
    """NOISE op: X/Y/Z rates follow the thresholds (t0,t1,t2)."""
    n, B = 2, 20000
    s = O.OracleState(n, B, seed=7)
    s.apply_hadamard_gate(1)  # qubit 1 in |+>: X noise invisible, Z flips
    p = 0.1
    t = [int(p * k * 2 ** 32) for k in (1, 2, 3)]
    s.run_program(np.array([[O.ORC_NOISE, 0, 0, 0, *t, 0], [O.ORC_NOISE, 1, 0, 0, *t, 0],
                            [O.ORC_H, 1, 0, 0, 0, 0, 0, 0],
                            [O.ORC_M, 0, 0, 0, O.coin_threshold(0.5), 0, 0, 0],
                            [O.ORC_M, 1, 0, 1, O.coin_threshold(0.5), 0, 0, 0]], np.int32))
    r0, d0 = s.records[0]
    r1, d1 = s.records[1]
    assert d0.all() and d1.all()
    # qubit 0 flips on X or Y (rate 2p); qubit 1 flips on Z or Y (rate 2p)
    assert abs(r0.mean() - 2 * p) < 0.01
    assert abs(r1.mean() - 2 * p) < 0.01


# ---------------- flagged CX/CZ and random flags ----------------

def _random_state(n, B, seed, depth=60):
    """A state scrambled by random H/S/CX, identical in every shot."""
    rng = np.random.default_rng(seed)
    s = O.OracleState(n, B, seed=seed)
    for _ in range(depth):
        code = int(rng.choice([O.ORC_H, O.ORC_S, O.ORC_CX]))
        c, t = rng.choice(n, 2, replace=False)
        s._gate(code, int(c), int(t))
    return s


@pytest.mark.parametrize("code,gate", [(O.ORC_CX_FLAG, O.ORC_CX), (O.ORC_CZ_FLAG, O.ORC_CZ)])
def test_flagged_two_qubit_gate_is_the_gate_on_flagged_shots(code, gate):
    """A flagged shot ends exactly as under the unflagged gate (itself pinned
    by the statevector and identities above); an unflagged shot is untouched."""
    n, B = 5, 40
    flag = (np.arange(B) % 3 == 1).astype(np.int32)
    a = _random_state(n, B, 3)
    ref = _random_state(n, B, 3)
    ref._gate(gate, 1, 3)
    before = _random_state(n, B, 3)
    if code == O.ORC_CX_FLAG:
        a.apply_control_not_gate_with_apply_flag(1, 3, flag)
    else:
        a.apply_control_phase_gate_with_apply_flag(1, 3, flag)
    for b in range(B):
        want = ref if flag[b] else before
        assert np.array_equal(a.x[b], want.x[b])
        assert np.array_equal(a.z[b], want.z[b])
        assert np.array_equal(a.r[b], want.r[b])


def test_flagged_cx_makes_bell_pairs_only_on_flagged_shots():
    """Physics: |+>|0> then CX on flagged shots gives correlated outcomes there
    and a deterministic 0 on the target elsewhere."""
    n, B = 2, 256
    flag = (np.arange(B) % 2).astype(np.int32)
    s = O.OracleState(n, B, seed=11)
    s.apply_hadamard_gate(0)
    s.apply_control_not_gate_with_apply_flag(0, 1, flag)
    m0 = s.measure_qubit(0)
    m1 = s.measure_qubit(1)
    r0, _ = s.records[m0]
    r1, d1 = s.records[m1]
    f = flag.astype(bool)
    assert np.array_equal(r0[f], r1[f]) and d1[f].all()   # Bell pair: second is determined
    assert (r1[~f] == 0).all() and d1[~f].all()            # target untouched
    assert 0.3 < r0.mean() < 0.7


def _prog(rows):
    return np.array(rows, np.int64)


@pytest.mark.parametrize("code", [O.ORC_H_FLAG, O.ORC_S_FLAG, O.ORC_X_FLAG, O.ORC_Y_FLAG,
                                  O.ORC_Z_FLAG, O.ORC_CX_FLAG, O.ORC_CZ_FLAG])
def test_random_flag_extremes(code):
    """Probability 0 is the identity, probability 1 the unflagged gate; either
    way the op takes one noise number."""
    gate = {O.ORC_H_FLAG: O.ORC_H, O.ORC_S_FLAG: O.ORC_S, O.ORC_X_FLAG: O.ORC_X,
            O.ORC_Y_FLAG: O.ORC_Y, O.ORC_Z_FLAG: O.ORC_Z, O.ORC_CX_FLAG: O.ORC_CX,
            O.ORC_CZ_FLAG: O.ORC_CZ}[code]
    n, B = 4, 70
    for T, applied in ((0, False), (1 << 24, True)):
        a = _random_state(n, B, 8)
        a.run_program(_prog([[code, 2, 0, O.ORC_FLAG_RANDOM, T, 0, 0, 0]]))
        assert a.noise_seq == 1
        b = _random_state(n, B, 8)
        if applied:
            b._gate(gate, 2, 0)
        assert np.array_equal(a.x, b.x) and np.array_equal(a.z, b.z)
        assert np.array_equal(a.r, b.r)


def test_random_flag_rate_and_shot_independence():
    """X applied with probability p flips a |0> qubit on a fraction p of the
    shots, independently per shot; a shot's draw depends only on its global
    index (a sub-batch at shot0 reproduces the slice)."""
    n, B, p = 1, 20000, 0.3
    s = O.OracleState(n, B, seed=5)
    ops = _prog([[O.ORC_X_FLAG, 0, 0, O.ORC_FLAG_RANDOM, O.coin_threshold(p), 0, 0, 0],
                 [O.ORC_M, 0, 0, 0, O.coin_threshold(0.5), 0, 0, 0]])
    s.run_program(ops)
    r, d = s.records[0]
    assert d.all()
    assert abs(r.mean() - p) < 5 * np.sqrt(p * (1 - p) / B)
    sub = O.OracleState(n, 100, seed=5, shot0=700)
    sub.run_program(ops)
    assert np.array_equal(sub.records[0][0], r[700:800])


def test_random_flag_and_noise_share_one_sequence():
    """Random flags and NOISE ops draw from one sequence of noise numbers in
    program order: a NOISE after a random-flag op is the NOISE a state whose
    sequence already stands at 1 would draw."""
    n, B = 3, 512
    t = [int(0.2 * k * 2 ** 32) for k in (1, 2, 3)]
    a = _random_state(n, B, 9)
    a.run_program(_prog([[O.ORC_CX_FLAG, 0, 1, O.ORC_FLAG_RANDOM, 0, 0, 0, 0],
                         [O.ORC_NOISE, 2, 0, 0, *t, 0]]))
    b = _random_state(n, B, 9)
    b.noise_seq = 1
    b.run_program(_prog([[O.ORC_NOISE, 2, 0, 0, *t, 0]]))
    assert a.noise_seq == b.noise_seq == 2
    assert np.array_equal(a.r, b.r)
