"""Circuit text front end (no GPU): the engine's parser against the
reference's CliffordCircuit tests and its 59 benchmark ring circuits.

* Test_CliffordCircuit.hpp:21-99: an empty file has 0 operations; "M 0 1 2 3"
  is 1 operation; "H 0\\n S 0\\n S 0\\n H 0\\n M 0" is 5 operations.
* CliffordCircuit.hpp:44-46: an unsupported name (a blank line's empty name
  included) raises "Unsupported operation: <name>".
* Every ring circuit (benchmarks/circuits.tar.gz, tests/golden/ring_circuits.json)
  parsed and run through the oracle agrees, shot by shot, with a brute-force
  statevector: each measurement is determined exactly when the state gives
  probability 0 or 1 (with that value), else random with probability 1/2.
"""
import json
import os
import zlib

import numpy as np
import pytest

import oracle as O
from statevector import StateVector

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RING = json.load(open(os.path.join(ROOT, "tests", "golden", "ring_circuits.json")))["circuits"]


@pytest.fixture(scope="module")
def parse():
    import plaquette_simulator_bindings as b
    return b.parse_circuit


def test_reference_circuit_kats(parse):
    ops, lines, used = parse("")
    assert ops.shape == (0, 8) and lines == 0
    ops, lines, used = parse("M 0 1 2 3\n")
    assert lines == 1 and len(ops) == 4 and list(ops[:, 3]) == [0, 1, 2, 3] and used == 4
    ops, lines, used = parse("H 0\n S 0\n S 0\n H 0\n M 0")
    assert lines == 5 and list(ops[:, 0]) == [1, 2, 2, 1, 14]
    assert ops[4, 4] == O.coin_threshold(0.5)  # M lines measure with bias 1/2


def test_expansion_and_records(parse):
    ops, lines, used = parse("CX 0 1 2 3\nH 4 5\nM 1 0\nCZ 5 4\nM 2\n", first_record=7)
    assert lines == 5 and used == 6
    assert ops[:, :4].tolist() == [[6, 0, 1, 0], [6, 2, 3, 0], [1, 4, 0, 0], [1, 5, 0, 0],
                                   [14, 1, 0, 7], [14, 0, 0, 8], [7, 5, 4, 0], [14, 2, 0, 9]]


def test_stream_read_semantics(parse):
    # indices end at the first token that does not start with a number, and
    # a number with trailing characters ends the list after itself
    ops, _, _ = parse("H 1 2 x 3\nX 4,5 6\n  Z\t7  \r\n")
    assert ops[:, :2].tolist() == [[1, 1], [1, 2], [3, 4], [5, 7]]


@pytest.mark.parametrize("text,name", [("Q 0\n", "Q"), ("H 0\n\nM 0\n", ""), ("h 0", "h"),
                                       ("E_PAULI 0.1 0.1 0.1 0\n", "E_PAULI")])
def test_unsupported_operation(parse, text, name):
    with pytest.raises(RuntimeError) as e:
        parse(text)
    assert str(e.value) == "Unsupported operation: " + name


def test_odd_two_qubit_list_rejected(parse):
    with pytest.raises(RuntimeError):
        parse("CZ 0 1 2\n")


def test_ring_fixture_complete():
    assert len(RING) == 59
    widths = {int(k.split("_")[2]) for k in RING}
    assert widths == {4, 5, 6, 7, 8}


def _sv_apply(sv, code, q0, q1):
    {1: lambda: sv.h(q0), 2: lambda: sv.s(q0), 3: lambda: sv.x(q0), 4: lambda: sv.y(q0),
     5: lambda: sv.z(q0), 6: lambda: sv.cx(q0, q1), 7: lambda: sv.cz(q0, q1)}[code]()


@pytest.mark.parametrize("name", sorted(RING))
def test_ring_circuit_oracle_matches_statevector(parse, name):
    ops, lines, used = parse(RING[name])
    n = int(name.split("_")[2])
    assert used <= n
    B = 16
    st = O.OracleState(n, B, seed=zlib.crc32(name.encode()) & 0x7FFFFFFF)
    st.run_program(ops.astype(np.int32))
    nmeas = int((ops[:, 0] == 14).sum())
    assert len(st.records) == nmeas
    for b in range(0, B, 5):
        sv = StateVector(n)
        for code, q0, q1, slot in ops[:, :4]:
            if code == 14:
                res, det = st.get_measurement(int(slot), b)
                p1 = sv.prob_one(int(q0))
                if p1 < 1e-9 or p1 > 1 - 1e-9:
                    assert det == 1 and res == int(p1 > 0.5), (name, slot, b)
                else:
                    assert abs(p1 - 0.5) < 1e-9 and det == 0, (name, slot, b)
                sv.project(int(q0), res)
            else:
                _sv_apply(sv, int(code), int(q0), int(q1))
