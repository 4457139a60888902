"""Write reference_kats.json: the known-answer data held by the reference's
own C++ tests (plaquette_simulator/src/tests/Test_CliffordState.hpp and
Test_CliffordCircuit.hpp), transcribed as plain data.

Run in a container where the reference is mounted; the output JSON is
committed and is all the tests read at run time.  Line numbers below are the
reference's.
"""
import json
import os
import re
import sys

REF = os.environ.get("PQ_REFERENCE_ROOT", "/root/reference")
TEST = os.path.join(REF, "plaquette_simulator/src/tests/Test_CliffordState.hpp")


def lines(a, b):
    with open(TEST) as f:
        src = f.read().splitlines()
    return "\n".join(src[a - 1:b])


def ints(text):
    return [int(v) for v in re.findall(r"-?\d+", text)]


def flat_block(a, b, name):
    """Integers of `std::vector<int> name = {...};` inside lines a..b."""
    t = lines(a, b)
    m = re.search(name + r"\s*=\s*\{([^}]*)\}", t, re.S)
    return ints(m.group(1))


def tuples(a, b):
    t = lines(a, b)
    return [list(map(int, m)) for m in re.findall(r"make_tuple\((-?\d+),\s*(-?\d+),\s*(-?\d+)\)", t)]


def nested(a, b, name):
    t = lines(a, b)
    m = re.search(name + r"\s*=\s*\{(.*?)\};", t, re.S)
    rows = re.findall(r"\{([^{}]*)\}", m.group(1))
    return [ints(r) for r in rows]


def main(out):
    kats = {
        "source": "plaquette_simulator/src/tests/Test_CliffordState.hpp",
        # :29-64 initial tableau, n=3, B=2, scratch row r=1
        "init_n3": {
            "num_qubits": 3, "batch_size": 2,
            "x_rows": nested(40, 42, "tableau_x_check"),
            "z_rows": nested(44, 46, "tableau_z_check"),
            "r_rows": [0, 0, 0, 0, 0, 0, 1],
        },
        # :160-174 g table (destination row index i, source column j)
        "g_table": {
            "paulis": nested(162, 162, "paulis"),
            "expected": nested(163, 164, "expected"),
            "extra": [[1, 0, 1, 1, 1]],
        },
        # :176-202 checksum of a loaded 5-qubit tableau
        "checksum_5q": {
            "num_qubits": 5, "batch_size": 1,
            "r": flat_block(182, 182, "r_flat"),
            "x": flat_block(184, 187, "x_flat"),
            "z": flat_block(189, 192, "z_flat"),
        },
        # :204-342 RowProductSign(i, j); the test checks i, j < 5, i != j
        "row_product_sign_5q": {
            "num_qubits": 5, "batch_size": 1,
            "r": flat_block(212, 212, "r_flat"),
            "x": flat_block(214, 217, "x_flat"),
            "z": flat_block(219, 222, "z_flat"),
            "expected": tuples(243, 299),
            "checked_below": 5,
        },
        # :344-454 checksum after RowMultiply(i, j); checked for i, j < 5
        "row_multiply_5q": {
            "num_qubits": 5, "batch_size": 1,
            "r": flat_block(352, 352, "r_flat"),
            "x": flat_block(354, 357, "x_flat"),
            "z": flat_block(359, 362, "z_flat"),
            "expected_checksum": tuples(377, 433),
            "checked_below": 5,
        },
        # :456-511 MeasureDetermined(q) on fresh copies, q = 0..4
        "measure_determined_5q": {
            "num_qubits": 5, "batch_size": 1,
            "r": flat_block(464, 464, "r_flat"),
            "x": flat_block(466, 469, "x_flat"),
            "z": flat_block(471, 474, "z_flat"),
            "expected": flat_block(489, 489, "expected_results"),
        },
        # :513-542 MeasureRandom with forced coins: bias 0 -> 9, bias 1 -> 10
        "measure_random_forced": {
            "num_qubits": 2, "batch_size": 1,
            "ops": [["H", 1], ["S", 1], ["H", 0], ["CX", 0, 1]],
            "measure": 1,
            "cases": [[0.0, 9], [1.0, 10]],
        },
    }
    with open(out, "w") as f:
        json.dump(kats, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else
         os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json"))
