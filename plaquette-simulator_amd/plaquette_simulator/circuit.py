"""CliffordCircuit: the reference's circuit-file runner
(plaquette_simulator/src/CliffordCircuit.hpp:16-151), for Python users.

The file holds one operation per line, ``NAME q...`` with NAME in
H S X Y Z CX CZ M.  The whole file is parsed once by the engine's own parser
into an op list, and ``execute()`` runs it as one program on the GPU (one
persistent kernel by default, ``form="plain"`` for one launch per op,
``form="graph"`` for a hipGraph replay, ``form="frame"`` for the frame form,
which runs the shared X/Z once on the host).  Results and record ids are those of
the same operations issued one by one: each ``M`` line appends one record per
listed qubit (bias 1/2).

Errors mirror the reference: an unsupported name (a blank line's empty name
included) raises ``RuntimeError("Unsupported operation: <name>")`` and a file
that cannot be opened reads as an empty circuit.  Qubit indices are checked
when the circuit runs.
"""
import numpy as np

from plaquette_simulator_bindings import CliffordState, parse_circuit


class CliffordCircuit:
    def __init__(self, file_name, num_qubits, batch_size, seed=-1):
        try:
            with open(file_name, "r") as f:
                text = f.read()
        except OSError:
            text = ""
        self._init(text, num_qubits, batch_size, seed)

    @classmethod
    def from_text(cls, text, num_qubits, batch_size, seed=-1):
        c = cls.__new__(cls)
        c._init(text, num_qubits, batch_size, seed)
        return c

    def _init(self, text, num_qubits, batch_size, seed):
        ops, lines, _ = parse_circuit(text)
        self._ops = ops
        self._lines = int(lines)
        self._state = CliffordState(num_qubits, batch_size, seed, True)

    def get_clifford_state(self):
        return self._state

    def get_num_qubits(self):
        return self._state.get_num_qubits()

    def get_batch_size(self):
        return self._state.get_batch_size()

    def get_num_operations(self):
        """Lines read (CliffordCircuit.hpp:127-129)."""
        return self._lines

    @property
    def ops(self):
        """Engine op list (uint32 [nops, 8]) the lines expand to."""
        return self._ops

    def execute(self, form="persistent"):
        """Apply every operation once (CliffordCircuit.hpp:131-135); M lines
        append records after those the state already holds."""
        if len(self._ops) == 0:
            return
        ops = self._ops.copy()
        m = ops[:, 0] == 14
        ops[m, 3] += np.uint32(self._state.num_records)
        self._state.run_program(ops, form)
