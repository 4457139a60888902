"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (plain-C restatement of the reference
algorithm, reference int32 layout ``[B][2n+1][n]``) and exposes it through an
object whose method names follow the reference's Python/C++ API
(``plaquette_simulator/src/Bindings.cpp:13-31``), so tests that drive the
oracle read like the reference's own tests.  Only tests, the smoke check and
the CPU-baseline leg of the benchmark use this module, always as the checker.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ORC_H, ORC_S, ORC_X, ORC_Y, ORC_Z, ORC_CX, ORC_CZ = 1, 2, 3, 4, 5, 6, 7
ORC_H_FLAG, ORC_S_FLAG, ORC_NOISE = 8, 9, 10
ORC_X_FLAG, ORC_Y_FLAG, ORC_Z_FLAG, ORC_M = 11, 12, 13, 14
ORC_CX_FLAG, ORC_CZ_FLAG = 15, 16
ORC_FLAG_RANDOM = 0xFFFFFFFF  # slot of a flagged op with a Bernoulli flag

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64


def lib():
    """Load (building first if needed) the oracle shared library."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        import subprocess

        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    L = ctypes.CDLL(path)
    L.orc_g.restype = ctypes.c_int32
    L.orc_g.argtypes = [ctypes.c_int32] * 4
    L.orc_init_zero.argtypes = [_i64, _i64, _i32p, _i32p, _i32p]
    L.orc_checksum.restype = ctypes.c_int64
    L.orc_checksum.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, _i64, _i64]
    L.orc_gate.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, ctypes.c_int, _i64, _i64]
    L.orc_gate_flag.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, ctypes.c_int, _i64, _i32p]
    L.orc_gate2_flag.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, ctypes.c_int, _i64, _i64,
                                 _i32p]
    L.orc_pauli_flag.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, _i64, _i32p, _i32p]
    L.orc_row_product_sign.restype = ctypes.c_int32
    L.orc_row_product_sign.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, _i64, _i64, _i64]
    L.orc_row_multiply.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, _i64, _i64, _i64]
    L.orc_measure_determined.restype = ctypes.c_int32
    L.orc_measure_determined.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, _i64, _i64, _i64p]
    L.orc_measure_random.restype = ctypes.c_int32
    L.orc_measure_random.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, _i64, _i64, _i64,
                                     ctypes.c_int32, _i64p]
    L.orc_measure.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, _i64, _i32p, _i32p, _i32p, _i64p]
    L.orc_mix64.restype = ctypes.c_uint64
    L.orc_mix64.argtypes = [ctypes.c_uint64]
    L.orc_draw.restype = ctypes.c_uint64
    L.orc_draw.argtypes = [ctypes.c_uint64] * 3
    L.orc_coin_threshold.restype = ctypes.c_uint32
    L.orc_coin_threshold.argtypes = [ctypes.c_float]
    L.orc_run_program.restype = ctypes.c_int
    L.orc_run_program.argtypes = [_i64, _i64, _i32p, _i32p, _i32p, _i32p, _i64,
                                  ctypes.c_uint64, _i64, _i32p, _i64, _i64p, _i64p, _i64p]
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_get_threads.restype = ctypes.c_int
    _LIB = L
    return L


def _p32(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i32p)


DEFAULT_SEED = 213434232223  # CliffordState.hpp:135


class OracleState:
    """Reference-layout tableau driven by the oracle.

    ``x``/``z`` are int32 ``[B][2n+1][n]``; ``r`` is int32 ``[B][2n+1]``
    (CliffordState.hpp:113-120).  Coins for random outcomes come from
    ``coins`` (an int32 [B] array per call) or, when omitted, from the
    counter-based draw keyed by (seed, measurement sequence, shot).
    """

    def __init__(self, num_qubits, batch_size, seed=-1, init_to_zero=True, shot0=0):
        self.n = int(num_qubits)
        self.B = int(batch_size)
        self.R = 2 * self.n + 1
        self.seed = int(seed) if seed >= 0 else DEFAULT_SEED
        self.shot0 = int(shot0)
        self.x = np.zeros((self.B, self.R, self.n), np.int32)
        self.z = np.zeros((self.B, self.R, self.n), np.int32)
        self.r = np.zeros((self.B, self.R), np.int32)
        self.records: list[tuple[np.ndarray, np.ndarray]] = []
        self.meas_seq = 0
        self.noise_seq = 0
        self.stats = np.zeros(4, np.int64)
        if init_to_zero and self.n > 0 and self.B > 0:
            lib().orc_init_zero(self.n, self.B, _p32(self.x), _p32(self.z), _p32(self.r))

    def reset_clifford_state(self):
        """ResetCliffordState's intended semantics (CliffordState.hpp:309-316):
        the |0...0> tableau again; records and sequence numbers are kept."""
        if self.n > 0 and self.B > 0:
            lib().orc_init_zero(self.n, self.B, _p32(self.x), _p32(self.z), _p32(self.r))

    @classmethod
    def from_flat(cls, x, z, r, num_qubits, batch_size, seed=-1):
        s = cls(num_qubits, batch_size, seed, init_to_zero=False)
        s.x[...] = np.asarray(x, np.int32).reshape(s.x.shape)
        s.z[...] = np.asarray(z, np.int32).reshape(s.z.shape)
        s.r[...] = np.asarray(r, np.int32).reshape(s.r.shape)
        return s

    def _args(self):
        return (self.n, self.B, _p32(self.x), _p32(self.z), _p32(self.r))

    # -- gates (CliffordState.hpp:348-531) --
    def _gate(self, code, q0, q1=0):
        lib().orc_gate(*self._args(), code, int(q0), int(q1))

    def apply_hadamard_gate(self, q): self._gate(ORC_H, q)
    def apply_phase_gate(self, q): self._gate(ORC_S, q)
    def apply_pauli_x_gate(self, q): self._gate(ORC_X, q)
    def apply_pauli_y_gate(self, q): self._gate(ORC_Y, q)
    def apply_pauli_z_gate(self, q): self._gate(ORC_Z, q)
    def apply_control_not_gate(self, c, t): self._gate(ORC_CX, c, t)
    def apply_control_phase_gate(self, c, t): self._gate(ORC_CZ, c, t)

    def apply_hadamard_gate_with_apply_flag(self, q, flag):
        f = np.ascontiguousarray(flag, np.int32)
        lib().orc_gate_flag(*self._args(), ORC_H_FLAG, int(q), _p32(f))

    def apply_phase_gate_with_apply_flag(self, q, flag):
        f = np.ascontiguousarray(flag, np.int32)
        lib().orc_gate_flag(*self._args(), ORC_S_FLAG, int(q), _p32(f))

    def apply_control_not_gate_with_apply_flag(self, c, t, flag):
        f = np.ascontiguousarray(flag, np.int32)
        lib().orc_gate2_flag(*self._args(), ORC_CX_FLAG, int(c), int(t), _p32(f))

    def apply_control_phase_gate_with_apply_flag(self, c, t, flag):
        f = np.ascontiguousarray(flag, np.int32)
        lib().orc_gate2_flag(*self._args(), ORC_CZ_FLAG, int(c), int(t), _p32(f))

    def apply_pauli_flag(self, q, fx=None, fz=None):
        ax = np.ascontiguousarray(fx, np.int32) if fx is not None else None
        az = np.ascontiguousarray(fz, np.int32) if fz is not None else None
        lib().orc_pauli_flag(*self._args(), int(q),
                             _p32(ax) if ax is not None else None,
                             _p32(az) if az is not None else None)

    # -- measurement (CliffordState.hpp:587-623) --
    def coins_for(self, m, bias):
        L = lib()
        t = L.orc_coin_threshold(float(bias))
        return np.array([(L.orc_draw(self.seed, 2 * m, self.shot0 + b) >> 40) < t
                         for b in range(self.B)], np.int32)

    def measure_qubit(self, q, bias=0.5, measurement_id=None, coins=None):
        if coins is None:
            coins = self.coins_for(self.meas_seq, bias)
        self.meas_seq += 1
        coins = np.ascontiguousarray(coins, np.int32)
        res = np.zeros(self.B, np.int32)
        det = np.zeros(self.B, np.int32)
        lib().orc_measure(*self._args(), int(q), _p32(coins), _p32(res), _p32(det),
                          self.stats.ctypes.data_as(_i64p))
        if measurement_id is None:
            self.records.append((res, det))
            return len(self.records) - 1
        if measurement_id >= len(self.records):
            raise RuntimeError("Measurement id is out of range.")
        self.records[measurement_id] = (res, det)
        return measurement_id

    def get_measurement(self, mid, b):
        if mid < len(self.records) and b < self.B:
            res, det = self.records[mid]
            return int(res[b]), int(det[b])
        return None

    # -- per-shot helpers --
    def row_product_sign(self, b, i, k):
        return lib().orc_row_product_sign(*self._args(), b, i, k)

    def row_multiply(self, b, i, k):
        lib().orc_row_multiply(*self._args(), b, i, k)

    def measure_determined(self, b, q):
        return lib().orc_measure_determined(*self._args(), b, q, None)

    def check_sum(self, b0=0, b1=-1):
        return lib().orc_checksum(*self._args(), b0, b1)

    def device_to_host(self):
        return (self.x.reshape(-1).copy(), self.z.reshape(-1).copy(), self.r.reshape(-1).copy())

    # -- program interpreter --
    def run_program(self, ops: np.ndarray, nrec_total: int | None = None):
        """Run an op list (int32 [nops][8]); records are created as needed."""
        # fields are uint32 (a random flag's slot is 0xFFFFFFFF); the C side
        # reads the same bits as int32
        ops = np.ascontiguousarray(
            np.asarray(ops, np.int64).astype(np.uint32).view(np.int32)).reshape(-1, 8)
        need = len(self.records)
        for op in ops:
            if op[0] == ORC_M:
                need = max(need, int(op[3]) + 1)
        if nrec_total is not None:
            need = max(need, nrec_total)
        rec = np.zeros((need, 2, self.B), np.int32)
        for i, (res, det) in enumerate(self.records):
            rec[i, 0] = res
            rec[i, 1] = det
        ms = np.array([self.meas_seq], np.int64)
        ns = np.array([self.noise_seq], np.int64)
        rc = lib().orc_run_program(*self._args(), _p32(ops), len(ops), self.seed & (2**64 - 1),
                                   self.shot0, _p32(rec), need,
                                   ms.ctypes.data_as(_i64p), ns.ctypes.data_as(_i64p),
                                   self.stats.ctypes.data_as(_i64p))
        if rc != 0:
            raise RuntimeError(f"oracle program failed: {rc}")
        self.meas_seq = int(ms[0])
        self.noise_seq = int(ns[0])
        self.records = [(rec[i, 0].copy(), rec[i, 1].copy()) for i in range(need)]


def g(x1, z1, x2, z2):
    return lib().orc_g(x1, z1, x2, z2)


def mix64(v):
    return lib().orc_mix64(v)


def draw(seed, stream, idx):
    return lib().orc_draw(seed, stream, idx)


def coin_threshold(bias):
    return lib().orc_coin_threshold(bias)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()
