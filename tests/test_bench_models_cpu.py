"""The byte models bench.py divides by kernel time, pinned on the CPU.

Per-op model (`algorithmic_bytes`): every op reads and writes its planes
once, R = 2n+1 words per plane per 64-shot word group: H touches x_q, z_q
and r both ways (6 planes), S reads x_q, z_q, r and writes z_q, r (5), X and
Z read one column and r and write r (3), Y reads both columns (4), CX and CZ
read and write four columns and r (8); measurements from the engine's census
(MeasureRandom copies, clears and rowsums, MeasurementFunctors.hpp:225-265;
MeasureDetermined folds K' stabilizers into the scratch row, :189-201).

Fused-minimum model (`fused_minimum_bytes`): a gate segment (ops between
measurements) reads each column it uses once and writes each column it
changes once, plus r both ways.  These tests check small cases by hand, the
inequality fused <= per-op on random programs, and the d=9 figures the
bench line reports for its documented census."""
import numpy as np
import pytest

import bench
from plaquette_simulator import workloads as W

ZERO = {"random": 0, "determined": 0, "random_rows": 0, "determined_steps": 0}


def ops_of(build):
    b = W.OpList()
    build(b)
    return b.array()


@pytest.mark.parametrize("code,planes", [(W.OP_H, 6), (W.OP_S, 5), (W.OP_X, 3), (W.OP_Y, 4),
                                         (W.OP_Z, 3), (W.OP_CX, 8), (W.OP_CZ, 8)])
def test_per_op_gate_planes(code, planes):
    n, words = 5, 3
    R = 2 * n + 1
    ops = ops_of(lambda b: b.gate(code, 0, 1))
    got = bench.algorithmic_bytes(ops, n, words, ZERO)
    assert got["gates"] == planes * 8 * R * words
    assert got["total"] == got["gates"]


def test_per_op_flagged_noise_and_measurement_terms():
    n, words = 4, 2
    R = 2 * n + 1
    ops = ops_of(lambda b: (b.flagged(W.OP_H_FLAG, 0, 0), b.flagged(W.OP_S_FLAG, 1, 0),
                            b.flagged(W.OP_X_FLAG, 2, 0), b.noise(3, 0.1, 0.1, 0.1)))
    got = bench.algorithmic_bytes(ops, n, words, ZERO)
    # flagged H/S: their planes plus the flag word per group; conditional
    # Pauli: x, z, r read and r written plus the flag; noise: two mask words
    assert got["flagged"] == (48 * R * words + 8 * words) + (40 * R * words + 8 * words) + \
        (32 * R * words + 8 * words)
    assert got["noise"] == 32 * R * words + 16 * words
    # one random group-measurement rowsumming K rows; one determined with K' steps
    K, Kd = 3, 2
    c = {"random": 1, "determined": 1, "random_rows": K, "determined_steps": Kd}
    m = bench.algorithmic_bytes(np.zeros((0, 8), np.uint32), n, words, c)["measurements"]
    assert m == 8 * ((4 + 2 * K) * R + 2) + 8 * (2 * n + (Kd + 1) * R + 2)


def test_fused_segments_by_hand():
    n, words = 6, 1
    R = 2 * n + 1
    col = 8 * R * words
    # one segment: H(0) twice reads {x0, z0} and writes {x0, z0}: 4 columns + r twice
    one = ops_of(lambda b: (b.h(0), b.h(0)))
    assert bench.fused_minimum_bytes(one, n, words, ZERO)["total"] == 6 * col
    assert bench.algorithmic_bytes(one, n, words, ZERO)["total"] == 12 * col
    # a measurement splits segments: H(0) | M | CX(1,2) -> 6 + (4 read + 2 written + 2) columns
    two = ops_of(lambda b: (b.h(0), b.measure(3), b.cx(1, 2)))
    got = bench.fused_minimum_bytes(two, n, words, ZERO)
    assert got["segments"] == 2
    assert got["segment_columns"] == (6 + 8) * col
    # noise costs its two mask words per group, nothing else beyond the segment's columns
    nz = ops_of(lambda b: (b.cx(1, 2), b.noise(1, 0.01, 0.0, 0.0)))
    g = bench.fused_minimum_bytes(nz, n, words, ZERO)
    assert g["masks"] == 16 * words
    assert g["segment_columns"] == 8 * col


def test_fused_never_exceeds_per_op():
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(2, 40))
        ops = W.random_clifford(n, int(rng.integers(1, 300)), int(rng.integers(0, 1 << 30)),
                                meas_prob=float(rng.uniform(0, 0.3)))
        words = int(rng.integers(1, 5))
        f = bench.fused_minimum_bytes(ops, n, words, ZERO)["total"]
        a = bench.algorithmic_bytes(ops, n, words, ZERO)["total"]
        assert f <= a, (trial, f, a)


def test_d9_figures_for_the_documented_census():
    w = W.config_surface_d9()
    census = {"random": 20480, "determined": 184576, "random_rows": 74752,
              "determined_steps": 367872}
    words = (w.batch_size + 63) // 64
    a = bench.algorithmic_bytes(w.ops, w.num_qubits, words, census)
    f = bench.fused_minimum_bytes(w.ops, w.num_qubits, words, census)
    assert a["total"] == 36631785472
    assert f["total"] == 6192017408
