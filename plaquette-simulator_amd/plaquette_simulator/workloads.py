"""Circuit builders for the engine's op-list format (``pq_op``, 8 x uint32).

Each op row is ``(code, q0, q1, slot, t0, t1, t2, 0)`` with the op codes of
include/plaquette_amd.h.  These builders produce the workloads the engine is
measured on; they are plain NumPy (no GPU needed), so the same arrays feed
both the engine and the CPU oracle.

Workloads:
  * ``random_clifford``: uniform {H(q), S(q), CX(c,t)} with measurements
    either after every k-th op or as a fraction of ops.
  * ``surface_code_memory_z``: rotated surface code, memory-Z experiment.
    Data qubits sit on a d x d grid (index r*d + c); plaquette (r, c) for
    r, c in [-1, d-1] has corners NW (r,c), NE (r,c+1), SW (r+1,c),
    SE (r+1,c+1); it is X-type when r + c is even.  Bulk plaquettes (both
    coordinates in [0, d-2]) are kept, and on the boundary only X-type
    plaquettes of the top/bottom rows and Z-type of the left/right columns
    (weight 2).  Ancillas follow the data qubits, ordered by (r, c).
    CNOT layers use the hook-avoiding orders NW, NE, SW, SE for X-type
    (ancilla controls) and NW, SW, NE, SE for Z-type (ancilla is target).
    One round: H on X ancillas, 4 CX layers, H on X ancillas (each gate
    layer followed by a Pauli channel X/Y/Z with probability p each on every
    touched qubit), measure all ancillas in index order, then reset each
    ancilla with an X conditioned on its own outcome.  Final: measure all
    data qubits.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

OP_H, OP_S, OP_X, OP_Y, OP_Z, OP_CX, OP_CZ = 1, 2, 3, 4, 5, 6, 7
OP_H_FLAG, OP_S_FLAG, OP_NOISE = 8, 9, 10
OP_X_FLAG, OP_Y_FLAG, OP_Z_FLAG, OP_M = 11, 12, 13, 14
OP_CX_FLAG, OP_CZ_FLAG = 15, 16
FLAG_RANDOM = 0xFFFFFFFF  # slot of a flagged op whose flag is a Bernoulli draw
FLAGGED = {OP_H: OP_H_FLAG, OP_S: OP_S_FLAG, OP_X: OP_X_FLAG, OP_Y: OP_Y_FLAG,
           OP_Z: OP_Z_FLAG, OP_CX: OP_CX_FLAG, OP_CZ: OP_CZ_FLAG}
GATE_CODES = (OP_H, OP_S, OP_X, OP_Y, OP_Z, OP_CX, OP_CZ)


def coin_threshold(bias: float) -> int:
    """ceil(bias * 2^24) clamped to [0, 2^24] (coin = (draw >> 40) < T)."""
    t = float(np.float32(bias)) * 16777216.0
    if not t > 0.0:
        return 0
    if t >= 16777216.0:
        return 16777216
    return int(math.ceil(t))


def noise_thresholds(px: float, py: float, pz: float) -> tuple[int, int, int]:
    """Cumulative 32-bit thresholds (X if u < t0, Y if u < t1, Z if u < t2)."""
    def th(p):
        v = p * 4294967296.0
        return 0 if v <= 0 else (0xFFFFFFFF if v >= 4294967295.0 else int(v))
    return th(px), th(px + py), th(px + py + pz)


class OpList:
    """Incremental op-list builder."""

    def __init__(self):
        self.rows: list[tuple] = []
        self.nrec = 0

    def gate(self, code, q0, q1=0):
        self.rows.append((code, q0, q1, 0, 0, 0, 0, 0))

    def h(self, q): self.gate(OP_H, q)
    def s(self, q): self.gate(OP_S, q)
    def cx(self, c, t): self.gate(OP_CX, c, t)
    def cz(self, c, t): self.gate(OP_CZ, c, t)

    def flagged(self, code, q, slot, q1=0):
        """Flagged op `code` (H_FLAG..CZ_FLAG) with record `slot` as its flag."""
        self.rows.append((code, q, q1, slot, 0, 0, 0, 0))

    def with_prob(self, gate, q, prob, q1=0):
        """Gate `gate` (H..CZ) on each shot with probability `prob`."""
        self.rows.append((FLAGGED[gate], q, q1, FLAG_RANDOM, coin_threshold(prob), 0, 0, 0))

    def noise(self, q, px, py, pz):
        t0, t1, t2 = noise_thresholds(px, py, pz)
        self.rows.append((OP_NOISE, q, 0, 0, t0, t1, t2, 0))

    def measure(self, q, bias=0.5, slot=None):
        if slot is None:
            slot = self.nrec
            self.nrec += 1
        self.rows.append((OP_M, q, 0, slot, coin_threshold(bias), 0, 0, 0))
        return slot

    def array(self) -> np.ndarray:
        if not self.rows:
            return np.zeros((0, 8), np.uint32)
        return np.array(self.rows, dtype=np.uint32)


@dataclass
class Workload:
    name: str
    num_qubits: int
    batch_size: int
    seed: int
    ops: np.ndarray
    meta: dict

    @property
    def gates(self) -> int:
        return int(np.isin(self.ops[:, 0], GATE_CODES).sum())

    @property
    def measurements(self) -> int:
        return int((self.ops[:, 0] == OP_M).sum())

    @property
    def noise_ops(self) -> int:
        return int((self.ops[:, 0] == OP_NOISE).sum())


def random_clifford(n, nops, seed, meas_every=None, meas_prob=None, bias=0.5):
    """Uniform {H(q), S(q), CX(c,t)} ops, q/c/t uniform over [0, n), c != t.

    meas_every=k: a measurement on a uniform qubit follows every k-th op.
    meas_prob=p: each op is a measurement with probability p instead.
    Draws use numpy's PCG64 seeded with `seed`."""
    rng = np.random.default_rng(seed)
    b = OpList()
    for i in range(nops):
        if meas_prob is not None and rng.random() < meas_prob:
            b.measure(int(rng.integers(0, n)), bias)
            continue
        k = int(rng.integers(0, 3))
        if k == 0:
            b.h(int(rng.integers(0, n)))
        elif k == 1:
            b.s(int(rng.integers(0, n)))
        else:
            c = int(rng.integers(0, n))
            t = int(rng.integers(0, n - 1))
            t += t >= c
            b.cx(c, t)
        if meas_every and (i + 1) % meas_every == 0:
            b.measure(int(rng.integers(0, n)), bias)
    return b.array()


def surface_code_layout(d: int):
    """Data grid d x d and the d^2 - 1 plaquettes of a rotated surface code.

    Returns (plaquettes, n): each plaquette is (kind, ancilla, corners) with
    kind 'X' or 'Z' and corners a dict NW/NE/SW/SE -> data qubit."""
    plaq = []
    for r in range(-1, d):
        for c in range(-1, d):
            kind = "X" if (r + c) % 2 == 0 else "Z"
            bulk = 0 <= r <= d - 2 and 0 <= c <= d - 2
            top_bottom = (r == -1 or r == d - 1) and 0 <= c <= d - 2
            left_right = (c == -1 or c == d - 1) and 0 <= r <= d - 2
            if not (bulk or (top_bottom and kind == "X") or (left_right and kind == "Z")):
                continue
            corners = {}
            for name, (rr, cc) in (("NW", (r, c)), ("NE", (r, c + 1)),
                                   ("SW", (r + 1, c)), ("SE", (r + 1, c + 1))):
                if 0 <= rr < d and 0 <= cc < d:
                    corners[name] = rr * d + cc
            plaq.append((kind, corners))
    out = []
    for i, (kind, corners) in enumerate(plaq):
        out.append((kind, d * d + i, corners))
    return out, d * d + len(plaq)


ORDER_X = ("NW", "NE", "SW", "SE")
ORDER_Z = ("NW", "SW", "NE", "SE")


def surface_code_memory_z(d: int, rounds: int, p: float, bias: float = 0.5):
    """Op list of a rotated surface-code memory-Z experiment (see module doc).

    One round is built op by op; the remaining rounds are copies of it with
    their record slots shifted by the number of ancillas per round."""
    plaq, n = surface_code_layout(d)
    xs = [a for k, a, _ in plaq if k == "X"]
    ancillas = [a for _, a, _ in plaq]
    na = len(ancillas)
    b = OpList()

    def noisy(touched):
        if p > 0:
            for q in touched:
                b.noise(q, p, p, p)

    for a in xs:
        b.h(a)
    noisy(xs)
    for layer in range(4):
        touched = []
        for kind, a, corners in plaq:
            key = (ORDER_X if kind == "X" else ORDER_Z)[layer]
            if key not in corners:
                continue
            dq = corners[key]
            if kind == "X":
                b.cx(a, dq)
            else:
                b.cx(dq, a)
            touched += [a, dq]
        noisy(touched)
    for a in xs:
        b.h(a)
    noisy(xs)
    slots = [b.measure(a, bias) for a in ancillas]
    for a, sl in zip(ancillas, slots):
        b.flagged(OP_X_FLAG, a, sl)
    one = b.array()
    has_slot = np.isin(one[:, 0], (OP_M, OP_X_FLAG))
    blocks = []
    for r in range(rounds):
        blk = one.copy()
        blk[has_slot, 3] += np.uint32(r * na)
        blocks.append(blk)
    tail = OpList()
    tail.nrec = rounds * na
    for q in range(d * d):
        tail.measure(q, bias)
    blocks.append(tail.array())
    return np.concatenate(blocks, axis=0), n


# ---- the measured configurations ----

def config_random_n64(seed=123423, shots=1024):
    """n=64, 1024 shots, 2000 uniform H/S/CX ops, M on a uniform qubit after
    every 100th op, init |0...0>."""
    ops = random_clifford(64, 2000, seed, meas_every=100)
    return Workload("random_clifford_n64_b1024", 64, shots, seed, ops, {"ops": 2000})


def config_surface_d9(shots=16384, seed=7, p=1e-3, d=9, rounds=None):
    """Rotated surface-code memory-Z, d=9 (161 qubits), 9 rounds, 16384 shots,
    Pauli X/Y/Z noise p=1e-3 each after every gate layer, seed 7."""
    rounds = d if rounds is None else rounds
    ops, n = surface_code_memory_z(d, rounds, p)
    return Workload(f"surface_code_d{d}_r{rounds}_b{shots}", n, shots, seed, ops,
                    {"d": d, "rounds": rounds, "p": p})


def config_random_n1024(seed=42, shots=4096, nops=100000):
    """n=1024, 4096 shots, 100000 ops, 10% measurements on uniform qubits."""
    ops = random_clifford(1024, nops, seed, meas_prob=0.1)
    return Workload("random_clifford_n1024_b4096", 1024, shots, seed, ops, {"ops": nops})


def config_mixed_batch(count=256, seed=11, p=1e-3, dmin=3, dmax=25, smin=64, smax=4096):
    """`count` surface-code circuits, d uniform over odd {dmin..dmax},
    rounds = d, shots log-uniform in [smin, smax]; per-circuit state seeds
    drawn from the same generator."""
    rng = np.random.default_rng(seed)
    ds = list(range(dmin, dmax + 1, 2))
    out = []
    for i in range(count):
        d = int(rng.choice(ds))
        shots = int(round(math.exp(rng.uniform(math.log(smin), math.log(smax)))))
        cseed = int(rng.integers(0, 2 ** 31 - 1))
        ops, n = surface_code_memory_z(d, d, p)
        out.append(Workload(f"sc_d{d}_b{shots}_{i}", n, shots, cseed, ops,
                            {"d": d, "rounds": d, "p": p}))
    return out


def serving_request(rng, dmin=3, dmax=21, smin=256, smax=65536, pmin=5e-4, pmax=5e-3):
    """One circuit-sampling request of the serving stream: a surface-code
    memory-Z circuit with d uniform over odd {dmin..dmax}, rounds d, shots
    log-uniform in [smin, smax], p uniform in [pmin, pmax] and a state seed.
    Returns (d, shots, p, seed); the op list is surface_code_memory_z(d, d, p)."""
    d = int(rng.choice(list(range(dmin, dmax + 1, 2))))
    shots = int(round(math.exp(rng.uniform(math.log(smin), math.log(smax)))))
    p = float(rng.uniform(pmin, pmax))
    seed = int(rng.integers(0, 2 ** 31 - 1))
    return d, shots, p, seed


def config_serving_stream(lam, count, seed=2024):
    """`count` Poisson arrivals at rate `lam` per second of serving requests:
    a list of (arrival time in s, d, shots, p, seed)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / lam)
        out.append((t,) + serving_request(rng))
    return out
