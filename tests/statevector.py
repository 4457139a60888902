"""Brute-force statevector simulator for small n (test helper).

Independent of any tableau code: it pins the tableau gate rules, outcome
determinism and stabilizer signs through the mathematics (a Clifford
circuit's stabilizer rows must stabilize the simulated state).
Qubit j is bit j of the basis-state index.
"""
import numpy as np

_H = np.array([[1, 1], [1, -1]], complex) / np.sqrt(2)
_S = np.array([[1, 0], [0, 1j]], complex)
_X = np.array([[0, 1], [1, 0]], complex)
_Y = np.array([[0, -1j], [1j, 0]], complex)
_Z = np.array([[1, 0], [0, -1]], complex)


class StateVector:
    def __init__(self, n):
        self.n = n
        self.psi = np.zeros(2 ** n, complex)
        self.psi[0] = 1.0

    def _apply1(self, U, q):
        psi = self.psi.reshape([2] * self.n)  # axis k <-> qubit n-1-k
        ax = self.n - 1 - q
        psi = np.moveaxis(np.tensordot(U, np.moveaxis(psi, ax, 0), axes=(1, 0)), 0, ax)
        self.psi = psi.reshape(-1)

    def h(self, q): self._apply1(_H, q)
    def s(self, q): self._apply1(_S, q)
    def x(self, q): self._apply1(_X, q)
    def y(self, q): self._apply1(_Y, q)
    def z(self, q): self._apply1(_Z, q)

    def cx(self, c, t):
        idx = np.arange(2 ** self.n)
        sel = ((idx >> c) & 1) == 1
        out = self.psi.copy()
        out[idx[sel]] = self.psi[idx[sel] ^ (1 << t)]
        self.psi = out

    def cz(self, c, t):
        idx = np.arange(2 ** self.n)
        sel = (((idx >> c) & 1) == 1) & (((idx >> t) & 1) == 1)
        self.psi = np.where(sel, -self.psi, self.psi)

    def prob_one(self, q):
        idx = np.arange(2 ** self.n)
        return float(np.sum(np.abs(self.psi[((idx >> q) & 1) == 1]) ** 2))

    def project(self, q, outcome):
        idx = np.arange(2 ** self.n)
        keep = ((idx >> q) & 1) == outcome
        self.psi = np.where(keep, self.psi, 0)
        self.psi /= np.linalg.norm(self.psi)

    def pauli_expectation_sign(self, xbits, zbits, rbit):
        """<psi| (-1)^r P |psi> for the tableau row (x, z, r), P = prod_j
        {I,X,Z,Y}[x_j,z_j] (Aaronson-Gottesman row convention)."""
        phi = self.psi.copy()
        sv = StateVector(self.n)
        sv.psi = phi
        for j in range(self.n):
            if xbits[j] and zbits[j]:
                sv.y(j)
            elif xbits[j]:
                sv.x(j)
            elif zbits[j]:
                sv.z(j)
        val = np.vdot(self.psi, sv.psi)
        return ((-1) ** rbit) * val
