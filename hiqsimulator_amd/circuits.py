"""Gate matrices and seeded synthetic circuits for the state-vector engine.

Gate matrices follow ProjectQ's published definitions, which is where the
reference takes them from (``cmd.gate.matrix``, _simulator_mpi.py:476-488):
H = [[1,1],[1,-1]]/sqrt2, X, Y, Z, S, T, Rx(t) = exp(-i t X/2),
Ry(t) = exp(-i t Y/2), Rz(t) = diag(e^{-it/2}, e^{it/2}), R(p) = diag(1, e^{ip}).

A circuit is a list of ``(matrix, targets, controls)`` with matrix bit l
acting on ``targets[l]`` (targets[0] is the least significant bit), the
order SimulatorMPI::ApplyGate expects (SimulatorMPI.cpp:710-815).
"""
import math

import numpy as np

SQ2 = 1.0 / math.sqrt(2.0)
H = np.array([[SQ2, SQ2], [SQ2, -SQ2]], dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
S = np.array([[1, 0], [0, 1j]], dtype=np.complex128)
T = np.array([[1, 0], [0, np.exp(1j * math.pi / 4)]], dtype=np.complex128)


def Rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def Ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def Rz(t):
    return np.array([[np.exp(-0.5j * t), 0], [0, np.exp(0.5j * t)]], dtype=np.complex128)


def R(phi):
    return np.array([[1, 0], [0, np.exp(1j * phi)]], dtype=np.complex128)


def haar_unitary(dim, rng):
    """Haar-random unitary: QR of a complex Ginibre matrix with the phase fix."""
    g = (rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))) / math.sqrt(2.0)
    q, r = np.linalg.qr(g)
    d = np.diag(r)
    return q * (d / np.abs(d))


def random_layers_circuit(n_qubits, n_layers, seed, haar_every=0):
    """Random layered circuit on qubit ids 0..n-1.

    Each layer applies to every qubit one gate drawn uniformly from
    {H, Rx(t), Ry(t), Rz(t)} with t ~ U[0, 2pi), then CNOT(i -> i+1) for
    i = 0, 2, ... on even layers and i = 1, 3, ... on odd layers.  With
    ``haar_every`` > 0 a Haar-random 2-qubit unitary is added on a random
    pair after every ``haar_every``-th layer.
    """
    rng = np.random.default_rng(seed)
    gates = []
    for layer in range(n_layers):
        kinds = rng.integers(0, 4, size=n_qubits)
        thetas = rng.uniform(0.0, 2.0 * math.pi, size=n_qubits)
        c, s = np.cos(thetas / 2), np.sin(thetas / 2)
        m = np.zeros((n_qubits, 2, 2), dtype=np.complex128)
        m[kinds == 0] = H
        rx, ry, rz = kinds == 1, kinds == 2, kinds == 3
        m[rx, 0, 0] = m[rx, 1, 1] = c[rx]
        m[rx, 0, 1] = m[rx, 1, 0] = -1j * s[rx]
        m[ry, 0, 0] = m[ry, 1, 1] = c[ry]
        m[ry, 0, 1] = -s[ry]
        m[ry, 1, 0] = s[ry]
        m[rz, 0, 0] = np.exp(-0.5j * thetas[rz])
        m[rz, 1, 1] = np.exp(0.5j * thetas[rz])
        gates.extend(zip(list(m), ([q] for q in range(n_qubits)), ([] for _ in range(n_qubits))))
        gates.extend((X, [i + 1], [i]) for i in range(layer % 2, n_qubits - 1, 2))
        if haar_every and (layer + 1) % haar_every == 0 and n_qubits >= 2:
            a, b = rng.choice(n_qubits, size=2, replace=False)
            gates.append((haar_unitary(4, rng), [int(a), int(b)], []))
    return gates


def qft_like_circuit(n_qubits):
    """H on each qubit j (high to low), then controlled phases 2pi/2^(m-j+1)."""
    gates = []
    for j in reversed(range(n_qubits)):
        gates.append((H, [j], []))
        for m in reversed(range(j)):
            gates.append((R(2.0 * math.pi / 2 ** (j - m + 1)), [j], [m]))
    return gates


def random_permutation(ids, rng):
    """Uniform permutation of a list (Fisher-Yates via numpy)."""
    ids = list(ids)
    order = rng.permutation(len(ids))
    return [ids[i] for i in order]


def apply_circuit(sim, gates, flush_every=None):
    """Feed gates to a SimulatorMPI-like object, calling run() after every
    `flush_every` gates (None: once at the end)."""
    for i, (m, t, c) in enumerate(gates):
        sim.apply_controlled_gate(m, t, c)
        if flush_every and (i + 1) % flush_every == 0:
            sim.run()
    sim.run()


def apply_circuit_inorder(sim, gates, k):
    """Feed gates in program order, calling run() whenever the next gate
    would widen the pending fusion beyond k qubits (targets and controls),
    so every fused pass acts on at most k qubits."""
    cur = set()
    for m, t, c in gates:
        q = set(t) | set(c)
        if cur and len(cur | q) > k:
            sim.run()
            cur = set()
        sim.apply_controlled_gate(m, t, c)
        cur |= q
    sim.run()
