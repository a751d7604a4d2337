"""Time evolution exp(-i t H) for a Pauli sum H (the emulate_time_evolution
call of _simulator_mpi.py:469-475, never defined in the reference's C++),
against the closed form: scipy's expm of the dense Hamiltonian applied to
the same state, with and without controls, identity terms as a phase."""
import numpy as np
import pytest
from scipy.linalg import expm

from hiqsimulator_amd.circuits import apply_circuit_inorder, random_layers_circuit

pytestmark = pytest.mark.gpu

PAULI = {"X": np.array([[0, 1], [1, 0]], dtype=complex),
         "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
         "Z": np.array([[1, 0], [0, -1]], dtype=complex)}


def dense_term(term, n):
    """The Pauli string (list of (qubit, 'X'|'Y'|'Z')) on n qubits, index bit q = qubit q."""
    ops = {q: PAULI[p] for q, p in term}
    m = np.array([[1]], dtype=complex)
    for q in reversed(range(n)):  # kron order: highest bit first
        m = np.kron(m, ops.get(q, np.eye(2)))
    return m


def random_hamiltonian(n, nterms, rng):
    terms = [([], float(rng.normal()))]  # identity part
    for _ in range(nterms):
        w = int(rng.integers(1, min(4, n) + 1))
        qs = rng.choice(n, size=w, replace=False)
        terms.append(([(int(q), "XYZ"[int(rng.integers(3))]) for q in sorted(qs)], float(rng.normal())))
    return terms


def prepared(engine_mod, n, seed, precision="c128"):
    sim = engine_mod.SimulatorMPI(seed, n, 4, precision)
    sim.allocate_qureg(list(range(n)), 0)
    apply_circuit_inorder(sim, random_layers_circuit(n, 4, seed), 4)
    return sim


@pytest.mark.parametrize("t", [0.3, -1.7, 4.0])
def test_time_evolution_matches_expm(engine_mod, t):
    n = 7
    rng = np.random.default_rng(11)
    terms = random_hamiltonian(n, 6, rng)
    print("time evolution t=%g" % t, flush=True)
    sim = prepared(engine_mod, n, 3)
    psi = sim.get_state()
    H = sum(c * dense_term(term, n) for term, c in terms)
    ref = expm(-1j * t * H) @ psi
    sim.emulate_time_evolution(terms, t, list(range(n)), [])
    got = sim.get_state()
    assert np.abs(got - ref).max() <= 1e-10
    assert abs(np.vdot(got, got).real - 1.0) <= 1e-10


def test_time_evolution_with_controls_and_subset(engine_mod):
    """H on qubits (2, 5, 0) (mask bit i -> ids[i]) controlled by qubits 1 and 4."""
    n = 6
    rng = np.random.default_rng(5)
    terms = random_hamiltonian(3, 4, rng)
    ids, ctrls, t = [2, 5, 0], [1, 4], 0.9
    sim = prepared(engine_mod, n, 8)
    psi = sim.get_state()
    # dense H on the n-qubit register: term qubit i acts on qubit ids[i]
    H = sum(c * dense_term([(ids[q], p) for q, p in term], n) for term, c in terms)
    U = expm(-1j * t * H)
    ref = psi.copy()
    on = np.array([all((i >> c) & 1 for c in ctrls) for i in range(2 ** n)])
    ref[on] = (U @ np.where(on, psi, 0))[on]
    sim.emulate_time_evolution(terms, t, ids, ctrls)
    assert np.abs(sim.get_state() - ref).max() <= 1e-10


def test_time_evolution_c64_and_errors(engine_mod):
    n = 6
    rng = np.random.default_rng(2)
    terms = random_hamiltonian(n, 5, rng)
    sim = prepared(engine_mod, n, 4, "c64")
    psi = sim.get_state()
    H = sum(c * dense_term(term, n) for term, c in terms)
    ref = expm(-1j * 0.5 * H) @ psi
    sim.emulate_time_evolution(terms, 0.5, list(range(n)), [])
    got = sim.get_state()
    f = abs(np.vdot(ref, got)) ** 2 / (np.vdot(ref, ref).real * np.vdot(got, got).real)
    assert f >= 1 - 1e-6
    with pytest.raises(RuntimeError):
        sim.emulate_time_evolution(terms, 0.5, [0, 1], [])  # a term beyond the qureg
    with pytest.raises(RuntimeError):
        sim.emulate_time_evolution([([(0, "X")], 1.0)], 0.5, [0], [0])  # control is a target
