"""Backend-agnostic scenarios restating the reference test-suite's known
answers (hiq/projectq/backends/_sim/_simulator_mpi_test.py) at the level of
the `_cppsim_mpi.SimulatorMPI` API (no ProjectQ in this environment).

Each scenario takes ``make(seed, max_local, max_cluster)`` returning a
SimulatorMPI-like object: the oracle (CPU) or the MI355X engine.  Gates are
applied with one run() per gate, the reference engine's gate_fusion=False
path (_simulator_mpi.py:489-490), unless a scenario says otherwise.
"""
import math

import numpy as np
import pytest

from hiqsimulator_amd.circuits import H, Rx, Ry, Rz, S, X, Y, Z


def gate(sim, m, t, c=()):
    sim.apply_controlled_gate(np.asarray(m, dtype=np.complex128), list(t), list(c))
    sim.run()


def kqubit_gate(make):
    """test_simulator_kqubit_gate (:247-283): kron bit order and controls."""
    sim = make(1, 33, 4)
    m1, m2, m3 = Rx(0.3), Rx(0.8), Ry(0.1)
    m4 = Rz(0.9).dot(Ry(-0.1))
    m = np.kron(m4, np.kron(m3, np.kron(m2, m1)))
    sim.allocate_qureg([0, 1, 2, 3], 0)
    sim.allocate_qubit(4)
    gate(sim, Rx(-0.3), [0])
    gate(sim, Rx(-0.8), [1])
    gate(sim, Ry(-0.1), [2])
    gate(sim, Rz(-0.9), [3])
    gate(sim, Ry(0.1), [3])
    gate(sim, X, [4])
    gate(sim, m, [0, 1, 2, 3], [4])
    gate(sim, X, [4])
    gate(sim, m.conj().T, [0, 1, 2, 3], [4])
    assert sim.get_amplitude([0] * 5, [4, 0, 1, 2, 3]) == pytest.approx(1.0)
    # a 6-qubit gate cannot run (Run(): cannot apply 6 qubits gate)
    sim.apply_controlled_gate(np.eye(64), [0, 1, 2, 3, 4, 5], [])
    with pytest.raises(RuntimeError):
        sim.run()


def probability(make):
    """test_simulator_probability (:308-339)."""
    sim = make(1, 33, 4)
    q = list(range(6))
    sim.allocate_qureg(q, 0)
    for i in q:
        gate(sim, H, [i])
    bits = [0, 0, 1, 0, 1, 0]
    for i in range(6):
        assert sim.get_probability(bits[:i], q[:i]) == pytest.approx(0.5 ** i)
    with pytest.raises(RuntimeError):
        sim.get_probability([0], [99])
    with pytest.raises(RuntimeError):
        sim.get_probability([0, 1], [0])
    for i in q:
        gate(sim, H, [i])
    gate(sim, Ry(2 * math.acos(math.sqrt(0.3))), [0])
    assert sim.get_probability([0], [0]) == pytest.approx(0.3)
    gate(sim, Ry(2 * math.acos(math.sqrt(0.4))), [2])
    assert sim.get_probability([0], [2]) == pytest.approx(0.4)
    assert sim.get_probability([0, 0], [0, 2]) == pytest.approx(0.12)
    assert sim.get_probability([0, 1], [0, 2]) == pytest.approx(0.18)
    assert sim.get_probability([1, 0], [0, 2]) == pytest.approx(0.28)


def amplitude(make):
    """test_simulator_amplitude (:342-379)."""
    sim = make(1, 33, 4)
    q = list(range(6))
    sim.allocate_qureg(q, 0)
    for i in q:
        gate(sim, X, [i])
    for i in q:
        gate(sim, H, [i])
    assert sim.get_amplitude([0, 0, 1, 0, 1, 0], q) == pytest.approx(1.0 / 8.0)
    assert sim.get_amplitude([0, 0, 0, 0, 1, 0], q) == pytest.approx(-1.0 / 8.0)
    assert sim.get_amplitude([0, 1, 1, 0, 1, 0], q) == pytest.approx(-1.0 / 8.0)
    for i in q:
        gate(sim, H, [i])
    for i in q:
        gate(sim, X, [i])
    gate(sim, Ry(2 * math.acos(0.3)), [0])
    bits = [0] * 6
    assert sim.get_amplitude(bits, q) == pytest.approx(0.3)
    bits[0] = 1
    assert sim.get_amplitude(bits, q) == pytest.approx(math.sqrt(0.91))
    with pytest.raises(RuntimeError):
        sim.get_amplitude(bits, q[:-1])
    with pytest.raises(RuntimeError):
        sim.get_amplitude(bits, q[:-1] + [q[0]])
    sim.allocate_qubit(6)
    with pytest.raises(RuntimeError):
        sim.get_amplitude(bits, q)


def collapse(make):
    """test_simulator_collapse_wavefunction (:569-603)."""
    sim = make(1, 33, 4)
    q = [0, 1, 2, 3]
    with pytest.raises(RuntimeError):
        sim.collapse_wavefunction(q, [0] * 4)
    sim.allocate_qureg(q, 0)
    sim.collapse_wavefunction(q, [0] * 4)
    assert sim.get_probability([0] * 4, q) == pytest.approx(1.0)
    for i in q[1:]:
        gate(sim, H, [i])
    assert sim.get_probability([0] * 4, q) == pytest.approx(0.125)
    with pytest.raises(RuntimeError):
        sim.collapse_wavefunction(q, [1] + [0] * 3)
    with pytest.raises(RuntimeError):
        sim.collapse_wavefunction(q, [0, 0])
    sim.collapse_wavefunction(q[:-1], [0, 1, 0])
    assert sim.get_probability([0, 1, 0, 1], q) == pytest.approx(0.5)
    for i in q:
        sim.measure_qubits([i])
    for i in reversed(q):
        sim.deallocate_qubit(i)
    q = [4, 5, 6, 7]
    sim.allocate_qureg(q, 0)
    gate(sim, H, [4])
    gate(sim, X, [5], [4])
    sim.collapse_wavefunction([4], [1])
    assert sim.get_probability([1, 1], [4, 5]) == pytest.approx(1.0)


def no_uncompute(make):
    """test_simulator_no_uncompute_exception (:606-614)."""
    sim = make(1, 33, 4)
    sim.allocate_qubit(0)
    gate(sim, H, [0])
    with pytest.raises(RuntimeError):
        sim.deallocate_qubit(0)
    with pytest.raises(RuntimeError):
        sim.deallocate_qubit(17)


def entangle(make):
    """test_simulator_functional_entangle (:652-706), incl. the 4-control X
    that exceeds the cluster size and runs on the control-mask path."""
    sim = make(1, 33, 4)
    q = [0, 1, 2, 3, 4]
    sim.allocate_qureg(q, 0)
    gate(sim, H, [0])
    for b in q[1:]:
        gate(sim, X, [b], [0])
    vec = sim.cheat_local()[1]
    assert abs(vec[0]) ** 2 == pytest.approx(0.5)
    assert abs(vec[31]) ** 2 == pytest.approx(0.5)
    assert np.allclose(vec[1:31], 0)
    for b in q[2:]:
        gate(sim, X, [b], [0])
    for b in q[2:]:
        gate(sim, X, [b], [0, 1])
    vec = sim.cheat_local()[1]
    assert abs(vec[0]) ** 2 == pytest.approx(0.5)
    assert abs(vec[31]) ** 2 == pytest.approx(0.5)
    assert np.allclose(vec[1:31], 0)
    gate(sim, X, [4], [0, 1, 2, 3])
    gate(sim, X, [3], [0, 1, 2])
    gate(sim, X, [2], [0, 1])
    gate(sim, X, [1], [0])
    gate(sim, H, [0])
    vec = sim.cheat_local()[1]
    assert abs(vec[0]) ** 2 == pytest.approx(1.0)
    assert np.allclose(vec[1:], 0)


def ghz_measure(make, seed):
    """test_simulator_functional_measurement (:190-201): one call per qubit."""
    sim = make(seed, 33, 4)
    q = [0, 1, 2, 3, 4]
    sim.allocate_qureg(q, 0)
    gate(sim, H, [0])
    for b in q[1:]:
        gate(sim, X, [b], [0])
    res = [sim.measure_qubits([b])[0] for b in q]
    assert sum(res) in (0, 5)
    return res


def measure_mapped(make):
    """test_simulator_measure_mapped_qubit (:204-220) and cheat (:161-187)."""
    sim = make(1, 33, 4)
    ids, vec = sim.cheat_local()
    assert len(ids) == 0 and len(vec) == 1
    sim.allocate_qubit(1)
    ids, vec = sim.cheat_local()
    assert dict(ids) == {1: 0} and len(vec) == 2 and abs(vec[0]) == pytest.approx(1.0)
    gate(sim, X, [1])
    assert sim.measure_qubits([1]) == [True]
    gate(sim, X, [1])
    sim.deallocate_qubit(1)
    ids, vec = sim.cheat_local()
    assert len(ids) == 0 and len(vec) == 1


def expectation(make):
    """Commented-out test_simulator_expectation (:382-425)."""
    sim = make(1, 33, 4)
    q = [0, 1, 2]
    sim.allocate_qureg(q, 0)
    ev = sim.get_expectation_value
    assert ev([([(0, "Z")], 1.0)], q) == pytest.approx(1.0)
    gate(sim, X, [0])
    assert ev([([(0, "Z")], 1.0)], q) == pytest.approx(-1.0)
    gate(sim, H, [0])
    assert ev([([(0, "X")], 1.0)], q) == pytest.approx(-1.0)
    gate(sim, Z, [0])
    assert ev([([(0, "X")], 1.0)], q) == pytest.approx(1.0)
    for m in (X, S, Z, X):
        gate(sim, m, [0])
    assert ev([([(0, "Y")], 1.0)], q) == pytest.approx(1.0)
    gate(sim, Z, [0])
    assert ev([([(0, "Y")], 1.0)], q) == pytest.approx(-1.0)
    op_sum = [([(0, "Y"), (1, "X"), (2, "Z")], 1.0), ([(1, "X")], 1.0)]
    gate(sim, H, [1])
    gate(sim, X, [2])
    assert ev(op_sum, q) == pytest.approx(2.0)
    gate(sim, X, [2])
    assert ev(op_sum, q) == pytest.approx(0.0)
    assert ev([([], 0.4)], q) == pytest.approx(0.4)
    with pytest.raises(RuntimeError):
        ev([([(3, "Z")], 1.0)], q)


def apply_qubit_operator(make):
    """Commented-out test_simulator_applyqubitoperator (:441-478)."""
    sim = make(1, 33, 4)
    q = [0, 1, 2]
    sim.allocate_qureg(q, 0)
    sim.apply_qubit_operator([([(0, "X"), (1, "Y"), (2, "Z")], 1.0)], q)
    gate(sim, X, [0])
    gate(sim, Y, [1])
    gate(sim, Z, [2])
    assert sim.get_amplitude([0, 0, 0], q) == pytest.approx(1.0)
    gate(sim, H, [0])
    h = 1.0 / math.sqrt(2.0)
    sim.apply_qubit_operator([([(0, "X")], h), ([(0, "Z")], h)], [0])
    assert sim.get_amplitude([0, 0, 0], q) == pytest.approx(1.0)
    gate(sim, H, [0])
    sim.apply_qubit_operator([([], 0.5), ([(0, "Z")], 0.5)], [0])
    assert sim.get_amplitude([0, 0, 0], q) == pytest.approx(h)
    sim.apply_qubit_operator([([], 0.5), ([(0, "Z")], -0.5)], [0])
    assert sim.get_amplitude([0, 0, 0], q) == pytest.approx(0.0)


def set_wavefunction(make):
    """Commented-out test_simulator_set_wavefunction (:540-566)."""
    sim = make(1, 33, 4)
    sim.allocate_qureg([0, 1], 0)
    wf = [0.0, 0.0, math.sqrt(0.2), math.sqrt(0.8)]
    sim.set_wavefunction(wf, [0, 1])
    assert sim.get_probability([1], [0]) == pytest.approx(0.8)
    assert sim.get_probability([0, 1], [0, 1]) == pytest.approx(0.2)
    assert sim.get_probability([1], [1]) == pytest.approx(1.0)
    sim2 = make(1, 33, 4)
    sim2.allocate_qubit(0)
    sim2.set_wavefunction([1.0, 0.0], [0])
    gate(sim2, Y, [0])
    assert sim2.get_amplitude([1], [0]) == pytest.approx(1j)


def allocation_limits(make):
    """AllocateQubit beyond max(max_local, cluster) raises
    (SimulatorMPI.cpp:203-212); init only for the first qureg (:223-230)."""
    sim = make(1, 3, 2)
    sim.allocate_qureg([0, 1, 2], 0.5)
    assert np.allclose(sim.cheat_local()[1], 0.5)
    with pytest.raises(RuntimeError):
        sim.allocate_qubit(3)
    with pytest.raises(RuntimeError):
        sim.allocate_qureg([7], 0.25)


ALL_KNOWN_ANSWERS = [kqubit_gate, probability, amplitude, collapse, no_uncompute, entangle,
                     measure_mapped, expectation, apply_qubit_operator, set_wavefunction,
                     allocation_limits]
