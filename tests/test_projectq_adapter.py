"""The ProjectQ front end (hiqsimulator_amd.projectq): HiQMainEngine ->
GreedyScheduler -> SimulatorMPI, restating flows of the reference test suite
(_simulator_mpi_test.py).  Without ProjectQ installed, a minimal stand-in of
its engine interface (tests/fake_projectq) drives the adapters; the native
simulator is the oracle on the CPU and the MI355X engine under -m gpu."""
import math
import os
import sys

import numpy as np
import pytest

try:
    import projectq  # noqa: F401
    REAL_PROJECTQ = True
except ImportError:
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "fake_projectq"))
    REAL_PROJECTQ = False


def make_backend(kind, seed=1, gate_fusion=True):
    from projectq.cengines import BasicEngine

    from hiqsimulator_amd.projectq.backends import SimulatorMPI
    if kind == "engine":
        return SimulatorMPI(gate_fusion=gate_fusion, rnd_seed=seed)
    import oracle
    sim = SimulatorMPI.__new__(SimulatorMPI)
    BasicEngine.__init__(sim)
    sim._simulator = oracle.OracleSim(seed, 33, 4)
    sim._gate_fusion = gate_fusion
    return sim


BACKENDS = [pytest.param("oracle"), pytest.param("engine", marks=pytest.mark.gpu)]


def cnot(eng, c, t):
    from projectq.ops import Command, X
    eng.receive([Command(eng, X, ([t],), controls=[c])])


@pytest.mark.parametrize("kind", BACKENDS)
def test_ghz_measurement(kind):
    from projectq.ops import H, Measure

    from hiqsimulator_amd.projectq import GreedyScheduler, HiQMainEngine
    sums = set()
    for seed in range(8):
        sim = make_backend(kind, seed)
        eng = HiQMainEngine(sim, [GreedyScheduler()])
        q = eng.allocate_qureg(5)
        H | q[0]
        for b in q[1:]:
            cnot(eng, q[0], b)
        for b in q:
            Measure | b
        sums.add(sum(int(b) for b in q))
    assert sums <= {0, 5} and len(sums) == 2


@pytest.mark.parametrize("kind", BACKENDS)
def test_probability_amplitude_flow(kind):
    from projectq.ops import H, Measure, X

    from hiqsimulator_amd.projectq import GreedyScheduler, HiQMainEngine
    sim = make_backend(kind)
    eng = HiQMainEngine(sim, [GreedyScheduler()])
    q = eng.allocate_qureg(6)
    for b in q:
        X | b
    for b in q:
        H | b
    eng.flush()
    assert eng.backend.get_amplitude([0, 0, 1, 0, 1, 0], q) == pytest.approx(1.0 / 8.0)
    assert eng.backend.get_amplitude("000010", q) == pytest.approx(-1.0 / 8.0)
    for i in range(6):
        assert eng.backend.get_probability([1] * i, q[:i]) == pytest.approx(0.5 ** i)
    assert eng.backend.cheat()[0] == {b.id: b.id for b in q}
    for b in q:
        Measure | b


@pytest.mark.parametrize("kind", BACKENDS)
def test_deallocation_and_reallocation(kind):
    from projectq.ops import H, Measure

    from hiqsimulator_amd.projectq import GreedyScheduler, HiQMainEngine
    sim = make_backend(kind)
    eng = HiQMainEngine(sim, [GreedyScheduler()])
    q = eng.allocate_qureg(3)
    H | q[1]
    Measure | q[1]
    v = int(q[1])
    eng.deallocate_qubit(q[1])
    q2 = eng.allocate_qubit()
    eng.flush()
    ids = sim.get_qubits_ids()
    assert q2[0].id in ids and len(ids) == 3 and v in (0, 1)
    H | q[0]
    eng.flush()
    with pytest.raises(RuntimeError):
        eng.deallocate_qubit(q[0])
        eng.flush()


@pytest.mark.parametrize("kind", BACKENDS)
def test_controlled_kqubit_gate(kind):
    """test_simulator_kqubit_gate through the engines (controlled 4-qubit gate)."""
    from projectq.ops import Command, MatrixGate, X

    from hiqsimulator_amd.circuits import Rx, Ry, Rz
    from hiqsimulator_amd.projectq import GreedyScheduler, HiQMainEngine
    m = np.kron(Rz(0.9).dot(Ry(-0.1)), np.kron(Ry(0.1), np.kron(Rx(0.8), Rx(0.3))))
    sim = make_backend(kind)
    eng = HiQMainEngine(sim, [GreedyScheduler()])
    qureg = eng.allocate_qureg(4)
    qubit = eng.allocate_qubit()
    for g, b in ((Rx(-0.3), 0), (Rx(-0.8), 1), (Ry(-0.1), 2), (Rz(-0.9), 3), (Ry(0.1), 3)):
        MatrixGate(g) | qureg[b]
    X | qubit
    eng.receive([Command(eng, MatrixGate(m), (list(qureg),), controls=[qubit[0]])])
    X | qubit
    eng.receive([Command(eng, MatrixGate(m.conj().T), (list(qureg),), controls=[qubit[0]])])
    eng.flush()
    assert eng.backend.get_amplitude("0" * 5, qubit + qureg) == pytest.approx(1.0)


def test_is_available_rules():
    from projectq.ops import Allocate, BasicMathGate, Command, MatrixGate, Measure, TimeEvolution

    from hiqsimulator_amd.projectq import AllocateQuregGate
    sim = make_backend("oracle")
    assert sim.is_available(Command(None, Measure, ([],)))
    assert sim.is_available(Command(None, Allocate, ([],)))
    assert sim.is_available(Command(None, AllocateQuregGate(0), ([],)))
    assert not sim.is_available(Command(None, BasicMathGate(), ([],)))
    assert not sim.is_available(Command(None, TimeEvolution(), ([],)))
    assert sim.is_available(Command(None, MatrixGate(np.eye(32)), ([],)))
    assert not sim.is_available(Command(None, MatrixGate(np.eye(64)), ([],)))


def test_resource_counter_counts_qureg_allocation():
    """A ResourceCounter in a HiQ engine list counts an AllocateQuregGate as
    one allocation per qubit (the reference's ResourceCounter hook,
    hiq/projectq/ops/_gates.py:43-73): active qubits, widest point and the
    gate count; later gates on those qubits find their depth entries."""
    from projectq.backends import ResourceCounter
    from projectq.ops import H, Measure

    from hiqsimulator_amd.projectq import AllocateQuregGate, GreedyScheduler, HiQMainEngine
    counter = ResourceCounter()
    eng = HiQMainEngine(make_backend("oracle"), [counter, GreedyScheduler()])
    qureg = eng.allocate_qureg(5)
    assert counter._active_qubits == 5 and counter.max_width == 5
    assert sorted(counter._depth_of_qubit) == sorted(q.id for q in qureg)
    assert counter.gate_counts[(AllocateQuregGate(0), 0)] == 1
    assert counter.gate_class_counts[(AllocateQuregGate, 0)] == 1
    qubit = eng.allocate_qubit()
    assert counter._active_qubits == 6 and counter.max_width == 6
    H | qureg[2]
    assert counter._depth_of_qubit[qureg[2].id] == 1
    Measure | qubit
    eng.flush()
    eng.deallocate_qubit(qubit[0])
    assert counter._active_qubits == 5 and counter.max_width == 6


def test_resource_counter_without_hook_miscounts():
    """The counter's own _add_cmd (the hook bypassed) cannot count a register
    allocation: it finds no depth entry for the register's qubits."""
    from projectq.backends import ResourceCounter

    from hiqsimulator_amd.projectq import AllocateQuregGate
    from projectq.ops import Command
    from projectq.types import Qubit
    counter = ResourceCounter()
    cmd = Command(None, AllocateQuregGate(0), ([Qubit(None, 0), Qubit(None, 1)],))
    with pytest.raises(KeyError):
        ResourceCounter._hiq_own_add_cmd(counter, cmd)
    counter._add_cmd(cmd)
    assert counter._active_qubits == 2
