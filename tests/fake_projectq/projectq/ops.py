import numpy as np


class BasicGate:
    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))

    def generate_command(self, qubits):
        eng = qubits[0].engine
        return Command(eng, self, ([q for q in qubits],))

    def __or__(self, qubits):
        if isinstance(qubits, list) and qubits and isinstance(qubits[0], list):
            qubits = tuple(qubits)
        elif not isinstance(qubits, tuple):
            qubits = (list(qubits) if isinstance(qubits, list) else [qubits],)
        eng = qubits[0][0].engine
        eng.main_engine.receive([Command(eng.main_engine, self, qubits)])


class MatrixGate(BasicGate):
    def __init__(self, matrix):
        self._m = np.asarray(matrix, dtype=complex)

    @property
    def matrix(self):
        return self._m.copy()

    def __eq__(self, other):
        return isinstance(other, MatrixGate) and np.array_equal(self._m, other._m)

    def __hash__(self):
        return 0


class ClassicalInstructionGate(BasicGate):
    pass


class FastForwardingGate(ClassicalInstructionGate):
    pass


class FlushGate(FastForwardingGate):
    pass


class MeasureGate(FastForwardingGate):
    pass


class AllocateQubitGate(ClassicalInstructionGate):
    pass


class DeallocateQubitGate(FastForwardingGate):
    pass


class BasicMathGate(BasicGate):
    pass


class TimeEvolution(BasicGate):
    pass


class ZGate(MatrixGate):
    def __init__(self):
        MatrixGate.__init__(self, [[1, 0], [0, -1]])


Allocate = AllocateQubitGate()
Deallocate = DeallocateQubitGate()
Measure = MeasureGate()
H = MatrixGate(np.array([[1, 1], [1, -1]]) / np.sqrt(2))
X = MatrixGate([[0, 1], [1, 0]])
Z = ZGate()


class Command:
    def __init__(self, engine, gate, qubits, controls=(), tags=()):
        self.engine = engine
        self.gate = gate
        self.qubits = tuple(list(q) for q in qubits)
        self.control_qubits = list(controls)
        self.tags = list(tags)

    @property
    def all_qubits(self):
        return (self.control_qubits,) + self.qubits
