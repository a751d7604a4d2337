from .ops import Allocate, Command, FlushGate
from .types import Qubit, WeakQubitRef


class BasicEngine:
    def __init__(self):
        self.main_engine = None
        self.next_engine = None
        self.is_last_engine = False

    def send(self, command_list):
        self.next_engine.receive(command_list)


class MainEngine(BasicEngine):
    def __init__(self, backend, engine_list=()):
        BasicEngine.__init__(self)
        self.main_engine = self
        self.backend = backend
        self.mapper = None
        self.active_qubits = set()
        self.dirty_qubits = set()
        self._next_id = 0
        self._results = {}
        chain = list(engine_list) + [backend]
        self.next_engine = chain[0]
        for a, b in zip(chain, chain[1:]):
            a.next_engine = b
        for e in chain:
            e.main_engine = self
        backend.is_last_engine = True

    def get_new_qubit_id(self):
        self._next_id += 1
        return self._next_id - 1

    def allocate_qubit(self):
        qb = Qubit(self, self.get_new_qubit_id())
        self.active_qubits.add(qb)
        self.send([Command(self, Allocate, ([WeakQubitRef(self, qb.id)],))])
        return [qb]

    def receive(self, command_list):
        self.send(command_list)

    def set_measurement_result(self, qubit, value):
        self._results[qubit.id] = bool(value)

    def get_measurement_result(self, qubit):
        return self._results[qubit.id]

    def flush(self):
        self.send([Command(self, FlushGate(), ([WeakQubitRef(self, -1)],))])
