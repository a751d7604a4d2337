"""Stand-in for ProjectQ's ResourceCounter (projectq/backends/_resource.py):
the bookkeeping its _add_cmd keeps for allocations, deallocations and other
gates (active qubits, widest point, depth per qubit, gate and gate-class
counts).  Commands other than flushes are counted; all are passed on."""
from .cengines import BasicEngine
from .meta import get_control_count
from .ops import Allocate, Deallocate, FlushGate


class ResourceCounter(BasicEngine):
    def __init__(self):
        BasicEngine.__init__(self)
        self.gate_counts = {}
        self.gate_class_counts = {}
        self._active_qubits = 0
        self.max_width = 0
        self._depth_of_qubit = {}
        self._previous_max_depth = 0

    def is_available(self, cmd):
        if self.is_last_engine:
            return True
        return self.next_engine.is_available(cmd)

    @property
    def depth_of_dag(self):
        current = max(self._depth_of_qubit.values()) if self._depth_of_qubit else 0
        return max(current, self._previous_max_depth)

    def _add_cmd(self, cmd):
        if cmd.gate == Allocate:
            self._active_qubits += 1
            self._depth_of_qubit[cmd.qubits[0][0].id] = 0
        elif cmd.gate == Deallocate:
            self._active_qubits -= 1
            depth = self._depth_of_qubit.pop(cmd.qubits[0][0].id)
            self._previous_max_depth = max(self._previous_max_depth, depth)
        else:
            ids = set()
            for qureg in cmd.all_qubits:
                for qubit in qureg:
                    self._depth_of_qubit[qubit.id] += 1
                    ids.add(qubit.id)
            if len(ids) > 1:
                top = max(self._depth_of_qubit[i] for i in ids)
                for i in ids:
                    self._depth_of_qubit[i] = top
        self.max_width = max(self.max_width, self._active_qubits)
        nctrl = get_control_count(cmd)
        for table, key in ((self.gate_counts, (cmd.gate, nctrl)),
                           (self.gate_class_counts, (cmd.gate.__class__, nctrl))):
            table[key] = table.get(key, 0) + 1

    def receive(self, command_list):
        for cmd in command_list:
            if not isinstance(cmd.gate, FlushGate):
                self._add_cmd(cmd)
            if not self.is_last_engine:
                self.send([cmd])
