"""Meta gates used between the HiQ engines (reference: hiq/projectq/ops/_gates.py:20-72)."""
from projectq.ops import ClassicalInstructionGate, FastForwardingGate


class MetaSwapGate(FastForwardingGate):
    """Command to the backend to exchange qubit positions (the reference's
    global<->local swap; on one device a data permutation)."""

    def __str__(self):
        return "SW"


MetaSwap = MetaSwapGate()


class AllocateQuregGate(ClassicalInstructionGate):
    """Allocates a whole register at once, optionally setting every amplitude
    to `init` (first register only)."""

    def __init__(self, init=0):
        ClassicalInstructionGate.__init__(self)
        self.interchangeable_qubit_indices = []
        self._init = init

    @property
    def init(self):
        return self._init

    def __str__(self):
        return "AllocateQureg"


def _count_qureg_allocation(counter, cmd):
    """ResourceCounter._add_cmd for an AllocateQuregGate (_gates.py:43-70):
    every qubit of the register counts as one active qubit starting at depth
    0; the widest point, the gate count and the gate-class count are kept as
    ProjectQ's counter keeps them for any other gate."""
    from projectq.meta import get_control_count
    for qureg in cmd.qubits:
        for qubit in qureg:
            counter._active_qubits += 1
            counter._depth_of_qubit[qubit.id] = 0
    counter.max_width = max(counter.max_width, counter._active_qubits)
    nctrl = get_control_count(cmd)
    for table, key in ((counter.gate_counts, (cmd.gate, nctrl)),
                       (counter.gate_class_counts, (cmd.gate.__class__, nctrl))):
        table[key] = table.get(key, 0) + 1


def install_resource_counter_hook():
    """Makes ProjectQ's ResourceCounter count an AllocateQuregGate as one
    allocation per qubit, as the reference does by patching
    ResourceCounter._add_cmd at import (_gates.py:43-73).  Other commands go
    to the counter's own _add_cmd.  Idempotent; False when ProjectQ has no
    ResourceCounter."""
    try:
        from projectq.backends import ResourceCounter
    except ImportError:
        return False
    if getattr(ResourceCounter, "_hiq_own_add_cmd", None) is None:
        own = ResourceCounter._add_cmd

        def _add_cmd(self, cmd):
            if isinstance(cmd.gate, AllocateQuregGate):
                _count_qureg_allocation(self, cmd)
            else:
                own(self, cmd)

        ResourceCounter._hiq_own_add_cmd = own
        ResourceCounter._add_cmd = _add_cmd
    return True


install_resource_counter_hook()
