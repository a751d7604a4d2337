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
