class LogicalQubitIDTag:
    def __init__(self, logical_qubit_id):
        self.logical_qubit_id = logical_qubit_id


class DirtyQubitTag:
    pass


def get_control_count(cmd):
    return len(cmd.control_qubits)
