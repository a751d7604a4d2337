"""ProjectQ backend engine over the MI355X state-vector engine.

Same command handling as the reference's SimulatorMPI engine
(hiq/projectq/backends/_sim/_simulator_mpi.py:47-513): gates are handed to
the native engine, which fuses them until a FlushGate / FastForwardingGate
(or, without gate fusion, after every gate) triggers run().  No MPI: the
engine runs on one GPU, so the thread-level checks and the Allgather in
cheat() are gone.
"""
import math
import random

import numpy
from projectq.cengines import BasicEngine
from projectq.meta import LogicalQubitIDTag, get_control_count
from projectq.ops import (Allocate, BasicMathGate, Deallocate, FastForwardingGate, FlushGate,
                          Measure, TimeEvolution)
from projectq.types import WeakQubitRef

from .. import SimulatorMPI as _Backend
from .ops import AllocateQuregGate, MetaSwap


def _pauli_terms(qubit_operator, num_qubits):
    for term, _ in qubit_operator.terms.items():
        if term != () and term[-1][0] >= num_qubits:
            raise Exception("qubit_operator acts on more qubits than contained in the qureg.")
    return [(list(term), coeff) for term, coeff in qubit_operator.terms.items()]


class SimulatorMPI(BasicEngine):
    """ProjectQ backend running on one MI355X (see the module docstring).

    Args mirror the reference (gate_fusion, rnd_seed, num_local_qubits,
    max_fused_qubits) plus ``precision`` ("c128" or "c64") and ``device``.
    """

    def __init__(self, gate_fusion=False, rnd_seed=None, num_local_qubits=33,
                 max_fused_qubits=4, precision="c128", device=0):
        BasicEngine.__init__(self)
        if rnd_seed is None:
            rnd_seed = random.randint(0, 4294967295)
        self._simulator = _Backend(rnd_seed, num_local_qubits, max_fused_qubits, precision,
                                   device)
        self._gate_fusion = gate_fusion

    def is_available(self, cmd):
        if (cmd.gate == Measure or cmd.gate == Allocate or cmd.gate == Deallocate
                or isinstance(cmd.gate, AllocateQuregGate)):
            return True
        if isinstance(cmd.gate, (BasicMathGate, TimeEvolution)):
            return False
        try:
            return len(cmd.gate.matrix) <= 2 ** 5
        except Exception:
            return False

    def _convert_logical_to_mapped_qureg(self, qureg):
        mapper = self.main_engine.mapper
        if mapper is None:
            return qureg
        mapped = []
        for qubit in qureg:
            if qubit.id not in mapper.current_mapping:
                raise RuntimeError("Unknown qubit id. Please make sure you have called eng.flush().")
            mapped.append(WeakQubitRef(qubit.engine, mapper.current_mapping[qubit.id]))
        return mapped

    def get_expectation_value(self, qubit_operator, qureg):
        qureg = self._convert_logical_to_mapped_qureg(qureg)
        ops = _pauli_terms(qubit_operator, len(qureg))
        return self._simulator.get_expectation_value(ops, [qb.id for qb in qureg])

    def apply_qubit_operator(self, qubit_operator, qureg):
        qureg = self._convert_logical_to_mapped_qureg(qureg)
        ops = _pauli_terms(qubit_operator, len(qureg))
        return self._simulator.apply_qubit_operator(ops, [qb.id for qb in qureg])

    def get_probability(self, bit_string, qureg):
        qureg = self._convert_logical_to_mapped_qureg(qureg)
        bits = [bool(int(b)) for b in bit_string]
        return self._simulator.get_probability(bits, [qb.id for qb in qureg])

    def get_amplitude(self, bit_string, qureg):
        qureg = self._convert_logical_to_mapped_qureg(qureg)
        bits = [bool(int(b)) for b in bit_string]
        return self._simulator.get_amplitude(bits, [qb.id for qb in qureg])

    def set_wavefunction(self, wavefunction, qureg):
        qureg = self._convert_logical_to_mapped_qureg(qureg)
        self._simulator.set_wavefunction(numpy.asarray(wavefunction, dtype=complex),
                                         [qb.id for qb in qureg])

    def collapse_wavefunction(self, qureg, values):
        qureg = self._convert_logical_to_mapped_qureg(qureg)
        return self._simulator.collapse_wavefunction([qb.id for qb in qureg],
                                                     [bool(int(v)) for v in values])

    def cheat_local(self):
        return self._simulator.cheat_local()

    def cheat(self):
        """(id -> bit position, full state vector); one device holds it all."""
        return self._simulator.cheat_local()

    def get_qubits_ids(self):
        return self._simulator.get_qubits_ids()

    def get_local_qubits_ids(self):
        return self._simulator.get_local_qubits_ids()

    def get_global_qubits_ids(self):
        return self._simulator.get_global_qubits_ids()

    def set_qubits_perm(self, ids):
        self._simulator.set_qubits_perm(ids)

    def _do_swap(self, qubits):
        self._simulator.swap_qubits(qubits)

    def _handle(self, cmd):
        gate = cmd.gate
        if isinstance(gate, FlushGate):
            return
        if gate == MetaSwap:
            self._do_swap([qb.id for qr in cmd.all_qubits for qb in qr])
        elif gate == Measure:
            assert get_control_count(cmd) == 0
            ids = [qb.id for qr in cmd.qubits for qb in qr]
            out = self._simulator.measure_qubits(ids)
            tag = None
            for t in cmd.tags:
                if isinstance(t, LogicalQubitIDTag):
                    tag = t
            i = 0
            for qr in cmd.qubits:
                for qb in qr:
                    if tag is not None:
                        qb = WeakQubitRef(qb.engine, tag.logical_qubit_id)
                    self.main_engine.set_measurement_result(qb, out[i])
                    i += 1
        elif gate == Allocate:
            self._simulator.allocate_qubit(cmd.qubits[0][0].id)
        elif isinstance(gate, AllocateQuregGate):
            self._simulator.allocate_qureg([qb.id for qr in cmd.qubits for qb in qr], gate.init)
        elif gate == Deallocate:
            self._simulator.deallocate_qubit(cmd.qubits[0][0].id)
        elif isinstance(gate, BasicMathGate):
            self._simulator.emulate_math(gate.get_math_function(cmd.qubits),
                                         [[qb.id for qb in qr] for qr in cmd.qubits],
                                         [qb.id for qb in cmd.control_qubits])
        elif isinstance(gate, TimeEvolution):
            # _simulator_mpi.py:469-475 (is_available still declines the
            # gate, as the reference does, so ProjectQ decomposes it unless
            # a caller sends it here directly)
            op = [(list(term), coeff) for (term, coeff) in gate.hamiltonian.terms.items()]
            self._simulator.emulate_time_evolution(op, gate.time, [qb.id for qb in cmd.qubits[0]],
                                                   [qb.id for qb in cmd.control_qubits])
        elif len(gate.matrix) <= 2 ** 5:
            ids = [qb.id for qr in cmd.qubits for qb in qr]
            matrix = numpy.asarray(gate.matrix, dtype=complex)
            if 2 ** len(ids) != matrix.shape[0]:
                raise Exception("Simulator: Error applying {} gate: {}-qubit gate applied to {} "
                                "qubits.".format(str(gate), int(math.log(matrix.shape[0], 2)),
                                                 len(ids)))
            self._simulator.apply_controlled_gate(matrix, ids,
                                                  [qb.id for qb in cmd.control_qubits])
            if not self._gate_fusion:
                self._simulator.run()
        else:
            raise Exception("This simulator only supports controlled k-qubit gates with k < 6!\n"
                            "Please add an auto-replacer engine to your list of compiler engines.")

    def receive(self, command_list):
        for cmd in command_list:
            if isinstance(cmd.gate, (FlushGate, FastForwardingGate)):
                self._simulator.run()
            self._handle(cmd)
            if not self.is_last_engine:
                self.send([cmd])
