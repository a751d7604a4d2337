"""Compiler engines of the HiQ front end over the MI355X engine.

GreedyScheduler follows the staging semantics of the reference
(hiq/projectq/cengines/_greedyscheduler.py:31-265): commands are cached;
deallocations are deferred and sent (highest id first) after the next
scheduling; an allocation or a fast-forwarding gate (Measure, Flush,
MetaSwap) forces scheduling.  Scheduling picks the local qubit set once
(SwapScheduler -> set_qubits_perm), then repeatedly asks the
ClusterScheduler for the next cluster and sends it followed by a FlushGate
(one fused pass); whatever cannot run locally triggers a MetaSwap to the
next stage's local set.  Every gate is scheduled as non-diagonal, as in the
reference (which forces BasicGate.is_diagonal to False).  HiQMainEngine adds
the register-at-once allocation (_hiq_main_engine.py:22-81).
"""
from projectq import MainEngine
from projectq.cengines import BasicEngine
from projectq.ops import (AllocateQubitGate, Command, Deallocate, DeallocateQubitGate,
                          FastForwardingGate, FlushGate, ZGate)
from projectq.types import BasicQubit, Qubit, Qureg, WeakQubitRef

from .. import ClusterScheduler, SwapScheduler
from .ops import AllocateQuregGate, MetaSwap


class GreedyScheduler(BasicEngine):
    def __init__(self, supremacy_circuit=False, num_splits=10 ** 6, cluster_size=4):
        BasicEngine.__init__(self)
        self._cmds = []
        self._deallocs = []
        self._staged = False
        self._supremacy = supremacy_circuit
        self.NUM_SPLITS = num_splits
        self.CLUSTER_SIZE = cluster_size

    # -- backend queries ----------------------------------------------------
    def _backend(self):
        return self.main_engine.backend

    def _all_ids(self):
        b = self._backend()
        return list(b.qubits_ids) if hasattr(b, "qubits_ids") else list(b.get_qubits_ids())

    def _local_ids(self):
        return list(self._backend().get_local_qubits_ids())

    def _global_ids(self):
        return list(self._backend().get_global_qubits_ids())

    # -- gate lists -----------------------------------------------------------
    def _gate_lists(self):
        targets = [[qb.id for qr in c.qubits for qb in qr] for c in self._cmds]
        ctrls = [[qb.id for qb in c.control_qubits] for c in self._cmds]
        return targets, ctrls, [False] * len(self._cmds)

    def _check_commands(self):
        n_local = len(self._local_ids())
        for t in self._gate_lists()[0]:
            if len(t) > n_local:
                raise Exception("Can't apply {}-qubits gate (only {} local qubits)"
                                .format(len(t), n_local))
            if len(t) > 5:
                raise Exception("Can't apply {}-qubits gate (no more that 5 qubits allowed)"
                                .format(len(t)))

    def _remove_ending_cz(self):
        # A CZ whose qubits are all untouched afterwards cannot change
        # measurement statistics (supremacy circuits).
        used = set()
        keep = []
        for cmd in reversed(self._cmds):
            ids = [qb.id for qr in cmd.all_qubits for qb in qr]
            if isinstance(cmd.gate, ZGate) and not any(i in used for i in ids):
                continue
            used.update(ids)
            keep.append(cmd)
        self._cmds = list(reversed(keep))

    def _prepare_ctrlz(self):
        # A controlled Z is symmetric: put a global target on a local control.
        local, glob = set(self._local_ids()), set(self._global_ids())
        for cmd in self._cmds:
            if not isinstance(cmd.gate, ZGate):
                continue
            ctrls = list(cmd.control_qubits)
            target = list(cmd.qubits[0])
            if target[0].id in glob:
                for i, c in enumerate(ctrls):
                    if c.id in local:
                        ctrls[i], target[0] = target[0], ctrls[i]
                        break
            cmd.qubits = (target,)
            cmd.control_qubits = ctrls

    def _next_locals(self):
        local = self._local_ids()
        t, c, d = self._gate_lists()
        new = SwapScheduler(t, c, d, self.NUM_SPLITS, len(local), True).ScheduleSwap()
        if not new:
            new = SwapScheduler(t, c, d, self.NUM_SPLITS, len(local), False).ScheduleSwap()
        g_to_l = sorted(set(new) - set(local))
        l_to_g = sorted(set(local) - set(new))[:len(g_to_l)]
        return g_to_l, l_to_g

    def _run_clusters(self):
        self._prepare_ctrlz()
        local, glob = self._local_ids(), self._global_ids()
        while self._cmds:
            t, c, d = self._gate_lists()
            avail = ClusterScheduler(t, c, d, local, glob, self.CLUSTER_SIZE).ScheduleCluster()
            if not avail:
                return
            for i in avail:
                self.send([self._cmds[i]])
            self.send([Command(self, FlushGate(), ([WeakQubitRef(self, -1)],))])
            taken = set(avail)
            self._cmds = [cmd for i, cmd in enumerate(self._cmds) if i not in taken]

    def _force_scheduling(self):
        if not self._cmds:
            return
        self._check_commands()
        if self._supremacy:
            self._remove_ending_cz()
        if not self._staged:
            self._staged = True
            ids = self._all_ids()
            g_to_l, l_to_g = self._next_locals()
            for a, b in zip(g_to_l, l_to_g):
                p, q = ids.index(a), ids.index(b)
                ids[p], ids[q] = ids[q], ids[p]
            self._backend().set_qubits_perm(ids)
        self._run_clusters()
        while self._cmds:
            g_to_l, l_to_g = self._next_locals()
            assert g_to_l, "no stage can make progress"
            pairs = []
            for g, l in zip(g_to_l, l_to_g):
                pairs += [BasicQubit(self.main_engine, g), BasicQubit(self.main_engine, l)]
            self.send([MetaSwap.generate_command(pairs)])
            self._run_clusters()

    def _send_deallocations(self):
        for cmd in sorted(self._deallocs, key=lambda c: c.qubits[0][0].id, reverse=True):
            self.send([cmd])
        self._deallocs = []

    def receive(self, command_list):
        for cmd in command_list:
            if isinstance(cmd.gate, DeallocateQubitGate):
                self._deallocs.append(cmd)
            elif isinstance(cmd.gate, (AllocateQubitGate, AllocateQuregGate, FastForwardingGate)):
                self._force_scheduling()
                self._send_deallocations()
                self.send([cmd])
            else:
                if self._deallocs:
                    self._force_scheduling()
                    self._send_deallocations()
                self._cmds.append(cmd)


class HiQMainEngine(MainEngine):
    """MainEngine with register-at-once allocation (AllocateQuregGate)."""

    def allocate_qureg(self, n, init=0.0):
        cmd = Command(self, AllocateQuregGate(init), ())
        if not self.backend.is_available(cmd):
            return super(HiQMainEngine, self).allocate_qureg(n)
        qubits = Qureg()
        for _ in range(n):
            qb = Qubit(self, self.main_engine.get_new_qubit_id())
            qubits.append(qb)
            self.main_engine.active_qubits.add(qb)
        self.send([Command(self, AllocateQuregGate(init), (qubits,))])
        return qubits

    def deallocate_qubit(self, qubit):
        if qubit.id == -1:
            raise ValueError("Already deallocated.")
        is_dirty = qubit.id in self.dirty_qubits
        weak = WeakQubitRef(qubit.engine, qubit.id)
        if not isinstance(qubit, WeakQubitRef) and qubit in self.active_qubits:
            self.active_qubits.remove(qubit)
            qubit.id = -1
        from projectq.meta import DirtyQubitTag
        self.send([Command(self, Deallocate, (Qureg([weak]),),
                           tags=[DirtyQubitTag()] if is_dirty else [])])
