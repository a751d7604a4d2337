"""Gate-list staging for the engine: the GreedyScheduler's planning loop.

The reference's GreedyScheduler (hiq/projectq/cengines/_greedyscheduler.py)
caches ProjectQ commands and, on a flush, repeatedly asks the C++
ClusterScheduler for the next cluster among the cached gates, sends those
gates followed by a FlushGate (one fused pass), and removes them
(_call_cluster_scheduler, :119-137).  With every qubit local (one device)
the SwapScheduler step leaves the qubit order unchanged (:212-222) and no
MetaSwap is ever needed (:224-242).  Gate diagonality is reported as False
for every gate, as the reference forces it (`BasicGate.is_diagonal`,
:28).

This module runs the same loop on plain gate lists ``(matrix, targets,
controls)`` so that planning can happen ahead of time (outside a timed
region) and independently of ProjectQ.
"""
from . import ClusterScheduler, _cppsim_amd

PLAN_MODES = {"reference": 0, "lookahead": 1}


def plan_clusters(gates, local_ids, cluster_size, mode="reference"):
    """Return the fused clusters as lists of gate indices, in execution order.

    Within a cluster the indices are ascending (program order); gates are
    only moved past gates they share no qubit with.  ``mode="reference"``
    runs the GreedyScheduler loop below; ``mode="lookahead"`` the native
    one-scan packer (same commutation rule, a few percent more clusters,
    orders of magnitude faster on long circuits; every qubit local).
    """
    if mode == "lookahead":
        cluster_of = _cppsim_amd.plan_clusters([list(g[1]) for g in gates],
                                               [list(g[2]) for g in gates], int(cluster_size), 1)
        clusters = [[] for _ in range(max(cluster_of) + 1 if cluster_of else 0)]
        for i, c in enumerate(cluster_of):
            clusters[c].append(i)
        return clusters
    if mode != "reference":
        raise ValueError("mode must be 'reference' or 'lookahead'")
    remaining = list(range(len(gates)))
    clusters = []
    local_ids = list(local_ids)
    while remaining:
        targets = [list(gates[i][1]) for i in remaining]
        ctrls = [list(gates[i][2]) for i in remaining]
        avail = ClusterScheduler(targets, ctrls, [False] * len(remaining), local_ids, [],
                                 int(cluster_size)).ScheduleCluster()
        if not avail:
            raise RuntimeError("ClusterScheduler returned no cluster for %d gates"
                               % len(remaining))
        taken = set(avail)
        clusters.append([remaining[i] for i in avail])
        remaining = [g for j, g in enumerate(remaining) if j not in taken]
    return clusters


def apply_scheduled(sim, gates, cluster_size, local_ids=None, mode="reference"):
    """Apply a gate list cluster by cluster (one run() per cluster)."""
    if local_ids is None:
        local_ids = sim.get_qubits_ids()
    for cluster in plan_clusters(gates, local_ids, cluster_size, mode):
        for i in cluster:
            m, t, c = gates[i]
            sim.apply_controlled_gate(m, t, c)
        sim.run()
