"""Whole-circuit cluster planners (hiqsim_plan_clusters), CPU only.

* mode 0 must return exactly the clusters of the GreedyScheduler loop over
  ClusterScheduler (_greedyscheduler.py:119-137), restated in Python in
  hiqsimulator_amd.scheduling;
* mode 1 (lookahead packer) must be a valid reordering: every gate once,
  at most k qubits per cluster, every pair of gates that share a qubit kept
  in program order, and the oracle state after the clustered order equal to
  the in-order state.
"""
import numpy as np
import pytest

from hiqsimulator_amd import _cppsim_amd
from hiqsimulator_amd.circuits import X, apply_circuit_inorder, random_layers_circuit
from hiqsimulator_amd.scheduling import plan_clusters


def circuit(n, layers, seed):
    g = random_layers_circuit(n, layers, seed, haar_every=3)
    rng = np.random.default_rng(seed)
    for _ in range(3):  # a few gates with controls
        q = rng.choice(n, size=3, replace=False)
        g.insert(int(rng.integers(len(g))), (X, [int(q[0])], [int(q[1]), int(q[2])]))
    return g


def qubits(gate):
    return set(gate[1]) | set(gate[2])


@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_reference_mode_matches_python_loop(k):
    g = circuit(7, 6, 11 + k)
    ref = plan_clusters(g, list(range(7)), k, mode="reference")
    cluster_of = _cppsim_amd.plan_clusters([x[1] for x in g], [x[2] for x in g], k, 0)
    got = [[] for _ in range(max(cluster_of) + 1)]
    for i, c in enumerate(cluster_of):
        got[c].append(i)
    assert got == ref


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_lookahead_is_a_valid_reordering(oracle_mod, seed, k):
    n = 8
    g = circuit(n, 12, seed)
    clusters = plan_clusters(g, list(range(n)), k, mode="lookahead")
    flat = [i for c in clusters for i in c]
    assert sorted(flat) == list(range(len(g)))
    where = {i: ci for ci, c in enumerate(clusters) for i in c}
    for c in clusters:
        assert c == sorted(c)
        width = len(set().union(*(qubits(g[i]) for i in c)))
        assert width <= max(k, max(len(qubits(g[i])) for i in c))
    for i in range(len(g)):
        for j in range(i + 1, len(g)):
            if qubits(g[i]) & qubits(g[j]):
                assert where[i] <= where[j], (i, j)
    a = oracle_mod.OracleSim(1, n, 5)
    b = oracle_mod.OracleSim(1, n, 5)
    for s in (a, b):
        s.allocate_qureg(list(range(n)), 0)
    apply_circuit_inorder(a, g, 5)
    for c in clusters:
        for i in c:
            b.apply_controlled_gate(*g[i])
        b.run()
    assert np.abs(a.get_state() - b.get_state()).max() <= 1e-12


def test_lookahead_is_close_to_reference_count():
    g = random_layers_circuit(16, 20, 1234, haar_every=5)
    for k in (3, 4, 5):
        ref = len(plan_clusters(g, list(range(16)), k, mode="reference"))
        got = len(plan_clusters(g, list(range(16)), k, mode="lookahead"))
        assert got <= 1.1 * ref, (k, got, ref)


def test_plan_errors():
    with pytest.raises(RuntimeError):
        _cppsim_amd.plan_clusters([[0]], [[]], 2, 7)


@pytest.mark.parametrize("n,op_width", [(12, 4), (13, 3), (15, 2), (17, 4)])
def test_tile_plan_is_a_valid_reordering(oracle_mod, n, op_width):
    """Tile-capacity plans: every gate in one (pass, op); an op's bits lie in
    {0..5} + the pass's planned bits (six bits >= 6); gates that share a bit
    keep their order; the oracle state after (pass, op, program) order equals
    the in-order state."""
    g = circuit(n, 8, 31 + n)
    tg = [list(x[1]) for x in g]
    cg = [list(x[2]) for x in g]
    pass_of, op_of, hpos = _cppsim_amd.plan_tile_passes(tg, cg, n, op_width)
    npass = max(pass_of) + 1
    assert len(hpos) == 6 * npass
    for p in range(npass):
        hp = hpos[6 * p:6 * p + 6]
        assert hp == sorted(hp) and len(set(hp)) == 6 and all(6 <= b < n for b in hp)
    order = sorted(range(len(g)), key=lambda i: (pass_of[i], op_of[i], i))
    for i in range(len(g)):
        tile = set(range(6)) | set(hpos[6 * pass_of[i]:6 * pass_of[i] + 6])
        assert qubits(g[i]) <= tile
    ops = {}
    for i in range(len(g)):
        ops.setdefault((pass_of[i], op_of[i]), set()).update(qubits(g[i]))
    for key, bits in ops.items():
        members = [i for i in range(len(g)) if (pass_of[i], op_of[i]) == key]
        assert len(bits) <= max(op_width, max(len(qubits(g[i])) for i in members))
    rank = {i: r for r, i in enumerate(order)}
    for i in range(len(g)):
        for j in range(i + 1, len(g)):
            if qubits(g[i]) & qubits(g[j]):
                assert rank[i] < rank[j], (i, j)
    a = oracle_mod.OracleSim(1, n, 5)
    b = oracle_mod.OracleSim(1, n, 5)
    for s in (a, b):
        s.allocate_qureg(list(range(n)), 0)
    apply_circuit_inorder(a, g, 5)
    for i in order:
        b.apply_controlled_gate(*g[i])
        b.run()
    assert np.abs(a.get_state() - b.get_state()).max() <= 1e-12


def test_tile_plan_cuts_passes():
    g = random_layers_circuit(30, 40, 1234)
    pass_of, _, _ = _cppsim_amd.plan_tile_passes([x[1] for x in g], [x[2] for x in g], 30, 4)
    clusters = _cppsim_amd.plan_clusters([x[1] for x in g], [x[2] for x in g], 4, 1)
    assert max(pass_of) + 1 < (max(clusters) + 1) / 2.5


def test_tile_plan_errors():
    with pytest.raises(RuntimeError):
        _cppsim_amd.plan_tile_passes([[0]], [[]], 11, 4)  # fewer than 12 bits
    with pytest.raises(RuntimeError):
        _cppsim_amd.plan_tile_passes([[12]], [[]], 12, 4)  # bit out of range


# ---- brute-force pins of the scheduler objectives ---------------------------
def _random_gate_list(rng, nq, ngates):
    gates, ctrls, diag = [], [], []
    for _ in range(ngates):
        nt = int(rng.integers(1, 3))
        nc = int(rng.integers(0, 2))
        q = [int(x) for x in rng.choice(nq, size=nt + nc, replace=False)]
        gates.append(q[:nt])
        ctrls.append(q[nt:])
        diag.append(bool(rng.random() < 0.3))
    return gates, ctrls, diag


def _cluster_taken(gates, ctrls, diag, local, cluster):
    """Gates a ClusterScheduler cluster takes (cluster_scheduler.h:146-160):
    in order, a gate is taken if it touches no qubit of a skipped gate, its
    local qubits lie in the cluster and, unless diagonal, its targets are
    all local; a skipped gate's qubits block every later gate on them."""
    bad, taken = set(), []
    for i, (t, c) in enumerate(zip(gates, ctrls)):
        q = set(t) | set(c)
        if not (q & bad) and (q & local) <= cluster and (diag[i] or set(t) <= local):
            taken.append(i)
        else:
            bad |= q
    return taken


@pytest.mark.parametrize("seed", range(12))
def test_cluster_scheduler_objective_brute_force(seed):
    """ScheduleCluster returns the gates of a cluster (a set of at most
    cluster_size local qubits) that takes the most gates, and among those
    the fewest qubits (cluster_scheduler.h:76-92, .cpp:40-77); with no such
    cluster, the first gate takeable on all local qubits.  Every cluster is
    enumerated here; only the choice among equal clusters is the search
    order's."""
    import itertools
    import hiqsimulator_amd as hq
    rng = np.random.default_rng(seed)
    nq = 7
    gates, ctrls, diag = _random_gate_list(rng, nq, 12)
    local = set(int(x) for x in rng.choice(nq, size=int(rng.integers(4, 7)), replace=False))
    glob = sorted(set(range(nq)) - local)
    for k in (1, 2, 3, 4):
        got = hq.ClusterScheduler(gates, ctrls, diag, sorted(local), glob, k).ScheduleCluster()
        best, cands = 0, []
        for r in range(0, k + 1):
            for cl in itertools.combinations(sorted(local), r):
                n = len(_cluster_taken(gates, ctrls, diag, local, set(cl)))
                if n > best:
                    best, cands = n, [(r, cl)]
                elif n == best and n:
                    cands.append((r, cl))
        if best:
            rmin = min(r for r, _ in cands)
            allowed = [_cluster_taken(gates, ctrls, diag, local, set(cl)) for r, cl in cands if r == rmin]
            assert got in allowed, (seed, k, got, allowed)
        else:
            first = _cluster_taken(gates, ctrls, diag, local, local)
            assert got == first[:1]


def _swap_count(gates, ctrls, diag, local):
    """Gates kept when `local` is the local set (swap_scheduler.h:73-88, in-order
    rule): a gate touching a skipped gate's qubit is skipped; otherwise a
    diagonal gate is taken and a non-diagonal one iff its targets are local."""
    bad, n = set(), 0
    for i, (t, c) in enumerate(zip(gates, ctrls)):
        q = set(t) | set(c)
        if not (q & bad) and (diag[i] or set(t) <= local):
            n += 1
        else:
            bad |= q
    return n


@pytest.mark.parametrize("seed", range(12))
def test_swap_scheduler_objective_brute_force(seed):
    """With an unbounded split budget ScheduleSwap's backtracking is
    exhaustive (swap_scheduler.cpp:90-172), so the local set it returns keeps
    as many gates as the best set of at most num_locals qubits, found here by
    enumerating every such set (single-qubit fusion off: every gate weighs 1)."""
    import itertools
    import hiqsimulator_amd as hq
    rng = np.random.default_rng(100 + seed)
    nq = 6
    gates, ctrls, diag = _random_gate_list(rng, nq, 10)
    for num_locals in (2, 3, 4):
        got = hq.SwapScheduler(gates, ctrls, diag, 10 ** 6, num_locals, False).ScheduleSwap()
        assert len(got) <= num_locals and len(set(got)) == len(got)
        best = max(_swap_count(gates, ctrls, diag, set(l))
                   for r in range(num_locals + 1) for l in itertools.combinations(range(nq), r))
        assert _swap_count(gates, ctrls, diag, set(got)) == best, (seed, num_locals, got)
