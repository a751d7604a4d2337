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
