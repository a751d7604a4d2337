"""The bench's parity step, on the CPU: ``inverse_step(step, start)`` undoes a
step of fused passes and qubit permutations.  A random 12-qubit state run
through one step of the headline structure (layers with permutations between
them, planned by the GreedyScheduler loop) and then its inverse, on the
oracle, comes back to itself."""
import numpy as np

import bench


def run(sim, step):
    for kind, payload in step:
        if kind == "pass":
            for m, t, c in payload:
                sim.apply_controlled_gate(m, t, c)
            sim.run()
        else:
            sim.permute_qubits(payload)


def test_inverse_step_restores_the_state(oracle_mod):
    n = 12
    _, step, _ = bench.build_step(n, 8, 4, 21, 3)
    assert sum(1 for kind, _ in step if kind == "perm") == 3
    rng = np.random.default_rng(4)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    o = oracle_mod.OracleSim(1, n, 4)
    o.allocate_qureg(list(range(n)), 0)
    o.set_state(psi)
    start = list(o.get_qubits_ids())
    run(o, step)
    moved = o.get_state()
    assert np.abs(moved - psi).max() > 1e-3  # the step does something
    run(o, bench.inverse_step(step, start))
    assert list(o.get_qubits_ids()) == start
    assert np.abs(o.get_state() - psi).max() <= 1e-12


def test_inverse_step_of_a_permutation_only_step():
    step = [("perm", [2, 0, 1]), ("perm", [1, 2, 0])]
    inv = bench.inverse_step(step, [0, 1, 2])
    assert inv == [("perm", [2, 0, 1]), ("perm", [0, 1, 2])]
