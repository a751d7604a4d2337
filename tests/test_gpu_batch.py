"""Batched many-circuit path on the GPU.

Every circuit's final state is compared with the oracle (the FP64
SimulatorMPI restatement fed the same gates in program order): complex128
max-abs <= 1e-12, complex64 fidelity >= 1 - 1e-6.  The plans cover the three
execution classes: a workgroup-owned state held in LDS, a workgroup-owned
state streamed through L2/HBM, and layered many-workgroup launches.
"""
import numpy as np
import pytest

from hiqsimulator_amd.batching import batch_workload, fidelity, make_batch
from hiqsimulator_amd.circuits import X, apply_circuit_inorder, haar_unitary, random_layers_circuit

pytestmark = pytest.mark.gpu


def mixed_circuits(seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in (1, 2, 3, 5, 8, 11, 13, 14, 15, 16, 17):
        depth = int(rng.integers(3, 25))
        g = random_layers_circuit(n, depth, int(rng.integers(1 << 30)),
                                  haar_every=5 if n >= 2 else 0)
        if n >= 3:
            g.insert(len(g) // 2, (X, [0], [1, 2]))                        # Toffoli
            g.append((haar_unitary(4, rng), [n - 1, 0], [1]))              # controlled 2-qubit
        out.append((n, g))
    out.append((4, []))  # no gates: stays |0000>
    out.append((12, []))  # no gates on the layered and tile-program paths either
    return out


def oracle_state(oracle_mod, n, gates):
    o = oracle_mod.OracleSim(1, n, 5)
    o.allocate_qureg(list(range(n)), 0)
    apply_circuit_inorder(o, gates, 5)
    return o.get_state()


PLANS = {
    # name: (k_by_n, wg_max_qubits, tile_min_qubits, prog_min_qubits)
    "k4_split14": ({}, 14, 0, 31),
    "kvar_all_wg": ({n: 2 + n % 4 for n in range(1, 18)}, 24, 0, 31),
    "k5_all_layered": ({n: 5 for n in range(1, 18)}, 0, 0, 31),
    "k2_split16": ({n: 2 for n in range(1, 18)}, 16, 0, 31),
    "k3_layered_direct": ({n: 3 for n in range(1, 18)}, 10, 31, 31),
    "programs_from_12": ({n: 2 + n % 4 for n in range(1, 18)}, 8, 0, 12),
    "programs_from_15": ({}, 11, 0, 15),
}


@pytest.mark.parametrize("plan", sorted(PLANS))
def test_batch_c128_matches_oracle(engine_mod, oracle_mod, plan):
    k_by_n, wg_max, tile_min, prog_min = PLANS[plan]
    circuits = mixed_circuits(5)
    print("batch c128 plan", plan, flush=True)
    b = make_batch(circuits, "c128", k_by_n, wg_max, tile_min_qubits=tile_min,
                   prog_min_qubits=prog_min)
    st = b.stats()
    assert st["circuits"] == len(circuits)
    b.run()
    first = [b.get_state(i) for i in range(len(circuits))]
    for i, (n, g) in enumerate(circuits):
        ref = oracle_state(oracle_mod, n, g)
        err = np.abs(first[i] - ref).max()
        assert err <= 1e-12, (plan, i, n, err)
    b.run()  # replay re-initialises every state
    for i in range(len(circuits)):
        assert np.array_equal(b.get_state(i), first[i])


def test_batch_classes_are_all_exercised(engine_mod):
    circuits = mixed_circuits(5)
    b = make_batch(circuits, "c128", {}, 14)
    st = b.stats()
    # n <= 13 fits LDS (128 KiB of complex128); n = 14 streams; 15..17 layered
    assert st["lds_circuits"] == sum(1 for n, _ in circuits if n <= 13)
    assert st["wg_circuits"] == sum(1 for n, _ in circuits if n <= 14)
    assert st["layer_launches"] > 0


@pytest.mark.parametrize("prog_min", [31, 12])
def test_batch_c64_fidelity(engine_mod, oracle_mod, prog_min):
    circuits = mixed_circuits(9)
    print("batch c64 prog_min", prog_min, flush=True)
    b = make_batch(circuits, "c64", {}, 11, prog_min_qubits=prog_min)
    b.run()
    for i, (n, g) in enumerate(circuits):
        ref = oracle_state(oracle_mod, n, g)
        got = b.get_state(i)
        assert fidelity(ref, got) >= 1 - 1e-6, (i, n)
        assert np.abs(got - ref).max() <= 1e-5


def test_batch_many_small_circuits(engine_mod, oracle_mod):
    """More circuits than workgroups: several circuits per workgroup."""
    circuits = batch_workload(700, 3, n_range=(6, 12), depth_range=(2, 12))
    print("batch many small", flush=True)
    b = make_batch(circuits, "c128", {}, 16)
    assert b.stats()["wg_circuits"] == 700
    b.run()
    rng = np.random.default_rng(0)
    for i in rng.choice(len(circuits), size=25, replace=False):
        n, g = circuits[i]
        assert np.abs(b.get_state(int(i)) - oracle_state(oracle_mod, n, g)).max() <= 1e-12
    for i in range(0, len(circuits), 7):
        assert abs(np.linalg.norm(b.get_state(i)) - 1.0) < 1e-12


def test_batch_errors(engine_mod):
    import hiqsimulator_amd
    b = hiqsimulator_amd.Batch("c128")
    with pytest.raises(RuntimeError):
        b.run()  # not built
    with pytest.raises(RuntimeError):
        b.build()  # no circuits
    with pytest.raises(RuntimeError):
        b.add_circuit(3, [(X, [3], [])])  # qubit id out of range
    with pytest.raises(RuntimeError):
        b.add_circuit(8, [(X, [0], [1, 2, 3, 4, 5])])  # 6 qubits with controls
    with pytest.raises(RuntimeError):
        b.add_circuit(3, [(X, [0], [0])])  # repeated qubit
    with pytest.raises(RuntimeError):
        b.set_plan({4: 6}, 16)
    with pytest.raises(RuntimeError):
        b.set_plan({4: 4}, 16, 32)
    with pytest.raises(RuntimeError):
        hiqsimulator_amd.Batch("c32")
