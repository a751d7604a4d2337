"""Tile passes (apply_tiled): the cluster scheduler retargeted to the LDS
tile.  Final states against the oracle fed the same gates in program order:
complex128 max-abs <= 1e-12, complex64 fidelity >= 1 - 1e-6."""
import numpy as np
import pytest

from hiqsimulator_amd.batching import fidelity
from hiqsimulator_amd.circuits import X, apply_circuit_inorder, haar_unitary, random_layers_circuit

pytestmark = pytest.mark.gpu


def circuit(n, layers, seed):
    rng = np.random.default_rng(seed)
    g = random_layers_circuit(n, layers, seed, haar_every=3)
    for _ in range(4):
        q = rng.choice(n, size=3, replace=False)
        g.insert(int(rng.integers(len(g))), (X, [int(q[0])], [int(q[1]), int(q[2])]))
    q = rng.choice(n, size=3, replace=False)
    g.append((haar_unitary(4, rng), [int(q[0]), int(q[1])], [int(q[2])]))
    return g


def oracle_state(oracle_mod, n, gates, ids=None):
    o = oracle_mod.OracleSim(1, n, 5)
    o.allocate_qureg(list(range(n)) if ids is None else ids, 0)
    apply_circuit_inorder(o, gates, 5)
    return o


@pytest.mark.parametrize("n,op_width", [(12, 4), (14, 3), (16, 5), (18, 2), (20, 4), (9, 4)])
def test_tiled_c128_matches_oracle(engine_mod, oracle_mod, n, op_width):
    g = circuit(n, 10, 100 + n)
    print("tiled n=%d op_width=%d" % (n, op_width), flush=True)
    sim = engine_mod.SimulatorMPI(1, n, 5)
    sim.allocate_qureg(list(range(n)), 0)
    passes = sim.apply_tiled(g, op_width)
    assert passes >= 1
    ref = oracle_state(oracle_mod, n, g).get_state()
    assert np.abs(sim.get_state() - ref).max() <= 1e-12


@pytest.mark.parametrize("mfma", [1, 0])
@pytest.mark.parametrize("op_width", [3, 4, 5])
def test_tile_ops_on_matrix_cores_and_vector_units(engine_mod, oracle_mod, mfma, op_width):
    """Ops of at least 3 qubits on the matrix cores (mfma = 1) or every op
    on the vector units (0): complex128 equals the oracle to 1e-12 and
    complex64 keeps fidelity 1 - 1e-6."""
    import hiqsimulator_amd
    mod = hiqsimulator_amd._cppsim_amd
    old = mod.get_tuning("tile_mfma")
    n = 15
    g = circuit(n, 10, 300 + op_width)
    ref = oracle_state(oracle_mod, n, g).get_state()
    print("tile ops mfma=%d op_width=%d" % (mfma, op_width), flush=True)
    try:
        mod.set_tuning("tile_mfma", mfma)
        for prec in ("c128", "c64"):
            sim = engine_mod.SimulatorMPI(1, n, 5, prec)
            sim.allocate_qureg(list(range(n)), 0)
            sim.apply_tiled(g, op_width)
            got = sim.get_state()
            if prec == "c128":
                assert np.abs(got - ref).max() <= 1e-12
            else:
                assert fidelity(ref, got) >= 1 - 1e-6
    finally:
        mod.set_tuning("tile_mfma", old)


@pytest.mark.parametrize("threads", [256, 512])
def test_tile_workgroup_sizes(engine_mod, oracle_mod, threads):
    """Both workgroup sizes of the matrix-core tile programs (the default
    picks 256 threads for complex64 and for complex128 programs with 5-qubit
    ops, 512 otherwise): op widths 4 and 5, both precisions, against the
    oracle."""
    import hiqsimulator_amd
    mod = hiqsimulator_amd._cppsim_amd
    old = mod.get_tuning("tile_threads")
    n = 15
    try:
        mod.set_tuning("tile_threads", threads)
        for op_width in (4, 5):
            g = circuit(n, 10, 400 + op_width)
            ref = oracle_state(oracle_mod, n, g).get_state()
            for prec in ("c128", "c64"):
                print("tile threads=%d op_width=%d %s" % (threads, op_width, prec), flush=True)
                sim = engine_mod.SimulatorMPI(1, n, 5, prec)
                sim.allocate_qureg(list(range(n)), 0)
                sim.apply_tiled(g, op_width)
                got = sim.get_state()
                if prec == "c128":
                    assert np.abs(got - ref).max() <= 1e-12
                else:
                    assert fidelity(ref, got) >= 1 - 1e-6
    finally:
        mod.set_tuning("tile_threads", old)


def test_tiled_after_permutation_and_pending_gates(engine_mod, oracle_mod):
    """apply_tiled after permute_qubits with gates still pending: the pending
    fusion (at most 5 qubits, as the reference's scheduler keeps it) runs
    first, at the qubits' pre-tile positions."""
    n = 15
    g = circuit(n, 8, 7)
    rng = np.random.default_rng(3)
    perm = [int(x) for x in rng.permutation(n)]
    print("tiled after permutation n=%d" % n, flush=True)
    sim = engine_mod.SimulatorMPI(1, n, 5)
    sim.allocate_qureg(list(range(n)), 0)
    head, tail = g[:20], g[20:]
    pending = set()
    for m, t, c in head:  # run at 5-qubit cluster boundaries; the last cluster stays pending
        q = set(t) | set(c)
        if len(pending | q) > 5:
            sim.run()
            pending = set()
        sim.apply_controlled_gate(m, t, c)
        pending |= q
    assert pending
    sim.permute_qubits(perm)
    sim.apply_tiled(tail, 4)
    o = oracle_state(oracle_mod, n, g)
    got = sim.get_state()
    # bring the oracle state into the engine's bit order
    ids = sim.get_qubits_ids()
    ref = o.get_state()
    idx = np.arange(1 << n)
    src = np.zeros(1 << n, dtype=np.int64)
    for p, q in enumerate(ids):
        src |= ((idx >> p) & 1) << q
    assert np.abs(got - ref[src]).max() <= 1e-12


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_tiled_round_trip_30q(engine_mod, precision):
    """Size-independent property at the benchmark size: the 30-qubit, 40-layer
    benchmark circuit with tile passes, then its inverse, returns a random
    state to itself (compared on the device in FP64)."""
    n = 30
    gates = random_layers_circuit(n, 40, 1234)
    inv = [(m.conj().T, t, c) for (m, t, c) in reversed(gates)]
    print("tiled round trip 30q %s" % precision, flush=True)
    g = engine_mod.SimulatorMPI(1, n, 4, precision)
    g.allocate_qureg(list(range(n)), 0)
    g.init_random_state(30)
    g.snapshot()
    passes = g.apply_tiled(gates, 4) + g.apply_tiled(inv, 4)
    r = g.compare_snapshot()
    print("tiled round trip 30q %s: %d passes, max-abs %.3e, 1 - fidelity %.3e"
          % (precision, passes, r["max_abs_diff"], 1 - r["fidelity"]), flush=True)
    assert passes < 2 * 196  # fewer HBM passes than the fused form
    if precision == "c128":
        assert r["max_abs_diff"] <= 1e-12
        assert 1 - r["fidelity"] <= 1e-12
    else:
        assert r["fidelity"] >= 1 - 1e-6


def test_tiled_c64_fidelity(engine_mod, oracle_mod):
    n = 18
    g = circuit(n, 12, 9)
    sim = engine_mod.SimulatorMPI(1, n, 5, "c64")
    sim.allocate_qureg(list(range(n)), 0)
    sim.apply_tiled(g, 4)
    ref = oracle_state(oracle_mod, n, g).get_state()
    assert fidelity(ref, sim.get_state()) >= 1 - 1e-6


def test_tiled_errors(engine_mod):
    sim = engine_mod.SimulatorMPI(1, 12, 5)
    sim.allocate_qureg(list(range(12)), 0)
    with pytest.raises(RuntimeError):
        sim.apply_tiled([(X, [12], [])], 4)  # unknown qubit
    with pytest.raises(RuntimeError):
        sim.apply_tiled([(X, [0], [])], 6)  # op width
    assert sim.apply_tiled([], 4) == 0
