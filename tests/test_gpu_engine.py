"""Engine-level parity on the GPU: the reference test-suite's known answers,
and differential runs against the oracle on identical API call sequences.

complex128 states: max-abs <= 1e-12 against the FP64 oracle.
Measurement: identical outcomes for the same seed (same mt19937 u stream and
block-CDF search as the reference, SimulatorMPI.cpp:897-1008).
Permutation: bit-exact.  complex64: fidelity >= 1 - 1e-6.
"""
import math

import numpy as np
import pytest

import scenarios
from hiqsimulator_amd.circuits import (H, X, apply_circuit_inorder, haar_unitary,
                                       qft_like_circuit, random_layers_circuit,
                                       random_permutation)

pytestmark = pytest.mark.gpu


def make_engine(seed, max_local, max_cluster):
    import hiqsimulator_amd
    return hiqsimulator_amd.SimulatorMPI(seed, max_local, max_cluster)


@pytest.mark.parametrize("scenario", scenarios.ALL_KNOWN_ANSWERS, ids=lambda f: f.__name__)
def test_known_answers(engine_mod, scenario):
    print("scenario", scenario.__name__, flush=True)
    scenario(make_engine)


def test_ghz_measurement_matches_oracle(engine_mod, oracle_mod):
    seen = set()
    for seed in range(10):
        a = scenarios.ghz_measure(make_engine, seed)
        b = scenarios.ghz_measure(lambda s, ml, mc: oracle_mod.OracleSim(s, ml, mc), seed)
        assert a == b
        seen.add(sum(a))
    assert seen == {0, 5}


@pytest.mark.parametrize("mode", [0, 1], ids=["pass-each", "deferred"])
@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_config0_random_circuit(engine_mod, oracle_mod, k, mode):
    """16 qubits, 20 layers, |0..0>, one seed: full state vs oracle, with
    one fused pass per run() and with deferred clusters (tile passes)."""
    n, layers, seed = 16, 20, 2024
    gates = random_layers_circuit(n, layers, seed)
    g = make_engine(seed, n, k)
    g.set_pass_mode(mode)
    o = oracle_mod.OracleSim(seed, n, k)
    for s in (g, o):
        s.allocate_qureg(list(range(n)), 0)
        apply_circuit_inorder(s, gates, k)
    a, b = g.get_state(), o.get_state()
    assert np.abs(a - b).max() <= 1e-12
    assert abs(np.linalg.norm(a) - 1.0) < 1e-12


def test_fusion_with_folded_controls(engine_mod, oracle_mod):
    """Gates fused into one pass, with controls folded into the matrix, and
    a wide controlled gate taking the control-mask path."""
    rng = np.random.default_rng(5)
    n = 9
    g = make_engine(1, n, 4)
    o = oracle_mod.OracleSim(1, n, 4)
    seq = [(haar_unitary(2, rng), [3], [7]), (haar_unitary(4, rng), [1, 3], []),
           (H, [7], []), (haar_unitary(2, rng), [1], [3, 7]), ("run",),
           (X, [8], [0, 2, 4, 6]), ("run",), (haar_unitary(8, rng), [5, 0, 2], []),
           (haar_unitary(2, rng), [4], [5]), ("run",), (np.diag([1, 1j]), [6], [2, 3]),
           (np.diag([1j, -1]), [6], []), ("run",)]
    for s in (g, o):
        s.allocate_qureg(list(range(n)), 0)
        for q in range(n):
            s.apply_controlled_gate(H, [q], [])
            s.run()
        for item in seq:
            if isinstance(item[0], str):
                s.run()
            else:
                s.apply_controlled_gate(*item)
    assert np.abs(g.get_state() - o.get_state()).max() <= 1e-12


def test_measure_collapse_dealloc_sequence(engine_mod, oracle_mod):
    n = 12
    gates = random_layers_circuit(n, 8, seed=77)
    g = make_engine(4242, n, 4)
    o = oracle_mod.OracleSim(4242, n, 4)
    for s in (g, o):
        s.allocate_qureg(list(range(n)), 0)
        apply_circuit_inorder(s, gates, 4)
    for q in [3, 0, 11, 7]:
        assert g.measure_qubits([q]) == o.measure_qubits([q])
        assert np.abs(g.get_state() - o.get_state()).max() <= 1e-12
    assert g.measure_qubits([1, 2, 5]) == o.measure_qubits([1, 2, 5])
    for q in [11, 3, 0, 7, 1]:
        g.deallocate_qubit(q)
        o.deallocate_qubit(q)
        assert g.get_qubits_ids() == o.get_qubits_ids()
        assert np.abs(g.get_state() - o.get_state()).max() <= 1e-12
    assert g.entropy() == pytest.approx(o.entropy(), abs=1e-12)
    ids = g.get_qubits_ids()
    for bits in ([0] * len(ids), [1, 0, 1, 0, 1, 0, 1][:len(ids)]):
        assert g.get_probability(bits, ids) == pytest.approx(o.get_probability(bits, ids),
                                                            abs=1e-13)
        assert g.get_amplitude(bits, ids) == pytest.approx(o.get_amplitude(bits, ids), abs=1e-13)


def test_measurement_stream_many_seeds(engine_mod, oracle_mod):
    """Same seed, same u stream, same block-CDF search: identical outcomes on
    a 17-qubit state (more than 2^15 amplitudes, so blocks hold >1 entry)."""
    n = 17
    gates = random_layers_circuit(n, 4, seed=9)
    for seed in (0, 1, 99, 2 ** 32 + 5):
        g = make_engine(seed, n, 4)
        o = oracle_mod.OracleSim(seed, n, 4)
        for s in (g, o):
            s.allocate_qureg(list(range(n)), 0)
            apply_circuit_inorder(s, gates, 4)
        for q in range(0, n, 3):
            assert g.measure_qubits([q, (q + 1) % n]) == o.measure_qubits([q, (q + 1) % n])
        assert np.abs(g.get_state() - o.get_state()).max() <= 1e-12


def test_expectation_and_operator_vs_oracle(engine_mod, oracle_mod):
    rng = np.random.default_rng(12)
    n = 10
    gates = random_layers_circuit(n, 5, seed=12)
    g = make_engine(1, n, 4)
    o = oracle_mod.OracleSim(1, n, 4)
    for s in (g, o):
        s.allocate_qureg(list(range(n)), 0)
        apply_circuit_inorder(s, gates, 4)
    order = [int(x) for x in rng.permutation(n)]
    for t in range(8):
        w = int(rng.integers(1, 5))
        qs = rng.choice(n, size=w, replace=False)
        term = [(int(q), "XYZ"[int(rng.integers(0, 3))]) for q in qs]
        ops = [(term, complex(rng.normal(), rng.normal())), ([], 0.25)]
        assert g.get_expectation_value(ops, order) == pytest.approx(
            o.get_expectation_value(ops, order), abs=1e-12)
    ops = [([(0, "X"), (3, "Y")], 0.5), ([(2, "Z")], -0.25j), ([], 0.1)]
    g.apply_qubit_operator(ops, order)
    o.apply_qubit_operator(ops, order)
    assert np.abs(g.get_state() - o.get_state()).max() <= 1e-12


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_entropy_closed_forms(engine_mod, prec):
    """Entropy (SimulatorMPI.cpp:681-694) against closed forms: a basis state
    0; H on n qubits n; GHZ 1; a product of Ry(theta_q)|0> the sum of the
    binary entropies H2(sin^2(theta_q / 2)); probabilities (1/2, 1/4, 1/8,
    1/8) 1.75.  complex64 amplitudes carry about 1e-7 relative error."""
    tol = 1e-12 if prec == "c128" else 2e-6

    def fresh(n):
        s = engine_mod.SimulatorMPI(1, n, 4, prec)
        s.allocate_qureg(list(range(n)), 0)
        return s

    def ry(t):
        c, s = math.cos(t / 2), math.sin(t / 2)
        return np.array([[c, -s], [s, c]], dtype=complex)

    def h2(p):
        return sum(-x * math.log2(x) for x in (p, 1.0 - p) if x > 0)

    g = fresh(14)
    g.apply_controlled_gate(X, [9], [])
    g.run()
    assert abs(g.entropy()) < tol
    for n in (1, 5, 20):
        g = fresh(n)
        for q in range(n):
            g.apply_controlled_gate(H, [q], [])
            g.run()
        assert abs(g.entropy() - n) < tol * max(1, n)
    g = fresh(16)
    g.apply_controlled_gate(H, [0], [])
    g.run()
    for q in range(1, 16):
        g.apply_controlled_gate(X, [q], [q - 1])
        g.run()
    assert abs(g.entropy() - 1.0) < tol
    thetas = np.random.default_rng(17).uniform(0.1, math.pi - 0.1, size=13)
    g = fresh(13)
    for q, t in enumerate(thetas):
        g.apply_controlled_gate(ry(t), [q], [])
        g.run()
    assert abs(g.entropy() - sum(h2(math.sin(t / 2) ** 2) for t in thetas)) < tol * 13
    g = fresh(2)
    g.set_state(np.sqrt(np.array([0.5, 0.25, 0.125, 0.125])).astype(complex))
    assert abs(g.entropy() - 1.75) < tol


def test_run_rejects_wide_pending_fusion(engine_mod):
    """Gates left pending across 13 qubits: run() raises the reference's
    "cannot apply" error before building the 2^13 x 2^13 fused matrix
    (which would take 1 GiB and seconds of host work)."""
    import time
    sim = make_engine(1, 13, 5)
    sim.allocate_qureg(list(range(13)), 0)
    for q in range(13):
        sim.apply_controlled_gate(H, [q], [])
    t0 = time.perf_counter()
    with pytest.raises(RuntimeError, match="cannot apply 13 qubits gate"):
        sim.run()
    assert time.perf_counter() - t0 < 2.0


def test_permutations_keep_results(engine_mod, oracle_mod):
    """Random qubit permutations between circuit halves: state (read in the
    permuted order) and probabilities agree with the oracle, which also
    permutes, and the GPU data movement is bit-exact."""
    n = 14
    rng = np.random.default_rng(3)
    gates = random_layers_circuit(n, 10, seed=3)
    g = make_engine(1, n, 4)
    o = oracle_mod.OracleSim(1, n, 4)
    for s in (g, o):
        s.allocate_qureg(list(range(n)), 0)
    half = len(gates) // 2
    for part in (gates[:half], gates[half:]):
        for s in (g, o):
            apply_circuit_inorder(s, part, 4)
        perm = random_permutation(g.get_qubits_ids(), rng)
        before = g.get_state()
        g.permute_qubits(perm)
        o.permute_qubits(perm)
        after = g.get_state()
        assert np.array_equal(np.sort_complex(before), np.sort_complex(after))
        assert g.get_qubits_ids() == o.get_qubits_ids() == perm
        assert np.abs(g.get_state() - o.get_state()).max() <= 1e-12


@pytest.mark.parametrize("mode", [0, 1], ids=["pass-each", "deferred"])
def test_complex64_fidelity(engine_mod, oracle_mod, mode):
    n = 16
    gates = random_layers_circuit(n, 20, seed=5)
    g = engine_mod.SimulatorMPI(5, n, 4, "c64")
    g.set_pass_mode(mode)
    o = oracle_mod.OracleSim(5, n, 4)
    for s in (g, o):
        s.allocate_qureg(list(range(n)), 0)
        apply_circuit_inorder(s, gates, 4)
    a, b = g.get_state(), o.get_state()
    fid = abs(np.vdot(b, a)) ** 2
    print("c64 16q max-abs %.3e fidelity 1-%.3e" % (np.abs(a - b).max(), 1 - fid))
    assert fid >= 1 - 1e-6


def test_qft_equals_dft(engine_mod):
    """QFT-like circuit == normalised DFT up to a bit reversal of the output
    index (closed form), on a random normalised state."""
    n = 12
    rng = np.random.default_rng(8)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    st /= np.linalg.norm(st)
    g = make_engine(1, n, 4)
    g.allocate_qureg(list(range(n)), 0)
    g.set_state(st)
    apply_circuit_inorder(g, qft_like_circuit(n), 4)
    got = g.get_state()
    dft = np.fft.ifft(st, norm="ortho")
    rev = np.array([int(format(i, "0%db" % n)[::-1], 2) for i in range(1 << n)])
    assert np.abs(got[rev] - dft).max() <= 1e-12


def test_large_state_roundtrip_26q(engine_mod):
    """Size-independent properties at 26 qubits: a random circuit followed by
    its inverse returns |0..0>; a permutation and its inverse are bit-exact."""
    n = 26
    gates = random_layers_circuit(n, 4, seed=26)
    inv = [(m.conj().T, t, c) for (m, t, c) in reversed(gates)]
    g = make_engine(1, n, 5)
    g.allocate_qureg(list(range(n)), 0)
    apply_circuit_inorder(g, gates, 5)
    assert g.get_probability([0] * n, list(range(n))) < 0.5
    rng = np.random.default_rng(1)
    before = g.get_amplitude([1, 0] * 13, list(range(n)))
    perm = random_permutation(list(range(n)), rng)
    g.permute_qubits(perm)
    g.permute_qubits(list(range(n)))
    assert g.get_amplitude([1, 0] * 13, list(range(n))) == before
    apply_circuit_inorder(g, inv, 5)
    assert abs(g.get_amplitude([0] * n, list(range(n))) - 1.0) < 1e-10
    assert abs(g.get_probability([0] * n, list(range(n))) - 1.0) < 1e-10


def test_benchmark_circuit_round_trip_30q(engine_mod):
    """At the benchmark size: the 30-qubit, 20-layer complex128 circuit in
    fused passes of at most 4 qubits (the GreedyScheduler loop), a random
    qubit permutation, the inverse circuit and the inverse permutation return
    a random state to itself within 1e-12 (compared on the device in FP64)."""
    from hiqsimulator_amd.scheduling import apply_scheduled
    n = 30
    gates = random_layers_circuit(n, 20, 4321)
    inv = [(m.conj().T, t, c) for (m, t, c) in reversed(gates)]
    print("benchmark circuit round trip 30q", flush=True)
    g = make_engine(1, n, 4)
    g.allocate_qureg(list(range(n)), 0)
    g.init_random_state(31)
    g.snapshot()
    apply_scheduled(g, gates, 4)
    g.permute_qubits(random_permutation(list(range(n)), np.random.default_rng(2)))
    apply_scheduled(g, inv, 4)
    g.permute_qubits(list(range(n)))
    r = g.compare_snapshot()
    print("benchmark circuit round trip 30q: max-abs %.3e, 1 - fidelity %.3e"
          % (r["max_abs_diff"], 1 - r["fidelity"]), flush=True)
    assert r["max_abs_diff"] <= 1e-12
    assert 1 - r["fidelity"] <= 1e-12


def test_measurement_and_sampling_30q(engine_mod):
    """Size-independent checks at 30 qubits (16 GiB complex128), with the
    top bit involved: H on qubits 0, 15 and 29 of |0..0> gives 8 equally
    likely outcomes.  Probabilities are 1/8 each; 8192 samples fall only on
    those outcomes with frequencies that pass a chi-square test; measuring
    qubit 29 collapses it (probability 1 afterwards) and keeps the norm."""
    n = 30
    print("measurement and sampling 30q", flush=True)
    g = make_engine(77, n, 4)
    g.allocate_qureg(list(range(n)), 0)
    for q in (0, 15, 29):
        g.apply_controlled_gate(H, [q], [])
        g.run()
    for bits in range(8):
        vals = [(bits >> i) & 1 for i in range(3)]
        assert abs(g.get_probability(vals, [0, 15, 29]) - 0.125) < 1e-12
    assert abs(g.entropy() - 3.0) < 1e-12  # uniform over 8 outcomes
    s = np.asarray(g.sample(8192, 2024), dtype=np.uint64)
    allowed = {(b & 1) | ((b >> 1 & 1) << 15) | ((b >> 2 & 1) << 29) for b in range(8)}
    assert set(int(x) for x in np.unique(s)) <= allowed
    counts = np.array([np.count_nonzero(s == a) for a in sorted(allowed)])
    chi2 = float(((counts - 1024.0) ** 2 / 1024.0).sum())
    assert chi2 < 24.3  # 7 degrees of freedom, p = 0.001
    (bit,) = g.measure_qubits([29])
    assert abs(g.get_probability([bit], [29]) - 1.0) < 1e-12
    assert abs(g.get_probability([], []) - 1.0) < 1e-12
    assert abs(g.get_probability([0, 0], [0, 15]) - 0.25) < 1e-12


def test_allocation_and_deallocation_30q(engine_mod):
    """Allocation and deallocation at 30 qubits: a 29-qubit state grows by
    one qubit (16 GiB), the new qubit is entangled with qubit 28 and
    measured, and deallocating it shrinks the state back with its norm and
    the other qubit's collapsed value kept."""
    print("allocation and deallocation 30q", flush=True)
    g = make_engine(5, 30, 4)
    g.allocate_qureg(list(range(29)), 0)
    g.apply_controlled_gate(H, [28], [])
    g.run()
    g.allocate_qubit(29)
    g.apply_controlled_gate(X, [29], [28])  # (|00> + |11>) / sqrt(2) on 28, 29
    g.run()
    assert abs(g.get_probability([1, 1], [28, 29]) - 0.5) < 1e-12
    assert abs(g.get_probability([0, 1], [28, 29])) < 1e-12
    with pytest.raises(RuntimeError):
        g.deallocate_qubit(29)  # entangled
    (bit,) = g.measure_qubits([28])
    assert abs(g.get_probability([bit], [29]) - 1.0) < 1e-12
    g.deallocate_qubit(29)
    assert len(g.get_qubits_ids()) == 29
    assert abs(g.get_probability([], []) - 1.0) < 1e-12
    assert abs(g.get_probability([bit], [28]) - 1.0) < 1e-12
    assert abs(g.get_amplitude([0] * 28 + [bit], list(range(29))) - 1.0) < 1e-12


def test_ghz_expectations_30q(engine_mod):
    """Pauli expectations and a Pauli operator at 30 qubits against the
    closed forms of the GHZ state (|0..0> + |1..1>) / sqrt(2):
    <Z0 Z29> = 1, <X..X> = 1, <Y0 Y1 X..X> = -1, <Z0> = 0, and X on every
    qubit maps the state to itself."""
    n = 30
    print("ghz expectations 30q", flush=True)
    g = make_engine(3, n, 4)
    g.allocate_qureg(list(range(n)), 0)
    g.apply_controlled_gate(H, [0], [])
    g.run()
    for q in range(1, n):
        g.apply_controlled_gate(X, [q], [q - 1])
        g.run()
    ids = list(range(n))

    def ev(term):
        return g.get_expectation_value([(term, 1.0)], ids)

    assert abs(ev([(0, "Z"), (29, "Z")]) - 1.0) < 1e-12
    assert abs(ev([(q, "X") for q in range(n)]) - 1.0) < 1e-12
    assert abs(ev([(0, "Y"), (1, "Y")] + [(q, "X") for q in range(2, n)]) + 1.0) < 1e-12
    assert abs(ev([(0, "Z")])) < 1e-12
    assert abs(g.entropy() - 1.0) < 1e-12  # two outcomes of probability 1/2
    g.apply_qubit_operator([([(q, "X") for q in range(n)], 1.0)], ids)
    assert abs(g.get_probability([1] * n, ids) - 0.5) < 1e-12
    assert abs(g.get_amplitude([0] * n, ids) - 1.0 / math.sqrt(2.0)) < 1e-12


def test_time_evolution_30q(engine_mod):
    """exp(-i t H) at 30 qubits against a closed form: H = c1 Z0 + c2 X0 on
    |0>, controlled by qubit 29.  With qubit 29 set, P(qubit 0 = 1) =
    c2^2 / (c1^2 + c2^2) * sin^2(t * sqrt(c1^2 + c2^2)); with it clear the
    state is unchanged."""
    n = 30
    c1, c2, t = 0.3, 0.7, 1.9
    w = math.sqrt(c1 * c1 + c2 * c2)
    expect = (c2 / w) ** 2 * math.sin(t * w) ** 2
    print("time evolution 30q", flush=True)
    for ctrl in (1, 0):
        g = make_engine(4, n, 4)
        g.allocate_qureg(list(range(n)), 0)
        if ctrl:
            g.apply_controlled_gate(X, [29], [])
            g.run()
        g.emulate_time_evolution([([(0, "Z")], c1), ([(0, "X")], c2)], t, [0], [29])
        p1 = g.get_probability([1], [0])
        assert abs(p1 - (expect if ctrl else 0.0)) < 1e-10, (ctrl, p1, expect)
        assert abs(g.get_probability([], []) - 1.0) < 1e-10


def test_snapshot_restore(engine_mod):
    n = 18
    g = make_engine(1, n, 4)
    g.allocate_qureg(list(range(n)), 0)
    apply_circuit_inorder(g, random_layers_circuit(n, 3, seed=4), 4)
    before = g.get_state()
    g.snapshot()
    apply_circuit_inorder(g, random_layers_circuit(n, 2, seed=5), 4)
    g.permute_qubits(list(reversed(range(n))))
    assert not np.array_equal(g.get_state(), before)
    g.restore()
    assert g.get_qubits_ids() == list(range(n))
    assert np.array_equal(g.get_state(), before)


def test_deferred_clusters_flush_at_every_read(engine_mod, oracle_mod):
    """Clusters queued by run() (deferred mode, 14 qubits) are applied before
    every call that reads the state or changes the qubit set or order:
    probabilities, amplitudes, measurement, expectation, permutation,
    allocation and deallocation all see the same state as the oracle, which
    applies each cluster at once; a cluster with a 5-qubit control set takes
    the immediate control-mask pass after the queue."""
    n = 14
    gates = random_layers_circuit(n, 6, seed=31)
    g = make_engine(77, n + 1, 4)
    o = oracle_mod.OracleSim(77, n + 1, 4)
    assert g.pass_mode() == (1, 0)
    for s in (g, o):
        s.allocate_qureg(list(range(n)), 0)
        apply_circuit_inorder(s, gates[:60], 4)
    assert g.pass_mode()[1] > 0  # clusters are waiting
    ids = list(range(n))
    assert g.get_probability([1, 0, 1], [0, 5, 13]) == pytest.approx(
        o.get_probability([1, 0, 1], [0, 5, 13]), abs=1e-13)
    assert g.pass_mode()[1] == 0
    for s in (g, o):
        apply_circuit_inorder(s, gates[60:120], 4)
        s.apply_controlled_gate(X, [3], [0, 1, 2, 4, 5])  # 6 qubits: control-mask pass
        s.run()
    assert g.get_amplitude([0] * n, ids) == pytest.approx(o.get_amplitude([0] * n, ids), abs=1e-13)
    for s in (g, o):
        apply_circuit_inorder(s, gates[120:200], 4)
    assert g.measure_qubits([2, 7]) == o.measure_qubits([2, 7])
    for s in (g, o):
        apply_circuit_inorder(s, gates[200:260], 4)
    term = [([(1, "X"), (4, "Y"), (9, "Z")], 0.7)]
    assert g.get_expectation_value(term, ids) == pytest.approx(o.get_expectation_value(term, ids), abs=1e-12)
    perm = random_permutation(ids, np.random.default_rng(3))
    for s in (g, o):
        apply_circuit_inorder(s, gates[260:320], 4)
        s.permute_qubits(perm)
        apply_circuit_inorder(s, gates[320:], 4)
        s.allocate_qubit(n)
        s.apply_controlled_gate(H, [n], [])
        s.run()
    assert np.abs(g.get_state() - o.get_state()).max() <= 1e-12
    (bit,) = g.measure_qubits([n])
    assert o.measure_qubits([n]) == [bit]
    for s in (g, o):
        s.deallocate_qubit(n)
    assert np.abs(g.get_state() - o.get_state()).max() <= 1e-12
    assert g.entropy() == pytest.approx(o.entropy(), abs=1e-12)
