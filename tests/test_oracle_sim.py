"""Pins the oracle's SimulatorMPI restatement (CPU only).

* the reference test-suite's known answers (tests/scenarios.py);
* fusion: the fused matrix equals the sequential product of the gates
  embedded by brute force (numpy kron / permutation), including control
  folding (fusion_mpi.hpp:101-226);
* measurement RNG: std::mt19937's 10000th output for the default seed is
  4123659995 (C++ standard [rand.predef]); the u stream equals an
  independent numpy restatement of MT19937 + generate_canonical;
* measurement sampling: the empirical outcome distribution matches |psi|^2.
"""
import math

import numpy as np
import pytest

import scenarios
from hiqsimulator_amd.circuits import H, X, haar_unitary, random_layers_circuit


def make_oracle(seed, max_local, max_cluster):
    import oracle
    return oracle.OracleSim(seed, max_local, max_cluster)


@pytest.mark.parametrize("scenario", scenarios.ALL_KNOWN_ANSWERS, ids=lambda f: f.__name__)
def test_known_answers(oracle_mod, scenario):
    scenario(make_oracle)


def test_ghz_measurement(oracle_mod):
    seen = set()
    for seed in range(12):
        res = scenarios.ghz_measure(make_oracle, seed)
        seen.add(sum(res))
    assert seen == {0, 5}


def embed(m, targets, n):
    """2^n x 2^n matrix of gate m acting on bit positions `targets`."""
    k = len(targets)
    dim = 1 << n
    full = np.zeros((dim, dim), dtype=np.complex128)
    for col in range(dim):
        j = 0
        for l, t in enumerate(targets):
            j |= ((col >> t) & 1) << l
        base = col
        for t in targets:
            base &= ~(1 << t)
        for i in range(1 << k):
            row = base
            for l, t in enumerate(targets):
                row |= ((i >> l) & 1) << t
            full[row, col] += m[i, j]
    return full


def controlled(m, nctrl):
    d = m.shape[0]
    big = np.eye(d << nctrl, dtype=np.complex128)
    big[-d:, -d:] = m
    return big


def test_fusion_against_bruteforce(oracle_mod):
    rng = np.random.default_rng(42)
    for trial in range(30):
        ids_pool = [3, 8, 11, 20, 21]
        gates = []
        for g in range(int(rng.integers(1, 7))):
            k = int(rng.integers(1, 3))
            ids = [int(x) for x in rng.choice(ids_pool, size=k, replace=False)]
            gates.append((haar_unitary(1 << k, rng), ids))
        f = oracle_mod.OracleFusion()
        for m, ids in gates:
            f.insert(m, ids)
        M, fids, ctrls, diag = f.perform()
        assert fids == sorted(set(q for _, ids in gates for q in ids))
        assert ctrls == [] and not diag
        n = len(fids)
        expect = np.eye(1 << n, dtype=np.complex128)
        for m, ids in gates:
            expect = embed(m, [fids.index(q) for q in ids], n) @ expect
        assert np.abs(M - expect).max() < 1e-13


def test_fusion_controls(oracle_mod):
    """First item's controls stay a mask; a later item without them folds
    them into every item as the highest bits (fusion_mpi.hpp:190-226)."""
    rng = np.random.default_rng(1)
    u = haar_unitary(2, rng)
    v = haar_unitary(2, rng)
    f = oracle_mod.OracleFusion()
    f.insert(u, [5], [9])
    M, ids, ctrls, diag = f.perform()
    assert ids == [5] and ctrls == [9] and np.allclose(M, u)
    f = oracle_mod.OracleFusion()
    f.insert(u, [5], [9])
    f.insert(v, [2], [])
    M, ids, ctrls, diag = f.perform()
    assert ids == [2, 5, 9] and ctrls == []
    cu = embed(controlled(u, 1), [1, 2], 3)  # u on id 5 (bit 1), control id 9 (bit 2)
    expect = embed(v, [0], 3) @ cu
    assert np.abs(M - expect).max() < 1e-13


def test_fusion_factor_and_diag(oracle_mod):
    f = oracle_mod.OracleFusion()
    f.insert(np.array([[2.0j]]), [])
    f.insert(np.diag([1.0, 1j]), [4])
    M, ids, ctrls, diag = f.perform()
    assert ids == [4] and diag
    assert np.allclose(M, np.diag([2.0j, -2.0]))


def _mt19937_numpy(seed, count):
    """MT19937 (Matsumoto & Nishimura 1998), 32-bit output."""
    mt = [0] * 624
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
    idx = 624
    out = []
    for _ in range(count):
        if idx >= 624:
            for i in range(624):
                y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        out.append(y)
    return out


def test_mt19937_known_answer(oracle_mod):
    assert oracle_mod.mt19937_nth(5489, 10000) == 4123659995


def test_uniform_stream_matches_generate_canonical(oracle_mod):
    # generate_canonical<double, 53>: k = 2 draws, (g0 + g1 * 2^32) / 2^64,
    # clamped below 1; uniform_real_distribution(0, 1) returns it unchanged.
    for seed in (1, 12345, 2 ** 32 + 7):
        raw = _mt19937_numpy(seed % 2 ** 32, 40)
        expect = []
        for i in range(20):
            v = (raw[2 * i] + raw[2 * i + 1] * 2.0 ** 32) / 2.0 ** 64
            expect.append(min(v, math.nextafter(1.0, 0.0)))
        got = oracle_mod.uniform_stream(seed, 20)
        assert list(got) == expect


def test_measurement_distribution(oracle_mod):
    """Outcome frequencies of one-qubit measurements follow |psi|^2."""
    counts = 0
    trials = 400
    for seed in range(trials):
        sim = oracle_mod.OracleSim(seed, 33, 4)
        sim.allocate_qureg([0, 1], 0)
        sim.apply_controlled_gate(np.array([[math.sqrt(0.3), -math.sqrt(0.7)],
                                            [math.sqrt(0.7), math.sqrt(0.3)]]), [1], [])
        sim.run()
        counts += sim.measure_qubits([1])[0]
    p = counts / trials
    assert abs(p - 0.7) < 4 * math.sqrt(0.21 / trials)


def test_permutation_closed_form(oracle_mod):
    rng = np.random.default_rng(9)
    n = 6
    sim = oracle_mod.OracleSim(1, 33, 4)
    sim.allocate_qureg(list(range(n)), 0)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sim.set_state(st)
    new = [int(x) for x in rng.permutation(n)]
    sim.permute_qubits(new)
    got = sim.get_state()
    for i in range(1 << n):
        j = 0
        for p, q in enumerate(new):
            j |= ((i >> q) & 1) << p  # old bit q (qubit q) -> new bit p
        assert got[j] == st[i]
    assert sim.get_qubits_ids() == new


def test_random_circuit_state_is_normalised(oracle_mod):
    sim = oracle_mod.OracleSim(3, 33, 4)
    sim.allocate_qureg(list(range(10)), 0)
    from hiqsimulator_amd.circuits import apply_circuit_inorder
    apply_circuit_inorder(sim, random_layers_circuit(10, 6, seed=3), 4)
    assert abs(np.linalg.norm(sim.get_state()) - 1.0) < 1e-12


def test_expectation_matches_dense_operator(oracle_mod):
    rng = np.random.default_rng(4)
    n = 5
    sim = oracle_mod.OracleSim(1, 33, 4)
    sim.allocate_qureg(list(range(n)), 0)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    st /= np.linalg.norm(st)
    sim.set_state(st)
    P = {"X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.array([[1, 0], [0, -1]]), "I": np.eye(2)}
    for trial in range(20):
        ops = ["IXYZ"[int(x)] for x in rng.integers(0, 4, size=n)]
        term = [(i, o) for i, o in enumerate(ops) if o != "I"]
        full = np.array([[1.0]])
        for o in reversed(ops):  # qubit n-1 is the most significant bit
            full = np.kron(full, P[o])
        expect = np.vdot(st, full @ st).real
        assert sim.get_expectation_value([(term, 1.0)], list(range(n))) == pytest.approx(expect)


def test_qft_closed_form(oracle_mod):
    """QFT-like circuit == orthonormal inverse DFT with the output index
    bit-reversed (fixed here on the oracle; the GPU test relies on it)."""
    from hiqsimulator_amd.circuits import apply_circuit_inorder, qft_like_circuit
    for n in (3, 6):
        rng = np.random.default_rng(n)
        st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        st /= np.linalg.norm(st)
        o = oracle_mod.OracleSim(1, n, 4)
        o.allocate_qureg(list(range(n)), 0)
        o.set_state(st)
        apply_circuit_inorder(o, qft_like_circuit(n), 4)
        rev = np.array([int(format(i, "0%db" % n)[::-1], 2) for i in range(1 << n)])
        assert np.abs(o.get_state()[rev] - np.fft.ifft(st, norm="ortho")).max() < 1e-13


def test_greedy_cluster_plan_is_equivalent(oracle_mod):
    """The GreedyScheduler loop (ClusterScheduler clusters, one run each)
    reorders gates only past disjoint ones: same state as program order."""
    from hiqsimulator_amd.circuits import apply_circuit
    from hiqsimulator_amd.scheduling import apply_scheduled, plan_clusters
    n = 10
    gates = random_layers_circuit(n, 8, seed=21)
    for k in (2, 3, 4, 5):
        clusters = plan_clusters(gates, list(range(n)), k)
        assert sorted(i for c in clusters for i in c) == list(range(len(gates)))
        for c in clusters:
            qs = set(q for i in c for q in list(gates[i][1]) + list(gates[i][2]))
            assert len(qs) <= k and c == sorted(c)
        a = oracle_mod.OracleSim(1, n, k)
        b = oracle_mod.OracleSim(1, n, k)
        for s in (a, b):
            s.allocate_qureg(list(range(n)), 0)
        apply_scheduled(a, gates, k)
        apply_circuit(b, gates, flush_every=1)
        assert np.abs(a.get_state() - b.get_state()).max() < 1e-12


def test_cluster_scheduler_known_structure():
    """Small hand-checked cases of the selection rule: most takeable gates,
    ties to fewer qubits, a blocked qubit stops later gates on it."""
    import hiqsimulator_amd as hq
    # gates: H0, H1, CX(0->2), H3; cluster size 2 -> {0,1}: gates 0,1 (2 gates)
    # vs {0,2}: gates 0,2 (2 gates) -> tie, same size, first found wins.
    cs = hq.ClusterScheduler([[0], [1], [2], [3]], [[], [], [0], []], [False] * 4,
                             [0, 1, 2, 3], [], 2)
    assert cs.ScheduleCluster() == [0, 1]
    # with size 3, {0,1,2} takes H0, H1, CX(0->2).
    cs = hq.ClusterScheduler([[0], [1], [2], [3]], [[], [], [0], []], [False] * 4,
                             [0, 1, 2, 3], [], 3)
    assert cs.ScheduleCluster() == [0, 1, 2]
    # a skipped gate on qubit 1 blocks later gates on qubit 1.
    cs = hq.ClusterScheduler([[1], [0], [1]], [[2], [], []], [False] * 3, [0, 1, 2], [], 1)
    assert cs.ScheduleCluster() == [1]
    # SwapScheduler with every qubit local keeps all of them.
    ss = hq.SwapScheduler([[0], [1], [2]], [[], [0], []], [False] * 3, 10 ** 6, 3, True)
    assert ss.ScheduleSwap() == [0, 1, 2]
