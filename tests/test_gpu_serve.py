"""Request loop (continuous batching) on the GPU.

* Each request's result is the same alone and in the mix: samples equal
  bit for bit, expectations equal (the passes, their arithmetic and the
  Philox stream depend on the request alone).
* Results against independent paths: expectations against the oracle
  (complex128 <= 1e-12, complex64 <= 1e-5), samples equal to
  SimulatorMPI.sample on the same circuit and seed (complex128), sample
  frequencies against the oracle's probabilities.
* Admission under memory pressure: a state pool that holds only a few
  states at a time still completes every request, with unchanged results.
"""
import numpy as np
import pytest

from hiqsimulator_amd.circuits import apply_circuit_inorder, random_layers_circuit
from hiqsimulator_amd.scheduling import plan_clusters
from hiqsimulator_amd.serving import (make_server, random_pauli_strings, result, run_server,
                                      serving_workload, submit)

pytestmark = pytest.mark.gpu


def small_requests(seed, count=14):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(rng.integers(3, 15))
        gates = random_layers_circuit(n, int(rng.integers(2, 30)), int(rng.integers(1 << 30)))
        if i % 2 == 0:
            reqs.append(dict(arrival=0.0, n=n, gates=gates, kind="samples", shots=1024,
                             seed=1000 + i, terms=[]))
        else:
            reqs.append(dict(arrival=0.0, n=n, gates=gates, kind="expectation", shots=0, seed=0,
                             terms=random_pauli_strings(n, 8, rng)))
    return reqs


def serve_all(reqs, precision, **kw):
    s = make_server(precision, state_bytes=kw.pop("state_bytes", 1 << 30),
                    matrix_bytes=256 << 20, **kw)
    ids = [submit(s, r) for r in reqs]
    while not s.idle():
        s.step(block=True)
    return [result(s, rid, r) for rid, r in zip(ids, reqs)]


def oracle_state_sim(oracle_mod, r):
    o = oracle_mod.OracleSim(1, r["n"], 5)
    o.allocate_qureg(list(range(r["n"])), 0)
    apply_circuit_inorder(o, r["gates"], 5)
    return o


@pytest.mark.parametrize("precision,prog_min", [("c128", 31), ("c64", 31), ("c128", 12)])
def test_same_alone_and_in_mix(engine_mod, oracle_mod, precision, prog_min):
    reqs = small_requests(3)
    print("serve mix", precision, "prog_min", prog_min, flush=True)
    mix = serve_all(reqs, precision, passes_per_step=3, prog_min_qubits=prog_min)
    for i, r in enumerate(reqs):
        alone = serve_all([r], precision, passes_per_step=5, prog_min_qubits=prog_min)[0]
        if r["kind"] == "samples":
            assert np.array_equal(mix[i], alone), i
        else:
            assert np.array_equal(mix[i], alone), i
            o = oracle_state_sim(oracle_mod, r)
            ref = np.array([o.expectation_term(x, z) for x, z in r["terms"]])
            tol = 1e-12 if precision == "c128" else 1e-5
            assert np.abs(mix[i] - ref).max() <= tol, (i, np.abs(mix[i] - ref).max())


@pytest.mark.parametrize("prog_min", [31, 12])
def test_edge_requests(engine_mod, prog_min):
    """Requests with no gates, 1-qubit requests and a 12-qubit one with a
    single gate, on the fused-pass and tile-program classes: basis states
    give one outcome per shot, and the expectations are exact."""
    from hiqsimulator_amd.circuits import H, X
    reqs = [
        dict(arrival=0.0, n=1, gates=[], kind="samples", shots=64, seed=5, terms=[]),
        dict(arrival=0.0, n=13, gates=[], kind="samples", shots=64, seed=6, terms=[]),
        dict(arrival=0.0, n=12, gates=[(X, [11], [])], kind="samples", shots=64, seed=7, terms=[]),
        dict(arrival=0.0, n=1, gates=[(H, [0], [])], kind="expectation", shots=0, seed=0,
             terms=[(0, 1), (1, 0), (1, 1)]),  # Z, X, Y on |+>
        dict(arrival=0.0, n=13, gates=[], kind="expectation", shots=0, seed=0,
             terms=[(0, 1 << 12), (1 << 12, 0)]),  # Z, X on |0>
    ]
    print("serve edge requests prog_min", prog_min, flush=True)
    out = serve_all(reqs, "c128", passes_per_step=4, prog_min_qubits=prog_min)
    assert np.array_equal(out[0], np.zeros(64, dtype=np.uint64))
    assert np.array_equal(out[1], np.zeros(64, dtype=np.uint64))
    assert np.array_equal(out[2], np.full(64, 1 << 11, dtype=np.uint64))
    assert np.abs(np.asarray(out[3]) - np.array([0.0, 1.0, 0.0])).max() <= 1e-12
    assert np.abs(np.asarray(out[4]) - np.array([1.0, 0.0])).max() <= 1e-12


def test_samples_match_engine_sampler(engine_mod):
    reqs = [r for r in small_requests(4) if r["kind"] == "samples"]
    got = serve_all(reqs, "c128", prog_min_qubits=31)
    for r, g in zip(reqs, got):
        n = r["n"]
        sim = engine_mod.SimulatorMPI(1, n, 4)
        ids = list(range(n))
        sim.allocate_qureg(ids, 0)
        for cl in plan_clusters(r["gates"], ids, 4, mode="lookahead"):
            for j in cl:
                sim.apply_controlled_gate(*r["gates"][j])
            sim.run()
        assert np.array_equal(sim.sample(r["shots"], r["seed"]), g)


def test_sample_frequencies(engine_mod, oracle_mod):
    n = 5
    gates = random_layers_circuit(n, 6, 77)
    r = dict(arrival=0.0, n=n, gates=gates, kind="samples", shots=1 << 16, seed=5, terms=[])
    got = serve_all([r], "c128")[0]
    p = np.abs(oracle_state_sim(oracle_mod, r).get_state()) ** 2
    counts = np.bincount(got.astype(np.int64), minlength=1 << n)
    assert counts[p < 1e-15].sum() == 0
    expected = p * r["shots"]
    mask = expected > 5
    chi2 = (((counts - expected) ** 2)[mask] / expected[mask]).sum()
    # 32 cells: the 0.999 quantile of chi-square with 31 degrees of freedom is ~61
    assert chi2 < 61, chi2
    sim = engine_mod.SimulatorMPI(1, n, 4)
    sim.allocate_qureg(list(range(n)), 0)
    apply_circuit_inorder(sim, gates, 4)
    a = sim.sample(4096, 9)
    b = sim.sample(4096, 9)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, sim.sample(4096, 10))
    assert np.abs(sim.get_state() - oracle_state_sim(oracle_mod, r).get_state()).max() <= 1e-12


def test_admission_under_memory_pressure(engine_mod):
    reqs = small_requests(6, count=10)
    free = serve_all(reqs, "c64")
    # room for about two 14-qubit states at a time (2^14 * 8 B = 128 KiB each)
    tight = serve_all(reqs, "c64", state_bytes=300 << 10, passes_per_step=2)
    for a, b in zip(free, tight):
        assert np.array_equal(a, b)


def test_poisson_stream_small(engine_mod):
    reqs = serving_workload(40, rate=400.0, seed=8, n_range=(8, 16), depth_range=(5, 20))
    s = make_server("c64", state_bytes=1 << 30, matrix_bytes=256 << 20)
    summary, results = run_server(s, reqs, keep_results=True)
    print("poisson small", summary, flush=True)
    assert len(results) == len(reqs)
    assert summary["p99_ms"] >= summary["p50_ms"] > 0
    assert s.stats()["completed"] == len(reqs)


def test_server_errors(engine_mod):
    s = make_server("c64", state_bytes=1 << 20, matrix_bytes=1 << 20)
    with pytest.raises(RuntimeError):
        s.submit_samples(20, [], 1024, 1)  # state does not fit the pool
    with pytest.raises(RuntimeError):
        s.submit_samples(4, [], 0, 1)  # no shots
    with pytest.raises(RuntimeError):
        s.submit_expectation(4, [], [1 << 5], [0])  # term beyond the register
    rid = s.submit_samples(3, [], 16, 1)
    with pytest.raises(RuntimeError):
        s.samples(rid)  # not done yet
    while not s.idle():
        s.step(block=True)
    assert np.array_equal(s.samples(rid), np.zeros(16, dtype=np.uint64))
    with pytest.raises(RuntimeError):
        s.expectations(rid)


@pytest.mark.parametrize("heavy", [14, 31])
def test_heavy_requests_one_at_a_time(engine_mod, heavy):
    """Requests of at least heavy_qubits qubits are admitted one at a time;
    the results do not depend on the limit."""
    rng = np.random.default_rng(21)
    reqs = []
    for i in range(3):
        gates = random_layers_circuit(14, 20, int(rng.integers(1 << 30)))
        reqs.append(dict(arrival=0.0, n=14, gates=gates, kind="samples", shots=256, seed=50 + i,
                         terms=[]))
    print("serve heavy limit", heavy, flush=True)
    # fused passes (planned at submit), so that every request can be admitted at the first step
    s = make_server("c64", state_bytes=1 << 30, matrix_bytes=256 << 20, passes_per_step=2,
                    heavy_qubits=heavy, prog_min_qubits=31)
    ids = [submit(s, r) for r in reqs]
    while not s.idle():
        s.step(block=True)
    got = [result(s, rid, r) for rid, r in zip(ids, reqs)]
    peak = s.stats()["peak_active"]
    if heavy == 14:
        assert peak == 1
    else:
        assert peak == 3
    ref = serve_all(reqs, "c64", passes_per_step=2, heavy_qubits=31, prog_min_qubits=31)
    for a, b in zip(got, ref):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    with pytest.raises(RuntimeError):
        s.set_heavy_qubits(0)
