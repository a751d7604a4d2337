"""The metric's workloads on the GPU at their real shapes.

* 30-qubit complex128, 40 layers, 8 uniform qubit permutations (the headline
  step of bench.py), in fused passes of at most k = 2, 3, 4, 5 qubits
  (k = 5 both as one-op tile passes on the matrix cores and in the dense
  forms): one step and its exact inverse return a random state to itself
  within 1e-12 (compared on the device in FP64).
* The same step through the drop-in flow with deferred clusters (run()
  queues, tile passes apply): equal to the one-pass-per-run() result within
  1e-12; the circuit alone takes at most 60 HBM passes instead of 196.
* Qubit permutations at 28 qubits, index by index: every amplitude carries
  its own index as a tag and lands where the closed form out[sigma(i)] = in[i]
  puts it (bit-exact, both precisions).
* 4096-circuit batch workload shape (n ~ U{10..22}, depth ~ U{10..60}, Haar
  2-qubit unitaries every 5th layer) with the stored autotuned plans: every
  circuit of 18 or more qubits and every third smaller one against the
  oracle (complex128 <= 1e-12, complex64 fidelity >= 1 - 1e-6).
* Serving shape in complex64 on the tile-program route (requests of 12 to
  20 qubits, depth 20..100): every served final state against the oracle,
  fidelity >= 1 - 1e-6; the same request alone gives the same state bit for
  bit.
"""
import numpy as np
import pytest

import bench
from hiqsimulator_amd.circuits import random_layers_circuit, random_permutation

pytestmark = pytest.mark.gpu

N30 = 30


def fidelity(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


@pytest.mark.parametrize("k,dense5_tile", [(2, 1), (3, 1), (4, 1), (5, 1), (5, 0)],
                         ids=["k2", "k3", "k4", "k5-tile", "k5-dense"])
def test_headline_step_round_trip_30q(engine_mod, k, dense5_tile):
    from hiqsimulator_amd import _cppsim_amd
    gates, step, _ = bench.build_step(N30, 40, k, 1234, 8)
    n_pass = sum(1 for kind, _ in step if kind == "pass")
    print("headline step round trip 30q k=%d dense5_tile=%d: %d passes" % (k, dense5_tile, n_pass),
          flush=True)
    old = _cppsim_amd.get_tuning("dense5_tile")
    _cppsim_amd.set_tuning("dense5_tile", dense5_tile)
    try:
        g = engine_mod.SimulatorMPI(1, N30, k)
        g.set_pass_mode(0)
        g.allocate_qureg(list(range(N30)), 0)
        g.init_random_state(7)
        g.snapshot()
        start = g.get_qubits_ids()
        g.set_timing(True)
        bench.run_step(g, step)
        counts = [g.kernel_time(c)[0] for c in range(5)]
        g.set_timing(False)
        assert sum(counts) == n_pass and counts[k - 1] > 0, counts
        bench.run_step(g, bench.inverse_step(step, start))
        assert g.get_qubits_ids() == start
        r = g.compare_snapshot()
    finally:
        _cppsim_amd.set_tuning("dense5_tile", old)
    print("  max-abs %.3e, 1 - fidelity %.3e" % (r["max_abs_diff"], 1 - r["fidelity"]), flush=True)
    assert r["max_abs_diff"] <= 1e-12
    assert 1 - r["fidelity"] <= 1e-12


def test_drop_in_deferred_step_matches_eager_30q(engine_mod):
    """run() in the default (deferred) mode: the same calls as the eager
    path, clusters applied as tile passes, same state within 1e-12."""
    gates, step, _ = bench.build_step(N30, 40, 4, 1234, 8)
    n_pass = sum(1 for kind, _ in step if kind == "pass")
    print("drop-in deferred step 30q", flush=True)
    g = engine_mod.SimulatorMPI(1, N30, 4)
    assert g.pass_mode()[0] == 1
    g.allocate_qureg(list(range(N30)), 0)
    g.set_pass_mode(0)
    g.init_random_state(7)
    bench.run_step(g, step)
    g.snapshot()
    # back to the identity order first: the random state is drawn per index,
    # so both runs start from the same state only in the same qubit order
    g.permute_qubits(list(range(N30)))
    g.init_random_state(7)
    g.set_pass_mode(1)
    g.set_timing(True)
    bench.run_step(g, step)
    tiles, _ = g.kernel_time(engine_mod.KERNEL_CLASSES.index("tile"))
    dense = sum(g.kernel_time(c)[0] for c in range(5))
    g.set_timing(False)
    r = g.compare_snapshot()
    print("  %d tile passes for %d clusters, max-abs %.3e" % (tiles, n_pass, r["max_abs_diff"]),
          flush=True)
    assert dense == 0 and 0 < tiles < n_pass
    assert r["max_abs_diff"] <= 1e-12
    assert 1 - r["fidelity"] <= 1e-12


def test_drop_in_circuit_hbm_passes_30q(engine_mod):
    """The 1780-gate circuit through the GreedyScheduler loop (one run() per
    cluster, 196 clusters) in the default mode: at most 60 HBM passes, and
    the inverse circuit through the same flow restores the random state."""
    from hiqsimulator_amd.scheduling import apply_scheduled
    gates = random_layers_circuit(N30, 40, 1234)
    inv = [(m.conj().T, t, c) for (m, t, c) in reversed(gates)]
    print("drop-in circuit passes 30q", flush=True)
    g = engine_mod.SimulatorMPI(1, N30, 4)
    g.allocate_qureg(list(range(N30)), 0)
    g.init_random_state(9)
    g.snapshot()
    g.set_timing(True)
    apply_scheduled(g, gates, 4)
    tiles, _ = g.kernel_time(engine_mod.KERNEL_CLASSES.index("tile"))
    g.set_timing(False)
    apply_scheduled(g, inv, 4)
    r = g.compare_snapshot()
    print("  %d tile passes, round trip max-abs %.3e" % (tiles, r["max_abs_diff"]), flush=True)
    assert tiles <= 60
    assert r["max_abs_diff"] <= 1e-12


def tags(n):
    i = np.arange(1 << n, dtype=np.int64)
    return (i & 0x3FFF).astype(np.float64) + 1j * (i >> 14).astype(np.float64)


def untag(v):
    return v.real.astype(np.int64) + (v.imag.astype(np.int64) << 14)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_permutation_tagged_amplitudes_28q(engine_mod, precision):
    """out index j holds the amplitude of in index i(j), bit p of i(j) = bit
    sigma[p] of j's source position: compared index by index for two random
    permutations in a row (the tags are exact in both precisions)."""
    n = 28
    print("tagged permutation 28q", precision, flush=True)
    rng = np.random.default_rng(28)
    g = engine_mod.SimulatorMPI(1, n, 4, precision)
    g.allocate_qureg(list(range(n)), 0)
    g.set_state(tags(n))
    order = list(range(n))
    for _ in range(2):
        perm = random_permutation(order, rng)
        g.permute_qubits(perm)
        order = perm
        got = untag(g.get_state())
        chunk = 1 << 24
        for j0 in range(0, 1 << n, chunk):
            j = np.arange(j0, j0 + chunk, dtype=np.int64)
            want = np.zeros_like(j)
            for p, q in enumerate(order):  # index bit p now holds qubit q (= its first position)
                want |= ((j >> p) & 1) << q
            assert np.array_equal(got[j0:j0 + chunk], want), (j0, precision)
        del got


def test_batch_workload_against_oracle(engine_mod, oracle_mod):
    from hiqsimulator_amd.batching import batch_workload, load_tuning, make_batch
    circuits = batch_workload(48, 2024, n_range=(10, 22))
    check = [i for i, (n, _) in enumerate(circuits) if n >= 18]
    check += [i for i, (n, _) in enumerate(circuits) if n < 18][::3]
    print("batch workload: %d circuits, %d checked (n %s)" % (len(circuits), len(check),
          sorted(circuits[i][0] for i in check)), flush=True)
    ref = oracle_mod.circuit_states([circuits[i] for i in check])
    print("  oracle states done", flush=True)
    for prec in ("c128", "c64"):
        t = load_tuning(prec)
        assert t is not None, "stored tuning table missing for " + prec
        b = make_batch(circuits, prec, t["k_by_n"], t["wg_max_qubits"],
                       tile_min_qubits=t["tile_min_qubits"], prog_min_qubits=t["prog_min_qubits"])
        b.run()
        st = b.stats()
        worst_abs, worst_fid = 0.0, 1.0
        for i, r in zip(check, ref):
            s = b.get_state(i)
            worst_abs = max(worst_abs, float(np.abs(s - r).max()))
            worst_fid = min(worst_fid, fidelity(r, s))
        print("  %s: %d tile-program passes, max-abs %.3e, min fidelity 1-%.3e"
              % (prec, st["tile_passes"], worst_abs, 1 - worst_fid), flush=True)
        if prec == "c128":
            assert worst_abs <= 1e-12
        else:
            assert worst_fid >= 1 - 1e-6
        del b


def test_serving_c64_tile_programs_against_oracle(engine_mod, oracle_mod):
    from hiqsimulator_amd.serving import make_server, result, serving_workload, submit
    reqs = serving_workload(20, rate=1000.0, seed=44, n_range=(12, 20), depth_range=(20, 100))
    for r in reqs:
        r["kind"] = "state"
    print("serving c64 tile programs: n %s" % sorted(r["n"] for r in reqs), flush=True)
    ref = oracle_mod.circuit_states([(r["n"], r["gates"]) for r in reqs])
    print("  oracle states done", flush=True)
    s = make_server("c64", state_bytes=1 << 30, matrix_bytes=256 << 20, prog_min_qubits=12)
    ids = [submit(s, r) for r in reqs]
    while not s.idle():
        s.step(block=True)
    got = [result(s, rid, r) for rid, r in zip(ids, reqs)]
    fids = [fidelity(a, b) for a, b in zip(ref, got)]
    print("  min fidelity 1-%.3e, max-abs %.3e"
          % (1 - min(fids), max(float(np.abs(a - b).max()) for a, b in zip(ref, got))), flush=True)
    assert min(fids) >= 1 - 1e-6
    for i in (0, len(reqs) // 2, len(reqs) - 1):
        alone = make_server("c64", state_bytes=1 << 30, matrix_bytes=256 << 20, prog_min_qubits=12)
        rid = submit(alone, reqs[i])
        while not alone.idle():
            alone.step(block=True)
        assert np.array_equal(result(alone, rid, reqs[i]), got[i]), i
