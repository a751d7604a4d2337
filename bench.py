"""Benchmark: 30-qubit complex128 random circuit, fused k-qubit HBM passes.

Workload (one "step"): the 30-qubit, 40-layer random circuit (each layer: one
of {H, Rx, Ry, Rz} per qubit with random angles, then CNOT ladders), 1780
gates, fused into passes of at most k qubits (default k=4), with 8 uniformly
random qubit-index permutations (the in-HBM replacement of the reference's
qswap) at fixed layer boundaries.  The state (16 GiB) stays resident in HBM;
scheduling/fusion-cluster planning happens before the timed region.

metric value = algorithmic HBM bytes of all passes in the step (32 B per
amplitude per fused or permutation pass) / step wall time.
roofline     = the dominant kernel (dense k-qubit pass): 32 * 2^30 B per
               launch / its average launch duration from HIP events recorded
               on the engine's stream over the timed region.
cpu_baseline = the reference's own AVX2 kernels (oracle/_ref, OpenMP with
               OMP_PROC_BIND=spread, OMP_PLACES=cores) in a child process
               started before this one touches the GPU: per k = 2..5 the
               median of 5 passes on the 2^n-amplitude complex128 state, and
               the first 36 of the headline circuit's fused passes (about
               11 s), the whole circuit's time projected from them.
Also reported: single k-qubit passes at random placements in both precisions
(k_sweep_gbs, k_sweep_gbs_c64), the step at every fusion width (circuit_k_sweep), the drop-in
flow (run() per cluster with deferred tile passes), tile programs, the
QFT-like precision round trip, the 4096-circuit batch and the serving
streams (p50/p99 with 95% bootstrap intervals, served-state fidelity).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--k 4] [--qubits 30]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def plan_passes(gates, k):
    """In program order: a new pass starts when the next gate would widen the
    current one beyond k qubits (targets + controls)."""
    passes, cur, cur_q = [], [], set()
    for g in gates:
        q = set(g[1]) | set(g[2])
        if cur and len(cur_q | q) > k:
            passes.append(cur)
            cur, cur_q = [], set()
        cur.append(g)
        cur_q |= q
    if cur:
        passes.append(cur)
    return passes


def plan_greedy(gates, k, n):
    """The GreedyScheduler loop: ClusterScheduler clusters, one pass each."""
    from hiqsimulator_amd.scheduling import plan_clusters
    return [[gates[i] for i in c] for c in plan_clusters(gates, list(range(n)), k)]


def build_step(n, layers, k, seed, n_perm, planner="greedy"):
    from hiqsimulator_amd.circuits import random_layers_circuit, random_permutation
    rng = np.random.default_rng(seed + 1)
    gates = random_layers_circuit(n, layers, seed)
    # Permutations after every (layers // n_perm)-th layer boundary.
    bounds = set()
    if n_perm:
        stride = max(1, layers // n_perm)
        bounds = {stride * (i + 1) for i in range(n_perm)}
    ops, idx = [], 0
    ids = list(range(n))
    # Walk layers (gate counts per layer alternate: n + n//2, n + (n-1)//2).
    for layer in range(layers):
        cnt = n + len(range(layer % 2, n - 1, 2))
        chunk = gates[idx:idx + cnt]
        idx += cnt
        ops.append(("gates", chunk))
        if (layer + 1) in bounds:
            perm = random_permutation(ids, rng)
            ops.append(("perm", perm))
    assert idx == len(gates), (idx, len(gates))
    # Fuse within each stretch between permutations.
    plan = (lambda g: plan_greedy(g, k, n)) if planner == "greedy" else (lambda g: plan_passes(g, k))
    step = []
    stretch = []
    for kind, payload in ops:
        if kind == "gates":
            stretch.extend(payload)
        else:
            step.extend(("pass", p) for p in plan(stretch))
            stretch = []
            step.append(("perm", payload))
    step.extend(("pass", p) for p in plan(stretch))
    # the same circuit for tile passes: each stretch between permutations is
    # one apply_tiled call
    tiled, stretch = [], []
    for kind, payload in ops:
        if kind == "gates":
            stretch.extend(payload)
        else:
            tiled.append(("tiled", stretch))
            stretch = []
            tiled.append(("perm", payload))
    tiled.append(("tiled", stretch))
    return gates, step, tiled


def run_step(sim, step, op_width=4):
    passes = 0
    for kind, payload in step:
        if kind == "pass":
            for m, t, c in payload:
                sim.apply_controlled_gate(m, t, c)
            sim.run()
            passes += 1
        elif kind == "tiled":
            passes += sim.apply_tiled(payload, op_width)
        else:
            sim.permute_qubits(payload)
    return passes


def inverse_step(step, start_order):
    """The exact inverse of a step that starts with the qubits in
    start_order: passes inverted (gates reversed and conjugate-transposed) in
    reverse order, and each permutation undone by permuting back to the order
    before it."""
    orders = [list(start_order)] + [list(p) for kind, p in step if kind == "perm"]
    out, j = [], len(orders) - 1
    for kind, payload in reversed(step):
        if kind == "pass":
            out.append(("pass", [(m.conj().T, t, c) for (m, t, c) in reversed(payload)]))
        elif kind == "perm":
            j -= 1
            out.append(("perm", orders[j]))
        else:
            raise ValueError("inverse_step: unexpected entry %r" % kind)
    return out


def kernel_sweep(hq, n, reps=24, precision="c128"):
    """Per-pass GB/s of the dense kernel for k = 1..5, averaged over `reps`
    random placements (a pass's time depends on where its targets lie)."""
    from hiqsimulator_amd.circuits import haar_unitary
    rng = np.random.default_rng(5)
    sim = hq.SimulatorMPI(1, n, 5, precision)
    sim.allocate_qureg(list(range(n)), 0)
    amp_bytes = 16 if precision == "c128" else 8
    out = {}
    for k in range(1, 6):
        mats = [haar_unitary(1 << k, rng) for _ in range(reps)]
        poss = [[int(p) for p in rng.permutation(n)[:k]] for _ in range(reps)]
        sim.apply_matrix(mats[0], poss[0])  # warm
        sim.synchronize()
        sim.set_timing(True)
        for m, p in zip(mats, poss):
            sim.apply_matrix(m, p)
        cnt, ms = sim.kernel_time(k - 1)
        sim.set_timing(False)
        out[str(k)] = round(2.0 * amp_bytes * 2 ** n / (ms / cnt * 1e-3) / 1e9, 1)
    del sim
    return out


def cpu_info():
    """CPU model, sockets, physical cores and the cores this process may use."""
    model, phys = None, set()
    pid = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name" and model is None:
                    model = val
                elif key == "physical id":
                    pid = val
                elif key == "core id":
                    core = val
                elif not key and pid is not None:
                    phys.add((pid, core))
                    pid = core = None
        if pid is not None:
            phys.add((pid, core))
    except OSError:
        pass
    return {"model": model, "sockets": len({p for p, _ in phys}) or None,
            "physical_cores": len(phys) or None, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline_child(args):
    """The reference CPU path, run in a child process whose OpenMP runtime
    starts with OMP_PROC_BIND=spread and OMP_PLACES=cores (the reference's
    advice, _simulator_mpi.py:52-59): the reference's own AVX2 kernels
    (oracle/_ref, built with its release flags) dispatched as its Run() does.
    * per k = 2..5: one warm pass, then the median of --cpu-reps passes at
      random placements on the 2^n-amplitude complex128 state;
    * the headline circuit's fused passes (the reference's fusion, then one
      kernel pass per cluster), a bounded prefix of --cpu-circuit-passes of
      them, with the whole circuit's time projected from it; the step's
      qubit permutations have no counterpart on the reference's one-node path."""
    import oracle
    from hiqsimulator_amd.circuits import haar_unitary
    info = cpu_info()  # before OpenMP binds this thread to one core
    n = args.qubits
    if not oracle.RefKernels.available("intrin"):
        return {"value": None, "unit": "GB/s", "error": "oracle/_ref not built"}
    ref = oracle.RefKernels("intrin")
    t0 = time.perf_counter()
    h, _ = ref.new_state(1 << n)
    ref.fill(h, 2.0 ** (-n / 2))
    alloc_s = time.perf_counter() - t0
    rng = np.random.default_rng(3)
    per_k = {}
    for k in range(2, 6):
        times = []
        for rep in range(args.cpu_reps + 1):
            m = haar_unitary(1 << k, rng)
            pos = [int(p) for p in rng.permutation(n)[:k]]
            t0 = time.perf_counter()
            ref.apply(h, m, pos)
            if rep:
                times.append(time.perf_counter() - t0)
        med = float(np.median(times))
        per_k[str(k)] = {"median_s": round(med, 4), "gbs": round(32.0 * 2 ** n / med / 1e9, 2),
                         "passes": len(times)}
        print("cpu k=%d: %s" % (k, per_k[str(k)]), file=sys.stderr, flush=True)
    circuit = None
    if not args.no_cpu_circuit:
        t0 = time.perf_counter()
        gates, step, _ = build_step(n, args.layers, args.k, args.seed, args.perms, args.planner)
        plan_s = time.perf_counter() - t0
        all_passes = [payload for kind, payload in step if kind == "pass"]
        passes = all_passes[:max(1, args.cpu_circuit_passes)]
        ref.fill(h, 2.0 ** (-n / 2))
        t0 = time.perf_counter()
        for cl in passes:
            f = oracle.OracleFusion()
            for m, t, c in cl:
                mm = np.asarray(m, dtype=np.complex128)
                if c:  # controls folded as the highest matrix bits (fusion_mpi.hpp:167-187)
                    d = mm.shape[0]
                    full = np.eye(d << len(c), dtype=np.complex128)
                    full[-d:, -d:] = mm
                    mm = full
                f.insert(mm, list(t) + list(c))
            M, ids, _, diag = f.perform(cap_qubits=5)
            ref.apply(h, M, ids, 0, diag)
        el = time.perf_counter() - t0
        circuit = {"seconds": round(el, 2), "fused_passes_timed": len(passes),
                   "fused_passes": len(all_passes), "gates": len(gates),
                   "projected_circuit_seconds": round(el * len(all_passes) / len(passes), 2),
                   "fusion_k": args.k, "plan_seconds": round(plan_s, 2),
                   "gbs": round(32.0 * 2 ** n * len(passes) / el / 1e9, 2),
                   "note": "the first %d of the headline step's %d fused passes (reference fusion + "
                           "kernels), whole circuit projected; its %d qubit permutations have no "
                           "counterpart on the reference's one-node path"
                           % (len(passes), len(all_passes), sum(1 for kind, _ in step if kind == "perm"))}
        print("cpu circuit: %s" % circuit, file=sys.stderr, flush=True)
    ref.free_state(h)
    return {"value": per_k[str(args.k)]["gbs"], "unit": "GB/s", "cores": ref.num_threads(),
            "kind": "reference",
            "sample": "reference AVX2 kernels (oracle/_ref, -O3 -ffast-math -mavx2 -mfma, OpenMP) on a "
                      "2^%d-amplitude complex128 state: per k = 2..5 the median of %d passes at random "
                      "placements; value = k=%d" % (n, args.cpu_reps, args.k),
            "per_k": per_k, "circuit": circuit, "cpu": info,
            "omp": {v: os.environ.get(v) for v in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "state_alloc_s": round(alloc_s, 2)}


def cpu_baseline(args):
    """Runs cpu_baseline_child in a fresh Python process (before this process
    touches the GPU), so that OpenMP starts with the thread binding set."""
    import subprocess
    env = dict(os.environ)
    env["OMP_PROC_BIND"] = "spread"
    env["OMP_PLACES"] = "cores"
    if "OMP_NUM_THREADS" not in env:
        phys = cpu_info()["physical_cores"] or os.cpu_count() or 1
        env["OMP_NUM_THREADS"] = str(min(phys, len(os.sched_getaffinity(0))))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--qubits", str(args.qubits),
           "--layers", str(args.layers), "--k", str(args.k), "--seed", str(args.seed), "--perms",
           str(args.perms), "--planner", args.planner, "--cpu-reps", str(args.cpu_reps),
           "--cpu-circuit-passes", str(args.cpu_circuit_passes)]
    if args.no_cpu_circuit:
        cmd.append("--no-cpu-circuit")
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True)
    if r.returncode != 0:
        return {"value": None, "unit": "GB/s", "error": "cpu baseline exited with %d" % r.returncode}
    return json.loads(r.stdout.strip().splitlines()[-1])


def bench_precision(hq, n, log_fn):
    """The QFT-like circuit (H on each qubit, controlled phases 2*pi/2^k on
    every pair) and its inverse on a random normalised complex Gaussian
    state, in complex128 and complex64; the round trip is compared with the
    initial state on the device in FP64 (max-abs, fidelity, norm drift)."""
    from hiqsimulator_amd.circuits import qft_like_circuit
    from hiqsimulator_amd.scheduling import apply_scheduled
    gates = qft_like_circuit(n)
    inv = [(m.conj().T, t, c) for (m, t, c) in reversed(gates)]
    out = {"workload": "%dq QFT-like circuit (%d gates) and its inverse on a random state" % (n, len(gates))}
    for prec in ("c128", "c64"):
        g = hq.SimulatorMPI(1, n, 4, prec)
        g.allocate_qureg(list(range(n)), 0)
        g.init_random_state(n)
        g.snapshot()
        t0 = time.perf_counter()
        apply_scheduled(g, gates, 4)
        apply_scheduled(g, inv, 4)
        g.synchronize()
        el = time.perf_counter() - t0
        r = g.compare_snapshot()
        out[prec] = {"max_abs": r["max_abs_diff"], "one_minus_fidelity": max(0.0, 1.0 - r["fidelity"]),
                     "norm_drift": abs(r["state_norm2"] - r["ref_norm2"]), "seconds": round(el, 2)}
        del g
        log_fn("precision %s: %s" % (prec, out[prec]))
    return out


def bench_batch(hq, count, seed, log_fn, plain_sample=64):
    """Many-circuit batch: ``count`` circuits, n ~ U{10..22}, depth ~ U{10..60},
    a Haar 2-qubit unitary every 5th layer, all in one graph launch, complex64
    and complex128; plan from the stored autotuner table when present."""
    from hiqsimulator_amd.batching import (batch_workload, fidelity, load_tuning, make_batch,
                                           one_at_a_time)
    t0 = time.perf_counter()
    circuits = batch_workload(count, seed)
    log_fn("batch: %d circuits generated (%.1f s)" % (count, time.perf_counter() - t0))
    out = {"workload": "%d circuits, n~U{10..22}, depth~U{10..60}, Haar 2q every 5th layer"
                       % count, "gates": int(sum(len(g) for _, g in circuits))}
    # effective_gbs counts 2 * B * 2^n bytes per HBM pass (fused or tile program)
    sample_idx = list(range(0, count, max(1, count // 64)))[:64]
    states = {}
    for prec in ("c64", "c128"):
        tune = load_tuning(prec)
        k_by_n = tune["k_by_n"] if tune else {}
        wg_max = tune["wg_max_qubits"] if tune else 16
        tile_min = tune["tile_min_qubits"] if tune else 0
        prog_min = tune["prog_min_qubits"] if tune else 31
        t0 = time.perf_counter()
        b = make_batch(circuits, prec, k_by_n, wg_max, tile_min_qubits=tile_min,
                       prog_min_qubits=prog_min)
        build_s = time.perf_counter() - t0
        b.run()
        ms = min(b.run() for _ in range(3))
        st = b.stats()
        ab = 16 if prec == "c128" else 8
        states[prec] = [b.get_state(i) for i in sample_idx]
        out[prec] = {"circuits_per_s": round(count / (ms * 1e-3), 1), "batch_ms": round(ms, 3),
                     "effective_gbs": round(st["amp_passes"] * 2 * ab / (ms * 1e-3) / 1e9, 1),
                     "hbm_passes": st["passes"], "wg_circuits": st["wg_circuits"],
                     "lds_circuits": st["lds_circuits"], "layer_launches": st["layer_launches"],
                     "arena_gib": round(st["arena_bytes"] / 2 ** 30, 2), "build_s": round(build_s, 2),
                     "plan": "autotuned table" if tune else "default (k=4, wg_max_qubits=16)",
                     "wg_max_qubits": wg_max, "tile_min_qubits": tile_min,
                     "prog_min_qubits": prog_min, "tile_program_passes": st["tile_passes"],
                     "tile_program_ops": st["tile_ops"],
                     # algorithmic flops of the tile ops run on the matrix cores
                     "mfma_tflops": round(st["tile_mfma_flops"] / (ms * 1e-3) / 1e12, 2)}
        log_fn("batch %s: %s" % (prec, out[prec]))
        del b
    fids = [fidelity(a, b) for a, b in zip(states["c128"], states["c64"])]
    errs = [float(np.abs(a - b).max()) for a, b in zip(states["c128"], states["c64"])]
    out["c64_vs_c128"] = {"sample": len(fids), "min_fidelity": min(fids), "max_abs": max(errs)}
    sample = [circuits[i] for i in sample_idx[:plain_sample]]
    for mode in ("fused", "deferred", "tiled"):
        plain_s = one_at_a_time(sample, "c64", mode=mode)
        out["plain_one_at_a_time_c64_" + mode] = {"sample": len(sample),
                                                 "circuits_per_s": round(len(sample) / plain_s, 1)}
        log_fn("batch: plain path (%s) %s" % (mode, out["plain_one_at_a_time_c64_" + mode]))
    best = max(out["plain_one_at_a_time_c64_" + m]["circuits_per_s"] for m in ("fused", "deferred", "tiled"))
    out["c64_vs_best_plain"] = round(out["c64"]["circuits_per_s"] / best, 2)
    return out


def bootstrap_ci(lat_ms, q, reps=2000, seed=0):
    """95% bootstrap interval of the q-th percentile of lat_ms."""
    lat = np.asarray(lat_ms)
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, lat.size, size=(reps, lat.size))
    vals = np.percentile(lat[idx], q, axis=1)
    return [round(float(np.percentile(vals, 2.5)), 2), round(float(np.percentile(vals, 97.5)), 2)]


def served_fidelity(hq, reqs, log_fn, count=8, max_qubits=22):
    """A sample of a stream's circuits served again as state requests (in
    one mix, complex64 tile programs) against the complex128 engine on the
    same circuit (one fused pass per cluster): fidelity and max-abs."""
    from hiqsimulator_amd.scheduling import plan_clusters
    from hiqsimulator_amd.serving import make_server, result, submit
    pick = [r for r in reqs if r["n"] <= max_qubits][:count]
    s = make_server("c64", state_bytes=8 << 30, matrix_bytes=1 << 30)
    ids = [submit(s, dict(r, kind="state")) for r in pick]
    while not s.idle():
        s.step(block=True)
    fids, errs = [], []
    for rid, r in zip(ids, pick):
        got = result(s, rid, dict(r, kind="state"))
        n = r["n"]
        sim = hq.SimulatorMPI(1, n, 4, "c128")
        sim.set_pass_mode(0)
        sim.allocate_qureg(list(range(n)), 0)
        for cl in plan_clusters(r["gates"], list(range(n)), 4, mode="lookahead"):
            for j in cl:
                sim.apply_controlled_gate(*r["gates"][j])
            sim.run()
        ref = sim.get_state()
        del sim
        fids.append(float(abs(np.vdot(ref, got)) ** 2 / (np.vdot(ref, ref).real * np.vdot(got, got).real)))
        errs.append(float(np.abs(ref - got).max()))
    out = {"sample": len(pick), "qubits": [r["n"] for r in pick], "min_fidelity": min(fids),
           "one_minus_fidelity_max": max(0.0, 1.0 - min(fids)), "max_abs": max(errs),
           "reference": "complex128 engine, one fused pass per cluster", "target": "fidelity >= 1 - 1e-6"}
    log_fn("served fidelity: %s" % out)
    return out


def bench_serving(hq, count, rates, seed, log_fn):
    """Poisson request streams, n log-uniform in 12..28, depth ~ U{20..100},
    1024 samples or 8 Pauli expectations each, complex64; the request loop
    (tile programs from 12 qubits) against one-circuit-at-a-time
    SimulatorMPI (apply_tiled) on the same stream.  p50/p99 carry 95%
    bootstrap intervals."""
    from hiqsimulator_amd.serving import make_server, run_plain, run_server, serving_workload
    out = []
    for rate in rates:
        reqs = serving_workload(count, rate, seed)
        server = make_server("c64", state_bytes=64 << 30, matrix_bytes=4 << 30)
        srv, _ = run_server(server, reqs)
        st = server.stats()
        del server
        log_fn("serving rate %g: server %s" % (rate, {k: v for k, v in srv.items() if k != "lat_ms"}))
        plain, _ = run_plain(reqs, "c64", tiled=True)
        log_fn("serving rate %g: plain %s" % (rate, {k: v for k, v in plain.items() if k != "lat_ms"}))
        entry = {"rate_per_s": rate, "requests": count, "peak_active": st["peak_active"], "steps": st["steps"]}
        for name, res in (("server", srv), ("plain", plain)):
            lat = res.pop("lat_ms")
            entry[name] = {k: round(v, 3) for k, v in res.items()}
            entry[name]["p50_ci95_ms"] = bootstrap_ci(lat, 50)
            entry[name]["p99_ci95_ms"] = bootstrap_ci(lat, 99)
        out.append(entry)
    fid = served_fidelity(hq, serving_workload(count, rates[0], seed), log_fn)
    return {"streams": out, "fidelity": fid}


def bench_drop_in(hq, n, layers, k, seed, steps, log_fn):
    """The drop-in flow: the headline circuit through the GreedyScheduler
    loop (clusters planned beforehand), apply_controlled_gate per gate and
    run() per cluster, as the reference's Python engine drives its C++
    object.  In the default (deferred) mode the clusters run as tile passes;
    with pass mode 0 each is one fused pass, as in the reference."""
    from hiqsimulator_amd.circuits import random_layers_circuit
    from hiqsimulator_amd.scheduling import plan_clusters
    gates = random_layers_circuit(n, layers, seed)
    clusters = plan_clusters(gates, list(range(n)), k)
    out = {"workload": "%dq %d-layer circuit (%d gates), GreedyScheduler loop (%d clusters), "
                       "apply_controlled_gate + run() per cluster" % (n, layers, len(gates), len(clusters))}
    sim = hq.SimulatorMPI(seed, n, k)
    sim.allocate_qureg(list(range(n)), 0)

    def feed():
        for cl in clusters:
            for i in cl:
                m, t, c = gates[i]
                sim.apply_controlled_gate(m, t, c)
            sim.run()
        sim.synchronize()

    for mode, name in ((1, "deferred"), (0, "eager")):
        sim.set_pass_mode(mode)
        feed()  # warm
        sim.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            feed()
        el = (time.perf_counter() - t0) / steps
        tiles = sim.kernel_time(hq.KERNEL_CLASSES.index("tile"))[0] // steps
        dense = sum(sim.kernel_time(c)[0] for c in range(6)) // steps
        sim.set_timing(False)
        out[name] = {"ms_per_circuit": round(el * 1e3, 2), "hbm_passes_per_circuit": int(tiles + dense)}
    out["drop_in_ms_per_circuit"] = out["deferred"]["ms_per_circuit"]
    out["speedup_vs_eager"] = round(out["eager"]["ms_per_circuit"] / out["deferred"]["ms_per_circuit"], 3)
    # the deferred result equals the eager one: the circuit, then its inverse
    # eagerly, returns a random state to itself
    inv = [(m.conj().T, t, c) for (m, t, c) in reversed(gates)]
    sim.init_random_state(5)
    sim.snapshot()
    sim.set_pass_mode(1)
    for cl in clusters:
        for i in cl:
            sim.apply_controlled_gate(*gates[i])
        sim.run()
    sim.set_pass_mode(0)
    for cl in reversed(clusters):
        for i in reversed(cl):
            sim.apply_controlled_gate(*inv[len(gates) - 1 - i])
        sim.run()
    r = sim.compare_snapshot()
    out["deferred_then_eager_inverse_max_abs"] = float(r["max_abs_diff"])
    del sim
    log_fn("drop-in: %s" % out)
    return out


def bench_circuit_k_sweep(hq, n, layers, seed, perms, planner, ks, log_fn):
    """The headline step at every fusion width k: ms per step, fused passes,
    average pass time (HIP events) and GB/s, and the error of one step and
    its exact inverse (compared on the device in FP64)."""
    out = {}
    sim = hq.SimulatorMPI(seed, n, 5)
    sim.set_pass_mode(0)
    sim.allocate_qureg(list(range(n)), 0)
    for k in ks:
        gates, step, _ = build_step(n, layers, k, seed, perms, planner)
        n_pass = sum(1 for kind, _ in step if kind == "pass")
        sim.set_qubits_perm(list(range(n)))
        sim.init_random_state(k)
        run_step(sim, step)  # warm
        sim.synchronize()
        sim.set_timing(True)
        t0 = time.perf_counter()
        run_step(sim, step)
        sim.synchronize()
        el = time.perf_counter() - t0
        cnt = sum(sim.kernel_time(c)[0] for c in range(6))
        ms = sum(sim.kernel_time(c)[1] for c in range(6))
        sim.set_timing(False)
        start = list(sim.get_qubits_ids())
        sim.snapshot()
        run_step(sim, step)
        run_step(sim, inverse_step(step, start))
        r = sim.compare_snapshot()
        out[str(k)] = {"ms_per_step": round(el * 1e3, 1), "fused_passes": n_pass,
                       "avg_pass_ms": round(ms / cnt, 4), "pass_gbs": round(32.0 * 2 ** n / (ms / cnt * 1e-3) / 1e9, 1),
                       "step_gbs": round(32.0 * 2 ** n * (n_pass + perms) / el / 1e9, 1),
                       "round_trip_max_abs": float(r["max_abs_diff"])}
        log_fn("circuit k=%d: %s" % (k, out[str(k)]))
    del sim
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE/WRITE_SIZE, calibrated on a plain copy)."""
    path = os.path.join(ROOT, "profiles", "dense_pass_30q_c128_hbm_traffic.md")
    if kernel != "dense_k4" or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return int(line.split(":")[1].split()[0])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--perms", type=int, default=8)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--planner", choices=["greedy", "inorder"], default="greedy",
                    help="greedy: the GreedyScheduler/ClusterScheduler loop; inorder: program order")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-only", action="store_true",
                    help="run only the CPU baseline and print its JSON (the child process bench.py starts)")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed CPU passes per k (median)")
    ap.add_argument("--cpu-circuit-passes", type=int, default=36,
                    help="fused passes of the headline circuit timed on the CPU (about 11 s at 30 qubits)")
    ap.add_argument("--no-cpu-circuit", action="store_true",
                    help="skip the whole-circuit CPU timing (about a minute at 30 qubits)")
    ap.add_argument("--no-drop-in", action="store_true")
    ap.add_argument("--no-circuit-sweep", action="store_true")
    ap.add_argument("--circuit-ks", default="2,3,4,5",
                    help="fusion widths of the whole-step sweep")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the step-and-inverse error check of the headline circuit")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-tile", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--precision-qubits", type=int, default=32)
    ap.add_argument("--tile-op-width", type=int, default=None,
                    help="op width of the tile programs (default: --k)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-circuits", type=int, default=4096)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rates", default="20,60,120")
    ap.add_argument("--block-swizzle", type=int, default=None,
                    help="low block-index bits reversed in the streaming passes (library default 8)")
    ap.add_argument("--block-swizzle-wide", type=int, default=None,
                    help="the same for dense passes of k >= 4 (library default 6)")
    ap.add_argument("--tile-high-target", type=int, default=None,
                    help="dense passes with every target at or above this bit use the direct form")
    ap.add_argument("--direct-block", type=int, default=None,
                    help="threads per block of the direct dense form of 4-qubit passes (256, 512, 1024)")
    ap.add_argument("--tile-mfma", type=int, default=None,
                    help="tile-program op engine: 1 ops of 3+ qubits on the matrix cores, 0 all on "
                         "the vector units")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain state copy (snapshot) for counter calibration")
    args = ap.parse_args()
    if args.cpu_baseline_only:
        print(json.dumps(cpu_baseline_child(args)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    # The CPU baseline runs first, in a child process, before this process
    # touches the GPU (its OpenMP runtime starts with the thread binding set).
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args)
        log("cpu baseline: %s" % cpu)

    import hiqsimulator_amd as hq
    if args.block_swizzle is not None:
        hq._cppsim_amd.set_tuning("block_swizzle", args.block_swizzle)
    if args.block_swizzle_wide is not None:
        hq._cppsim_amd.set_tuning("block_swizzle_wide", args.block_swizzle_wide)
    if args.tile_high_target is not None:
        hq._cppsim_amd.set_tuning("tile_high_target", args.tile_high_target)
    if args.direct_block is not None:
        hq._cppsim_amd.set_tuning("direct_block", args.direct_block)
    if args.tile_mfma is not None:
        hq._cppsim_amd.set_tuning("tile_mfma", args.tile_mfma)
    tuning = {name: hq._cppsim_amd.get_tuning(name)
              for name in ("block_swizzle", "block_swizzle_wide", "tile_high_target", "tile_high_target_wide", "direct_block",
                           "tile_mfma")}
    n, k = args.qubits, args.k
    t0 = time.perf_counter()
    gates, step, _ = build_step(n, args.layers, k, args.seed, args.perms, args.planner)
    n_pass = sum(1 for s in step if s[0] == "pass")
    n_perm = sum(1 for s in step if s[0] == "perm")
    plan_s = time.perf_counter() - t0
    log("plan: %d gates -> %d fused passes + %d permutations (%.2f s)"
        % (len(gates), n_pass, n_perm, plan_s))

    sim = hq.SimulatorMPI(args.seed, n, k)
    sim.set_pass_mode(0)  # the reference's one fused pass per run()
    sim.allocate_qureg(list(range(n)), 0)
    sim.synchronize()
    if args.calibrate_copy:
        sim.snapshot()
        sim.synchronize()
        log("calibration copy: %d bytes read + %d bytes written" % (16 * 2 ** n, 16 * 2 ** n))
    for w in range(args.warmup):
        run_step(sim, step)
        sim.synchronize()
        log("warmup %d done" % w)
    sim.set_timing(True)
    sim.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        run_step(sim, step)
    sim.synchronize()
    el = time.perf_counter() - t0
    log("timed %d steps in %.3f s" % (args.steps, el))
    times = {name: sim.kernel_time(i) for i, name in enumerate(hq.KERNEL_CLASSES)}
    sim.set_timing(False)
    norm = sim.get_probability([], [])
    # Error of the timed step: one more step and its exact inverse from the
    # current state, compared on the device with a snapshot in FP64.
    parity = None
    if not args.no_parity:
        start = list(sim.get_qubits_ids())
        sim.snapshot()
        run_step(sim, step)
        run_step(sim, inverse_step(step, start))
        r = sim.compare_snapshot()
        parity = {"workload": "one step and its inverse from the state after the timed steps",
                  "max_abs": float(r["max_abs_diff"]), "one_minus_fidelity": float(1.0 - r["fidelity"]),
                  "tolerance": 1e-12}
        log("parity: %s" % parity)
    bytes_step = 32.0 * 2 ** n * (n_pass + n_perm)
    ms_step = el / args.steps * 1e3
    value = bytes_step * args.steps / el / 1e9

    dom = "dense_k%d" % k
    cnt, tot = times[dom]
    per_launch_ms = tot / cnt if cnt else float("nan")
    achieved = 32.0 * 2 ** n / (per_launch_ms * 1e-3) / 1e9
    kernel_share = {name: round(ms, 2) for name, (c, ms) in times.items() if c}

    dense_ms_step = sum(ms for name, (c, ms) in times.items() if name.startswith("dense")) / args.steps
    del sim

    # The same gates with tile passes (the cluster scheduler targeting the
    # LDS tile): fewer HBM passes, each applying several fused ops.  Run on a
    # fresh state without the permutations; compared with the fused passes'
    # kernel time of the run above.
    tile = None
    if not args.no_tile:
        tw = args.tile_op_width or k
        ts = hq.SimulatorMPI(args.seed, n, k)
        ts.allocate_qureg(list(range(n)), 0)
        for w in range(args.warmup):
            ts.apply_tiled(gates, tw)
        ts.synchronize()
        ts.set_timing(True)
        t0 = time.perf_counter()
        tp = 0
        for s in range(args.steps):
            tp = ts.apply_tiled(gates, tw)
        ts.synchronize()
        el_t = time.perf_counter() - t0
        tcnt, ttot = ts.kernel_time(hq.KERNEL_CLASSES.index("tile"))
        ts.set_timing(False)
        tile_ms = ttot / tcnt if tcnt else float("nan")
        # Algorithmic flops of the ops (8 * 2^k per amplitude per op; ops of
        # 3+ qubits run on the matrix cores when tile_mfma is on), from the
        # same tile plan apply_tiled makes on this fresh state.
        pass_of, op_of, _ = hq._cppsim_amd.plan_tile_passes([g[1] for g in gates], [g[2] for g in gates], n, tw)
        support = {}
        for gi, (p, o) in enumerate(zip(pass_of, op_of)):
            m = 0
            for q in list(gates[gi][1]) + list(gates[gi][2]):
                m |= 1 << q
            support[(p, o)] = support.get((p, o), 0) | m
        widths = [bin(m).count("1") for m in support.values()]
        mfma_flop = sum(8.0 * 2 ** w * 2 ** n for w in widths if w >= 3)
        tile = {"hbm_passes_per_circuit": tp, "fused_passes_per_circuit": n_pass,
                "ms_per_circuit": round(el_t / args.steps * 1e3, 2),
                "fused_pass_kernel_ms_per_circuit": round(dense_ms_step, 2),
                "speedup_vs_fused_kernels": round(dense_ms_step / (el_t / args.steps * 1e3), 3),
                "tile_pass_avg_ms": round(tile_ms, 4),
                "tile_pass_hbm_gbs": round(32.0 * 2 ** n / (tile_ms * 1e-3) / 1e9, 1),
                "op_width": tw, "ops_per_circuit": len(widths),
                "mfma_flop_per_circuit": mfma_flop,
                "mfma_tflops": round(mfma_flop / (tile_ms * 1e-3 * tp) / 1e12, 2) if tp else None,
                "final_norm": ts.get_probability([], [])}
        del ts
        log("tile passes: %s" % tile)

    sweep = None if args.no_sweep else kernel_sweep(hq, n)
    if sweep:
        log("per-pass GB/s by k: %s" % sweep)
    sweep64 = None if args.no_sweep else kernel_sweep(hq, n, precision="c64")
    if sweep64:
        log("per-pass GB/s by k, complex64: %s" % sweep64)
    drop_in = None if args.no_drop_in else bench_drop_in(hq, n, args.layers, k, args.seed, args.steps, log)
    circuit_sweep = None
    if not args.no_circuit_sweep:
        circuit_sweep = bench_circuit_k_sweep(hq, n, args.layers, args.seed, args.perms, args.planner,
                                              [int(x) for x in args.circuit_ks.split(",") if x], log)
    precision = None if args.no_precision else bench_precision(hq, args.precision_qubits, log)
    batch = None if args.no_batch else bench_batch(hq, args.batch_circuits, args.seed, log)
    serving = None
    if not args.no_serving:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        serving = bench_serving(hq, args.serve_requests, rates, args.seed, log)

    line = {
        "metric": "30q fp64 fused-pass HBM GB/s vs peak; mixed-size circuits/s + p99 latency",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random circuit)",
        "config": {"workload": "%dq complex128 random circuit, %d layers, %d gates, fused k=%d "
                               "(%s planner, %d passes) + %d qubit permutations" %
                               (n, args.layers, len(gates), k, args.planner, n_pass, n_perm),
                   "planner": args.planner,
                   "qubits": n, "layers": args.layers, "fusion_k": k, "passes_per_step": n_pass,
                   "permutations_per_step": n_perm, "plan_seconds": round(plan_s, 3),
                   "final_norm": norm, "kernel_tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom) if n == 30 else None,
                     "kernel": dom, "launches": cnt, "avg_launch_ms": round(per_launch_ms, 4)},
        "kernel_ms_total": kernel_share,
        "rates": {"fused_passes_per_s": round(n_pass / (ms_step * 1e-3), 1),
                  "gates_per_s": round(len(gates) / (ms_step * 1e-3), 1),
                  "permutations_per_s": round(n_perm / (ms_step * 1e-3), 2)},
        "k_sweep_gbs": sweep,
        "k_sweep_gbs_c64": sweep64,
        "circuit_k_sweep": circuit_sweep,
        "tile_planner": tile,
        "drop_in": drop_in,
        "cpu_baseline": cpu,
        "parity": parity,
        "precision": precision,
        "batch": batch,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
