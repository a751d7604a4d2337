"""Serving: Poisson request streams against the request loop (``Server``) and
against the plain one-circuit-at-a-time path (``SimulatorMPI``).

A request is a random layered circuit on n qubits (n log-uniform in
[n_min, n_max], depth uniform) followed either by ``shots`` computational-
basis samples under a fixed Philox seed or by the expectation values of a
few random Pauli strings.  Both paths are driven in real time: a request
cannot start before its arrival time, and its latency runs from the arrival
time to the moment its result is on the host.
"""
import math
import time

import numpy as np

from . import Server, SimulatorMPI
from .circuits import random_layers_circuit
from .scheduling import plan_clusters

PAULI = "XYZ"


def random_pauli_strings(n, count, rng, wmin=1, wmax=4):
    """``count`` Pauli strings of weight wmin..wmax on n qubits as (xmask, zmask)."""
    out = []
    for _ in range(count):
        w = int(rng.integers(wmin, min(wmax, n) + 1))
        qs = rng.choice(n, size=w, replace=False)
        x = z = 0
        for q in qs:
            p = PAULI[int(rng.integers(3))]
            if p in "XY":
                x |= 1 << int(q)
            if p in "YZ":
                z |= 1 << int(q)
        out.append((x, z))
    return out


def serving_workload(count, rate, seed, n_range=(12, 28), depth_range=(20, 100), shots=1024,
                     n_terms=8, sample_fraction=0.5):
    """Requests with Poisson arrivals at ``rate`` per second.

    Returns a list of dicts: arrival (s from the stream start), n, gates,
    kind ("samples" | "expectation"; the request loop also takes "state",
    which returns the final state), shots, seed, terms [(xmask, zmask)].
    """
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    lo, hi = math.log(n_range[0]), math.log(n_range[1] + 1)
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = min(n_range[1], int(math.exp(rng.uniform(lo, hi))))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        gates = random_layers_circuit(n, depth, int(rng.integers(1 << 31)))
        if rng.random() < sample_fraction:
            reqs.append(dict(arrival=t, n=n, gates=gates, kind="samples", shots=shots,
                             seed=int(rng.integers(1 << 62)), terms=[]))
        else:
            reqs.append(dict(arrival=t, n=n, gates=gates, kind="expectation", shots=0, seed=0,
                             terms=random_pauli_strings(n, n_terms, rng)))
    return reqs


def submit(server, req):
    if req["kind"] == "samples":
        return server.submit_samples(req["n"], req["gates"], req["shots"], req["seed"])
    if req["kind"] == "state":
        return server.submit_state(req["n"], req["gates"])
    xs = [x for x, _ in req["terms"]]
    zs = [z for _, z in req["terms"]]
    return server.submit_expectation(req["n"], req["gates"], xs, zs)


def result(server, rid, req):
    if req["kind"] == "samples":
        return server.samples(rid)
    if req["kind"] == "state":
        return server.state(rid)
    return server.expectations(rid)


def latency_summary(lat, makespan, count):
    """p50/p99/mean latency (ms), circuits/s and makespan; ``lat_ms`` holds
    every request's latency (ms, sorted)."""
    lat = np.sort(np.asarray(lat))
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "circuits_per_s": count / makespan,
            "makespan_s": makespan, "lat_ms": [float(x) * 1e3 for x in lat]}


def run_server(server, reqs, keep_results=False):
    """Drives ``server`` with ``reqs`` in real time; returns (summary, results)."""
    order = sorted(range(len(reqs)), key=lambda i: reqs[i]["arrival"])
    ids = {}
    arrival_abs = {}
    lat = {}
    results = {}
    t0 = time.monotonic()
    nxt = 0
    while len(lat) < len(reqs):
        now = time.monotonic()
        while nxt < len(order) and t0 + reqs[order[nxt]]["arrival"] <= now:
            i = order[nxt]
            rid = submit(server, reqs[i])
            ids[rid] = i
            arrival_abs[rid] = t0 + reqs[i]["arrival"]
            nxt += 1
        pending_arrival = nxt < len(order)
        done = server.step(block=not pending_arrival)
        for rid in done:
            t_done = server.times(rid)[2]
            lat[rid] = t_done - arrival_abs[rid]
            if keep_results:
                results[ids[rid]] = result(server, rid, reqs[ids[rid]])
            server.forget(rid)
        if not done and server.idle() and pending_arrival:
            wait = t0 + reqs[order[nxt]]["arrival"] - time.monotonic()
            if wait > 0:
                time.sleep(wait)
    makespan = time.monotonic() - t0
    return latency_summary(list(lat.values()), makespan, len(reqs)), results


def run_plain(reqs, precision="c64", k=4, keep_results=False, device=0, tiled=True):
    """The same stream through one SimulatorMPI per request, one request at a
    time in arrival order: the gates go through ``apply_tiled`` (tile
    passes, from 12 qubits on) or, with ``tiled=False`` or below 12 qubits,
    cluster by cluster through apply_controlled_gate and run(); then
    ``sample`` or one expectation per Pauli string."""
    order = sorted(range(len(reqs)), key=lambda i: reqs[i]["arrival"])
    lat = []
    results = {}
    t0 = time.monotonic()
    for i in order:
        r = reqs[i]
        wait = t0 + r["arrival"] - time.monotonic()
        if wait > 0:
            time.sleep(wait)
        n = r["n"]
        sim = SimulatorMPI(1, n, k, precision, device)
        ids = list(range(n))
        sim.allocate_qureg(ids, 0)
        if tiled and n >= 12:
            sim.apply_tiled(r["gates"], k)
        else:
            for cl in plan_clusters(r["gates"], ids, k, mode="lookahead"):
                for g in cl:
                    m, t, c = r["gates"][g]
                    sim.apply_controlled_gate(m, t, c)
                sim.run()
        if r["kind"] == "samples":
            out = sim.sample(r["shots"], r["seed"])
        else:
            out = np.array([sim.expectation_complex([(pauli_term(x, z), 1.0)], ids)
                            for x, z in r["terms"]])
        lat.append(time.monotonic() - (t0 + r["arrival"]))
        if keep_results:
            results[i] = out
        del sim
    makespan = time.monotonic() - t0
    return latency_summary(lat, makespan, len(reqs)), results


def pauli_term(x, z):
    """(xmask, zmask) -> ProjectQ term [(index, 'X'|'Y'|'Z'), ...]."""
    term = []
    q = 0
    m = x | z
    while m >> q:
        if (m >> q) & 1:
            bx, bz = (x >> q) & 1, (z >> q) & 1
            term.append((q, "Y" if bx and bz else ("X" if bx else "Z")))
        q += 1
    return term


def make_server(precision="c64", state_bytes=32 << 30, matrix_bytes=2 << 30, passes_per_step=8,
                max_active=1024, k_by_n=None, device=0, prog_min_qubits=12, heavy_qubits=24):
    """A request loop with its pools; requests of at least prog_min_qubits
    qubits run as tile programs (31 = never), requests of at least
    heavy_qubits qubits are admitted one at a time (31 = no limit)."""
    s = Server(precision, device, int(state_bytes), int(matrix_bytes), int(passes_per_step),
               int(max_active))
    if k_by_n:
        s.set_widths(dict(k_by_n))
    s.set_prog_min_qubits(int(prog_min_qubits))
    s.set_heavy_qubits(int(heavy_qubits))
    return s
