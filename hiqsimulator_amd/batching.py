"""Batches of independent circuits: workload generator, builder and autotuner.

A ``Batch`` (native, ``hiqsimulator_amd.Batch``) runs many circuits, each
from |0...0> on its own qubits 0..n-1, in one graph launch.  Its plan has two
knobs, both picked here by measurement:

* the fusion width k per qubit count n (every fused pass is a dense
  2^k x 2^k matrix over at most k qubits);
* ``wg_max_qubits``: circuits of at most this many qubits are each owned by
  one workgroup of a persistent launch (LDS-resident up to 128 KiB of
  state), larger ones are advanced by layered launches with many workgroups
  per state;
* ``tile_min_qubits``: layered passes of circuits from this many qubits on
  may take the wave-tile form (LDS regrouping, non-temporal streaming);
  smaller ones use the direct form;
* ``prog_min_qubits``: layered circuits from this many qubits on (at least
  12) run as tile programs: each HBM pass loads 4096-amplitude tiles into
  LDS and applies several fused ops there (the cluster scheduler targeting
  the tile), with k_by_n[n] as the op width.
"""
import json
import os
import time

import numpy as np

from . import Batch
from .circuits import random_layers_circuit

DEFAULT_WG_MAX_QUBITS = 16
TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")


def tuning_path(precision, arch="gfx950"):
    return os.path.join(TUNING_DIR, "%s_batch_%s.json" % (arch, precision))


def load_tuning(precision, arch="gfx950"):
    """The stored autotuner result for this precision, or None."""
    path = tuning_path(precision, arch)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    t["k_by_n"] = {int(n): int(k) for n, k in t["k_by_n"].items()}
    t.setdefault("tile_min_qubits", 0)
    t.setdefault("prog_min_qubits", 31)
    return t


def save_tuning(result, precision, arch="gfx950"):
    os.makedirs(TUNING_DIR, exist_ok=True)
    with open(tuning_path(precision, arch), "w") as f:
        json.dump({"k_by_n": {str(n): k for n, k in sorted(result["k_by_n"].items())},
                   "wg_max_qubits": result["wg_max_qubits"],
                   "tile_min_qubits": result.get("tile_min_qubits", 0),
                   "prog_min_qubits": result.get("prog_min_qubits", 31),
                   "probe_ms": result.get("probe_ms", {})}, f, indent=1, sort_keys=True)


def batch_workload(count, seed, n_range=(10, 22), depth_range=(10, 60), haar_every=5):
    """``count`` random layered circuits with n ~ U{n_range}, depth ~ U{depth_range}
    and a Haar-random 2-qubit unitary on a random pair after every
    ``haar_every``-th layer.  Returns ``[(n, gates), ...]``."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    ds = rng.integers(depth_range[0], depth_range[1] + 1, size=count)
    seeds = rng.integers(0, 2**31 - 1, size=count)
    return [(int(n), random_layers_circuit(int(n), int(d), int(s), haar_every=haar_every))
            for n, d, s in zip(ns, ds, seeds)]


def make_batch(circuits, precision="c128", k_by_n=None, wg_max_qubits=DEFAULT_WG_MAX_QUBITS,
               device=0, build=True, tile_min_qubits=0, prog_min_qubits=31):
    """A ``Batch`` holding ``circuits`` ([(n, gates), ...]), planned and built."""
    b = Batch(precision, device)
    for n, gates in circuits:
        b.add_circuit(n, gates)
    b.set_plan(dict(k_by_n or {}), int(wg_max_qubits), int(tile_min_qubits), int(prog_min_qubits))
    if build:
        b.build()
    return b


def _time_batch(b, reps):
    b.run()  # warm
    return min(b.run() for _ in range(reps))


def autotune(ns, precision="c128", widths=(2, 3, 4, 5), depth=20, amp_budget=1 << 24,
             reps=3, seed=7, device=0, wg_limit=20, layer_floor=12, log=None):
    """Pick the fusion width per qubit count and the workgroup/layered split.

    For every n in ``ns`` a probe batch of copies of one depth-``depth``
    random circuit (enough copies to hold about ``amp_budget`` amplitudes,
    8..2048) is timed for every width in ``widths``, on the persistent workgroup
    path (n <= ``wg_limit``) and on the layered path (n >= ``layer_floor``)
    as direct passes, wave-tile passes and (n >= 12) tile programs.
    Returns ``{"k_by_n": {n: k}, "wg_max_qubits": t, "tile_min_qubits": m,
    "prog_min_qubits": p, "probe_ms": {...}}`` where t is the largest n whose
    best workgroup time beats its best layered time (the split is monotone
    in n by construction), p the smallest n from which tile programs win
    for every larger layered n, and m the same for the wave-tile form among
    the remaining layered n.
    """
    ns = sorted(set(int(n) for n in ns))
    best = {}
    probe = {}
    for n in ns:
        copies = max(8, min(2048, amp_budget >> n))
        b = Batch(precision, device)
        b.add_circuit(n, random_layers_circuit(n, depth, seed + n, haar_every=5))
        b.add_copies(0, copies - 1)
        for path, limit, tmin, pmin in (("wg", 24, 31, 31), ("layer", 0, 31, 31), ("tile", 0, 0, 31),
                                        ("prog", 0, 31, 0)):
            if path == "wg" and n > wg_limit:
                continue
            if path != "wg" and n < layer_floor:
                continue
            if path == "prog" and n < 12:
                continue
            for k in widths:
                if k > n:
                    continue
                b.set_plan({n: k}, limit, tmin, pmin)
                b.build()
                ms = _time_batch(b, reps)
                probe[(n, path, k)] = ms
                if log:
                    log("autotune n=%d %s k=%d: %.3f ms (%d circuits)" % (n, path, k, ms, copies))
                if (n, path) not in best or ms < best[(n, path)][1]:
                    best[(n, path)] = (k, ms)
        del b
    def layered(n, forms=("layer", "tile", "prog")):
        cands = [best[(n, p)] + (p,) for p in forms if (n, p) in best]
        return min(cands, key=lambda c: c[1]) if cands else None

    wg_max = 0
    for n in ns:
        wg = best.get((n, "wg"))
        ly = layered(n)
        if wg and (ly is None or wg[1] <= ly[1]):
            wg_max = n
        else:
            break
    prog_min = 31
    for n in reversed([n for n in ns if n > wg_max]):
        ly = layered(n)
        if ly is not None and ly[2] == "prog":
            prog_min = n
        else:
            break
    tile_min = 31
    for n in reversed([n for n in ns if wg_max < n < prog_min]):
        ly = layered(n, ("layer", "tile"))
        if ly is not None and ly[2] == "tile":
            tile_min = n
        else:
            break
    k_by_n = {}
    for n in ns:
        if n <= wg_max:
            form = "wg"
        elif n >= prog_min:
            form = "prog"
        else:
            form = "tile" if n >= tile_min else "layer"
        k_by_n[n] = best[(n, form)][0] if (n, form) in best else 4
    return {"k_by_n": k_by_n, "wg_max_qubits": wg_max, "tile_min_qubits": tile_min,
            "prog_min_qubits": prog_min,
            "probe_ms": {"%d/%s/k%d" % key: round(v, 4) for key, v in probe.items()}}


def one_at_a_time(circuits, precision="c128", k=4, device=0, mode="fused"):
    """The plain path: one SimulatorMPI per circuit, the final state
    synchronised.  Returns seconds.  ``mode``:
    * "fused": apply_controlled_gate with run() per fused cluster of at most
      k qubits (lookahead plan), one HBM pass per cluster (pass mode 0);
    * "deferred": the same calls in the default pass mode (clusters of
      12+-qubit states run as tile passes);
    * "tiled": the whole gate list through apply_tiled (op width k)."""
    from . import SimulatorMPI
    from .scheduling import plan_clusters
    if mode not in ("fused", "deferred", "tiled"):
        raise ValueError("mode must be 'fused', 'deferred' or 'tiled'")
    plans = [plan_clusters(g, list(range(n)), k, mode="lookahead") for n, g in circuits]
    t0 = time.perf_counter()
    for (n, gates), clusters in zip(circuits, plans):
        sim = SimulatorMPI(1, n, k, precision, device)
        if mode == "fused":
            sim.set_pass_mode(0)
        sim.allocate_qureg(list(range(n)), 0)
        if mode == "tiled":
            sim.apply_tiled(gates, k)
        else:
            for cl in clusters:
                for i in cl:
                    m, t, c = gates[i]
                    sim.apply_controlled_gate(m, t, c)
                sim.run()
        sim.synchronize()
        del sim
    return time.perf_counter() - t0


def fidelity(a, b):
    """|<a|b>|^2 / (|a|^2 |b|^2)."""
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def main(argv=None):
    """``python -m hiqsimulator_amd.batching``: builds the batch workload with
    the stored tuning table and replays it; prints the plan and the replay
    times (a short driver for profiling the batch kernels on their own).
    With ``--autotune`` it first re-measures the table for the precision
    (qubit counts 10..22) and stores it."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--circuits", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--precision", choices=["c64", "c128"], default="c64")
    ap.add_argument("--replays", type=int, default=2)
    ap.add_argument("--tile-mfma", type=int, default=None,
                    help="tile-program op engine: 1 ops of 3+ qubits on the matrix cores, 0 all on "
                         "the vector units")
    ap.add_argument("--autotune", action="store_true",
                    help="re-measure and store the tuning table before replaying")
    a = ap.parse_args(argv)
    from . import _cppsim_amd
    if a.tile_mfma is not None:
        _cppsim_amd.set_tuning("tile_mfma", a.tile_mfma)
    if a.autotune:
        res = autotune(range(10, 23), a.precision, amp_budget=1 << 23, reps=2,
                       log=lambda m: print(m, flush=True))
        save_tuning(res, a.precision)
        print("autotune %s: %s -> %s" % (a.precision, json.dumps(
            {k: v for k, v in res.items() if k != "probe_ms"}), tuning_path(a.precision)), flush=True)
    circ = batch_workload(a.circuits, a.seed)
    t = load_tuning(a.precision)
    if t is None:
        raise SystemExit("no tuning table for %s at %s" % (a.precision, tuning_path(a.precision)))
    b = make_batch(circ, a.precision, t["k_by_n"], t["wg_max_qubits"],
                   tile_min_qubits=t["tile_min_qubits"], prog_min_qubits=t["prog_min_qubits"])
    print("plan", json.dumps(b.stats()), flush=True)
    for i in range(a.replays):
        print("replay %d: %.3f ms" % (i, b.run()), flush=True)


if __name__ == "__main__":
    main()
