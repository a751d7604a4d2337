"""Batch planning on the host (no device): Batch.prepare() plans and fuses
every circuit, lays out the arena and the launches, and checks every
launch's blocks and addresses against its buffers (the kernels map blocks
to items through the items' first blocks, so the items must tile each
launch exactly).  Sizes here include launches of more than 65536 blocks,
whose 2-D grids are padded."""
import pytest

import hiqsimulator_amd as hq
from hiqsimulator_amd.batching import batch_workload
from hiqsimulator_amd.scheduling import plan_clusters


def make(circuits, precision, k_by_n, wg_max, prog_min=31):
    b = hq.Batch(precision)
    for n, g in circuits:
        b.add_circuit(n, g)
    b.set_plan(k_by_n, wg_max, 0, prog_min)
    return b


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("wg_max", [0, 14, 24])
def test_prepare_checks_a_mixed_workload(precision, wg_max):
    circuits = batch_workload(48, 5, n_range=(3, 21), depth_range=(2, 14))
    k_by_n = {n: 1 + n % 5 for n in range(1, 23)}
    b = make(circuits, precision, k_by_n, wg_max)
    b.prepare()
    st = b.stats()
    expect = sum(len(plan_clusters(g, list(range(n)), k_by_n[n], mode="lookahead"))
                 for n, g in circuits)
    assert st["passes"] == expect
    assert st["wg_circuits"] == sum(1 for n, _ in circuits if n <= wg_max)
    ab = 16 if precision == "c128" else 8
    assert st["arena_bytes"] == sum(ab << n for n, _ in circuits)
    lds_n = 13 if precision == "c128" else 14
    assert st["lds_circuits"] == sum(1 for n, _ in circuits if n <= min(wg_max, lds_n))
    if wg_max < 21:
        assert st["layer_launches"] > 0


@pytest.mark.parametrize("prog_min", [12, 16])
def test_prepare_tile_programs(prog_min):
    circuits = batch_workload(40, 6, n_range=(10, 20), depth_range=(3, 16))
    b = make(circuits, "c128", {}, 11, prog_min)
    b.prepare()
    st = b.stats()
    progs = [(n, g) for n, g in circuits if n > 11 and n >= prog_min]
    assert (st["tile_passes"] > 0) == bool(progs)
    expect = flops = 0
    for n, g in progs:
        po, oo, _ = hq._cppsim_amd.plan_tile_passes([x[1] for x in g], [x[2] for x in g], n, 4)
        expect += max(po) + 1 if po else 0
        # each op acts on the union of its gates' qubits (controls folded in)
        ops = {}
        for i, key in enumerate(zip(po, oo)):
            ops.setdefault(key, set()).update(list(g[i][1]) + list(g[i][2]))
        flops += sum(8 * 2 ** len(q) * 2 ** n for q in ops.values() if len(q) >= 3)
    assert st["tile_passes"] == expect
    assert st["tile_ops"] == sum(
        len(set(zip(*hq._cppsim_amd.plan_tile_passes([x[1] for x in g], [x[2] for x in g], n, 4)[:2])))
        for n, g in progs)
    assert st["tile_mfma_flops"] == flops


def test_prepare_large_launches():
    # 24 layered 22-qubit circuits: launches of 24 * 2^18 / 256 blocks and more
    circuits = batch_workload(24, 9, n_range=(22, 22), depth_range=(2, 3))
    b = make(circuits, "c64", {22: 2}, 16)
    b.prepare()
    st = b.stats()
    assert st["layer_launches"] > 0
    # more blocks than one 65536-wide grid row and not a multiple of it: the
    # kernels must skip the padding blocks
    assert st["padded_launches"] > 0


def test_prepare_errors():
    b = hq.Batch("c128")
    with pytest.raises(RuntimeError):
        b.prepare()  # no circuits
    with pytest.raises(RuntimeError):
        b.run()  # not built
